"""PML layers (mjw_set_pml, Montjoie's first-order LEAP_FROG_PML scheme) on
the GPU: the Montjoie fixtures made with AddPML / DampingPML / LEAP_FROG_PML
(2-D and 3-D, receivers inside the layers) and the oracle, FP64; FP32 state;
graph replay; the refusals of the second-order-only entry points."""
import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from montjoie_amd._lib import MjwError

pytestmark = pytest.mark.gpu

PMLS = [n for n in CS.golden_names() if n.endswith("_pml")]


def _problem(c, prec="fp64", nsteps=None):
    """The extended mesh of the case handed to the library as a caller's mesh."""
    nsteps = c.nsteps if nsteps is None else nsteps
    P = CS.gpu_problem(c, prec)
    P.set_pulse(M.time_source_table(c.f0, 0.0, c.dt, nsteps, 0, 0.5 * c.dt))
    P.leapfrog_init(c.dt)
    return P


@pytest.mark.parametrize("name", PMLS)
def test_pml_fp64_matches_montjoie_and_oracle(name):
    c = CS.load(name)
    print(f"pml {name}", flush=True)
    P = _problem(c)
    tr = P.run(c.nsteps)[:, 0, :]
    _, _, otr = CS.oracle_run(c)
    e_ref = CS.rel_l2(tr, c.traces)
    e_or = CS.rel_l2(tr, otr)
    print(f"{name}: rel-L2 vs Montjoie {e_ref:.2e}, vs oracle {e_or:.2e}")
    assert e_ref < 1e-10 and e_or < 1e-12
    # the PML receivers on their own (small amplitudes)
    for k in range(len(c.rec)):
        if np.abs(c.traces[:, k]).max() > 1e-9:
            assert CS.rel_l2(tr[:, k], c.traces[:, k]) < 1e-9, k


@pytest.mark.parametrize("prec", ["fp32", "fp32_g64"])
def test_pml_fp32_close_to_fp64(prec):
    c = CS.load("ref2d_q4_pml")
    t64 = _problem(c).run(c.nsteps)[:, 0, :]
    t32 = _problem(c, prec).run(c.nsteps)[:, 0, :]
    err = CS.rel_l2(t32, t64)
    print(f"{prec} vs fp64: {err:.2e}")
    assert err < 1e-5


def test_pml_graph_equals_eager_and_refusals():
    c = CS.load("ref3d_q3_pml")
    a = _problem(c, nsteps=70).run(70, graph=True)
    b = _problem(c, nsteps=70).run(70, graph=False)
    assert np.array_equal(a, b)
    P = _problem(c, nsteps=70)
    with pytest.raises(MjwError):
        P.energy()
    with pytest.raises(MjwError):
        P.leapfrog_init(c.dt, u0=np.ones(P.mesh.N))
    P.advance(5, traces=False)
    u, u1 = P.state()
    assert u1 is None and np.isfinite(u).all()
