"""Fourth-order modified-equation leapfrog (mjw_set_time_order(4),
OrderTimeScheme = 4 LEAP_FROG).  Whole runs of the Montjoie fixtures
ref*_order4 are in test_gpu_parity.py and test_deck.py; here: the Taylor
start from u0, du0 and the damped three-level step against the oracle, graph
replays across the 6-step cycle of the levels, interleaved shots, FP32, and
the entry points that refuse the scheme."""
import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from montjoie_amd._lib import MjwError
from montjoie_amd.batch import Batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tables(f0, dt, n):
    return tuple(O.source_table("RICKER", f0, 0.0, 0.0, dt, n, k) for k in (1, 2, 3))


def _gpu_tables(f0, dt, n):
    return tuple(M.temporal_source_table("RICKER", f0, 0.0, 0.0, dt, n, k) for k in (1, 2, 3))


@pytest.mark.parametrize("dim,ncomp,damped", [(2, 1, False), (2, 1, True), (3, 1, True),
                                              (3, 3, False), (2, 2, True)])
def test_order4_matches_oracle(dim, ncomp, damped):
    """Initial data u0, du0 (the Taylor start) plus a Dirac source, with and
    without damping, acoustic and elastic: traces and the final two levels
    against or_leapfrog4 (FP64, 1e-12)."""
    m = M.regular_mesh(dim, 3, 3, 0.1, seed=13)
    o = O.mesh(dim, 3, 3, 0.1, 13)
    E = m.E
    mu = 1.0 + np.arange(E) % 3
    sigma = np.where(np.arange(E) % 2 == 0, 5.0, 0.0)
    src = [0.45, 0.52, 0.48][:dim]
    rng = np.random.default_rng(dim * 10 + ncomp)
    rec = rng.uniform(0.15, 0.85, size=(4, dim))
    if ncomp == 1:
        P = M.WaveProblem(m, 1.0, mu)
        OP = O.acoustic_problem(o, mu, 1.0)
        P.set_sources([src])
        b = O.dirac(o, src)
        P.set_receivers(rec)
        recs, comp = rec, None
    else:
        lam, rho = 2.0 + np.arange(E) % 2, 1.0 + 0.5 * (np.arange(E) % 3)
        P = M.WaveProblem(m, rho, mu, lam=lam, ncomp=ncomp)
        OP = O.elastic_problem(o, lam, mu, rho)
        pol = [0.3, -0.5, 0.8][:dim]
        P.set_sources([src], pol=[pol])
        b = O.dirac(o, src, pol)
        recs = np.vstack([rec] * ncomp)
        comp = np.repeat(np.arange(ncomp, dtype=np.int32), len(rec))
        P.set_receivers(recs, comp=comp)
    if damped:
        P.set_damping(sigma)
    S = O.damping(o, sigma) if damped else None
    f0, dt, n = 8.0, 2e-3, 70
    P.set_time_order(4)
    P.set_pulse(*_gpu_tables(f0, dt, n))
    u0 = 1e-2 * rng.standard_normal(ncomp * m.N)
    du0 = rng.standard_normal(ncomp * m.N)
    u0[np.tile(o.dirichlet.astype(bool), ncomp)] = 0.0
    P.leapfrog_init(dt, u0=u0, du0=du0)
    tr = P.run(n)[:, 0, :]
    un, un1 = (x[0] for x in P.state())
    oN, oN1, otr = O.leapfrog4(OP, b, _tables(f0, dt, n), dt, n, recs, comp, u0=u0, du0=du0, S=S)
    err = CS.rel_l2(tr, otr)
    print(f"order 4 dim {dim} ncomp {ncomp} damped {damped}: traces rel-L2 vs oracle {err:.2e}")
    assert err < 1e-12
    assert CS.rel_l2(un, oN) < 1e-12 and CS.rel_l2(un1, oN1) < 1e-12


def test_order4_graph_equals_eager_across_the_cycle():
    """Graph replays (24 steps, four turns of the 6-step cycle of the three
    levels) give the eager steps' bits, entered at every phase of the cycle
    and with traces recorded."""
    c = CS.load("ref2d_q4_order4_damped")
    n = 120
    runs = []
    for phases in ([n], [1, n - 1], [5, 40, n - 45], [3, 3, 3, n - 9]):
        for graph in (True, False):
            P = CS.gpu_problem(c)
            CS.gpu_set_pulse(c, P, n)
            P.leapfrog_init(c.dt)
            parts = [P.run(k, graph=graph)[1:] for k in phases]
            runs.append(np.concatenate(parts))
    assert np.abs(runs[0]).max() > 0
    assert all(np.array_equal(runs[0], r) for r in runs[1:])


def test_order4_interleaved_shots_equal_single_shots():
    """3-D, 3 interleaved shots (the shot-interleaved layout in the node-wise
    steps) equal three single-shot runs (FP64, 1e-13); FP32 state within
    1e-5 of FP64."""
    c = CS.load("ref3d_q3_order4")
    srcs = [c.src, c.src + 0.05, c.src - 0.04]
    _, tm = CS.gpu_run(c, "fp64", nshots=3, srcs=srcs)
    for s, x in enumerate(srcs):
        _, ts = CS.gpu_run(c, "fp64", srcs=[x])
        assert CS.rel_l2(tm[:, s, :], ts[:, 0, :]) < 1e-13
    _, t32 = CS.gpu_run(c, "fp32_g64")
    _, t64 = CS.gpu_run(c, "fp64")
    err = CS.rel_l2(t32, t64)
    print(f"order 4 FP32 state vs FP64: {err:.2e}")
    assert err < 1e-5


def test_order4_refusals():
    """The energy of the second-order scheme, mjw_profile, the adjoint, the
    batches and the PML layers refuse the fourth-order scheme; switching back
    to order 2 restores the second-order run bit for bit."""
    from montjoie_amd.adjoint import AdjointSolver
    m = M.regular_mesh(3, 3, 3, 0.1, seed=4)
    P = M.WaveProblem(m, 1.0, 1.0)
    P.set_sources([[0.5, 0.5, 0.5]])
    P.set_receivers([[0.4, 0.5, 0.6]])
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, 2e-3, 40))
    P.leapfrog_init(2e-3)
    ref = P.run(30)
    P.set_time_order(4)
    P.leapfrog_init(2e-3)
    P.run(4)
    with pytest.raises(MjwError, match="second-order"):
        P.energy()
    with pytest.raises(MjwError):
        P.profile(2)
    with pytest.raises(MjwError):
        Batch([P])
    with pytest.raises(MjwError):
        AdjointSolver(P, 2e-3, 4).forward()
    P.set_time_order(2)
    P.leapfrog_init(2e-3)
    assert np.array_equal(P.run(30), ref)
