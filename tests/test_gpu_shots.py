"""Several shots of one mesh in one interleaved launch (k_shots3d,
k_shots_el) against the oracle, shot by shot: 16 shots, as in the 16-shot
workloads (elastic Q4 with random positions and polarisations, acoustic Q5),
FP64 to 1e-12 and the FP32 paths to 1e-5."""
import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NS = 16
TOL = {"fp64": 1e-12, "fp32_g64": 1e-5, "fp32": 1e-5}


def _shots(seed, dim=3):
    rng = np.random.default_rng(seed)
    pts = 0.25 + 0.5 * rng.random((NS, dim))
    v = rng.standard_normal((NS, 3))
    return pts, v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("prec", ["fp64", "fp32_g64", "fp32"])
def test_elastic_16_shots_match_oracle(prec):
    """The Q4 elastic fixture's medium (Montjoie-checked through
    ref3d_el_q4), 16 shots with seeded positions and unit polarisations."""
    c = CS.load("ref3d_el_q4")
    n = 80
    pts, pol = _shots(2027)
    m = M.regular_mesh(3, c.nx, c.r, c.pert, c.seed)
    P = M.WaveProblem(m, c.rho_e, c.mu_e, lam=c.lam_e, ncomp=3, nshots=NS, precision=prec)
    P.set_sources(pts, pol=pol, shot=np.arange(NS))
    rec = np.vstack([c.rec] * 3)
    comp = np.repeat(np.arange(3, dtype=np.int32), len(c.rec))
    P.set_receivers(rec, comp=comp)
    P.set_pulse(M.ricker_pulse_table(c.f0, 0.0, c.dt, n))
    P.leapfrog_init(c.dt)
    tr = P.run(n)
    o = O.mesh(3, c.nx, c.r, c.pert, c.seed)
    OP = O.elastic_problem(o, c.lam_e, c.mu_e, c.rho_e)
    pulse = O.pulse_table(c.f0, 0.0, c.dt, n)
    worst = 0.0
    for s in range(NS):
        _, _, otr = O.leapfrog(OP, O.dirac(o, pts[s], pol[s]), pulse, c.dt, n, rec, comp)
        worst = max(worst, CS.rel_l2(tr[:, s, :], otr))
    print(f"elastic Q4 {NS} shots {prec}: worst shot rel-L2 vs oracle {worst:.2e}")
    assert worst < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32_g64", "fp32"])
def test_acoustic_16_shots_match_oracle(prec):
    """Q5 perturbed layered cube (the headline medium at 3^3), 16 shots."""
    m = M.regular_mesh(3, 3, 5, 0.1, 2024)
    o = O.mesh(3, 3, 5, 0.1, 2024)
    mu = O.layered(o, [1.0, 4.0, 2.25])
    pts, _ = _shots(2028)
    rec = np.array([[0.3, 0.4, 0.6], [0.7, 0.55, 0.35], [0.5, 0.2, 0.8]])
    dt, n = 2e-3, 80
    P = M.WaveProblem(m, 1.0, mu, nshots=NS, precision=prec)
    P.set_sources(pts, shot=np.arange(NS))
    P.set_receivers(rec)
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, n))
    P.leapfrog_init(dt)
    tr = P.run(n)
    OP = O.acoustic_problem(o, mu, 1.0)
    pulse = O.pulse_table(8.0, 0.0, dt, n)
    worst = 0.0
    for s in range(NS):
        _, _, otr = O.leapfrog(OP, O.dirac(o, pts[s]), pulse, dt, n, rec)
        worst = max(worst, CS.rel_l2(tr[:, s, :], otr))
    print(f"acoustic Q5 {NS} shots {prec}: worst shot rel-L2 vs oracle {worst:.2e}")
    assert worst < TOL[prec]
