"""Damped leapfrog (mjw_set_damping): rho u'' + sigma u' - div(mu grad u) = f
(elastic: - div(C : grad u)) with S_h = sigma J w lumped like the mass.  Whole runs of the damped golden
cases against Montjoie's traces are in test_gpu_parity.py and test_deck.py
(ref*_damped); here: non-zero initial velocity (the -S du0 term of the first
step) against the oracle, FP32, interleaved shots, graph replay, and the
paths that refuse damped problems."""
import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from montjoie_amd._lib import MjwError
from montjoie_amd.adjoint import AdjointSolver
from montjoie_amd.batch import Batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _problem(dim, prec="fp64", nshots=1):
    m = M.regular_mesh(dim, 3, 4, 0.1, seed=5)
    o = O.mesh(dim, 3, 4, 0.1, 5)
    mu = 1.0 + np.arange(m.E) % 3
    sigma = np.where(np.arange(m.E) % 2 == 0, 6.0, 0.0)
    P = M.WaveProblem(m, 1.0, mu, precision=prec, nshots=nshots)
    P.set_damping(sigma)
    return m, o, mu, sigma, P


@pytest.mark.parametrize("dim", [2, 3])
def test_damped_initial_velocity_matches_oracle(dim):
    """u0, du0 given: U^1 = U^0 + dt du0 + dt^2/2 D^{-1}(-A U^0 - S du0), then
    the damped steps; traces and final state against the oracle (FP64)."""
    m, o, mu, sigma, P = _problem(dim)
    rng = np.random.default_rng(dim)
    rec = rng.uniform(0.1, 0.9, size=(5, dim))
    P.set_receivers(rec)
    u0 = rng.standard_normal(m.N) * 1e-2
    du0 = rng.standard_normal(m.N)
    u0[o.dirichlet.astype(bool)] = 0.0
    dt, nsteps = 2e-3, 60
    P.leapfrog_init(dt, u0=u0, du0=du0)
    tr = P.run(nsteps)[:, 0, :]
    uN, _ = P.state()
    OP = O.acoustic_problem(o, mu, 1.0)
    S = O.damping(o, sigma)
    oN, _, otr = O.leapfrog(OP, np.zeros(o.N), None, dt, nsteps, rec, u0=u0, du0=du0, S=S)
    assert CS.rel_l2(tr, otr) < 1e-12
    assert CS.rel_l2(uN[0], oN) < 1e-12


def test_damped_fp32_and_shots():
    """FP32 damped run within 1e-5 of FP64; 3 interleaved shots equal the
    three single-shot runs (FP64, 1e-13)."""
    c = CS.load("ref3d_q3_damped")
    _, t64 = CS.gpu_run(c, "fp64")
    _, t32 = CS.gpu_run(c, "fp32")
    assert CS.rel_l2(t32, t64) < 1e-5
    srcs = [c.src, c.src + 0.05, c.src - 0.04]
    _, tm = CS.gpu_run(c, "fp64", nshots=3, srcs=srcs)
    for s, x in enumerate(srcs):
        _, ts = CS.gpu_run(c, "fp64", srcs=[x])
        assert CS.rel_l2(tm[:, s, :], ts[:, 0, :]) < 1e-13


@pytest.mark.parametrize("prec", ["fp64", "fp32_g64"])
def test_damped_elastic_shots_match_oracle(prec):
    """Damped elastic media (sigma I in the mass tensor): 4 interleaved shots
    of the 3-D fixture ref3d_el_q3_damped with their own sources and
    polarisations, each against the oracle's damped run (FP64 1e-12, FP32
    state 1e-5), after the initial velocity's -S du0 term is exercised by a
    single-shot run from (u0, du0)."""
    c = CS.load("ref3d_el_q3_damped")
    rng = np.random.default_rng(31)
    srcs = 0.3 + 0.4 * rng.random((4, 3))
    pol = rng.standard_normal((4, 3))
    m = M.regular_mesh(3, c.nx, c.r, c.pert, c.seed)
    o = O.mesh(3, c.nx, c.r, c.pert, c.seed)
    P = M.WaveProblem(m, c.rho_e, c.mu_e, lam=c.lam_e, ncomp=3, nshots=4, precision=prec)
    P.set_damping(c.sigma_e)
    P.set_sources(srcs, pol=pol, shot=np.arange(4))
    rec = np.vstack([c.rec] * 3)
    comp = np.repeat(np.arange(3, dtype=np.int32), len(c.rec))
    P.set_receivers(rec, comp=comp)
    n = 100
    P.set_pulse(M.ricker_pulse_table(c.f0, 0.0, c.dt, n))
    P.leapfrog_init(c.dt)
    tr = P.run(n)
    OP = O.elastic_problem(o, c.lam_e, c.mu_e, c.rho_e)
    S = O.damping(o, c.sigma_e)
    pulse = O.pulse_table(c.f0, 0.0, c.dt, n)
    tol = 1e-12 if prec == "fp64" else 1e-5
    for s in range(4):
        _, _, otr = O.leapfrog(OP, O.dirac(o, srcs[s], pol[s]), pulse, c.dt, n, rec, comp, S=S)
        err = CS.rel_l2(tr[:, s, :], otr)
        print(f"damped elastic shot {s} {prec}: rel-L2 vs oracle {err:.2e}")
        assert err < tol
    if prec != "fp64":
        return
    Q = M.WaveProblem(m, c.rho_e, c.mu_e, lam=c.lam_e, ncomp=3, precision=prec)
    Q.set_damping(c.sigma_e)
    Q.set_receivers(rec, comp=comp)
    u0 = 1e-2 * rng.standard_normal(3 * m.N)
    du0 = rng.standard_normal(3 * m.N)
    u0[np.tile(o.dirichlet.astype(bool), 3)] = 0.0
    Q.leapfrog_init(c.dt, u0=u0, du0=du0)
    tq = Q.run(40)[:, 0, :]
    _, _, oq = O.leapfrog(OP, np.zeros(3 * o.N), None, c.dt, 40, rec, comp, u0=u0, du0=du0, S=S)
    assert CS.rel_l2(tq, oq) < 1e-12


def test_damped_graph_equals_eager():
    c = CS.load("ref2d_q4_damped")
    _, a = CS.gpu_run(c, "fp64", graph=True)
    _, b = CS.gpu_run(c, "fp64", graph=False)
    assert np.array_equal(a, b)


def test_damping_refusals_and_reset():
    """Adjoint and batches refuse damped problems; all-zero sigma restores
    the undamped run bit for bit."""
    m, o, mu, sigma, P = _problem(3)
    P.set_receivers([[0.4, 0.5, 0.6]])
    P.set_sources([[0.5, 0.5, 0.5]])
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, 2e-3, 40))
    P.leapfrog_init(2e-3)
    with pytest.raises(MjwError):
        AdjointSolver(P, 2e-3, 40, 4).forward()
    with pytest.raises(MjwError):
        Batch([P])
    P.set_damping(np.zeros(m.E))
    P.leapfrog_init(2e-3)
    a = P.run(40)
    Q = M.WaveProblem(m, 1.0, mu)
    Q.set_receivers([[0.4, 0.5, 0.6]])
    Q.set_sources([[0.5, 0.5, 0.5]])
    Q.set_pulse(M.ricker_pulse_table(8.0, 0.0, 2e-3, 40))
    Q.leapfrog_init(2e-3)
    assert np.array_equal(a, Q.run(40))
