"""Isotropic elastodynamics on quadrilaterals (two components, k_elastic2d)
against the oracle's restatement of Montjoie's product and leapfrog, and
every launch shape against each other.  Montjoie's own 2-D elastic
solutions are compared in test_gpu_parity.py (fixtures ref2d_el_*)."""
import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from montjoie_amd import tuner
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("r,nx", [(1, 5), (2, 4), (3, 3), (4, 3), (5, 2), (6, 2), (7, 2)])
def test_elastic2d_apply_matches_oracle(r, nx):
    m = M.regular_mesh(2, nx, r, 0.12, 8)
    o = O.mesh(2, nx, r, 0.12, 8)
    lam = 1.0 + (np.arange(m.E) % 3)
    mu = 2.0 + (np.arange(m.E) % 2)
    rho = 1.0 + 0.5 * (np.arange(m.E) % 2)
    P = M.WaveProblem(m, rho, mu, lam=lam, ncomp=2)
    Po = O.elastic_problem(o, lam, mu, rho)
    u = np.random.default_rng(r).standard_normal(2 * m.N)
    y0 = np.random.default_rng(r + 1).standard_normal(2 * m.N)
    ref = -O.apply(Po, u)
    assert CS.rel_l2(P.apply_kh(u), ref) < 1e-12
    assert CS.rel_l2(P.apply_kh(u, alpha=2.0, beta=0.5, Y=y0), 0.5 * y0 + 2 * ref) < 1e-12
    np.testing.assert_allclose(P.mass(), Po.M, rtol=1e-13)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32_g64", 1e-5), ("fp32", 1e-5)])
@pytest.mark.parametrize("r", [3, 6])
def test_elastic2d_leapfrog_matches_oracle(prec, tol, r):
    """Layered Lame parameters and density, Dirac with a polarisation,
    receivers on both components; traces and final state."""
    nx, pert, seed = 4, 0.1, 9
    m = M.regular_mesh(2, nx, r, pert, seed)
    lay = (np.arange(m.E) // nx) * 3 // nx
    lam = np.array([1.5, 3.0, 2.2])[lay]
    mu = np.array([1.0, 2.5, 1.7])[lay]
    rho = np.array([1.0, 1.6, 1.3])[lay]
    P = M.WaveProblem(m, rho, mu, lam=lam, ncomp=2, precision=prec)
    src, pol = np.array([0.47, 0.54]), np.array([0.6, -0.8])
    rec = np.array([[0.3, 0.4], [0.7, 0.55], [0.52, 0.21]])
    pts = np.vstack([rec, rec])
    comp = np.repeat(np.arange(2), len(rec))
    f0, nsteps = 10.0, 120
    dt = 0.4 * 2.0 / np.sqrt(P.lambda_max(60, 1))
    P.set_sources(src, pol=pol)
    P.set_receivers(pts, comp=comp)
    P.set_pulse(M.ricker_pulse_table(f0, 0.0, dt, nsteps))
    P.leapfrog_init(dt)
    tr = P.run(nsteps)[:, 0, :]
    uN = P.state()[0][0]
    o = O.mesh(2, nx, r, pert, seed)
    Po = O.elastic_problem(o, lam, mu, rho)
    oN, _, otr = O.leapfrog(Po, O.dirac(o, src, pol), O.pulse_table(f0, 0.0, dt, nsteps), dt,
                            nsteps, pts, comp.astype(np.int32))
    e_tr, e_u = CS.rel_l2(tr, otr), CS.rel_l2(uN, oN)
    print(f"Q{r} {prec}: traces {e_tr:.2e}, state {e_u:.2e}")
    assert e_tr < tol and e_u < tol


def test_elastic2d_every_tile_and_graph_bitwise():
    m = M.regular_mesh(2, 6, 4, 0.1, 3)
    P = M.WaveProblem(m, 1.2, 2.0, lam=1.5, ncomp=2, nshots=2)
    P.set_sources(np.array([[0.45, 0.52], [0.6, 0.3]]), pol=np.array([[1.0, 0.0], [0.3, 0.7]]),
                  shot=np.arange(2))
    P.set_receivers(np.array([[0.3, 0.4], [0.3, 0.4]]), comp=np.array([0, 1]))
    dt = 0.4 * 2.0 / np.sqrt(P.lambda_max(40, 1))
    P.set_pulse(M.ricker_pulse_table(10.0, 0.0, dt, 64))
    out = []
    for tile in tuner.tile_choices(P):
        P.set_tile(tile)
        for graph in (True, False):
            P.leapfrog_init(dt)
            out.append(P.run(64, graph=graph))
    assert len(out) >= 4
    for tr in out[1:]:
        assert np.array_equal(tr, out[0])
    # the two shots are independent problems: each equals its one-shot run
    for s in range(2):
        Q = M.WaveProblem(m, 1.2, 2.0, lam=1.5, ncomp=2)
        Q.set_sources(np.array([[0.45, 0.52], [0.6, 0.3]])[s],
                      pol=np.array([[1.0, 0.0], [0.3, 0.7]])[s])
        Q.set_receivers(np.array([[0.3, 0.4], [0.3, 0.4]]), comp=np.array([0, 1]))
        Q.set_pulse(M.ricker_pulse_table(10.0, 0.0, dt, 64))
        Q.leapfrog_init(dt)
        assert np.array_equal(Q.run(64)[:, 0, :], out[0][:, s, :])
        Q.close()


def test_elastic2d_energy_conserved():
    m = M.regular_mesh(2, 5, 4, 0.1, 4)
    P = M.WaveProblem(m, 1.3, 2.0, lam=1.0, ncomp=2)
    P.set_sources(np.array([0.5, 0.5]), pol=np.array([0.8, 0.6]))
    dt = 0.4 * 2.0 / np.sqrt(P.lambda_max(40, 1))
    P.set_pulse(M.ricker_pulse_table(10.0, 0.0, dt, 400))
    P.leapfrog_init(dt)
    P.advance(300, traces=False)  # the pulse ends at t = 2.5 / f0
    e1 = P.energy()
    P.advance(100, traces=False)
    e2 = P.energy()
    assert abs(e2 - e1) < 1e-12 * abs(e1)
