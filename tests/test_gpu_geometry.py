"""The two geometry modes of the 3-D acoustic kernels.

STORED reads the per-node tensor J w mu DF^{-1} DF^{-T} that mjw_create
computed from the node positions (isoparametric Q_r map); TRILINEAR rebuilds
it in the kernel from the trilinear map of the element's 8 vertices, which
is the same map whenever the element is straight-sided.  Both are checked
against the oracle (which maps every element by its 8 vertices, as Montjoie
does for straight elements), against each other, and on a curved mesh,
where only STORED applies."""
import dataclasses

import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from montjoie_amd._lib import MjwError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def curve(xyz, amp=0.03):
    """Smooth non-trilinear displacement that keeps the cube's faces."""
    s = np.prod(np.sin(np.pi * xyz), axis=1)
    out = xyz.copy()
    out[:, 0] += amp * s
    out[:, 1] -= 0.5 * amp * s
    return out


def test_modes_detected():
    m3 = M.regular_mesh(3, 2, 3, 0.1, seed=1)
    P = M.WaveProblem(m3, 1.0, 1.0)
    assert P.trilinear_available and P.geometry == "trilinear"
    P.set_geometry("stored")
    assert P.geometry == "stored"
    P.set_geometry("trilinear")
    assert P.geometry == "trilinear"
    P2 = M.WaveProblem(M.regular_mesh(2, 3, 3, 0.1, seed=1), 1.0, 1.0)
    assert not P2.trilinear_available and P2.geometry == "stored"
    Pe = M.WaveProblem(m3, 1.0, 1.0, lam=1.0, ncomp=3)
    assert Pe.trilinear_available and Pe.geometry == "trilinear"
    mc = dataclasses.replace(m3, xyz=curve(m3.xyz))
    Pc = M.WaveProblem(mc, 1.0, 1.0)
    assert not Pc.trilinear_available and Pc.geometry == "stored"
    with pytest.raises(MjwError):
        Pc.set_geometry("trilinear")


@pytest.mark.parametrize("r,nx", [(1, 4), (2, 3), (3, 3), (4, 2), (5, 2), (6, 2), (7, 2)])
@pytest.mark.parametrize("mode", ["trilinear", "stored"])
def test_apply_matches_oracle_both_modes(r, nx, mode):
    m = M.regular_mesh(3, nx, r, 0.1, seed=8)
    o = O.mesh(3, nx, r, 0.1, 8)
    mu = 1.0 + (np.arange(m.E) % 3)
    P = M.WaveProblem(m, 1.0, mu)
    P.set_geometry(mode)
    G, _ = O.geometry(o, mu, 1.0)
    u = np.random.default_rng(r).standard_normal(m.N)
    ref = -O.stiffness(o, G, u, O.color(o))
    assert CS.rel_l2(P.apply_kh(u), ref) < 1e-12


@pytest.mark.parametrize("r,nx", [(1, 3), (2, 2), (3, 2), (4, 2), (5, 1), (6, 1), (7, 1)])
@pytest.mark.parametrize("mode", ["trilinear", "stored"])
def test_elastic_apply_both_modes(r, nx, mode):
    m = M.regular_mesh(3, nx, r, 0.1, seed=6)
    o = O.mesh(3, nx, r, 0.1, 6)
    lam = 1.0 + (np.arange(m.E) % 3)
    mu = 2.0 + (np.arange(m.E) % 2)
    P = M.WaveProblem(m, 1.0, mu, lam=lam, ncomp=3)
    P.set_geometry(mode)
    Po = O.elastic_problem(o, lam, mu, 1.0)
    u = np.random.default_rng(r).standard_normal(3 * m.N)
    assert CS.rel_l2(P.apply_kh(u), -O.apply(Po, u)) < 1e-12


@pytest.mark.parametrize("prec", ["fp32", "fp32_g64"])
def test_fp32_trilinear_against_oracle(prec):
    m = M.regular_mesh(3, 3, 5, 0.1, seed=5)
    o = O.mesh(3, 3, 5, 0.1, 5)
    P = M.WaveProblem(m, 1.0, 2.0, precision=prec)
    assert P.geometry == "trilinear"
    G, _ = O.geometry(o, 2.0, 1.0)
    u = np.random.default_rng(0).standard_normal(m.N)
    ref = -O.stiffness(o, G, u, O.color(o))
    assert CS.rel_l2(P.apply_kh(u), ref) < 3e-6


def test_curved_mesh_stored_geometry_identities():
    """Curved elements (the oracle maps elements by their vertices only, so
    it does not cover them): the isoparametric operator must satisfy
    A 1 = 0, symmetry, and u^T A u = mu |g|^2 sum(M)/rho for u = g.x,
    since the discrete gradient of the interpolated linear field is exact
    for any isoparametric map."""
    m = M.regular_mesh(3, 2, 4, 0.0, seed=0)
    mc = dataclasses.replace(m, xyz=curve(m.xyz))
    mu = 1.5
    P = M.WaveProblem(mc, 1.0, mu, dirichlet=False)
    assert P.geometry == "stored"
    one = np.ones(mc.N)
    assert np.abs(P.apply_kh(one)).max() < 1e-12
    rng = np.random.default_rng(4)
    u, v = rng.standard_normal(mc.N), rng.standard_normal(mc.N)
    uAv, vAu = u @ P.apply_kh(v), v @ P.apply_kh(u)
    assert abs(uAv - vAu) < 1e-12 * abs(uAv)
    g = np.array([0.3, -1.1, 0.7])
    lin = mc.xyz @ g
    e = -lin @ P.apply_kh(lin)
    assert abs(e - mu * (g @ g) * P.mass().sum()) < 1e-12 * e
    # the curved operator is not the straight one
    Pf = M.WaveProblem(m, 1.0, mu, dirichlet=False)
    assert CS.rel_l2(Pf.apply_kh(u), P.apply_kh(u)) > 1e-3


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_leapfrog_modes_agree(prec, tol):
    m = M.regular_mesh(3, 4, 4, 0.1, seed=3)
    out = {}
    for mode in ("stored", "trilinear"):
        P = M.WaveProblem(m, 1.0, M.layered_values(m, [1.0, 4.0, 2.25]), precision=prec)
        P.set_geometry(mode)
        P.set_sources([(0.5, 0.5, 0.5)])
        P.set_receivers(np.array([[0.3, 0.4, 0.6], [0.7, 0.2, 0.5]]))
        dt = 1e-3
        P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, 300))
        P.leapfrog_init(dt)
        tr = P.run(299)
        out[mode] = (tr, P.state()[1])
    assert CS.rel_l2(out["trilinear"][0], out["stored"][0]) < tol
    assert CS.rel_l2(out["trilinear"][1], out["stored"][1]) < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_elastic_leapfrog_modes_agree(prec, tol):
    m = M.regular_mesh(3, 3, 3, 0.1, seed=4)
    out = {}
    for mode in ("stored", "trilinear"):
        P = M.WaveProblem(m, 1.0, M.layered_values(m, [1.0, 2.0, 1.5]),
                          lam=M.layered_values(m, [2.0, 1.0, 3.0]), ncomp=3, nshots=2,
                          precision=prec)
        P.set_geometry(mode)
        P.set_sources([(0.5, 0.5, 0.5), (0.4, 0.6, 0.55)], pol=[(0, 0, 1), (1, 0.5, 0)],
                      shot=[0, 1])
        P.set_receivers(np.array([[0.3, 0.4, 0.6], [0.7, 0.2, 0.5]]), comp=[0, 2])
        dt = 1e-3
        P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, 200))
        P.leapfrog_init(dt)
        out[mode] = (P.run(199), P.state()[1])
    assert CS.rel_l2(out["trilinear"][0], out["stored"][0]) < tol
    assert CS.rel_l2(out["trilinear"][1], out["stored"][1]) < tol


@pytest.mark.parametrize("mode", ["trilinear", "stored"])
def test_set_mu_equals_fresh_problem(mode):
    m = M.regular_mesh(3, 3, 3, 0.1, seed=2)
    mu0 = np.full(m.E, 1.0)
    mu1 = 1.0 + np.random.default_rng(1).random(m.E)
    P = M.WaveProblem(m, 1.0, mu0)
    P.set_geometry(mode)
    P.set_mu(mu1)
    Q = M.WaveProblem(m, 1.0, mu1)
    Q.set_geometry(mode)
    u = np.random.default_rng(2).standard_normal(m.N)
    assert CS.rel_l2(P.apply_kh(u), Q.apply_kh(u)) < 1e-14


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_solver_node_order_and_device_vectors(prec):
    """Device vectors live in the solver node order (and, for several 3-D
    acoustic shots, the interleaved layout); host entry points and
    mjw_apply_kh_dev take the caller's order and layout."""
    import torch
    m = M.regular_mesh(3, 3, 4, 0.1, seed=7)
    P = M.WaveProblem(m, 1.0, 1.0 + (np.arange(m.E) % 2), precision=prec, nshots=2)
    idx = P.node_order()
    assert np.array_equal(np.sort(idx), np.arange(m.N))
    dt = np.float64 if prec == "fp64" else np.float32
    rng = np.random.default_rng(3)
    u = rng.standard_normal((2, m.N))
    y = rng.standard_normal((2, m.N))
    ref = np.stack([P.apply_kh(u[s], alpha=1.5, beta=0.5, Y=y[s]) for s in range(2)])
    U = torch.tensor(u.astype(dt), device="cuda")
    Y = torch.tensor(y.astype(dt), device="cuda")
    P.apply_kh_dev(U.data_ptr(), Y.data_ptr(), alpha=1.5, beta=0.5)
    torch.cuda.synchronize()
    tol = 1e-13 if prec == "fp64" else 1e-5
    assert CS.rel_l2(Y.cpu().numpy().astype(np.float64), ref) < tol
    # state_dev: solver order
    P.set_sources([(0.5, 0.5, 0.5), (0.4, 0.6, 0.5)], shot=[0, 1])
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, 1e-3, 40))
    P.leapfrog_init(1e-3)
    P.advance(30, traces=False)
    un, un1 = P.state()
    a, _ = P.state_dev()
    # 3-D with several shots: interleaved, u[g*S + s]
    dev = device_array(a, 2 * m.N, dt).reshape(m.N, 2).T
    np.testing.assert_array_equal(dev[:, idx].astype(np.float64), un)


class _DevPtr:
    """__cuda_array_interface__ view of a raw device pointer."""
    def __init__(self, ptr, n, dt):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": np.dtype(dt).str,
                                         "data": (ptr, False), "version": 2}


def device_array(ptr, n, dt):
    import torch
    t = torch.as_tensor(_DevPtr(ptr, n, dt), device="cuda")
    torch.cuda.synchronize()
    return t.cpu().numpy()
