"""The oracle pinned: against the reference's own known-answer tables and
solutions (tests/golden, made from Montjoie's sources), and against what the
mathematics fixes.  CPU only."""
import dataclasses
import json
import os
import sys

import numpy as np
import pytest

import cases as CS
from oracle import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import decks  # noqa: E402

GOLD = CS.GOLD


# ---------------------------------------------------------------- quadrature
def test_gl_closed_forms():
    # GaussLobattoPoints.cxx:62-85
    x, w = O.gl(1)
    assert np.array_equal(x, [0, 1]) and np.array_equal(w, [0.5, 0.5])
    x, w = O.gl(2)
    assert np.allclose(x, [0, 0.5, 1], rtol=0, atol=0) and np.allclose(w, [1 / 6, 2 / 3, 1 / 6],
                                                                         rtol=1e-16)
    x, w = O.gl(3)
    a = (5 - np.sqrt(5)) / 10
    assert np.allclose(x, [0, a, 1 - a, 1], atol=1e-15)
    assert np.allclose(w, [1 / 12, 5 / 12, 5 / 12, 1 / 12], atol=1e-15)


def test_gl_against_solin_table():
    """Known-answer table shipped with the reference (Table_Quadrature.txt)."""
    tab = json.load(open(os.path.join(GOLD, "gl_solin.json")))["rules"]
    assert len(tab) >= 10
    for n, rows in tab.items():
        r = int(n) - 1
        t = np.array([float(a) for a, _ in rows])[::-1]
        wt = np.array([float(b) for _, b in rows])[::-1]
        x, w = O.gl(r)
        np.testing.assert_allclose(x, 0.5 * (1 + t), atol=2e-16)
        np.testing.assert_allclose(w, 0.5 * wt, atol=3e-16)


@pytest.mark.parametrize("r", range(1, 9))
def test_gl_exactness(r):
    """GL with r+1 points integrates x^k exactly for k <= 2r-1
    (QuadratureTest.cc:144-171 pattern); x^{2r} is not exact."""
    x, w = O.gl(r)
    for k in range(2 * r):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14
    assert abs(np.dot(w, x ** (2 * r)) - 1.0 / (2 * r + 1)) > 1e-12


@pytest.mark.parametrize("r", range(1, 9))
def test_derivative_matrix(r):
    """D exact on polynomials of degree <= r, rows sum to 0, and equal to a
    finite-difference derivative of the product-form basis."""
    x, _ = O.gl(r)
    D = O.gl_deriv(r)
    assert np.abs(D.sum(axis=1)).max() < 1e-11
    for k in range(r + 1):
        np.testing.assert_allclose(D @ x ** k, k * x ** max(k - 1, 0) if k else 0 * x, atol=1e-10)
    h = 1e-6
    for b in range(r + 1):
        fd = [(O.phi(r, b, t + h) - O.phi(r, b, t - h)) / (2 * h) for t in x]
        np.testing.assert_allclose(D[:, b], fd, atol=1e-6 * max(1, np.abs(D).max()))


def test_pulse_matches_closed_form_ricker():
    """The GL-24 interpolant's derivative tracks the analytic Ricker
    (TimeSource.cxx:55-67) to interpolation accuracy and is 0 after 2.5/f0."""
    f0, dt, n = 10.0, 1e-3, 400
    p = O.pulse_table(f0, 0.0, dt, n)
    t = np.arange(n) * dt + dt
    exact = np.array([O.ricker(f0, s) for s in t])
    live = t < 2.5 / f0
    assert np.abs(p[live] - exact[live]).max() < 1e-6
    assert np.all(p[~live] == 0.0)


# ---------------------------------------------------------------- operators
@pytest.mark.parametrize("dim,r,pert", [(2, 1, 0.0), (2, 4, 0.15), (3, 2, 0.1), (3, 5, 0.1)])
def test_stiffness_invariants(dim, r, pert):
    """A 1 = 0 without Dirichlet, symmetry u^T A v = v^T A u, and
    u^T A u = int mu |grad u|^2 exactly for linear u."""
    m = O.mesh(dim, 3, r, pert, 3)
    mu = 1.0 + np.arange(m.E) % 3
    G, _ = O.geometry(m, mu, 1.0)
    cols = O.color(m)
    one = np.ones(m.N)
    assert np.abs(O.stiffness(m, G, one, cols)).max() < 1e-11
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal((2, m.N))
    Au, Av = O.stiffness(m, G, u, cols), O.stiffness(m, G, v, cols)
    assert abs(v @ Au - u @ Av) < 1e-11 * np.abs(u) @ np.abs(Av)
    a = np.arange(1, dim + 1) * 0.7
    lin = m.xyz @ a
    # exact energy: sum_e mu_e |a|^2 vol_e
    _, Mloc = O.geometry(m, mu, 1.0)
    vol_e = Mloc.reshape(m.E, -1).sum(axis=1)
    exact = np.sum(mu * vol_e) * (a @ a)
    assert abs(lin @ O.stiffness(m, G, lin, cols) - exact) < 1e-11 * exact


@pytest.mark.parametrize("dim,r", [(2, 1), (2, 3), (3, 1), (3, 2)])
def test_stiffness_equals_assembled_matrix(dim, r):
    """Matrix-free product == column-by-column assembled matrix built from
    brute-force element integrals with dense Lagrange derivatives
    (test_matrix_vector_product.cc:111-200 pattern), affine mesh, mu = 1."""
    nx = 2
    m = O.mesh(dim, nx, r)
    G, _ = O.geometry(m, 1.0, 1.0)
    h = 1.0 / nx
    x, w = O.gl(r)
    D = O.gl_deriv(r)
    n = r + 1
    # reference-element stiffness on [0,h]^dim: sum_q w_q h^dim grad phi_a . grad phi_b
    nd = n ** dim
    idx = np.array(np.unravel_index(np.arange(nd), (n,) * dim, order="F")).T
    K = np.zeros((nd, nd))
    for q in range(nd):
        wq = np.prod(w[idx[q]]) * h ** dim
        grads = np.zeros((nd, dim))
        for a in range(nd):
            for d in range(dim):
                g = D[idx[q][d], idx[a][d]] / h
                for d2 in range(dim):
                    if d2 != d:
                        g *= 1.0 if idx[a][d2] == idx[q][d2] else 0.0
                grads[a, d] = g
        K += wq * grads @ grads.T
    A = np.zeros((m.N, m.N))
    for e in range(m.E):
        en = m.elem_nodes[e]
        A[np.ix_(en, en)] += K
    cols = O.color(m)
    for j in range(0, m.N, max(1, m.N // 7)):
        ej = np.zeros(m.N)
        ej[j] = 1.0
        np.testing.assert_allclose(O.stiffness(m, G, ej, cols), A[:, j], atol=1e-11)


def test_colouring_is_conflict_free():
    m = O.mesh(3, 4, 2, 0.1, 1)
    c, nc = O.color(m)
    assert nc == 8
    for col in range(nc):
        nodes = m.elem_nodes[c == col].ravel()
        assert len(np.unique(nodes)) == len(nodes)


def test_mass_sums_to_volume():
    m = O.mesh(3, 3, 3, 0.12, 4)
    _, Mloc = O.geometry(m, 1.0, 2.5)
    M, _ = O.mass(m, Mloc, dirichlet=False)
    assert abs(M.sum() - 2.5) < 1e-13


def test_dirac_at_vertex_is_unit_vector_and_partition_of_unity():
    m = O.mesh(2, 8, 4)
    b = O.dirac(m, (0.5, 0.5))
    assert np.count_nonzero(b) == 1 and abs(b.max() - 1.0) < 1e-15
    m3 = O.mesh(3, 3, 3, 0.1, 2)
    b3 = O.dirac(m3, (0.47, 0.51, 0.53))
    assert abs(b3.sum() - 1.0) < 1e-13
    # exact for polynomials: sum_i b_i p(x_i) = p(x0) for p of degree <= r
    p = lambda X: 1 + X[..., 0] ** 2 - 2 * X[..., 1] * X[..., 2]  # noqa: E731
    assert abs(b3 @ p(m3.xyz) - p(np.array([0.47, 0.51, 0.53]))) < 1e-12


# ---------------------------------------------------------------- leapfrog
def test_energy_conserved_without_source():
    """E^{n+1/2} = 1/2|(U^{n+1}-U^n)/dt|_M^2 + 1/2 U^{n+1} A U^n is constant."""
    m = O.mesh(2, 6, 3, 0.1, 8)
    P = O.acoustic_problem(m, 1.0 + np.arange(m.E) % 2, 1.0)
    rng = np.random.default_rng(1)
    u0 = rng.standard_normal(m.N) * (1 - m.dirichlet)
    lam = O.lambda_max(P)
    dt = 0.5 * 2 / np.sqrt(lam)
    Es = []
    for ns in (5, 60, 200):
        uN, uN1, _ = O.leapfrog(P, None, None, dt, ns, None, None, u0)
        E = 0.5 * np.sum(P.M * (uN1 - uN) ** 2) / dt ** 2 + 0.5 * uN1 @ O.apply(P, uN)
        Es.append(E)
    assert max(Es) - min(Es) < 1e-12 * max(Es)


def test_leapfrog_second_order():
    """Error against a 4x finer time step shrinks like dt^2
    (explicit_scheme.cc:632-749 pattern)."""
    m = O.mesh(2, 4, 3)
    P = O.acoustic_problem(m, 1.0, 1.0)
    u0 = np.sin(np.pi * m.xyz[:, 0]) * np.sin(np.pi * m.xyz[:, 1])
    T = 0.2

    def run(dt):
        n = int(round(T / dt))
        uN, _, _ = O.leapfrog(P, None, None, dt, n, None, None, u0)
        return uN

    ref = run(T / 1600)
    e1 = np.linalg.norm(run(T / 100) - ref)
    e2 = np.linalg.norm(run(T / 200) - ref)
    assert abs(np.log2(e1 / e2) - 2.0) < 0.05


@pytest.mark.parametrize("name", CS.golden_names())
def test_oracle_matches_montjoie(name):
    """Traces of Montjoie's own solver (tests/golden, 15 significant digits)."""
    c = CS.load(name)
    _, _, tr = CS.oracle_run(c)
    ref = CS.reference_traces(c)
    assert tr.shape == ref.shape
    assert CS.rel_l2(tr, ref) < 1e-12


def _pulse_closed_form(kind, f0, par, t):
    """d/dt of each TemporalSource's function, differentiated by hand (the
    EvaluateDerivative formulas of Source/TimeSource.cxx agree)."""
    pi = np.pi
    if kind == "DERIVATIVE_RICKER":
        a = f0 * t - 1.0
        q = (pi * a) ** 2
        v = 2 * pi ** 2 * f0 * a * (3 - 2 * q) * np.exp(-q)
        return np.where((t <= -0.9 / f0) | (t >= 2.9 / f0), 0.0, v)
    if kind == "MODIFIED_RICKER":
        g = 2 * (pi * f0 * (t - par)) ** 2
        return (2 * g - 1) * np.exp(-g)
    if kind == "GAUSSIAN":
        s = np.where(t < 1 / f0, t, t - par)
        u = pi * (f0 * s - 1)
        v = -2 * pi * f0 * u * np.exp(-u * u)
        return np.where((t >= 1 / f0) & (t < 1 / f0 + par), 0.0, v)
    if kind == "HARMONIC":
        tt = np.clip(t, 1e-12, par - 1e-12)
        a, b = (f0 * tt) ** 2, (f0 * (tt - par)) ** 2
        e = np.exp(-1 / (pi * a)) * np.exp(-1 / (pi * b))
        w = 2 * pi * f0
        v = e * (w * np.cos(w * tt) + np.sin(w * tt) * 2 * (1 / (pi * a * tt) +
                                                              1 / (pi * b * (tt - par))))
        return np.where((t <= 0) | (t >= par), 0.0, v)
    if kind == "SINUS_GAUSSIAN":
        toff = np.sqrt(np.log(1e6) / par)
        w = 2 * pi * f0
        return np.exp(-par * (t - toff) ** 2) * (w * np.cos(w * t) - 2 * par * (t - toff) *
                                                 np.sin(w * t))
    raise ValueError(kind)


@pytest.mark.parametrize("kind,par,tol", [("DERIVATIVE_RICKER", 0.0, 1e-12),
                                          ("MODIFIED_RICKER", 0.06, 1e-12),
                                          ("GAUSSIAN", 0.05, 1e-11),
                                          ("HARMONIC", 0.2, 3e-5),
                                          ("SINUS_GAUSSIAN", 2000.0, 1e-11)])
def test_temporal_source_pulse_closed_forms(kind, par, tol):
    """The leapfrog's table of each TemporalSource (the derivative of the
    degree-24 interpolant over 0.5/f0 windows, at t_n + dt) against the
    function's derivative by hand, relative to its peak, before the cut-off;
    the cut-offs are GetNewTimeSource's.  The HARMONIC's onset
    exp(-1/(pi f0^2 t^2)) is smooth but not analytic at t = 0, so the
    polynomial of the first window departs from it by 1e-5 of the peak there
    (Montjoie interpolates the same way)."""
    f0, dt, n = 10.0, 1e-3, 300
    tab = O.source_table(kind, f0, par, 0.0, dt, n)
    t = np.arange(n) * dt + dt
    lim = {"DERIVATIVE_RICKER": lambda: 2.5 / f0,
           "MODIFIED_RICKER": lambda: (2.5 + par) / (np.sqrt(2) * f0),
           "GAUSSIAN": lambda: 2.5 / f0 + par, "HARMONIC": lambda: par,
           "SINUS_GAUSSIAN": lambda: 2 * np.sqrt(np.log(1e6) / par)}[kind]()
    assert O.source_limit(kind, f0, par) == pytest.approx(lim, rel=1e-15)
    live = t < lim
    ref = _pulse_closed_form(kind, f0, par, t)
    assert np.abs(tab[live] - ref[live]).max() <= tol * np.abs(ref[live]).max()
    assert not tab[~live].any()


def test_order4_scheme_is_fourth_order_in_time():
    """The oracle's OrderTimeScheme = 4 leapfrog (or_leapfrog4) on a smooth
    initial field: halving dt divides the error at t = 0.5 by 16 (measured
    order 4.0 +- 0.15), where the second-order scheme divides it by 4; and
    it stays stable at 1.5 times the second-order CFL limit, below its own
    2 sqrt(3)."""
    m = O.mesh(2, 4, 4, 0.1, 3)
    P = O.acoustic_problem(m, 1.0 + np.arange(m.E) % 2, 1.0)
    X = m.xyz.reshape(m.N, 2)
    u0 = np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])
    T = 0.5

    def run(k, order):
        dt = T / k
        if order == 4:
            return O.leapfrog4(P, np.zeros(m.N), (None, None, None), dt, k, u0=u0)[0]
        return O.leapfrog(P, np.zeros(m.N), None, dt, k, u0=u0)[0]

    ref = run(3200, 4)
    for order in (2, 4):
        e = [np.linalg.norm(run(k, order) - ref) for k in (100, 200)]
        assert abs(np.log2(e[0] / e[1]) - order) < 0.15, (order, e)
    lam = O.lambda_max(P, 200, 5)
    dt = 1.5 * 2.0 / np.sqrt(lam)
    n = 400
    u4 = O.leapfrog4(P, np.zeros(m.N), (None, None, None), dt, n, u0=u0)[0]
    u2 = O.leapfrog(P, np.zeros(m.N), None, dt, n, u0=u0)[0]
    assert np.abs(u4).max() < 2 * np.abs(u0).max()
    assert not np.abs(u2).max() < 1e3  # grows without bound (overflows)


def test_young_poisson_2d_is_plane_stress():
    """YOUNG_POISSON media of a 2-D deck: Montjoie's traces are matched by the
    plane-stress Lame coefficient lambda = E nu / (1 - nu^2) the fixture
    stores, and missed by the 3-D (plane-strain) E nu / ((1 + nu)(1 - 2 nu));
    undamping the layers (sigma 0, 4, 15) misses them too."""
    c = CS.load("ref2d_el_q4_young_damped")
    z = np.load(os.path.join(CS.GOLD, c.name + ".npz"))
    E, nu = z["young_E"], z["young_nu"]
    np.testing.assert_allclose(z["lam_layers"], E * nu / (1 - nu * nu), rtol=1e-15)
    np.testing.assert_allclose(z["layers"], 0.5 * E / (1 + nu), rtol=1e-15)
    ref = CS.reference_traces(c)
    _, _, tr = CS.oracle_run(c, 150)
    assert CS.rel_l2(tr, ref[:151]) < 1e-12
    e = np.arange(c.nx ** c.dim)
    lay = ((e // c.nx ** (c.dim - 1)) * len(E)) // c.nx
    np.testing.assert_array_equal(c.lam_e, z["lam_layers"][lay])
    strain = (E * nu / ((1 + nu) * (1 - 2 * nu)))[lay]
    _, _, tr3 = CS.oracle_run(dataclasses.replace(c, lam_e=strain), 150)
    _, _, tr0 = CS.oracle_run(dataclasses.replace(c, sigma_e=None), 150)
    assert CS.rel_l2(tr3, ref[:151]) > 1e-3 and CS.rel_l2(tr0, ref[:151]) > 1e-3


def test_lambda_max_bounds_cfl():
    """dt just above 2/sqrt(lambda_max) blows up, just below stays bounded."""
    m = O.mesh(2, 4, 3)
    P = O.acoustic_problem(m, 1.0, 1.0)
    lam = O.lambda_max(P)
    rng = np.random.default_rng(2)
    u0 = rng.standard_normal(m.N) * (1 - m.dirichlet)
    dmax = 2 / np.sqrt(lam)
    ok, _, _ = O.leapfrog(P, None, None, 0.98 * dmax, 3000, None, None, u0)
    bad, _, _ = O.leapfrog(P, None, None, 1.02 * dmax, 3000, None, None, u0)
    assert np.abs(ok).max() < 1e3 * np.abs(u0).max()
    assert not (np.abs(bad).max() < 1e6 * np.abs(u0).max())  # inf / nan count as blow-up


def test_elastic_rigid_motions_and_symmetry():
    m = O.mesh(3, 2, 2, 0.1, 5)
    P = O.elastic_problem(m, 2.0, 1.5, 1.0)
    # rigid translation and infinitesimal rotation are in the kernel
    t = np.concatenate([np.full(m.N, 1.0), np.full(m.N, -2.0), np.full(m.N, 0.5)])
    X = m.xyz
    rot = np.concatenate([-X[:, 1], X[:, 0], 0 * X[:, 0]])
    for v in (t, rot):
        assert np.abs(O.apply(P, v)).max() < 1e-11
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal((2, 3 * m.N))
    assert abs(v @ O.apply(P, u) - u @ O.apply(P, v)) < 1e-10


@pytest.mark.parametrize("r", [1, 2, 4, 6])
def test_elastic2d_kernel_invariants(r):
    """Quadrilaterals, two components: rigid translations and the
    infinitesimal rotation are in the kernel, A is symmetric, and for the
    linear field u = G x the energy u^T A u equals the area times
    lambda tr(e)^2 + 2 mu e:e with e = sym(G) (GL quadrature is exact for
    the bilinear Jacobian).  A transposed DF^{-1}, a dropped symmetrisation
    or a missing lambda term fails one of them."""
    m = O.mesh(2, 3, r, 0.12, 5)
    lam, mu = 2.0, 1.5
    P = O.elastic_problem(m, lam, mu, 1.0)
    assert P.ncomp == 2
    X = m.xyz
    t = np.concatenate([np.full(m.N, 1.0), np.full(m.N, -2.0)])
    rot = np.concatenate([-X[:, 1], X[:, 0]])
    for v in (t, rot):
        assert np.abs(O.apply(P, v)).max() < 1e-11
    rng = np.random.default_rng(r)
    u, v = rng.standard_normal((2, 2 * m.N))
    assert abs(v @ O.apply(P, u) - u @ O.apply(P, v)) < 1e-10 * np.abs(u @ O.apply(P, u))
    G = np.array([[0.7, -0.3], [0.45, 1.1]])
    ul = np.concatenate([X @ G[0], X @ G[1]])
    e = 0.5 * (G + G.T)
    area = 1.0  # the unit square: perturbation moves interior vertices only
    ref = area * (lam * np.trace(e) ** 2 + 2 * mu * np.sum(e * e))
    assert abs(ul @ O.apply(P, ul) - ref) < 1e-12 * ref


def test_oracle_adjoint_gradient_matches_finite_differences():
    """The FP64 reference gradient is pinned by central differences of the
    oracle's own misfit (the reference has no adjoint to compare with)."""
    m = O.mesh(2, 4, 3, 0.1, 3)
    mu0 = 1.0 + 0.5 * (np.arange(m.E) % 3)
    b = O.dirac(m, (0.43, 0.52))
    dt, N = 4e-3, 120
    pulse = O.pulse_table(10.0, 0.0, dt, N)
    rec = np.array([[0.3, 0.3], [0.7, 0.6], [0.55, 0.2]])
    dobs = O.forward_traces(O.acoustic_problem(m, mu0 * 1.1, 1.0), b, pulse, dt, N, rec)

    def J(mu):
        tr = O.forward_traces(O.acoustic_problem(m, mu, 1.0), b, pulse, dt, N, rec)
        return 0.5 * np.sum((tr - dobs) ** 2), tr

    _, tr0 = J(mu0)
    g = O.adjoint_gradient(O.acoustic_problem(m, mu0, 1.0), b, pulse, dt, N, rec, tr0 - dobs)
    rng = np.random.default_rng(0)
    for _ in range(3):
        dm = rng.standard_normal(m.E)
        eps = 1e-4
        fd = (J(mu0 + eps * dm)[0] - J(mu0 - eps * dm)[0]) / (2 * eps)
        assert abs(fd - g @ dm) < 1e-6 * abs(fd)
    # a planted sign error in the recursion would flip or scale the result
    assert abs(g @ g) > 0


def test_sinus_media_closed_forms():
    """PhysicalMedia SINUS: no cut-off gives offset + amplitude prod sin(2 pi
    x / L); outside the box the offset; on a face with a cut-off the offset;
    the cut-off factor is prod exp(-c / d^2) over both faces (the whole field
    is pinned against Montjoie by the ref*_sinus fixtures)."""
    x = np.array([[0.3, 0.7], [0.1, 0.45], [1.2, 0.5], [0.25, 0.6]])
    f = O.sinus_media(x, [0, 0], [1, 1], [0, 0], 2.0, 0.5, [0.8, 1.3])
    want = 2.0 + 0.5 * np.sin(2 * np.pi * x[:, 0] / 0.8) * np.sin(2 * np.pi * x[:, 1] / 1.3)
    want[2] = 2.0
    np.testing.assert_allclose(f, want, rtol=1e-15)
    c = [0.01, 0.02]
    g = O.sinus_media(x, [0.25, 0.2], [0.9, 0.8], c, 2.0, 0.5, [0.8, 1.3])
    p = x[0]
    cut = np.prod([np.exp(-c[a] / (p[a] - lo) ** 2) * np.exp(-c[a] / (p[a] - hi) ** 2)
                   for a, (lo, hi) in enumerate([(0.25, 0.9), (0.2, 0.8)])])
    np.testing.assert_allclose(g[0], 2.0 + 0.5 * np.sin(2 * np.pi * 0.3 / 0.8)
                               * np.sin(2 * np.pi * 0.7 / 1.3) * cut, rtol=1e-14)
    assert g[1] == 2.0 and g[2] == 2.0  # outside the box
    assert g[3] == 2.0  # on the face x = 0.25, where the cut-off vanishes


@pytest.mark.parametrize("name", ["ref2d_q4_plane", "ref3d_q3_plane"])
@pytest.mark.parametrize("kind", ["plane", "line"])
def test_oracle_snapshots_match_montjoie(name, kind):
    """SismoPlane / SismoLine files written by Montjoie: the oracle's state at
    each snapshot step, interpolated at the oracle's grid points (0 outside the
    mesh), equals the stored values; the header carries the grid type."""
    c = CS.load(name)
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    if kind == "plane":
        pts = O.plane_points(c.dim, z["plane"])
        gtype = 0 if c.dim == 2 else 2
    else:
        pts = O.line_points(c.dim, z["line"])
        gtype = 1 if c.dim == 2 else 3
    steps = CS.plane_steps(c, z[kind + "_times"])
    files = z[kind + "_files"]
    assert len(steps) == files.shape[0]
    m = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    outside = 0
    for k, s in enumerate(steps):
        head, ref = CS.plane_file(files[k], len(pts))
        assert list(np.frombuffer(head[:16], np.int32)) == [1, c.dim, 1, gtype]
        if s == 0:
            assert not ref.any()
            continue
        uN, _, _ = CS.oracle_run(c, s)
        got = O.snapshot(m, uN, pts)
        outside = int((got == 0).sum())
        # zeros agree, up to points on the boundary that one side places
        # inside (a Dirichlet value of rounding size) and the other outside
        tiny = 1e-12 * np.abs(ref).max()
        assert np.all(np.abs(ref[got == 0]) <= tiny) and np.all(np.abs(got[ref == 0]) <= tiny)
        assert CS.rel_l2(got, ref) < 1e-10
    assert outside > 0  # the grids reach outside the mesh


def _montjoie_dofs(c: CS.Case, regular: bool):
    """The oracle's Montjoie DOF number of each oracle node for a golden case."""
    m = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    corners, bnd = decks.medit_topology(c.dim, c.nx)
    if regular:
        bnd = O.regular_boundary(c.dim, c.nx)
    return m, O.montjoie_dofs(c.dim, c.r, corners, bnd, m.elem_nodes.reshape(m.E, -1), m.N)


@pytest.mark.parametrize("name", [n for n in CS.golden_names() if n.endswith(
    ("_reprise", "_reprise_regular"))])
def test_oracle_checkpoints_match_montjoie(name):
    """SaveReprise files written by Montjoie: IterationAcousSave[New] hold the
    last two checkpoints (int32 size 3, then nt, -1, b), and SolAcous<b> the
    oracle's U^nt at those steps, put in Montjoie's DOF order by the oracle's
    numbering (component-blocked for elastic), to rel-L2 1e-10."""
    c = CS.load(name)
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    saves = O.reprise_saves(c.nsteps, int(z["save_reprise"]))
    assert len(saves) >= 2
    ints = lambda k: list(np.frombuffer(z["reprise_" + k].tobytes(), np.int32))  # noqa: E731
    assert ints("IterationAcousSaveNew0N1") == [3, saves[-1][0], -1, saves[-1][1]]
    assert ints("IterationAcousSave0N1") == [3, saves[-2][0], -1, saves[-2][1]]
    m, dof = _montjoie_dofs(c, bool(z["regular"]))
    assert np.array_equal(np.sort(dof), np.arange(m.N))
    ncomp = c.dim if c.elastic else 1
    for nt, b in saves[-2:]:
        raw = z[f"reprise_SolAcous{b}_P0N1"].tobytes()
        assert np.frombuffer(raw[:4], np.int32)[0] == ncomp * m.N
        ref = np.frombuffer(raw[4:], np.float64)
        uN, _, _ = CS.oracle_run(c, nt)
        got = np.zeros(ncomp * m.N)
        for k in range(ncomp):
            got[k * m.N + dof] = uN[k * m.N:(k + 1) * m.N]
        assert np.abs(ref).max() > 0
        assert CS.rel_l2(got, ref) < 1e-10, (nt, CS.rel_l2(got, ref))


def test_oracle_montjoie_dofs_tiny():
    """The numbering on the smallest meshes, against hand-derived values: one
    Q2 square (vertices, edges (0,1) (1,2) (3,2)->(2,3) (0,3), interior) and
    one Q2 cube's face and interior nodes."""
    en = np.arange(9).reshape(1, 9)
    d = O.montjoie_dofs(2, 2, [[0, 1, 2, 3]], [], en, 9)
    # local q = i + 3 j: corners 0 2 8 6 -> vertices 0 1 2 3; edges (0,1)=4,
    # (1,2)=5, (2,3)=6 walked from 2 to 3 (i = 1 on the top edge is its middle),
    # (3,0) = 7 stored as (0,3); interior 8
    assert list(d) == [0, 4, 1, 7, 8, 5, 3, 6, 2]
    en3 = np.arange(27).reshape(1, 27)
    d3 = O.montjoie_dofs(3, 2, [[0, 1, 2, 3, 4, 5, 6, 7]], [], en3, 27)
    # 8 vertices, 12 edges (8..19), 6 faces (20..25) in the cube's face order
    # x=0, y=0, z=0, z=1, y=1, x=1, interior 26
    assert d3[13] == 26
    assert [d3[q] for q in (12, 10, 4, 22, 16, 14)] == [20, 21, 22, 23, 24, 25]
    assert sorted(d3) == list(range(27))


@pytest.mark.parametrize("name", [n for n in CS.golden_names() if "_el_" in n])
def test_oracle_elastic_stress_matches_montjoie(name):
    """Elastic SismoPoint rows end with the stress tensor of each point (d*d
    values, row-major): the oracle's gradient receivers turned into
    lambda tr(grad u) I + mu (grad u + grad u^T) match them to rel-L2 1e-10
    on every row, in 2-D and 3-D."""
    c = CS.load(name)
    npt, d = len(c.rec), c.dim
    ref = c.traces[:, d * npt:]
    assert ref.shape[1] == d * d * npt
    got = CS.oracle_stress(c)
    assert got.shape == ref.shape
    assert np.abs(ref).max() > 0
    assert CS.rel_l2(got, ref) < 1e-10, CS.rel_l2(got, ref)


def test_oracle_gradient_weights_exact_on_linear_fields():
    """Gradient weights reproduce the gradient of a linear field exactly in a
    perturbed element (straight-sided hexahedra contain the linear fields)
    and sum to zero."""
    m = O.mesh(3, 3, 4, 0.15, 5)
    pts = np.array([[0.31, 0.52, 0.47], [0.77, 0.12, 0.66]])
    el, W, Gw = O.receiver_gradients(m, pts)
    X = m.xyz.reshape(m.N, 3)
    en = m.elem_nodes.reshape(m.E, -1)
    g = np.array([0.3, -1.2, 2.5])
    for k in range(len(pts)):
        u = X[en[el[k]]] @ g + 0.7
        np.testing.assert_allclose(Gw[k] @ u, g, rtol=0, atol=1e-12)
        np.testing.assert_allclose(Gw[k].sum(axis=1), 0.0, atol=1e-12)
        np.testing.assert_allclose(W[k] @ u, pts[k] @ g + 0.7, rtol=0, atol=1e-12)


@pytest.mark.parametrize("name,kind", [("ref3d_el_q3_plane", "plane"), ("ref3d_el_q3_plane", "line"),
                                       ("ref2d_el_q4_plane", "plane"), ("ref2d_el_q4_plane", "line"),
                                       ("ref3d_el_q3_grad", "line"), ("ref2d_el_q4_grad", "plane")])
def test_oracle_elastic_snapshots_match_montjoie(kind, name):
    """Elastic SismoPlane / SismoLine files written by Montjoie: one file per
    displacement component (_U<m>) and per stress component (_dU<k>, k = d i
    + j) at each snapshot step; the oracle's fields at the grid points match
    them to rel-L2 1e-10 with the same zeros outside the mesh (hexahedra and
    quadrilaterals)."""
    c = CS.load(name)
    d = c.dim
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    pts = O.plane_points(d, z["plane"]) if kind == "plane" else O.line_points(d, z["line"])
    gtype = (2 if kind == "plane" else 3) if d == 3 else (0 if kind == "plane" else 1)
    steps = CS.plane_steps(c, z[kind + "_times"])
    files, names = z[kind + "_files"], list(z[kind + "_names"])
    prefix = "totalPlane" if kind == "plane" else "totalLine"
    suffixes = [f"_U{m}" for m in range(d)] + [f"_dU{k}" for k in range(d * d)]
    nf = len(suffixes)
    assert names == [f"{prefix}{k:04d}{sf}.dat" for k in range(len(steps)) for sf in suffixes]
    m = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    for k, s in enumerate(steps):
        uN, _, _ = CS.oracle_run(c, s)
        disp, sig = O.snapshot_elastic(m, uN, pts, c.lam_e, c.mu_e, c.display_stress)
        fields = np.vstack([disp, sig])
        for f in range(nf):
            head, ref = CS.plane_file(files[nf * k + f], len(pts))
            assert list(np.frombuffer(head[:16], np.int32)) == [1, d, 1, gtype]
            got = fields[f]
            if not ref.any():
                assert not got.any()
                continue
            tiny = 1e-12 * np.abs(ref).max()
            assert np.all(np.abs(ref[got == 0]) <= tiny) and np.all(np.abs(got[ref == 0]) <= tiny)
            assert CS.rel_l2(got, ref) < 1e-10, (s, names[nf * k + f], CS.rel_l2(got, ref))


@pytest.mark.parametrize("name", [n for n in CS.golden_names() if "circle" in n])
def test_oracle_circle_matches_montjoie(name):
    """SismoCircle files written by Montjoie (every step): the oracle's
    iterate at or_circle_points equals the stored values (0 at the points
    outside the mesh, which the one fixture has), so the point placement is
    pinned to the interpolation's sensitivity; the header carries grid type 3
    and N."""
    c = CS.load(name)
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    cx, cy, rad, n = (float(v) for v in z["circle"])
    n = int(n)
    pts = O.circle_points(cx, cy, rad, n)
    assert np.allclose(np.hypot(pts[:, 0] - cx, pts[:, 1] - cy), rad, rtol=1e-15, atol=1e-15)
    m = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    el, _, _ = O.receiver_gradients(m, pts)
    inside = el >= 0
    c.rec = pts[inside]
    _, _, tr = CS.oracle_run(c)
    files = z["circle_files"]
    assert files.shape[0] == c.nsteps + 1
    got = np.zeros((c.nsteps + 1, n))
    got[:, inside] = tr
    ref = np.stack([CS.plane_file(f, n)[1] for f in files])
    head = CS.plane_file(files[0], n)[0]
    assert list(np.frombuffer(head[:16], np.int32)) == [1, 2, 1, 3]
    assert list(np.frombuffer(head[-8:], np.int32)) == [n, n]
    assert not ref[:, ~inside].any()
    err = CS.rel_l2(got, ref)
    print(f"{name}: {int(inside.sum())}/{n} points inside, rel-L2 {err:.2e}")
    assert err < 1e-10


PMLS = [n for n in CS.golden_names() if n.endswith("_pml")]


@pytest.mark.parametrize("name", PMLS)
def test_oracle_pml_matches_montjoie(name):
    """AddPML layers + DampingPML + OrderTimeScheme = 2 LEAP_FROG_PML run by
    Montjoie: the oracle's first-order split scheme on its extruded mesh
    matches the seismograms (receivers inside the layers included) tighter
    than the generic 1e-10 golden test."""
    c = CS.load(name)
    pm, _, _, _ = CS.pml_setup(c)
    assert int(pm.pml.sum()) == pm.mesh.E - c.nx ** c.dim
    _, _, tr = CS.oracle_run(c)
    err = CS.rel_l2(tr, c.traces)
    print(f"{name}: {pm.mesh.E} elements ({int(pm.pml.sum())} in the PML), rel-L2 {err:.2e}")
    assert err < 1e-12


def test_time_table_derivatives():
    """or_time_table: nb_deriv 1 at offset dt is the leapfrog pulse table;
    nb_deriv 0 and 1 follow the closed-form Ricker primitive and pulse to the
    interpolation error; zero from 2.5/f0 on."""
    f0, dt, n = 8.0, 1e-3, 500
    assert np.array_equal(O.time_table(f0, 0.0, dt, n, 1, dt), O.pulse_table(f0, 0.0, dt, n))
    t = np.arange(n) * dt + 0.5 * dt
    a = f0 * t - 1.0
    g = np.exp(-(np.pi * a) ** 2)
    live = t < 2.5 / f0
    p0 = O.time_table(f0, 0.0, dt, n, 0, 0.5 * dt)
    p1 = O.time_table(f0, 0.0, dt, n, 1, 0.5 * dt)
    assert np.abs(p0[live] - (-a * g / f0)[live]).max() < 1e-11
    assert np.abs(p1[live] - ((2 * np.pi ** 2 * a * a - 1) * g)[live]).max() < 1e-9
    assert not p0[~live].any() and not p1[~live].any()
    # second and third derivatives: f0 2 pi^2 a (3 - 2 pi^2 a^2) g and
    # f0^2 2 pi^2 (3 - 12 pi^2 a^2 + 4 pi^4 a^4) g, to the interpolation error
    pi2 = np.pi ** 2
    p2 = O.time_table(f0, 0.0, dt, n, 2, 0.5 * dt)
    p3 = O.time_table(f0, 0.0, dt, n, 3, 0.5 * dt)
    e2 = f0 * 2 * pi2 * a * (3 - 2 * pi2 * a * a) * g
    e3 = f0 ** 2 * 2 * pi2 * (3 - 12 * pi2 * a * a + 4 * pi2 * pi2 * a ** 4) * g
    assert np.abs(p2[live] - e2[live]).max() < 1e-9 * np.abs(e2).max()
    assert np.abs(p3[live] - e3[live]).max() < 1e-8 * np.abs(e3).max()
    assert not p2[~live].any() and not p3[~live].any()


def test_pml_mesh_geometry():
    """The extruded layers: vertex coordinates of every layer, kinds and
    media of the corner elements, profile zero on the box faces and
    dsig a^2 on the outer faces."""
    pm = O.pml_mesh(2, 4, 3, 0.1, 5, 0.3, 3)
    m = pm.mesh
    assert m.nx == 10 and pm.pml.sum() == 100 - 16
    xs = np.unique(np.round(m.vert.reshape(-1, 2)[:, 0], 12))
    for k in range(1, 4):
        assert np.any(np.isclose(xs, -0.1 * k)) and np.any(np.isclose(xs, 1 + 0.1 * k))
    assert tuple(pm.kind[0]) == (-1, -1) and pm.phys[0] == 0
    assert tuple(pm.kind[99]) == (1, 1) and pm.phys[99] == 15
    tau = O.pml_tau(pm, 2.0, 1.5)
    dsig = np.log(1000.0) * 3 * 2.0 * 1.5 / (2 * 0.3 ** 3)
    X = m.xyz[m.elem_nodes]
    outer = np.isclose(X[:, :, 0], -0.3) & (pm.kind[:, 0] == -1)[:, None]
    assert np.allclose(tau[:, :, 0][outer], dsig * 0.09)
    face = np.isclose(X[:, :, 0], 0.0)
    assert np.all(tau[:, :, 0][face] == 0.0)
    assert np.all(tau[pm.pml == 0] == 0.0)
