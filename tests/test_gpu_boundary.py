"""The operator-level boundary on the GPU: Montjoie's own leapfrog scheme
(ModifiedEquationIterator + RunSecondOrderScheme, compiled from the reference
into oracle/_ref/mjref_drive_hip.x) stepping through libmjwave.so by way of
include/mjwave_montjoie.hpp; the same call sequence replayed through the C++
adapter (tests/adapter_replay.cpp); the mass/damping operators, the time
source and the Dirichlet entry against the oracle and closed forms; several
sources of one shot; the energy against the oracle's M and A."""
import os
import subprocess

import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVE = os.path.join(ROOT, "oracle", "_ref", "mjref_drive_hip.x")
DAMPED = ["ref2d_q4_damped", "ref3d_q3_damped"]


def _write_input(path, c):
    m = M.regular_mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    sig = c.sigma_e if c.sigma_e is not None else None
    with open(path, "wb") as f:
        np.array([c.dim, c.r, c.nsteps, int(sig is not None)], np.int32).tofile(f)
        np.array([m.N, m.E], np.int64).tofile(f)
        np.array([c.dt, c.f0], np.float64).tofile(f)
        np.ascontiguousarray(m.xyz, np.float64).tofile(f)
        np.ascontiguousarray(m.elem_nodes, np.int32).tofile(f)
        np.ascontiguousarray(m.dirichlet, np.uint8).tofile(f)
        np.ascontiguousarray(c.rho_e, np.float64).tofile(f)
        np.ascontiguousarray(c.mu_e, np.float64).tofile(f)
        if sig is not None:
            np.ascontiguousarray(sig, np.float64).tofile(f)
        np.ascontiguousarray(c.src, np.float64).tofile(f)
    return m


def _traces_of_states(c, states_path, N):
    """Receiver values of every stored iterate, through the oracle's
    interpolation weights (rows n = 0..nsteps)."""
    U = np.fromfile(states_path, np.float64).reshape(-1, N)
    o = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    el, W = O.receivers(o, c.rec)
    en = o.elem_nodes
    return np.stack([U[:, en[el[r]]] @ W[r] for r in range(len(el))], axis=1)


@pytest.mark.parametrize("name", DAMPED)
def test_montjoie_scheme_drives_library(name, tmp_path):
    """Montjoie's ModifiedEquationIterator calls HipOdeSystem's operators (Kh,
    Sh, Dh, DhMinusdtSh, DhPlusdtSh, AddScalarTimeSource, Dirichlet) on a
    damped deck; the iterates it hands to GiveIterate match Montjoie's own
    seismograms."""
    if not os.path.exists(DRIVE):
        pytest.skip("oracle/_ref/mjref_drive_hip.x not built (needs the reference sources)")
    c = CS.load(name)
    m = _write_input(tmp_path / "in.bin", c)
    print(f"drive {name}", flush=True)
    subprocess.run([DRIVE, str(tmp_path / "in.bin"), str(tmp_path / "st.bin")], check=True,
                   timeout=110)
    tr = _traces_of_states(c, tmp_path / "st.bin", m.N)
    assert tr.shape == c.traces.shape
    err = CS.rel_l2(tr, c.traces)
    print(f"{name}: Montjoie's scheme over libmjwave vs Montjoie rel-L2 {err:.2e}")
    assert err < 1e-10


@pytest.fixture(scope="module")
def replay_exe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("replay") / "adapter_replay"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "adapter_replay.cpp"), "-o", str(exe),
                    "-L", os.path.join(ROOT, "montjoie_amd"), "-lmjwave",
                    "-Wl,-rpath," + os.path.join(ROOT, "montjoie_amd")], check=True)
    return exe


@pytest.mark.parametrize("name", DAMPED)
def test_adapter_replay_matches_montjoie(name, tmp_path, replay_exe):
    """ModifiedEquation.cxx:99-124,295-333 replayed call by call through the
    C++ adapter on host vectors, damped media: Montjoie's traces to 1e-10."""
    c = CS.load(name)
    m = _write_input(tmp_path / "in.bin", c)
    print(f"replay {name}", flush=True)
    subprocess.run([str(replay_exe), str(tmp_path / "in.bin"), str(tmp_path / "st.bin")],
                   check=True, timeout=110)
    tr = _traces_of_states(c, tmp_path / "st.bin", m.N)
    err = CS.rel_l2(tr, c.traces)
    print(f"{name}: adapter replay vs Montjoie rel-L2 {err:.2e}")
    assert err < 1e-10


@pytest.mark.parametrize("dim", [2, 3])
def test_mass_operators_match_oracle(dim):
    """D, S, D - dt/2 S, D^{-1}, (D + dt/2 S)^{-1} on host vectors against the
    oracle's assembled lumped mass and damping (Dirichlet rows: D = 1, S = 0)."""
    m = M.regular_mesh(dim, 3, 4, 0.1, seed=9)
    o = O.mesh(dim, 3, 4, 0.1, 9)
    mu = 1.0 + np.arange(m.E) % 3
    rho = 1.0 + 0.5 * (np.arange(m.E) % 2)
    sigma = np.where(np.arange(m.E) % 3 == 0, 7.0, 0.5)
    P = M.WaveProblem(m, rho, mu)
    P.set_damping(sigma)
    Mo = O.acoustic_problem(o, mu, rho).M
    So = O.damping(o, sigma)
    So[o.dirichlet.astype(bool)] = 0.0
    rng = np.random.default_rng(dim)
    u, y = rng.standard_normal(m.N), rng.standard_normal(m.N)
    dt = 3e-3
    close = lambda a, b: CS.rel_l2(a, b) < 1e-13  # noqa: E731
    assert close(P.apply_dh(u, 2.0, 0.5, y), 0.5 * y + 2.0 * Mo * u)
    assert close(P.apply_sh(u, -1.5, 1.0, y), y - 1.5 * So * u)
    assert close(P.apply_dh_minus_dt_sh(dt, u, -1.0, dt * dt, y),
                 dt * dt * y - (Mo - 0.5 * dt * So) * u)
    assert close(P.solve_dh(y), y / Mo)
    assert close(P.solve_dh_plus_dt_sh(dt, y), y / (Mo + 0.5 * dt * So))
    z = P.set_dirichlet_condition(y)
    d = o.dirichlet.astype(bool)
    assert np.all(z[d] == 0.0) and np.array_equal(z[~d], y[~d])
    # without damping the S terms vanish
    Q = M.WaveProblem(m, rho, mu)
    assert np.array_equal(Q.apply_sh(u), np.zeros(m.N))
    assert np.array_equal(Q.solve_dh_plus_dt_sh(dt, y), Q.solve_dh(y))


def _ricker(f0, t, k):
    """Closed forms of the Ricker primitive (k = 0), the pulse (1) and its
    derivative (2): a = f0 t - 1, P = -a e^{-pi^2 a^2} / f0."""
    a = f0 * t - 1.0
    g = np.exp(-(np.pi * a) ** 2)
    return [-a * g / f0, (2 * np.pi ** 2 * a * a - 1) * g,
            2 * np.pi ** 2 * f0 * a * (3 - 2 * np.pi ** 2 * a * a) * g][k]


def test_add_scalar_time_source():
    """AddScalarTimeSource: at the step times the pulse equals the leapfrog
    table entry; the interpolated primitive and second derivative follow
    the closed forms; nothing from t = 2.5/f0 on; only the shot's sources."""
    m = M.regular_mesh(3, 3, 3, 0.1, seed=4)
    o = O.mesh(3, 3, 3, 0.1, 4)
    P = M.WaveProblem(m, 1.0, 1.0, nshots=2)
    src = np.array([[0.43, 0.52, 0.47], [0.61, 0.38, 0.55]])
    P.set_sources(src, shot=[0, 1])
    f0, dt, n = 8.0, 2e-3, 200
    # the windows only move forward: shot 1 at an early time first
    P.set_time_source(f0, 0.0, n * dt * 1e-15)
    y1 = P.add_scalar_time_source(0.05, None, alpha=2.0, shot=1)
    assert CS.rel_l2(y1, 2.0 * _ricker(f0, 0.05, 1) * O.dirac(o, src[1])) < 1e-9
    P.set_time_source(f0, 0.0, n * dt * 1e-15)
    tab = O.pulse_table(f0, 0.0, dt, n)
    b0 = O.dirac(o, src[0])
    peak = max(abs(_ricker(f0, t, 2)) for t in np.linspace(0, 2.5 / f0, 400))
    for k in range(n):
        t = k * dt + dt
        y = P.add_scalar_time_source(t, None, alpha=1.0, nb_deriv=1, shot=0)
        assert CS.rel_l2(y, tab[k] * b0) < 1e-14 or (tab[k] == 0 and not y.any())
        if t < 2.5 / f0:
            y0 = P.add_scalar_time_source(t, None, nb_deriv=0, shot=0)
            y2 = P.add_scalar_time_source(t, None, nb_deriv=2, shot=0)
            i = int(np.argmax(np.abs(b0)))
            assert abs(y0[i] / b0[i] - _ricker(f0, t, 0)) < 1e-9 / f0
            assert abs(y2[i] / b0[i] - _ricker(f0, t, 2)) < 1e-7 * peak
    y = P.add_scalar_time_source(2.5 / f0, np.ones(m.N), nb_deriv=1)
    assert np.array_equal(y, np.ones(m.N))


@pytest.mark.parametrize("kind,par", [("DERIVATIVE_RICKER", 0.0), ("MODIFIED_RICKER", 0.06),
                                      ("GAUSSIAN", 0.05), ("HARMONIC", 0.2),
                                      ("SINUS_GAUSSIAN", 2000.0)])
def test_add_scalar_time_source_every_kind(kind, par):
    """AddScalarTimeSource with each TemporalSource kind
    (mjw_set_temporal_source): at the step times the pulse times the source
    vector equals the oracle's table entry; nothing from the kind's cut-off
    on."""
    m = M.regular_mesh(2, 4, 3, 0.1, seed=8)
    o = O.mesh(2, 4, 3, 0.1, 8)
    P = M.WaveProblem(m, 1.0, 1.0)
    src = np.array([[0.43, 0.52]])
    P.set_sources(src)
    f0, dt, n = 10.0, 1e-3, 320
    P.set_time_source(f0, 0.0, n * dt * 1e-15, kind=kind, param=par)
    tab = O.source_table(kind, f0, par, 0.0, dt, n)
    b0 = O.dirac(o, src[0])
    scale = np.abs(tab).max()
    for k in range(n):
        y = P.add_scalar_time_source(k * dt + dt, None, alpha=1.0, nb_deriv=1)
        assert np.abs(y - tab[k] * b0).max() <= 1e-13 * scale * np.abs(b0).max()
        if tab[k] == 0 and k * dt + dt >= O.source_limit(kind, f0, par):
            assert not y.any()
    assert (np.arange(n) * dt + dt >= O.source_limit(kind, f0, par)).any()


def test_two_sources_in_one_element_of_one_shot():
    """Two Dirac points of one shot in the same element: the shot's source
    vector is their sum (one entry per node, no write conflict); traces
    match the oracle run with the summed source, and graph replays are
    bitwise equal to eager steps."""
    m = M.regular_mesh(3, 3, 4, 0.1, seed=12)
    o = O.mesh(3, 3, 4, 0.1, 12)
    mu = 1.0 + np.arange(m.E) % 2
    pts = np.array([[0.45, 0.5, 0.52], [0.52, 0.44, 0.48]])
    rec = np.array([[0.3, 0.4, 0.6], [0.7, 0.55, 0.35], [0.5, 0.5, 0.5]])
    dt, n = 3e-3, 120
    runs = []
    for graph in (True, False):
        P = M.WaveProblem(m, 1.0, mu)
        P.set_sources(pts, shot=[0, 0])
        P.set_receivers(rec)
        P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, n))
        P.leapfrog_init(dt)
        runs.append(P.run(n, graph=graph)[:, 0, :])
    assert np.array_equal(runs[0], runs[1])
    b = O.dirac(o, pts[0]) + O.dirac(o, pts[1])
    _, _, otr = O.leapfrog(O.acoustic_problem(o, mu, 1.0), b, O.pulse_table(8.0, 0.0, dt, n),
                           dt, n, rec)
    assert CS.rel_l2(runs[0], otr) < 1e-12


@pytest.mark.parametrize("dim,ncomp", [(2, 1), (3, 1), (2, 2), (3, 3)])
def test_energy_matches_oracle(dim, ncomp):
    """mjw_energy = 1/2 |(U^{n+1} - U^n)/dt|_M^2 + 1/2 (U^{n+1})^T A U^n with the
    oracle's M and A, after 60 steps of a source run (FP64)."""
    m = M.regular_mesh(dim, 3, 3, 0.1, seed=21)
    o = O.mesh(dim, 3, 3, 0.1, 21)
    mu = 1.0 + np.arange(m.E) % 3
    src = [0.45] * dim
    dt, n = 2e-3, 60
    if ncomp == 1:
        P = M.WaveProblem(m, 1.0, mu)
        OP = O.acoustic_problem(o, mu, 1.0)
        P.set_sources(src)
    else:
        lam, rho = 2.0 + np.arange(m.E) % 2, 1.0 + 0.5 * (np.arange(m.E) % 3)
        P = M.WaveProblem(m, rho, mu, lam=lam, ncomp=dim)
        OP = O.elastic_problem(o, lam, mu, rho)
        P.set_sources(src, pol=[0.3, -0.5, 0.8][:dim])
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, n))
    P.leapfrog_init(dt)
    P.advance(n, traces=False)
    un, un1 = (x[0] for x in P.state())
    Mv = np.tile(OP.M, ncomp)
    e = 0.5 * np.sum(Mv * ((un1 - un) / dt) ** 2) + 0.5 * un1 @ O.apply(OP, un)
    got = P.energy()[0]
    print(f"dim {dim} ncomp {ncomp}: energy {got:.15e} oracle {e:.15e}")
    assert abs(got - e) < 1e-12 * abs(e)
