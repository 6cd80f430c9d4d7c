"""Parity of the HIP path (through the C ABI) with the oracle and with
Montjoie's own solutions (tests/golden).  Tolerances: FP64 paths 1e-10
relative L2 against Montjoie (the requirement) and 1e-12 against the
oracle; FP32 paths are compared with the FP64 result and their error is
bounded explicitly."""
import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,r,nx,pert", [(2, 1, 5, 0.0), (2, 4, 4, 0.15), (2, 7, 2, 0.1),
                                           (3, 1, 4, 0.1), (3, 2, 3, 0.1), (3, 3, 3, 0.1),
                                           (3, 4, 2, 0.1), (3, 5, 2, 0.1), (3, 6, 2, 0.1),
                                           (3, 7, 2, 0.05)])
def test_apply_matches_oracle_fp64(dim, r, nx, pert):
    m = M.regular_mesh(dim, nx, r, pert, seed=5)
    o = O.mesh(dim, nx, r, pert, 5)
    mu = 1.0 + (np.arange(m.E) % 3)
    rho = 1.0 + 0.5 * (np.arange(m.E) % 2)
    P = M.WaveProblem(m, rho, mu)
    G, Mloc = O.geometry(o, mu, rho)
    cols = O.color(o)
    rng = np.random.default_rng(r)
    u = rng.standard_normal(m.N)
    y0 = rng.standard_normal(m.N)
    ref = -O.stiffness(o, G, u, cols)
    got = P.apply_kh(u)
    assert CS.rel_l2(got, ref) < 1e-12
    # beta / alpha handling: Y = 0.5 Y + 2 K u
    got2 = P.apply_kh(u, alpha=2.0, beta=0.5, Y=y0)
    assert CS.rel_l2(got2, 0.5 * y0 + 2 * ref) < 1e-12
    Mo, _ = O.mass(o, Mloc)
    np.testing.assert_allclose(P.mass(), Mo, rtol=1e-13)


@pytest.mark.parametrize("prec,tol", [("fp32", 3e-6), ("fp32_g64", 3e-6)])
def test_apply_fp32(prec, tol):
    m = M.regular_mesh(3, 3, 5, 0.1, seed=5)
    o = O.mesh(3, 3, 5, 0.1, 5)
    P = M.WaveProblem(m, 1.0, 2.0, precision=prec)
    G, _ = O.geometry(o, 2.0, 1.0)
    u = np.random.default_rng(0).standard_normal(m.N)
    ref = -O.stiffness(o, G, u, O.color(o))
    assert CS.rel_l2(P.apply_kh(u), ref) < tol


@pytest.mark.parametrize("r,nx", [(1, 3), (2, 2), (3, 2), (4, 2), (5, 1), (6, 1), (7, 1)])
def test_elastic_apply_matches_oracle(r, nx):
    m = M.regular_mesh(3, nx, r, 0.1, seed=6)
    o = O.mesh(3, nx, r, 0.1, 6)
    lam = 1.0 + (np.arange(m.E) % 3)
    mu = 2.0 + (np.arange(m.E) % 2)
    P = M.WaveProblem(m, 1.0, mu, lam=lam, ncomp=3)
    Po = O.elastic_problem(o, lam, mu, 1.0)
    u = np.random.default_rng(r).standard_normal(3 * m.N)
    assert CS.rel_l2(P.apply_kh(u), -O.apply(Po, u)) < 1e-12


@pytest.mark.parametrize("name", CS.golden_names())
def test_leapfrog_fp64_matches_montjoie(name):
    c = CS.load(name)
    if not len(c.rec):
        pytest.skip("no SismoPoint receivers (SismoCircle files: test_deck.py)")
    P, tr = CS.gpu_run(c, "fp64")
    ref = CS.reference_traces(c)
    err_ref = CS.rel_l2(tr[:, 0, :], ref)
    _, _, otr = CS.oracle_run(c)
    err_or = CS.rel_l2(tr[:, 0, :], otr)
    print(f"{name}: rel-L2 vs Montjoie {err_ref:.2e}, vs oracle {err_or:.2e}")
    assert err_ref < 1e-10
    assert err_or < 1e-12


def test_leapfrog_final_state_matches_oracle():
    c = CS.load("ref3d_q5_pert")
    P, _ = CS.gpu_run(c, "fp64")
    uN, uN1 = P.state()
    oN, oN1, _ = CS.oracle_run(c)
    assert CS.rel_l2(uN[0], oN) < 1e-12
    assert CS.rel_l2(uN1[0], oN1) < 1e-12


def test_graph_equals_eager_bitwise():
    c = CS.load("ref3d_q3_pert")
    _, a = CS.gpu_run(c, "fp64", graph=True)
    _, b = CS.gpu_run(c, "fp64", graph=False)
    assert np.array_equal(a, b)


def test_run_to_run_deterministic():
    c = CS.load("ref2d_q4_n8")
    _, a = CS.gpu_run(c, "fp32")
    _, b = CS.gpu_run(c, "fp32")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp32", "fp32_g64"])
def test_fp32_error_against_fp64(prec):
    c = CS.load("ref3d_q5_pert")
    _, ref = CS.gpu_run(c, "fp64")
    _, tr = CS.gpu_run(c, prec)
    err = CS.rel_l2(tr, ref)
    print(f"{prec}: traces rel-L2 vs FP64 {err:.2e}")
    assert err < 1e-5


def test_multishot_equals_single_shots():
    c = CS.load("ref3d_q3_pert")
    srcs = [c.src, (0.3, 0.6, 0.5), (0.55, 0.35, 0.62), (0.5, 0.5, 0.3)]
    _, multi = CS.gpu_run(c, "fp64", nshots=4, srcs=srcs, nsteps=60)
    for s, x in enumerate(srcs):
        _, one = CS.gpu_run(c, "fp64", nshots=1, srcs=[x], nsteps=60)
        assert np.array_equal(multi[:, s, :], one[:, 0, :])


def test_elastic_multishot_equals_single_shots():
    """Interleaved elastic shots (k_shots_el) against one-shot runs
    (k_elastic3d): bit for bit."""
    c = CS.load("ref3d_el_q3")
    srcs = [c.src, (0.3, 0.6, 0.5), (0.55, 0.35, 0.62)]
    _, multi = CS.gpu_run(c, "fp64", nshots=3, srcs=srcs, nsteps=50)
    for s, x in enumerate(srcs):
        _, one = CS.gpu_run(c, "fp64", nshots=1, srcs=[x], nsteps=50)
        assert np.array_equal(multi[:, s, :], one[:, 0, :])


def test_every_tile_gives_identical_results():
    c = CS.load("ref3d_q5_pert")
    P = CS.gpu_problem(c, "fp64")
    u = np.random.default_rng(1).standard_normal(P.ndof)
    outs = []
    for epb in (7, 14, 16, -1, -2, -4, 1007, 1014):  # column, thread per node, persistent
        P.set_tile(epb)
        outs.append(P.apply_kh(u))
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_every_shot_tile_gives_identical_results(prec):
    """Interleaved shots: every shot-lane count computes the same numbers."""
    from montjoie_amd import tuner
    m = M.regular_mesh(3, 3, 4, 0.1, seed=11)
    P = M.WaveProblem(m, 1.0, 1.0 + (np.arange(m.E) % 3), precision=prec, nshots=11)
    rng = np.random.default_rng(5)
    P.set_sources(0.25 + 0.5 * rng.random((11, 3)), shot=np.arange(11))
    P.set_receivers(0.2 + 0.6 * rng.random((5, 3)))
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, 1e-3, 40))
    outs = []
    for t in tuner.tile_choices(P):
        if t == MFMA_TILE:  # sums the divergence in another order: test_mfma_shot_tile_*
            continue
        P.set_tile(t)
        P.leapfrog_init(1e-3)
        outs.append(P.run(30))
    assert len(outs) >= 3
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


MFMA_TILE = 1000


@pytest.mark.parametrize("prec", ["fp32", "fp32_g64"])
@pytest.mark.parametrize("order", [2, 4, 5, 6])
def test_mfma_shot_tile_matches_vector_tiles(prec, order):
    """Interleaved acoustic shots with the contractions on the matrix cores
    (v_mfma_f32_4x4x1f32): 19 shots (a full and a partial tile of 16), every
    shot against the FP64 path (rel-L2 <= 1e-5, or no further than the vector
    kernel's own FP32 error where that is larger) and against the vector
    kernel; the interleaved operator application through the energy."""
    from montjoie_amd import tuner
    m = M.regular_mesh(3, 3, order, 0.1, seed=13)
    mu = 1.0 + (np.arange(m.E) % 3)
    rng = np.random.default_rng(8)
    src = 0.25 + 0.5 * rng.random((19, 3))
    rec = 0.2 + 0.6 * rng.random((6, 3))

    def make(p):
        Q = M.WaveProblem(m, 1.0, mu, precision=p, nshots=19)
        Q.set_sources(src, shot=np.arange(19))
        Q.set_receivers(rec)
        Q.set_pulse(M.ricker_pulse_table(8.0, 0.0, 1e-3, 80))
        return Q

    P = make(prec)
    tiles = tuner.tile_choices(P)
    assert MFMA_TILE in tiles
    vec = max(t for t in tiles if 0 < t <= 16)  # widest vector shot tile
    P.set_tile(vec)
    P.leapfrog_init(1e-3)
    ref = P.run(60)
    P.set_tile(MFMA_TILE)
    P.leapfrog_init(1e-3)
    got = P.run(60)
    P64 = make("fp64")
    P64.leapfrog_init(1e-3)
    r64 = P64.run(60)
    assert np.abs(ref).max() > 0
    for s in range(19):
        e_vec = CS.rel_l2(got[:, s], ref[:, s])
        e_64 = CS.rel_l2(got[:, s], r64[:, s])
        e_ref = CS.rel_l2(ref[:, s], r64[:, s])  # the vector kernel's own FP32 error
        tol = max(1e-5, 1.5 * e_ref)
        assert e_vec < 2 * tol and e_64 < tol, (s, e_vec, e_64, e_ref)
    # the interleaved operator application (energy of every shot)
    e_mfma = np.asarray(P.energy())
    P.set_tile(vec)
    e_vec = np.asarray(P.energy())
    assert np.abs(e_vec).max() > 0
    np.testing.assert_allclose(e_mfma, e_vec, rtol=1e-5, atol=1e-7 * np.abs(e_vec).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32_g64"])
def test_every_elastic_shot_tile_gives_identical_results(prec):
    """Interleaved elastic shots: the column mapping (tile ST) and the
    thread-per-node mapping (tile -ST) at every shot-lane count compute the
    same numbers."""
    from montjoie_amd import tuner
    m = M.regular_mesh(3, 3, 4, 0.1, seed=12)
    E = m.E
    P = M.WaveProblem(m, 1.0 + (np.arange(E) % 2), 1.0 + (np.arange(E) % 3),
                      lam=2.0 + (np.arange(E) % 4), ncomp=3, precision=prec, nshots=11)
    rng = np.random.default_rng(6)
    pol = rng.standard_normal((11, 3))
    P.set_sources(0.25 + 0.5 * rng.random((11, 3)), pol=pol, shot=np.arange(11))
    P.set_receivers(np.repeat(0.2 + 0.6 * rng.random((4, 3)), 3, axis=0),
                    np.tile(np.arange(3, dtype=np.int32), 4))
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, 1e-3, 40))
    tiles = tuner.tile_choices(P)
    assert any(t < 0 for t in tiles) and any(t > 0 for t in tiles)
    outs = []
    for t in tiles:
        print(f"elastic shot tile {t}", flush=True)
        P.set_tile(t)
        P.leapfrog_init(1e-3)
        outs.append(P.run(30))
    assert np.abs(outs[0]).max() > 0
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_energy_conserved_fp64():
    m = M.regular_mesh(3, 3, 4, 0.1, seed=9)
    P = M.WaveProblem(m, 1.0, 1.0 + np.arange(m.E) % 2)
    lam = P.lambda_max(100)
    dt = 0.5 * 2 / np.sqrt(lam)
    u0 = np.random.default_rng(4).standard_normal(m.N) * (1 - m.dirichlet)
    P.leapfrog_init(dt, u0=u0)
    E = []
    for _ in range(4):
        P.advance(50, traces=False)
        E.append(P.energy()[0])
    assert max(E) - min(E) < 1e-11 * max(E)


def test_lambda_max_matches_oracle():
    m = M.regular_mesh(2, 16, 4)
    P = M.WaveProblem(m, 1.0, 1.0)
    o = O.mesh(2, 16, 4)
    lo = O.lambda_max(O.acoustic_problem(o, 1.0, 1.0), 150, 3)
    lg = P.lambda_max(150)
    assert abs(lg - lo) < 1e-7 * lo


def test_xcd_block_order_is_result_neutral():
    """The XCD-aware block order changes which die runs an element, not the
    arithmetic: results are bitwise identical."""
    c = CS.load("ref3d_q5_pert")
    P = CS.gpu_problem(c, "fp64")
    u = np.random.default_rng(2).standard_normal(P.ndof)
    P.set_xcd_order(1)
    a = P.apply_kh(u)
    P.set_xcd_order(8)
    b = P.apply_kh(u)
    P.set_xcd_order(3)
    d = P.apply_kh(u)
    assert np.array_equal(a, b) and np.array_equal(a, d)


def test_cpp_adapter_runs_on_gpu(tmp_path):
    """The C++ adapter (HipWaveSystem + HipLeapFrogScheme) reproduces the
    oracle's traces (2-D Q4 8x8, 200 steps), with the second- and the
    fourth-order scheme."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "adapter_example"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "adapter_example.cpp"), "-o", str(exe),
                    "-L", os.path.join(root, "montjoie_amd"), "-lmjwave",
                    "-Wl,-rpath," + os.path.join(root, "montjoie_amd")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=100, check=True)
    got = np.array([[float(v) for v in line.split()] for line in out.stdout.strip().splitlines()])
    m = O.mesh(2, 8, 4)
    P = O.acoustic_problem(m, 1.0, 1.0)
    _, _, tr = O.leapfrog(P, O.dirac(m, (0.5, 0.5)), O.pulse_table(10.0, 0.0, 1e-3, 200), 1e-3,
                          200, [[0.8, 0.5], [0.3, 0.4]])
    np.testing.assert_allclose(got[:3], tr[100:103], rtol=1e-12, atol=1e-20)
    tabs = tuple(O.source_table("RICKER", 10.0, 0.0, 0.0, 1e-3, 200, k) for k in (1, 2, 3))
    _, _, tr4 = O.leapfrog4(P, O.dirac(m, (0.5, 0.5)), tabs, 1e-3, 200, [[0.8, 0.5], [0.3, 0.4]])
    np.testing.assert_allclose(got[3:], tr4[100:103], rtol=1e-12, atol=1e-20)
    assert not np.allclose(got[3:], got[:3], rtol=1e-9, atol=0)
