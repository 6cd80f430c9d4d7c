"""Deck driver (montjoie_amd/mjwave_run): Montjoie data files and Medit
meshes of the golden cases (tests/golden/decks.py writes them from the
parameters stored with each fixture).  CPU: the mesh the driver assembles
(--dump-mesh, no GPU call) is node for node the oracle's mesh.  GPU: the
seismogram file it writes matches the one Montjoie wrote for the same deck."""
import os
import subprocess
import sys

import numpy as np
import pytest

import cases as CS
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "montjoie_amd", "mjwave_run")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import decks  # noqa: E402


def _write(c: CS.Case, d):
    z = np.load(os.path.join(CS.GOLD, c.name + ".npz"))
    decks.write_deck(str(d), c.dim, c.nx, c.r, c.pert, c.seed, z["layers"], c.f0, c.dt,
                     c.nsteps, c.src, c.rec, elastic=c.elastic,
                     lam_layers=z["lam_layers"] if c.elastic else None,
                     rho_layers=z["rho_layers"] if c.elastic else None,
                     pol=c.pol, regular=bool(z["regular"]),
                     sinus_rho=z["sinus_rho"] if "sinus_rho" in z.files else None,
                     sinus_mu=z["sinus_mu"] if "sinus_mu" in z.files else None,
                     plane=z["plane"] if "plane" in z.files else None,
                     plane_times=z["plane_times"] if "plane_times" in z.files else None,
                     line=z["line"] if "line" in z.files else None,
                     line_times=z["line_times"] if "line_times" in z.files else None,
                     mesh_media=[(nm, int(z[f"meshmedia_{nm}_order"]),
                                  float(z[f"meshmedia_{nm}_offset"]),
                                  float(z[f"meshmedia_{nm}_amplitude"]), z[f"meshmedia_{nm}_nu"])
                                 for nm in ("mu", "rho") if f"meshmedia_{nm}_nu" in z.files],
                     save_reprise=int(z["save_reprise"]) if "save_reprise" in z.files else 0,
                     sigma_layers=z["sigma_layers"] if "sigma_layers" in z.files else None,
                     sinus_sigma=z["sinus_sigma"] if "sinus_sigma" in z.files else None,
                     output_format=str(z["output_format"]) if "output_format" in z.files
                     else "BINARY DOUBLE",
                     circle=z["circle"] if "circle" in z.files and len(z["circle"]) else None,
                     pml=CS.pml_deck(c) if c.pml is not None else None,
                     young=(z["young_E"], z["young_nu"])
                     if "young_E" in z.files and len(z["young_E"]) else None,
                     temporal=c.temporal, display_stress=c.display_stress,
                     time_order=c.time_order,
                     volume=(z["volume_refs"], z["volume_centre"], *z["volume_radius"])
                     if "volume_kind" in z.files and str(z["volume_kind"]) else None)


def _dump(path):
    raw = open(path, "rb").read()
    head = np.frombuffer(raw[:40], np.int64)
    dim, order, ncomp, N, E = (int(v) for v in head)
    nd = (order + 1) ** dim
    off = 40
    out = {"dim": dim, "order": order, "ncomp": ncomp, "N": N, "E": E}
    for name, dt, n in (("xyz", np.float64, N * dim), ("en", np.int32, E * nd),
                        ("dir", np.uint8, N), ("rho", np.float64, E), ("mu", np.float64, E),
                        ("lam", np.float64, E if ncomp > 1 else 0),
                        ("sigma", np.float64, E)):
        size = np.dtype(dt).itemsize * n
        out[name] = np.frombuffer(raw[off:off + size], dt)
        off += size
    # media varying inside the elements (PhysicalMedia): rho, mu per quadrature
    # point; then Montjoie's DOF number of every node (int64)
    nq = (len(raw) - off - 8 * N) // 16
    assert nq in (0, E * nd) and off + 16 * nq + 8 * N == len(raw)
    out["rho_q"] = np.frombuffer(raw[off:off + 8 * nq], np.float64).reshape(-1, nd)
    out["mu_q"] = np.frombuffer(raw[off + 8 * nq:off + 16 * nq], np.float64).reshape(-1, nd)
    out["mj_dof"] = np.frombuffer(raw[off + 16 * nq:], np.int64)
    out["xyz"] = out["xyz"].reshape(N, dim)
    out["en"] = out["en"].reshape(E, nd)
    return out


def _check_pml_mesh(c: CS.Case, g, pml_path):
    """AddPML deck: the driver's extruded mesh (its elements in its own order)
    against the oracle's extended structured mesh: the same elements (matched
    by centre), each holding the same GL nodes (matched by position), the same
    Dirichlet nodes, media of the element each layer was extruded from, the
    same PML elements and sigma at every quadrature point."""
    pm, rho, mu, tau = CS.pml_setup(c)
    o = pm.mesh
    nd = (c.r + 1) ** c.dim
    assert (g["N"], g["E"]) == (o.N, o.E)
    raw = open(pml_path, "rb").read()
    flag = np.frombuffer(raw[:o.E], np.uint8)
    gtau = np.frombuffer(raw[o.E:], np.float64).reshape(o.E, nd, c.dim)
    oen = o.elem_nodes.reshape(o.E, nd)
    oxyz = o.xyz.reshape(o.N, c.dim)
    key = lambda X: tuple(np.round(X.mean(axis=0), 9))  # noqa: E731
    oi = {key(oxyz[oen[e]]): e for e in range(o.E)}
    assert len(oi) == o.E
    for e in range(g["E"]):
        Xg = g["xyz"][g["en"][e]]
        eo = oi[key(Xg)]
        Xo = oxyz[oen[eo]]
        # local node q of the driver's element -> local node of the oracle's
        d2 = ((Xg[:, None, :] - Xo[None, :, :]) ** 2).sum(-1)
        lq = d2.argmin(axis=1)
        assert np.unique(lq).size == nd and d2[np.arange(nd), lq].max() < 1e-26, e
        assert g["rho"][e] == rho[eo] and g["mu"][e] == mu[eo], e
        assert flag[e] == pm.pml[eo], e
        np.testing.assert_allclose(gtau[e], tau[eo][lq], rtol=1e-13, atol=1e-13)
    # Dirichlet nodes: the outer boundary of the extended box
    gd = {tuple(np.round(x, 9)) for x in g["xyz"][g["dir"] == 1]}
    od = {tuple(np.round(x, 9)) for x in oxyz[o.dirichlet.astype(bool)]}
    assert gd == od


@pytest.mark.parametrize("name", CS.golden_names())
def test_driver_mesh_is_the_oracle_mesh(name, tmp_path):
    if not os.path.exists(RUNNER):
        pytest.fail("montjoie_amd/mjwave_run is not built (python -m montjoie_amd.build)")
    c = CS.load(name)
    _write(c, tmp_path)
    r = subprocess.run([RUNNER, "case.ini", "--dump-mesh", "mesh.bin"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    g = _dump(tmp_path / "mesh.bin")
    if c.pml is not None:
        _check_pml_mesh(c, g, tmp_path / "mesh.bin.pml")
        return
    o = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    assert (g["N"], g["E"]) == (o.N, o.E)
    oen = o.elem_nodes.reshape(o.E, -1)
    # same conforming numbering up to a permutation of the global nodes
    perm = np.full(o.N, -1)
    perm[oen.ravel()] = g["en"].ravel()
    assert np.array_equal(perm[oen], g["en"]) and np.unique(perm).size == o.N
    np.testing.assert_allclose(g["xyz"][perm], o.xyz.reshape(o.N, -1), rtol=0, atol=1e-14)
    assert np.array_equal(g["dir"][perm], o.dirichlet.astype(np.uint8))
    if c.mu_e.ndim == 2 or c.rho_e.ndim == 2:
        # PhysicalMedia SINUS evaluated by the driver at its quadrature points
        # against the oracle's field at the oracle's (same elements and
        # local order; positions equal to 1e-14); elastic decks keep mu per
        # element
        nd = g["rho_q"].shape[1]
        mu_q = c.mu_e if c.mu_e.ndim == 2 else np.repeat(c.mu_e[:, None], nd, axis=1)
        assert g["mu_q"].shape == mu_q.shape
        np.testing.assert_allclose(g["mu_q"], mu_q, rtol=1e-13, atol=0)
        np.testing.assert_allclose(g["rho_q"], c.rho_e, rtol=1e-13, atol=0)
    else:
        assert g["mu_q"].size == 0
        np.testing.assert_array_equal(g["mu"], c.mu_e)
        np.testing.assert_array_equal(g["rho"], c.rho_e)
    if c.elastic:
        np.testing.assert_array_equal(g["lam"], c.lam_e)
    # MateriauDielec's damping (a PhysicalMedia sigma field overrides it at
    # the quadrature points of its medium)
    if c.sigma_e is None:
        assert not g["sigma"].any()
    elif c.sigma_e.ndim == 1:
        np.testing.assert_array_equal(g["sigma"], c.sigma_e)
    # Montjoie's DOF numbering (SaveReprise files) equals the oracle's
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    corners, bnd = decks.medit_topology(c.dim, c.nx)
    if bool(z["regular"]):
        bnd = O.regular_boundary(c.dim, c.nx)
    od = O.montjoie_dofs(c.dim, c.r, corners, bnd, oen, o.N)
    assert np.array_equal(g["mj_dof"][perm], od)


def test_driver_rejects_unsupported_decks(tmp_path):
    c = CS.load("ref2d_q4_n8")
    _write(c, tmp_path)
    ini = (tmp_path / "case.ini").read_text()
    (tmp_path / "bad.ini").write_text(ini.replace("2 LEAP_FROG", "4 RUNGE_KUTTA"))
    r = subprocess.run([RUNNER, "bad.ini", "--dump-mesh", "x.bin"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode != 0 and "LEAP_FROG" in r.stderr


@pytest.mark.parametrize("line,word", [("LoadReprise = YES", "LoadReprise"),
                                       ("SaveReprise = 0", "SaveReprise")])
def test_driver_rejects_unsupported_outputs(tmp_path, line, word):
    """Restarts, raw-gradient elastic output and a non-positive checkpoint
    period stop the run with a message naming the keyword (before any GPU
    call: --dump-mesh runs on the CPU)."""
    c = CS.load("ref2d_q4_n8")
    _write(c, tmp_path)
    ini = (tmp_path / "case.ini").read_text()
    (tmp_path / "bad.ini").write_text(ini + line + "\n")
    r = subprocess.run([RUNNER, "bad.ini", "--dump-mesh", "x.bin"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode != 0 and word in r.stderr


def test_driver_rejects_media_on_another_mesh(tmp_path):
    c = CS.load("ref3d_q3_meshmedia")
    _write(c, tmp_path)
    ini = (tmp_path / "case.ini").read_text()
    (tmp_path / "bad.ini").write_text(ini.replace("SAME_MESH mesh.mesh", "MESH other.mesh"))
    r = subprocess.run([RUNNER, "bad.ini", "--dump-mesh", "x.bin"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode != 0 and "computational mesh" in r.stderr
    (tmp_path / "mu.don").write_bytes((tmp_path / "mu.don").read_bytes()[:-8])
    r = subprocess.run([RUNNER, "case.ini", "--dump-mesh", "x.bin"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode != 0 and "truncated" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", CS.golden_names())
def test_driver_seismograms_match_montjoie(name, tmp_path):
    """FP64 run of the deck: the written totalPoint.dat against Montjoie's,
    every column (elastic: displacement and stress; rel-L2 <= 1e-10, the
    requirement; times exact)."""
    c = CS.load(name)
    _write(c, tmp_path)
    print(f"deck {name}", flush=True)
    r = subprocess.run([RUNNER, "case.ini"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    if not len(c.rec):
        pytest.skip("no SismoPoint: the deck's SismoCircle files are checked by "
                    "test_driver_circle_matches_montjoie")
    out = np.loadtxt(tmp_path / "totalPoint.dat", ndmin=2)
    np.testing.assert_allclose(out[:, 0], c.times, rtol=1e-15, atol=1e-18)
    ref = c.traces
    assert out.shape == (ref.shape[0], ref.shape[1] + 1)
    if c.elastic:
        # displacement columns, then the stress tensor of each point
        nu = c.dim * len(c.rec)
        assert CS.rel_l2(out[:, 1:nu + 1], ref[:, :nu]) < 1e-10
        assert CS.rel_l2(out[:, nu + 1:], ref[:, nu:]) < 1e-10
    else:
        assert CS.rel_l2(out[:, 1:], ref) < 1e-10


def _dense_lambda_max(c: CS.Case) -> float:
    """Largest eigenvalue of D^{-1/2} K D^{-1/2} on the free nodes, dense."""
    m = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    P = O.acoustic_problem(m, c.mu_e, c.rho_e)
    free = np.flatnonzero(m.dirichlet == 0)
    K = np.zeros((free.size, free.size))
    for a, g in enumerate(free):
        e = np.zeros(m.N)
        e[g] = 1.0
        K[:, a] = O.apply(P, e)[free]
    s = 1.0 / np.sqrt(P.M[free])
    return float(np.linalg.eigvalsh(s[:, None] * K * s[None, :]).max())


@pytest.mark.gpu
def test_driver_time_step_auto(tmp_path):
    """TimeStep = AUTO: dt = 0.999 * 2 / sqrt(lambda_max(D^-1 K)) (Montjoie's
    rule; its ARPACK is not in the build, so lambda_max is checked against a
    dense eigensolver of the oracle's matrix), and the run at that dt against
    the oracle."""
    c = CS.load("ref2d_q4_n8")
    _write(c, tmp_path)
    ini = (tmp_path / "case.ini").read_text()
    sub = {"TimeStep": "TimeStep = AUTO",
           "ParametersOutputPoint": "ParametersOutputPoint = 0.0 1.0 0.0"}  # every step
    (tmp_path / "case.ini").write_text(
        "\n".join(sub.get(ln.split(" ")[0], ln) for ln in ini.splitlines()) + "\n")
    print("deck ref2d_q4_n8 TimeStep = AUTO", flush=True)
    r = subprocess.run([RUNNER, "case.ini"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    out = np.loadtxt(tmp_path / "totalPoint.dat", ndmin=2)
    dt = out[1, 0] - out[0, 0]
    want = 0.999 * 2.0 / np.sqrt(_dense_lambda_max(c))
    assert abs(dt - want) <= 1e-12 * want
    T = (c.nsteps - 0.5) * c.dt  # TimeInterval of the deck
    c.dt = float(out[1, 0])
    c.nsteps = int(np.ceil(T / c.dt - 1e-9))
    assert out.shape[0] == c.nsteps + 1
    _, _, tr = CS.oracle_run(c)
    assert CS.rel_l2(out[:, 1:], tr[:, :out.shape[1] - 1]) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ref2d_q4_plane", "ref3d_q3_plane", "ref3d_el_q3_plane",
                                  "ref2d_el_q4_plane", "ref3d_el_q3_grad", "ref2d_el_q4_grad"])
def test_driver_snapshots_match_montjoie(name, tmp_path):
    """SismoPlane and SismoLine: the driver's totalPlane%04d.dat and
    totalLine%04d.dat files (elastic: one per displacement and stress
    component, totalPlane%04d_U<m>.dat / _dU<k>.dat) against Montjoie's, the
    same names, byte for byte in the header, values to rel-L2 1e-10 with the
    same zeros (points outside the mesh)."""
    c = CS.load(name)
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    _write(c, tmp_path)
    print(f"deck {name} SismoPlane + SismoLine", flush=True)
    r = subprocess.run([RUNNER, "case.ini"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    for kind, prefix in (("plane", "totalPlane"), ("line", "totalLine")):
        files = sorted(f.name for f in tmp_path.iterdir() if f.name.startswith(prefix))
        ref = z[kind + "_files"]
        assert len(files) == ref.shape[0]
        g = z[kind]
        if not len(g):  # the deck has no grid of this kind
            continue
        npts = int(g[-1] * g[-2]) if kind == "plane" else int(g[-1])
        names = (list(z[kind + "_names"]) if kind + "_names" in z.files
                 else [f"{prefix}{k:04d}.dat" for k in range(len(files))])
        for k, f in enumerate(files):
            assert f == names[k]
            raw = (tmp_path / f).read_bytes()
            assert len(raw) == ref.shape[1]
            h, v = CS.plane_file(raw, npts)
            hr, vr = CS.plane_file(ref[k], npts)
            assert h == hr
            if vr.any():
                # zeros agree up to boundary points of rounding-size value
                tiny = 1e-12 * np.abs(vr).max()
                assert np.all(np.abs(vr[v == 0]) <= tiny) and np.all(np.abs(v[vr == 0]) <= tiny)
                assert CS.rel_l2(v, vr) < 1e-10
            else:
                assert not v.any()


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in CS.golden_names() if n.endswith(
    ("_reprise", "_reprise_regular"))])
def test_driver_checkpoints_match_montjoie(name, tmp_path):
    """SaveReprise: the driver's IterationAcousSave[New]0N1.dat equal
    Montjoie's byte for byte, its SolAcous<b>_P0N1.dat carry the same size and
    the same iterates in Montjoie's DOF order, to rel-L2 1e-10."""
    c = CS.load(name)
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    _write(c, tmp_path)
    print(f"deck {name} SaveReprise = {int(z['save_reprise'])}", flush=True)
    r = subprocess.run([RUNNER, "case.ini"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    for f in ("IterationAcousSaveNew0N1", "IterationAcousSave0N1"):
        assert (tmp_path / (f + ".dat")).read_bytes() == z["reprise_" + f].tobytes()
    for b in (0, 1):
        raw = (tmp_path / f"SolAcous{b}_P0N1.dat").read_bytes()
        ref = z[f"reprise_SolAcous{b}_P0N1"].tobytes()
        assert len(raw) == len(ref) and raw[:4] == ref[:4]
        v, vr = np.frombuffer(raw[4:], np.float64), np.frombuffer(ref[4:], np.float64)
        assert CS.rel_l2(v, vr) < 1e-10, CS.rel_l2(v, vr)


@pytest.mark.gpu
def test_driver_float_output_matches_montjoie(tmp_path):
    """OutputFormat = BINARY FLOAT: snapshot files with data type 0 and float
    grid parameters and values, and seismograms with 8 significant digits,
    as Montjoie writes them.  Headers and names exact; values to float
    rounding (1e-6 relative to the largest value)."""
    c = CS.load("fmt3d_q3_plane_float")
    z = np.load(os.path.join(CS.GOLD, "fmt3d_q3_plane_float.npz"))
    _write(c, tmp_path)
    print("deck fmt3d_q3_plane_float", flush=True)
    r = subprocess.run([RUNNER, "case.ini"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    row = (tmp_path / "totalPoint.dat").read_text().splitlines()[5].split("\t")
    assert all(len(v.split("e")[0].lstrip("-")) == 9 for v in row)  # d.ddddddd
    out = np.loadtxt(tmp_path / "totalPoint.dat", ndmin=2)
    assert out.shape == (c.traces.shape[0], c.traces.shape[1] + 1)
    assert np.abs(out[:, 1:] - c.traces).max() <= 1e-6 * np.abs(c.traces).max()
    for kind, prefix in (("plane", "totalPlane"), ("line", "totalLine")):
        files = sorted(f.name for f in tmp_path.iterdir() if f.name.startswith(prefix))
        ref = z[kind + "_files"]
        assert files == list(z[kind + "_names"])
        g = z[kind]
        npts = int(g[-1] * g[-2]) if kind == "plane" else int(g[-1])
        for k, f in enumerate(files):
            raw = (tmp_path / f).read_bytes()
            assert len(raw) == ref.shape[1]
            n = len(raw) - 4 * npts
            assert np.frombuffer(raw[8:12], np.int32)[0] == 0
            assert raw[:n] == bytes(ref[k][:n])
            v = np.frombuffer(raw[n:], np.float32)
            vr = np.frombuffer(bytes(ref[k][n:]), np.float32)
            assert np.abs(v - vr).max() <= 1e-6 * max(np.abs(vr).max(), 1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["fmt2d_q4_plane_ascii", "fmt3d_q3_plane_ascii_float"])
def test_driver_ascii_output_matches_montjoie(name, tmp_path):
    """OutputFormat = ASCII DOUBLE / FLOAT: text snapshot files, one field per
    line, header lines (grid count, dimension, data type, grid type, grid
    parameters at 15 or 7 significant digits) identical to Montjoie's, values
    to 1e-10 (DOUBLE) or float rounding (FLOAT) with the same zeros."""
    c = CS.load(name)
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    dbl = str(z["output_format"]).endswith("DOUBLE")
    _write(c, tmp_path)
    print(f"deck {name}", flush=True)
    r = subprocess.run([RUNNER, "case.ini"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    for kind, prefix in (("plane", "totalPlane"), ("line", "totalLine")):
        files = sorted(f.name for f in tmp_path.iterdir() if f.name.startswith(prefix))
        assert files == list(z[kind + "_names"])
        g = z[kind]
        npts = int(g[-1] * g[-2]) if kind == "plane" else int(g[-1])
        for k, f in enumerate(files):
            mine = (tmp_path / f).read_text().split("\n")
            ref = bytes(z[kind + "_files"][k][:z[kind + "_lens"][k]]).decode().split("\n")
            assert len(mine) == len(ref)
            nh = len(ref) - 1 - npts
            assert mine[:nh] == ref[:nh]
            v = np.array(mine[nh:-1], float)
            vr = np.array(ref[nh:-1], float)
            if not vr.any():
                assert not v.any()
                continue
            tiny = 1e-12 * np.abs(vr).max()
            assert np.all(np.abs(vr[v == 0]) <= tiny) and np.all(np.abs(v[vr == 0]) <= tiny)
            if dbl:
                assert CS.rel_l2(v, vr) < 1e-10
            else:
                assert np.abs(v - vr).max() <= 1e-6 * np.abs(vr).max()


CIRCLES = [n for n in CS.golden_names() if "circle" in n]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CIRCLES)
def test_driver_circle_matches_montjoie(name, tmp_path):
    """SismoCircle = x0 y0 r N with FileOutputCircle: the driver's
    totalCircle%04d.dat files against Montjoie's (one per output step, grid
    type 3), header byte for byte, values to rel-L2 1e-10, zeros at the points
    outside the mesh."""
    c = CS.load(name)
    z = np.load(os.path.join(CS.GOLD, name + ".npz"))
    _write(c, tmp_path)
    print(f"deck {name} SismoCircle", flush=True)
    r = subprocess.run([RUNNER, "case.ini"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    files = sorted(f.name for f in tmp_path.iterdir() if f.name.startswith("totalCircle"))
    ref = z["circle_files"]
    names = list(z["circle_names"])
    assert len(files) == len(names) and all(a == b for a, b in zip(files, names)), \
        (files[:3], names[:3])
    npts = int(z["circle"][3])
    got, want = [], []
    for k, f in enumerate(files):
        h, v = CS.plane_file((tmp_path / f).read_bytes(), npts)
        hr, vr = CS.plane_file(ref[k], npts)
        assert h == hr
        got.append(v)
        want.append(vr)
    got, want = np.stack(got), np.stack(want)
    # the same zeros (points outside the mesh; steps before the wave arrives,
    # where either side may hold values of underflow size)
    tiny = 1e-12 * np.abs(want).max()
    assert np.all(np.abs(want[got == 0]) <= tiny) and np.all(np.abs(got[want == 0]) <= tiny)
    err = CS.rel_l2(got, want)
    print(f"{name}: {len(files)} circle files, rel-L2 {err:.2e}")
    assert err < 1e-10


def _pml_dump(tmp_path, ini_text, name="p.ini"):
    (tmp_path / name).write_text(ini_text)
    r = subprocess.run([RUNNER, name, "--dump-mesh", name + ".bin"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    g = _dump(tmp_path / (name + ".bin"))
    raw = open(tmp_path / (name + ".bin.pml"), "rb").read()
    flag = np.frombuffer(raw[:g["E"]], np.uint8)
    return g, flag, np.frombuffer(raw[g["E"]:], np.float64)


def test_driver_pml_auto_layers_and_one_sided_axes(tmp_path):
    """AddPML = ... AUTO: ceil(thickness / mean boundary edge length) layers
    (0.25 over edges of 1/8: 2, the fixture's count, so the same mesh and
    sigma); X+ Y-: layers on those two sides only, the corner element once,
    sigma zero along the axes an element does not extend beyond."""
    c = CS.load("ref2d_q4_pml")
    _write(c, tmp_path)
    ini = (tmp_path / "case.ini").read_text()
    assert "AddPML = YES XY 0.25 2" in ini
    g2, f2, t2 = _pml_dump(tmp_path, ini, "two.ini")
    ga, fa, ta = _pml_dump(tmp_path, ini.replace("AddPML = YES XY 0.25 2",
                                                 "AddPML = YES XY 0.25 AUTO"), "auto.ini")
    assert (ga["N"], ga["E"]) == (g2["N"], g2["E"])
    assert np.array_equal(ga["xyz"], g2["xyz"]) and np.array_equal(ga["en"], g2["en"])
    assert np.array_equal(fa, f2) and np.array_equal(ta, t2)
    g1, f1, t1 = _pml_dump(tmp_path, ini.replace("AddPML = YES XY 0.25 2",
                                                 "AddPML = YES X+Y- 0.25 2"), "one.ini")
    nx, L, nd = c.nx, 2, (c.r + 1) ** 2
    assert g1["E"] == (nx + L) ** 2 and int(f1.sum()) == (nx + L) ** 2 - nx ** 2
    X = g1["xyz"][g1["en"]]                      # (E, nd, 2)
    tau = t1.reshape(g1["E"], nd, 2)
    cen = X.mean(axis=1)
    assert cen[:, 0].min() > 0 and cen[:, 1].max() < 1      # no layer at x < 0, y > 1
    in_x = cen[:, 0] > 1
    in_y = cen[:, 1] < 0
    assert np.array_equal(f1.astype(bool), in_x | in_y)
    assert np.all(tau[~in_x, :, 0] == 0) and np.all(tau[~in_y, :, 1] == 0)
    assert np.all(tau[in_x, :, 0] >= 0) and tau[in_x, :, 0].max() > 0
    # sigma grows away from the box: parabola dsig (x - 1)^2 in the x+ layers
    x = X[in_x, :, 0].ravel()
    s = tau[in_x, :, 0].ravel()
    k = np.argmax(x)
    np.testing.assert_allclose(s / s[k], (x - 1) ** 2 / (x[k] - 1) ** 2, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("edit,word", [
    (lambda s: s.replace("2 LEAP_FROG_PML", "2 LEAP_FROG"), "LEAP_FROG_PML"),
    (lambda s: s.replace("AddPML = YES XY 0.25 2", "AddPML = YES R 0.25 2"), "radial"),
    (lambda s: s.replace("AddPML = YES XY 0.25 2", "AddPML = 1 2 0.25 2"), "AddPML"),
    (lambda s: s + "SaveReprise = 10\n", "SaveReprise"),
    (lambda s: s.replace("TimeStep = ", "TimeStep = AUTO #"), "AUTO"),
])
def test_driver_rejects_unsupported_pml_decks(tmp_path, edit, word):
    """The second-order PML formulation, radial and reference-based layers,
    checkpoints and TimeStep = AUTO with the first-order scheme stop the run
    with a message naming what is missing (before any GPU call)."""
    c = CS.load("ref2d_q4_pml")
    _write(c, tmp_path)
    ini = (tmp_path / "case.ini").read_text()
    (tmp_path / "bad.ini").write_text(edit(ini))
    r = subprocess.run([RUNNER, "bad.ini", "--dump-mesh", "x.bin"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode != 0 and word in r.stderr, r.stderr
