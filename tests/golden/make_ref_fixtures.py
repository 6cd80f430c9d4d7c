"""Generate golden receiver traces by running Montjoie's own explicit solvers.

Run in the build container only (it needs /root/reference to build the
solvers with ``make -C oracle ref``):

    python tests/golden/make_ref_fixtures.py [case ...]

For each case this writes a Montjoie data file (.ini) and, for perturbed 3-D
meshes, a Medit .mesh file under oracle/_ref/run/<case>/, runs
oracle/_ref/mjref_acous.x or mjref_elastic.x on it, parses the SismoPoint
output (ASCII, 15 significant digits, Output/GridInterpolation.cxx:3150-3152)
and stores inputs + traces in tests/golden/<case>.npz.  The meshes, media,
sources, receivers and time steps are those of the synthetic problems used by
the tests; the vertices come from oracle.mesh (the same generator as the
product's mjw_mesh_regular, checked bit-for-bit by the tests).
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
RUN = os.path.join(REF, "run")
GOLD = os.path.dirname(os.path.abspath(__file__))


def _f(x: float) -> str:
    return repr(float(x))


def write_medit(path, m: O.Mesh, elem_ref, bnd_ref=1):
    """Medit mesh of the structured (perturbed) mesh: vertices, hexahedra
    (quadrilaterals in 2-D) with their media reference, boundary faces."""
    dim, nx = m.dim, m.nx
    nv1 = nx + 1
    V = m.vert.reshape(-1, dim)
    lines = ["MeshVersionFormatted 1", "Dimension", str(dim), "Vertices", str(len(V))]
    for v in V:
        lines.append(" ".join(_f(c) for c in v) + " 0")
    vid = lambda a, b, c=0: 1 + a + nv1 * (b + nv1 * c)  # noqa: E731
    if dim == 3:
        lines += ["Hexahedra", str(nx ** 3)]
        e = 0
        for c in range(nx):
            for b in range(nx):
                for a in range(nx):
                    cs = [vid(a, b, c), vid(a + 1, b, c), vid(a + 1, b + 1, c), vid(a, b + 1, c),
                          vid(a, b, c + 1), vid(a + 1, b, c + 1), vid(a + 1, b + 1, c + 1),
                          vid(a, b + 1, c + 1)]
                    lines.append(" ".join(map(str, cs)) + f" {int(elem_ref[e])}")
                    e += 1
        faces = []
        for i in range(nx):
            for j in range(nx):
                faces.append([vid(i, j, 0), vid(i, j + 1, 0), vid(i + 1, j + 1, 0), vid(i + 1, j, 0)])
                faces.append([vid(i, j, nx), vid(i + 1, j, nx), vid(i + 1, j + 1, nx), vid(i, j + 1, nx)])
                faces.append([vid(i, 0, j), vid(i + 1, 0, j), vid(i + 1, 0, j + 1), vid(i, 0, j + 1)])
                faces.append([vid(i, nx, j), vid(i, nx, j + 1), vid(i + 1, nx, j + 1), vid(i + 1, nx, j)])
                faces.append([vid(0, i, j), vid(0, i, j + 1), vid(0, i + 1, j + 1), vid(0, i + 1, j)])
                faces.append([vid(nx, i, j), vid(nx, i + 1, j), vid(nx, i + 1, j + 1), vid(nx, i, j + 1)])
        lines += ["Quadrilaterals", str(len(faces))]
        lines += [" ".join(map(str, f)) + f" {bnd_ref}" for f in faces]
    else:
        lines += ["Quadrilaterals", str(nx * nx)]
        e = 0
        for b in range(nx):
            for a in range(nx):
                cs = [vid(a, b), vid(a + 1, b), vid(a + 1, b + 1), vid(a, b + 1)]
                lines.append(" ".join(map(str, cs)) + f" {int(elem_ref[e])}")
                e += 1
        edges = []
        for i in range(nx):
            edges += [[vid(i, 0), vid(i + 1, 0)], [vid(nx, i), vid(nx, i + 1)],
                      [vid(i + 1, nx), vid(i, nx)], [vid(0, i + 1), vid(0, i)]]
        lines += ["Edges", str(len(edges))]
        lines += [" ".join(map(str, f)) + f" {bnd_ref}" for f in edges]
    lines.append("End")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def run_case(name, dim, nx, r, pert, seed, layers, f0, dt, nsteps, src, rec, elastic=False,
             lam_layers=None, rho_layers=None, pol=None, regular=False):
    d = os.path.join(RUN, name)
    os.makedirs(d, exist_ok=True)
    m = O.mesh(dim, nx, r, pert, seed)
    nl = len(layers)
    e = np.arange(m.E)
    ez = e // (nx ** (dim - 1))
    lay = (ez * nl) // nx
    ini = []
    if dim == 2:
        ini.append("TypeElement = QUADRANGLE_LOBATTO")
    else:
        ini.append("TypeElement = HEXAHEDRON_LOBATTO")
    ini.append("TypeEquation = " + ("TIME_ELASTIC" if elastic else "ACOUSTIC"))
    ini.append("MeshPath = ./")
    ini.append(f"Frequency = {_f(f0)} 0.0")
    if regular:
        ini.append(f"FileMesh = REGULAR {nx + 1} 0.0 1.0 0.0 1.0" + (" 0.0 1.0" if dim == 3 else "")
                   + " 1" + " 1" * (2 * dim))
    else:
        write_medit(os.path.join(d, "mesh.mesh"), m, lay + 1, bnd_ref=nl + 1)
        ini.append("FileMesh = mesh.mesh")
    if dim == 3:
        ini.append("TypeMesh = HEXA")
    ini.append(f"OrderDiscretization = {r}")
    bref = 1 if regular else nl + 1
    ini.append(f"ConditionReference = {bref} DIRICHLET")
    for k in range(nl):
        if elastic:
            ini.append(f"MateriauDielec = {k + 1} ISOTROPE {_f(rho_layers[k])} "
                       f"{_f(lam_layers[k])} {_f(layers[k])}")
        else:
            ini.append(f"MateriauDielec = {k + 1} ISOTROPE 1.0 {_f(layers[k])} 0.0")
    ini.append("TypeSource = SRC_DIRAC")
    ini.append("OriginePhase = " + " ".join(_f(c) for c in src))
    if elastic:
        ini.append("Polarization = " + " ".join(_f(c) for c in pol))
    T = (nsteps - 0.5) * dt
    ini.append(f"ParametersOutputPoint = 0.0 {_f(nsteps * dt + 0.5 * dt)} {_f(dt)}")
    ini.append("FileOutputPoint = diffracPoint.dat totalPoint.dat")
    for p in rec:
        ini.append("SismoPoint = " + " ".join(_f(c) for c in p))
    ini.append("OutputFormat = BINARY DOUBLE")
    ini.append("ElectricOrMagnetic = " + ("-1" if elastic else "0"))
    ini.append(f"TimeStep = {_f(dt)}")
    ini.append(f"TimeInterval = 0.0 {_f(T)}")
    ini.append("OrderTimeScheme = 2 LEAP_FROG")
    ini.append("TemporalSource = RICKER")
    ini.append("PrintLevel = 0")
    with open(os.path.join(d, "case.ini"), "w") as fh:
        fh.write("\n".join(ini) + "\n")
    exe = os.path.join(REF, "mjref_elastic.x" if elastic else "mjref_acous.x")
    args = [exe, "case.ini"] + ([] if elastic else [str(dim)])
    for f in os.listdir(d):
        if f.startswith("total") or f.startswith("diffrac"):
            os.remove(os.path.join(d, f))
    r_ = subprocess.run(args, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r_.returncode != 0:
        raise RuntimeError(f"{name}: reference run failed\n{r_.stdout[-3000:]}")
    out = np.loadtxt(os.path.join(d, "totalPoint.dat"), ndmin=2)
    times, traces = out[:, 0], out[:, 1:]
    np.savez_compressed(
        os.path.join(GOLD, name + ".npz"), dim=dim, nx=nx, r=r, pert=pert, seed=seed,
        layers=np.asarray(layers, float), f0=f0, dt=dt, nsteps=nsteps, src=np.asarray(src, float),
        rec=np.asarray(rec, float), times=times, traces=traces, elastic=elastic,
        lam_layers=np.asarray(lam_layers if lam_layers is not None else [], float),
        rho_layers=np.asarray(rho_layers if rho_layers is not None else [], float),
        pol=np.asarray(pol if pol is not None else [], float), regular=regular)
    print(f"{name}: {traces.shape} rows, times {times[0]}..{times[-1]}")


def seeded_points(dim, n, seed, lo=0.1, hi=0.9):
    return np.array([[lo + (hi - lo) * O.uniform(seed, dim * i + a) for a in range(dim)]
                     for i in range(n)])


CASES = {}


def case(f):
    CASES[f.__name__] = f
    return f


@case
def ref2d_q4_n8():
    circ = O.circle_points(0.5, 0.5, 0.3, 16)
    run_case("ref2d_q4_n8", 2, 8, 4, 0.0, 0, [1.0], 10.0, 1e-3, 200, (0.5, 0.5),
             np.vstack([circ, [[0.3, 0.4], [0.71, 0.52]]]), regular=True)


@case
def ref2d_c1():
    """64x64 Q4 unit square, c = 1, dt = 0.5 * CFL limit, 1000 steps, 16
    receivers on the circle of radius 0.3 around the source."""
    m = O.mesh(2, 64, 4)
    P = O.acoustic_problem(m, 1.0, 1.0)
    lam = O.lambda_max(P, 200, 11)
    dt = 0.5 * 2.0 / np.sqrt(lam)
    circ = O.circle_points(0.5, 0.5, 0.3, 16)
    run_case("ref2d_c1", 2, 64, 4, 0.0, 0, [1.0], 10.0, dt, 1000, (0.5, 0.5), circ, regular=True)


@case
def ref3d_q5_pert():
    """4^3 Q5 perturbed hexahedra, three layers, Dirac at the centre."""
    c = [1.0 + 2.0 * O.uniform(2025, k) for k in range(3)]
    mu = [v * v for v in c]
    rec = seeded_points(3, 12, 77)
    run_case("ref3d_q5_pert", 3, 4, 5, 0.1, 2024, mu, 8.0, 4e-3, 150, (0.5, 0.5, 0.5), rec)


@case
def ref3d_q3_pert():
    c = [1.0 + 2.0 * O.uniform(2026, k) for k in range(3)]
    mu = [v * v for v in c]
    rec = seeded_points(3, 12, 78)
    run_case("ref3d_q3_pert", 3, 5, 3, 0.1, 99, mu, 8.0, 5e-3, 150, (0.43, 0.52, 0.47), rec)


@case
def ref3d_el_q3():
    """isotropic elastic, 4^3 Q3 perturbed, layered lambda, mu, rho."""
    lam = [1.0 + 3.0 * O.uniform(2030, k) for k in range(3)]
    mu = [1.0 + 3.0 * O.uniform(2031, k) for k in range(3)]
    rho = [1.0 + O.uniform(2032, k) for k in range(3)]
    rec = seeded_points(3, 8, 79)
    run_case("ref3d_el_q3", 3, 4, 3, 0.1, 7, mu, 8.0, 3e-3, 150, (0.46, 0.55, 0.51), rec,
             elastic=True, lam_layers=lam, rho_layers=rho, pol=(0.3, -0.5, 0.8))


if __name__ == "__main__":
    if not os.path.exists(os.path.join(REF, "mjref_acous.x")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    names = sys.argv[1:] or list(CASES)
    for n in names:
        CASES[n]()
