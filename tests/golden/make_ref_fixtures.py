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
from tests.golden.decks import pml_spec, write_deck  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
RUN = os.path.join(REF, "run")
GOLD = os.path.dirname(os.path.abspath(__file__))


def _pml_fields(pml, dim, nx):
    """npz fields of a PML case: pml = [thickness, layers, vsigma, profile,
    offset, transverse] (layers = the count used: AUTO decks get Montjoie's
    ceil(thickness / h), h = 1/nx the boundary edge length of the structured
    mesh, thicknesses chosen away from multiples of h), pml_axes, pml_auto."""
    if pml is None:
        return dict(pml=np.zeros(0))
    sp = pml_spec(pml, dim)
    auto = sp["layers"] == "AUTO"
    lay = max(1, int(np.ceil(sp["thickness"] * nx))) if auto else int(sp["layers"])
    return dict(pml=np.array([sp["thickness"], lay, sp["vsigma"], sp.get("profile", 0),
                              sp.get("offset", 0.0), sp.get("transverse", 0.0)], float),
                pml_axes=np.array(sp["axes"]), pml_auto=np.array(int(auto)))


def run_case(name, dim, nx, r, pert, seed, layers, f0, dt, nsteps, src, rec, elastic=False,
             lam_layers=None, rho_layers=None, pol=None, regular=False, sinus_rho=None,
             sinus_mu=None, plane=None, plane_times=None, line=None, line_times=None,
             mesh_media=None, save_reprise=0, sigma_layers=None, sinus_sigma=None,
             output_format="BINARY DOUBLE", circle=None, pml=None, young=None,
             temporal="RICKER", display_stress=True, time_order=2, volume=None):
    d = os.path.join(RUN, name)
    write_deck(d, dim, nx, r, pert, seed, layers, f0, dt, nsteps, src, rec, elastic=elastic,
               lam_layers=lam_layers, rho_layers=rho_layers, pol=pol, regular=regular,
               sinus_rho=sinus_rho, sinus_mu=sinus_mu, plane=plane, plane_times=plane_times,
               line=line, line_times=line_times, mesh_media=mesh_media,
               save_reprise=save_reprise, sigma_layers=sigma_layers, sinus_sigma=sinus_sigma,
               output_format=output_format, circle=circle, pml=pml, young=young,
               temporal=temporal, display_stress=display_stress, time_order=time_order,
               volume=volume)
    exe = os.path.join(REF, "mjref_elastic.x" if elastic else "mjref_acous.x")
    args = [exe, "case.ini", str(dim)]
    for f in os.listdir(d):
        if f.startswith(("total", "diffrac", "SolAcous", "IterationAcous")):
            os.remove(os.path.join(d, f))
    planes = np.zeros((0, 0), np.uint8)
    lines = np.zeros((0, 0), np.uint8)
    r_ = subprocess.run(args, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r_.returncode != 0:
        raise RuntimeError(f"{name}: reference run failed\n{r_.stdout[-3000:]}")
    out = np.loadtxt(os.path.join(d, "totalPoint.dat"), ndmin=2)
    times, traces = out[:, 0], out[:, 1:]

    names = {}

    def snap_files(prefix):
        # snapshot files <prefix>0000.dat, ... (elastic: <prefix>0000_U0.dat,
        # ..., <prefix>0000_dU8.dat), sorted by name, raw bytes, one row each
        files = sorted(f for f in os.listdir(d) if f.startswith(prefix))
        names[prefix] = np.array(files)
        raw = [np.frombuffer(open(os.path.join(d, f), "rb").read(), np.uint8) for f in files]
        # text files (ASCII formats) differ in length: zero-padded rows + lengths
        names[prefix + "_len"] = np.array([len(r) for r in raw])
        width = max(len(r) for r in raw)
        return np.stack([np.pad(r, (0, width - len(r))) for r in raw])
    if plane is not None:
        planes = snap_files("totalPlane")
    if line is not None:
        lines = snap_files("totalLine")
    circles = np.zeros((0, 0), np.uint8)
    if circle is not None:
        # SismoCircle output (FileOutputCircle): one snapshot file per output
        # time in the grid layout of WriteMatlab, 0 at points outside the mesh
        circles = snap_files("totalCircle")
    reprise = {}
    if save_reprise:
        # checkpoint files, raw bytes (SolAcous<b>_P0N1.dat, IterationAcousSave[New]0N1.dat)
        for f in ("SolAcous0_P0N1.dat", "SolAcous1_P0N1.dat", "IterationAcousSaveNew0N1.dat",
                  "IterationAcousSave0N1.dat"):
            if os.path.exists(os.path.join(d, f)):
                reprise["reprise_" + f.split(".")[0]] = np.frombuffer(
                    open(os.path.join(d, f), "rb").read(), np.uint8)
    np.savez_compressed(
        os.path.join(GOLD, name + ".npz"), dim=dim, nx=nx, r=r, pert=pert, seed=seed,
        layers=np.asarray(layers, float), f0=f0, dt=dt, nsteps=nsteps, src=np.asarray(src, float),
        rec=np.asarray(rec, float), times=times, traces=traces, elastic=elastic,
        lam_layers=np.asarray(lam_layers if lam_layers is not None else [], float),
        rho_layers=np.asarray(rho_layers if rho_layers is not None else [], float),
        pol=np.asarray(pol if pol is not None else [], float), regular=regular,
        sinus_rho=np.asarray(sinus_rho if sinus_rho is not None else [], float),
        sinus_mu=np.asarray(sinus_mu if sinus_mu is not None else [], float),
        plane=np.asarray(plane if plane is not None else [], float),
        plane_times=np.asarray(plane_times if plane_times is not None else [], float),
        plane_files=planes,
        line=np.asarray(line if line is not None else [], float),
        line_times=np.asarray(line_times if line_times is not None else [], float),
        line_files=lines,
        **{f"meshmedia_{nm}_{k}": v for nm, rm, off, amp, nu in (mesh_media or [])
           for k, v in (("order", rm), ("offset", off), ("amplitude", amp), ("nu", nu))},
        save_reprise=save_reprise, **reprise,
        sigma_layers=np.asarray(sigma_layers if sigma_layers is not None else [], float),
        sinus_sigma=np.asarray(sinus_sigma if sinus_sigma is not None else [], float),
        output_format=output_format,
        circle=np.asarray(circle if circle is not None else [], float),
        young_E=np.asarray(young[0] if young is not None else [], float),
        young_nu=np.asarray(young[1] if young is not None else [], float),
        temporal=temporal, display_stress=display_stress, time_order=time_order,
        volume_refs=np.asarray(volume[0] if volume else [], int),
        volume_centre=np.asarray(volume[1] if volume else [], float),
        volume_radius=np.asarray(volume[2:] if volume else [], float),
        volume_kind="" if not volume else ("GAUSSIAN" if len(volume[1]) else "UNIFORM"),
        **_pml_fields(pml, dim, nx),
        circle_files=circles,
        circle_names=names.get("totalCircle", np.zeros(0, "U1")),
        plane_names=names.get("totalPlane", np.zeros(0, "U1")),
        line_names=names.get("totalLine", np.zeros(0, "U1")),
        plane_lens=names.get("totalPlane_len", np.zeros(0, int)),
        line_lens=names.get("totalLine_len", np.zeros(0, int)))
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
def ref3d_el_q4():
    """isotropic elastic at the order of the 32^3 16-shot workload: 3^3 Q4
    perturbed hexahedra, layered lambda, mu, rho drawn like the workload's
    (lambda, mu in [1, 4], rho in [1, 2])."""
    lam = [1.0 + 3.0 * O.uniform(2033, k) for k in range(3)]
    mu = [1.0 + 3.0 * O.uniform(2034, k) for k in range(3)]
    rho = [1.0 + O.uniform(2035, k) for k in range(3)]
    rec = seeded_points(3, 6, 92)
    run_case("ref3d_el_q4", 3, 3, 4, 0.1, 17, mu, 8.0, 3e-3, 150, (0.47, 0.54, 0.49), rec,
             elastic=True, lam_layers=lam, rho_layers=rho, pol=(-0.4, 0.7, 0.5))


@case
def ref2d_el_q4():
    """isotropic elastic on quadrilaterals: 8x8 Q4 perturbed unit square,
    three layers of lambda, mu in [1, 4] and rho in [1, 2]."""
    lam = [1.0 + 3.0 * O.uniform(2036, k) for k in range(3)]
    mu = [1.0 + 3.0 * O.uniform(2037, k) for k in range(3)]
    rho = [1.0 + O.uniform(2038, k) for k in range(3)]
    rec = seeded_points(2, 8, 93)
    run_case("ref2d_el_q4", 2, 8, 4, 0.12, 21, mu, 10.0, 1e-3, 300, (0.46, 0.53), rec,
             elastic=True, lam_layers=lam, rho_layers=rho, pol=(0.6, -0.8))


@case
def ref2d_el_q7():
    """isotropic elastic on quadrilaterals at the highest order: 4x4 Q7
    perturbed, two layers."""
    lam = [1.0 + 3.0 * O.uniform(2039, k) for k in range(2)]
    mu = [1.0 + 3.0 * O.uniform(2040, k) for k in range(2)]
    rho = [1.0 + O.uniform(2041, k) for k in range(2)]
    rec = seeded_points(2, 6, 94)
    run_case("ref2d_el_q7", 2, 4, 7, 0.1, 22, mu, 8.0, 1e-3, 300, (0.52, 0.44), rec,
             elastic=True, lam_layers=lam, rho_layers=rho, pol=(-0.3, 0.9))


def young_poisson(dim, E, nu):
    """Lame coefficients (lambda, mu) of the YOUNG_POISSON media the deck
    states, as Montjoie converts them (ElasticPhysicalIndice::SetInputData,
    Elliptic/PhysicalProperty.cxx:3321-3345): plane stress in 2-D."""
    E, nu = np.asarray(E, float), np.asarray(nu, float)
    lam = E * nu / (1.0 - nu * nu) if dim == 2 else E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return lam, 0.5 * E / (1.0 + nu)


@case
def ref3d_el_q3_damped():
    """damped isotropic elastic: 4^3 Q3 perturbed hexahedra, two layers of
    ISOTROPE rho lambda mu sigma with sigma 2 and 9."""
    lam = [1.0 + 3.0 * O.uniform(2042, k) for k in range(2)]
    mu = [1.0 + 3.0 * O.uniform(2043, k) for k in range(2)]
    rho = [1.0 + O.uniform(2044, k) for k in range(2)]
    rec = seeded_points(3, 5, 95)
    run_case("ref3d_el_q3_damped", 3, 4, 3, 0.1, 23, mu, 8.0, 4e-3, 150, (0.46, 0.53, 0.5), rec,
             elastic=True, lam_layers=lam, rho_layers=rho, pol=(0.3, -0.5, 0.8),
             sigma_layers=[2.0, 9.0])


@case
def ref2d_el_q4_young_damped():
    """2-D elastic media given as YOUNG_POISSON rho E nu sigma (Montjoie's
    plane-stress conversion to lambda, mu): 8x8 Q4 perturbed, three layers,
    sigma 0, 4 and 15."""
    E = [2.0 + 4.0 * O.uniform(2045, k) for k in range(3)]
    nu = [0.1 + 0.3 * O.uniform(2046, k) for k in range(3)]
    rho = [1.0 + O.uniform(2047, k) for k in range(3)]
    lam, mu = young_poisson(2, E, nu)
    rec = seeded_points(2, 6, 96)
    run_case("ref2d_el_q4_young_damped", 2, 8, 4, 0.1, 24, list(mu), 10.0, 1e-3, 300,
             (0.48, 0.51), rec, elastic=True, lam_layers=list(lam), rho_layers=rho,
             pol=(0.8, 0.6), sigma_layers=[0.0, 4.0, 15.0], young=(E, nu))


@case
def ref3d_el_q3_grad():
    """DisplayStress = NO: the SismoPoint rows and the SismoLine files keep
    grad u; 3^3 Q3 perturbed, two layers, a line leaving the cube."""
    lam = [1.0 + 3.0 * O.uniform(2048, k) for k in range(2)]
    mu = [1.0 + 3.0 * O.uniform(2049, k) for k in range(2)]
    rho = [1.0 + O.uniform(2050, k) for k in range(2)]
    run_case("ref3d_el_q3_grad", 3, 3, 3, 0.1, 25, mu, 8.0, 4e-3, 120, (0.47, 0.52, 0.5),
             seeded_points(3, 4, 98), elastic=True, lam_layers=lam, rho_layers=rho,
             pol=(0.2, 0.9, -0.4), line=[0.1, 0.3, 0.2, 1.15, 0.6, 0.7, 9],
             line_times=[0.2, 0.4, 0.1], display_stress=False)


@case
def ref2d_el_q4_grad():
    """DisplayStress = NO on quadrilaterals: SismoPoint rows and SismoPlane
    files of grad u; 8x8 Q4 perturbed, two layers."""
    lam = [1.0 + 3.0 * O.uniform(2051, k) for k in range(2)]
    mu = [1.0 + 3.0 * O.uniform(2052, k) for k in range(2)]
    rho = [1.0 + O.uniform(2053, k) for k in range(2)]
    run_case("ref2d_el_q4_grad", 2, 8, 4, 0.1, 26, mu, 10.0, 1e-3, 200, (0.45, 0.55),
             seeded_points(2, 5, 99), elastic=True, lam_layers=lam, rho_layers=rho,
             pol=(0.5, 0.5), plane=[0.05, 0.95, 0.1, 0.9, 7, 6], plane_times=[0.05, 0.15, 0.05],
             display_stress=False)


@case
def ref2d_q4_order4():
    """OrderTimeScheme = 4 LEAP_FROG (fourth-order modified equation): 8x8 Q4
    perturbed, three layers, at a time step past the second-order CFL limit
    (dt = 1.5 x 2/sqrt(lambda_max), below 2 sqrt(3)/sqrt(lambda_max))."""
    m = O.mesh(2, 8, 4, 0.1, 49)
    lay = [1.0, 2.0, 1.5]
    e = np.arange(m.E)
    P = O.acoustic_problem(m, np.asarray(lay)[((e // 8) * 3) // 8], 1.0)
    dt = 1.5 * 2.0 / np.sqrt(O.lambda_max(P, 200, 5))
    run_case("ref2d_q4_order4", 2, 8, 4, 0.1, 49, lay, 10.0, dt, int(0.3 / dt), (0.47, 0.52),
             seeded_points(2, 5, 100), time_order=4)


@case
def ref2d_q4_src_volume():
    """TypeSource = SRC_VOLUME GAUSSIAN 0.47 0.52 0.15 0.25: a Gaussian volume
    source (radius 0.15, cut off at 0.25) on every element, 8x8 Q4 perturbed,
    two layers"""
    run_case("ref2d_q4_src_volume", 2, 8, 4, 0.1, 53, [1.0, 2.0], 10.0, 1e-3, 250,
             (0.47, 0.52), seeded_points(2, 5, 105), volume=([], (0.47, 0.52), 0.15, 0.25))


@case
def ref3d_q3_order4_src_volume():
    """the fourth-order scheme driven by a Gaussian volume source (f, f' and
    f'' times the same load vector): 4^3 Q3 perturbed hexahedra, two layers,
    a MODIFIED_RICKER pulse"""
    run_case("ref3d_q3_order4_src_volume", 3, 4, 3, 0.1, 57, [1.0, 2.5], 8.0, 2e-3, 150,
             (0.5, 0.48, 0.52), seeded_points(3, 5, 111), temporal="MODIFIED_RICKER 0.06",
             time_order=4, volume=([], (0.5, 0.48, 0.52), 0.2, 0.3))


@case
def ref3d_q3_src_volume_ref():
    """SRC_VOLUME 1 GAUSSIAN ...: the volume source restricted to the elements
    of reference 1 (the lower layer), its Gaussian crossing the layer
    interface; 4^3 Q3 perturbed, two layers"""
    run_case("ref3d_q3_src_volume_ref", 3, 4, 3, 0.1, 54, [1.0, 2.5], 8.0, 5e-3, 150,
             (0.5, 0.5, 0.5), seeded_points(3, 4, 106),
             volume=([1], (0.46, 0.52, 0.45), 0.2, 0.3))


@case
def ref2d_el_q4_src_volume():
    """elastic: the Gaussian volume source times the Polarization; 8x8 Q4,
    two layers"""
    lam = [1.0 + 3.0 * O.uniform(2057, k) for k in range(2)]
    mu = [1.0 + 3.0 * O.uniform(2058, k) for k in range(2)]
    rho = [1.0 + O.uniform(2059, k) for k in range(2)]
    run_case("ref2d_el_q4_src_volume", 2, 8, 4, 0.1, 55, mu, 10.0, 1e-3, 250, (0.5, 0.5),
             seeded_points(2, 4, 107), elastic=True, lam_layers=lam, rho_layers=rho,
             pol=(0.6, -0.8), volume=([], (0.52, 0.47), 0.12, 0.2))


@case
def ref2d_el_q4_sinus():
    """elastic PhysicalMedia: rho and sigma SINUS fields on the one medium
    (PhysicalMedia = rho 1 ... SINUS, sigma 1 ... SINUS; MateriauDielec's
    sigma 1 elsewhere in the box); 6x6 Q4 perturbed"""
    run_case("ref2d_el_q4_sinus", 2, 6, 4, 0.1, 57, [2.0], 10.0, 1e-3, 250, (0.48, 0.51),
             seeded_points(2, 4, 109), elastic=True, lam_layers=[3.0], rho_layers=[1.5],
             pol=(0.7, 0.3), sigma_layers=[1.0],
             sinus_rho=[0.05, 0.1, 0.9, 0.95, 0.02, 0.01, 1.5, 0.4, 0.6, 0.7],
             sinus_sigma=[0.1, 0.05, 0.8, 0.9, 0.01, 0.03, 3.0, 2.0, 0.5, 0.4])


@case
def ref2d_q4_rec_outside():
    """SismoPoint receivers outside the mesh (beyond x = 1 and below y = 0)
    between ones inside: Montjoie writes 0 for them and keeps the columns"""
    rec = np.array([[0.3, 0.4], [1.2, 0.5], [0.71, 0.52], [0.5, -0.1], [0.6, 0.8]])
    run_case("ref2d_q4_rec_outside", 2, 6, 4, 0.1, 58, [1.0, 2.0], 10.0, 1e-3, 200, (0.47, 0.52),
             rec)


@case
def ref2d_q3_src_uniform_ref():
    """TypeSource = SRC_VOLUME 2 UNIFORM: f = 1 on the elements of reference 2
    (the middle layer); 8x8 Q3 perturbed, three layers"""
    run_case("ref2d_q3_src_uniform_ref", 2, 8, 3, 0.1, 56, [1.0, 2.0, 1.5], 10.0, 1e-3, 250,
             (0.5, 0.5), seeded_points(2, 5, 108), volume=([2], [], 0.0, 0.0))


@case
def ref2d_q4_order4_reprise():
    """checkpoint files of the fourth-order scheme (SaveReprise = 60: the
    iterate U^n, as for the second-order one, WriteSnapshot ->
    SaveDataReprise, VarInstationary.cxx:797-801,917-979)"""
    run_case("ref2d_q4_order4_reprise", 2, 6, 4, 0.1, 52, [1.0, 2.0], 10.0, 1e-3, 200,
             (0.47, 0.52), seeded_points(2, 3, 104), save_reprise=60, time_order=4)


@case
def ref2d_q4_order4_damped():
    """the fourth-order scheme with damping (sigma 0, 3, 12 by layer): the
    S d^3U/dt^3 terms and the three time levels"""
    run_case("ref2d_q4_order4_damped", 2, 8, 4, 0.1, 50, [1.0, 2.0, 1.5], 10.0, 1e-3, 300,
             (0.47, 0.52), seeded_points(2, 5, 101), sigma_layers=[0.0, 3.0, 12.0], time_order=4)


@case
def ref3d_q3_order4():
    """the fourth-order scheme on 4^3 Q3 perturbed hexahedra, two layers"""
    run_case("ref3d_q3_order4", 3, 4, 3, 0.1, 51, [1.0, 2.5], 8.0, 5e-3, 150,
             (0.43, 0.52, 0.47), seeded_points(3, 4, 102), time_order=4)


@case
def ref3d_el_q3_order4():
    """the fourth-order scheme for isotropic elastodynamics: 3^3 Q3 perturbed
    hexahedra, two layers"""
    lam = [1.0 + 3.0 * O.uniform(2054, k) for k in range(2)]
    mu = [1.0 + 3.0 * O.uniform(2055, k) for k in range(2)]
    rho = [1.0 + O.uniform(2056, k) for k in range(2)]
    run_case("ref3d_el_q3_order4", 3, 3, 3, 0.1, 27, mu, 8.0, 4e-3, 120, (0.47, 0.52, 0.5),
             seeded_points(3, 4, 103), elastic=True, lam_layers=lam, rho_layers=rho,
             pol=(0.3, 0.6, -0.7), time_order=4)


@case
def ref2d_el_q5_order4_damped():
    """the fourth-order scheme for 2-D elastodynamics with damping: 6x6 Q5
    perturbed quadrilaterals, three layers, sigma 0, 3 and 9; one SismoPoint
    outside the mesh (0 in its displacement and stress columns)"""
    lam = [1.0 + 3.0 * O.uniform(2060, k) for k in range(3)]
    mu = [1.0 + 3.0 * O.uniform(2061, k) for k in range(3)]
    rho = [1.0 + O.uniform(2062, k) for k in range(3)]
    rec = np.vstack([seeded_points(2, 4, 107), [[1.3, 0.45]]])
    run_case("ref2d_el_q5_order4_damped", 2, 6, 5, 0.1, 29, mu, 8.0, 2e-3, 200, (0.47, 0.53),
             rec, elastic=True, lam_layers=lam, rho_layers=rho, pol=(0.7, -0.4),
             sigma_layers=[0.0, 3.0, 9.0], time_order=4)


def _temporal_case(name, temporal, nsteps=250):
    """8x8 Q3 perturbed quadrilaterals, two layers, Frequency 10, a Dirac
    source driven by the TemporalSource line `temporal`."""
    run_case(name, 2, 8, 3, 0.1, 48, [1.0, 2.0], 10.0, 1e-3, nsteps, (0.47, 0.52),
             seeded_points(2, 5, 97), temporal=temporal)


@case
def ref2d_q3_src_derivative_ricker():
    _temporal_case("ref2d_q3_src_derivative_ricker", "DERIVATIVE_RICKER")


@case
def ref2d_q3_src_modified_ricker():
    """centre time 0.06: the cut-off (2.5 + 0.06)/(sqrt(2) 10) = 0.181 lies
    past the pulse"""
    _temporal_case("ref2d_q3_src_modified_ricker", "MODIFIED_RICKER 0.06")


@case
def ref2d_q3_src_gaussian():
    """a plateau of 0.05 between the rise and the fall"""
    _temporal_case("ref2d_q3_src_gaussian", "GAUSSIAN 0.05")


@case
def ref2d_q3_src_harmonic():
    _temporal_case("ref2d_q3_src_harmonic", "HARMONIC 0.2")


@case
def ref2d_q3_src_sinus_gaussian():
    """alpha 2000: cut-off 2 sqrt(ln(1e6)/2000) = 0.166, centred at 0.083"""
    _temporal_case("ref2d_q3_src_sinus_gaussian", "SINUS_GAUSSIAN 2000")


@case
def ref2d_c1_circle():
    """The 64x64 Q4 / 1000-step / 0.5-CFL case with its 16 receivers given as
    Montjoie's own SismoCircle = 0.5 0.5 0.3 16 (no SismoPoint)."""
    m = O.mesh(2, 64, 4)
    P = O.acoustic_problem(m, 1.0, 1.0)
    lam = O.lambda_max(P, 200, 11)
    dt = 0.5 * 2.0 / np.sqrt(lam)
    run_case("ref2d_c1_circle", 2, 64, 4, 0.0, 0, [1.0], 10.0, dt, 1000, (0.5, 0.5),
             np.zeros((0, 2)), regular=True, circle=(0.5, 0.5, 0.3, 16))


@case
def ref2d_q4_circle_out():
    """8x8 Q4 perturbed quadrilaterals, two layers, a SismoCircle that crosses
    the boundary of the square (points outside the mesh), after two
    SismoPoint receivers."""
    run_case("ref2d_q4_circle_out", 2, 8, 4, 0.1, 45, [1.0, 2.0], 10.0, 1e-3, 200, (0.46, 0.53),
             np.array([[0.3, 0.4], [0.71, 0.52]]), circle=(0.75, 0.45, 0.4, 12))


@case
def ref2d_q4_pml():
    """8x8 Q4 perturbed quadrilaterals, two layers, PML of two layers of
    thickness 0.125 on every side (AddPML = YES XY 0.25 2, parabolic
    DampingPML = 3), first-order scheme LEAP_FROG_PML; two receivers inside
    the PML."""
    rec = np.array([[0.3, 0.4], [0.71, 0.52], [0.86, 0.18], [1.12, 0.5], [-0.1, -0.15]])
    run_case("ref2d_q4_pml", 2, 8, 4, 0.1, 46, [1.0, 2.0], 10.0, 1e-3, 400, (0.47, 0.52), rec,
             pml=(2, 0.25, 2, 3.0))


@case
def ref3d_q3_pml():
    """4^3 Q3 perturbed hexahedra, two layers, PML of two layers of thickness
    0.25 on every face (AddPML = YES XYZ 0.5 2, DampingPML = 2), LEAP_FROG_PML;
    one receiver inside the PML corner region."""
    rec = np.vstack([seeded_points(3, 4, 93), [[1.2, -0.15, 0.5], [-0.3, 1.3, 1.2]]])
    run_case("ref3d_q3_pml", 3, 4, 3, 0.1, 47, [1.0, 2.5], 8.0, 4e-3, 150, (0.46, 0.53, 0.49),
             rec, pml=(3, 0.5, 2, 2.0))


@case
def ref2d_q3_onesided_pml():
    """6x6 Q3 perturbed quadrilaterals, two layers; PML only beyond x = 1 and
    below y = 0 (AddPML = YES X+Y- 0.3 AUTO: ceil(0.3 / (1/6)) = 2 layers),
    linear damping profile (DampingPML = LINEAR 2.5); receivers in both
    layers, in their corner and near the undamped Dirichlet sides."""
    rec = np.array([[0.35, 0.45], [0.7, 0.6], [1.15, 0.5], [0.5, -0.2], [1.2, -0.2],
                    [0.05, 0.9]])
    run_case("ref2d_q3_onesided_pml", 2, 6, 3, 0.1, 48, [1.0, 2.0], 10.0, 1e-3, 400,
             (0.46, 0.53), rec,
             pml=dict(axes="X+Y-", thickness=0.3, layers="AUTO", vsigma=2.5, profile=2))


@case
def ref2d_q4_shifted_pml():
    """8x8 Q4 perturbed quadrilaterals, two layers, PML of 3 layers of 0.1 on
    every side with the shifted parabola (DampingPML = SHIFTED_PARABOLE 2.0
    0.2) and transverse damping (TransverseDampingPML = 0.5)."""
    rec = np.array([[0.3, 0.4], [0.71, 0.52], [1.2, 0.5], [-0.2, -0.2], [0.5, 1.25]])
    run_case("ref2d_q4_shifted_pml", 2, 8, 4, 0.1, 49, [1.0, 2.0], 10.0, 1e-3, 400,
             (0.47, 0.52), rec,
             pml=dict(axes="XY", thickness=0.3, layers=3, vsigma=2.0, profile=3, offset=0.2,
                      transverse=0.5))


@case
def ref3d_q2_onesided_pml():
    """4^3 Q2 perturbed hexahedra, two layers; PML below x = 0, above y = 1
    and below z = 0 (AddPML = YES X-Y+Z- 0.5 2), constant damping
    (DampingPML = CONSTANT 1.5)."""
    rec = np.vstack([seeded_points(3, 3, 94), [[-0.3, 0.5, 0.5], [0.5, 1.3, -0.2]]])
    run_case("ref3d_q2_onesided_pml", 3, 4, 2, 0.1, 50, [1.0, 2.5], 8.0, 4e-3, 150,
             (0.46, 0.53, 0.49), rec,
             pml=dict(axes="X-Y+Z-", thickness=0.5, layers=2, vsigma=1.5, profile=1))


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


@case
def ref2d_q4_sinus():
    """8x8 Q4 unit square, rho and mu varying inside the elements
    (PhysicalMedia SINUS): mu with a Gaussian cut-off on a box inside the
    domain, rho over the whole square."""
    circ = O.circle_points(0.5, 0.5, 0.3, 8)
    # xmin(2) xmax(2) cut-off(2) offset amplitude period(2)
    mu = [0.15, 0.2, 0.85, 0.9, 0.02, 0.03, 2.0, 1.2, 0.7, 0.9]
    rho = [0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.5, 0.4, 1.3, 2.1]
    run_case("ref2d_q4_sinus", 2, 8, 4, 0.0, 0, [1.0], 10.0, 1e-3, 200, (0.5, 0.5), circ,
             regular=True, sinus_rho=rho, sinus_mu=mu)


@case
def ref2d_q4_order4_sinus():
    """the fourth-order scheme with rho and mu varying inside the elements
    (PhysicalMedia SINUS) and damping by layer: 8x8 Q4 unit square"""
    mu = [0.1, 0.15, 0.9, 0.85, 0.03, 0.02, 1.8, 1.0, 0.6, 1.1]
    rho = [0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.4, 0.3, 0.9, 1.7]
    run_case("ref2d_q4_order4_sinus", 2, 8, 4, 0.0, 0, [1.0], 10.0, 1.5e-3, 160, (0.45, 0.55),
             seeded_points(2, 6, 109), regular=True, sinus_rho=rho, sinus_mu=mu,
             sigma_layers=[3.0], time_order=4)


@case
def ref3d_q3_sinus():
    """4^3 Q3 perturbed hexahedra, rho and mu varying inside the elements."""
    rec = seeded_points(3, 10, 80)
    # xmin(3) xmax(3) cut-off(3) offset amplitude period(3)
    mu = [0.1, 0.2, 0.15, 0.9, 0.85, 0.8, 0.01, 0.02, 0.015, 2.5, 1.5, 0.8, 1.1, 0.9]
    rho = [0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.2, 0.3, 1.7, 1.3, 2.2]
    run_case("ref3d_q3_sinus", 3, 4, 3, 0.1, 31, [1.0], 8.0, 4e-3, 150, (0.46, 0.52, 0.49), rec,
             sinus_rho=rho, sinus_mu=mu)


@case
def ref3d_q5_c2s():
    """The 48^3 headline problem scaled to 12^3: Q5, vertices perturbed by
    0.1 h (seed 2024), three layers c = 1 + 2 U (seed 2025), Dirac at the centre,
    Ricker f0 = 8, 400 steps; 16 seeded receivers."""
    c = [1.0 + 2.0 * O.uniform(2025, k) for k in range(3)]
    mu = [v * v for v in c]
    rec = seeded_points(3, 16, 82)
    run_case("ref3d_q5_c2s", 3, 12, 5, 0.1, 2024, mu, 8.0, 1e-3, 400, (0.5, 0.5, 0.5), rec)


def smooth_nodal(m, rm, seed, lo, hi):
    """Values in [lo, hi] of a seeded smooth field at the order-rm nodal points
    of every element (Montjoie's local numbering)."""
    X = O.mesh_media_points(m, rm)
    a = [1.0 + 2.0 * O.uniform(seed, k) for k in range(m.dim)]
    g = np.prod([np.sin(np.pi * a[k] * X[..., k] + 0.3 * k) for k in range(m.dim)], axis=0)
    return lo + (hi - lo) * 0.5 * (1.0 + g)


@case
def ref2d_q4_meshmedia():
    """8x8 Q4 perturbed quadrilaterals (Medit file), rho and mu from
    PhysicalMedia SAME_MESH data files: mu of order 3 (interpolated to the
    Q4 quadrature points), rho of order 4."""
    m = O.mesh(2, 8, 4, 0.1, 41)
    mu = smooth_nodal(m, 3, 3031, 0.5, 1.5)
    rho = smooth_nodal(m, 4, 3032, 0.6, 1.2)
    circ = O.circle_points(0.5, 0.5, 0.3, 8)
    run_case("ref2d_q4_meshmedia", 2, 8, 4, 0.1, 41, [1.0], 10.0, 1e-3, 200, (0.48, 0.52), circ,
             mesh_media=[("mu", 3, 1.0, 2.0, mu), ("rho", 4, 0.5, 1.0, rho)])


@case
def ref3d_q3_meshmedia():
    """4^3 Q3 perturbed hexahedra, mu of order 2 and rho of order 4 (above
    the discretisation order) from SAME_MESH data files."""
    m = O.mesh(3, 4, 3, 0.1, 42)
    mu = smooth_nodal(m, 2, 3033, 0.5, 1.5)
    rho = smooth_nodal(m, 4, 3034, 0.6, 1.2)
    rec = seeded_points(3, 8, 83)
    run_case("ref3d_q3_meshmedia", 3, 4, 3, 0.1, 42, [1.0], 8.0, 4e-3, 150, (0.47, 0.51, 0.53),
             rec, mesh_media=[("mu", 2, 1.0, 2.0, mu), ("rho", 4, 0.5, 1.0, rho)])


@case
def ref2d_q4_plane():
    """8x8 Q4 unit square with a SismoPlane snapshot grid that reaches outside
    the mesh (zeros there), every 50 steps, and a SismoLine every 40 steps."""
    circ = O.circle_points(0.5, 0.5, 0.3, 4)
    run_case("ref2d_q4_plane", 2, 8, 4, 0.0, 0, [1.0], 10.0, 1e-3, 200, (0.47, 0.52), circ,
             regular=True, plane=[-0.1, 0.9, 0.2, 1.15, 7, 6], plane_times=[0.0, 0.2, 0.05],
             line=[0.1, 1.3, 0.2, 0.7, 7], line_times=[0.0, 0.2, 0.04])


@case
def ref3d_q3_plane():
    """4^3 Q3 perturbed hexahedra, a tilted SismoPlane A, B, C partly outside
    the cube, every 30 steps, and a SismoLine leaving the cube, every 50 steps."""
    rec = seeded_points(3, 4, 81)
    run_case("ref3d_q3_plane", 3, 4, 3, 0.1, 99, [1.0, 2.5], 8.0, 5e-3, 150, (0.43, 0.52, 0.47),
             rec, plane=[0.05, 0.1, 0.5, 1.2, 0.1, 0.55, 0.05, 0.95, 0.6, 6, 5],
             plane_times=[0.0, 0.6, 0.15], line=[0.1, 1.3, 0.2, 0.7, 0.3, 0.6, 9],
             line_times=[0.0, 0.75, 0.25])


@case
def ref2d_q4_reprise():
    """8x8 Q4 perturbed quadrilaterals, two layers, SaveReprise every 40 of 160
    steps: checkpoints at 0, 40, ..., 160 (the final iterate included)."""
    circ = O.circle_points(0.5, 0.5, 0.3, 4)
    run_case("ref2d_q4_reprise", 2, 8, 4, 0.1, 41, [1.0, 2.0], 10.0, 1e-3, 160, (0.47, 0.52),
             circ, save_reprise=40)


@case
def ref3d_q3_reprise():
    """4^3 Q3 perturbed hexahedra, two layers, SaveReprise every 50 of 120 steps
    (last checkpoint at 100, before the final step)."""
    rec = seeded_points(3, 4, 85)
    run_case("ref3d_q3_reprise", 3, 4, 3, 0.1, 99, [1.0, 2.5], 8.0, 5e-3, 120, (0.43, 0.52, 0.47),
             rec, save_reprise=50)


@case
def ref3d_q4_reprise_regular():
    """3^3 Q4 REGULAR cube, SaveReprise every 30 of 90 steps."""
    rec = seeded_points(3, 4, 86)
    run_case("ref3d_q4_reprise_regular", 3, 3, 4, 0.0, 0, [1.0], 8.0, 5e-3, 90,
             (0.43, 0.52, 0.47), rec, regular=True, save_reprise=30)


@case
def ref3d_el_q3_reprise():
    """4^3 Q3 perturbed hexahedra, isotropic elastic, SaveReprise every 25 of
    100 steps."""
    rec = seeded_points(3, 4, 87)
    run_case("ref3d_el_q3_reprise", 3, 4, 3, 0.1, 7, [1.0], 6.0, 4e-3, 100, (0.46, 0.53, 0.49),
             rec, elastic=True, lam_layers=[2.0], rho_layers=[1.0], pol=(0.3, -0.5, 0.8),
             save_reprise=25)


@case
def ref2d_el_q4_order4_reprise():
    """2-D elastic fourth-order scheme with SaveReprise every 30 of 120 steps
    (the checkpoints hold U^n only): 6x6 Q4 perturbed, two layers"""
    lam = [1.5 + 2.0 * O.uniform(2064, k) for k in range(2)]
    rho = [1.0 + O.uniform(2065, k) for k in range(2)]
    run_case("ref2d_el_q4_order4_reprise", 2, 6, 4, 0.1, 31, [1.0, 2.0], 8.0, 2e-3, 120,
             (0.52, 0.47), seeded_points(2, 4, 113), elastic=True, lam_layers=lam,
             rho_layers=rho, pol=(-0.5, 0.9), save_reprise=30, time_order=4)


@case
def ref3d_el_q3_plane():
    """Elastic 4^3 Q3 perturbed hexahedra with a tilted SismoPlane partly
    outside the cube every 30 steps and a SismoLine every 45 steps: one file
    per displacement component and per stress component."""
    rec = seeded_points(3, 3, 88)
    run_case("ref3d_el_q3_plane", 3, 4, 3, 0.1, 11, [1.5], 6.0, 4e-3, 90, (0.44, 0.53, 0.5),
             rec, elastic=True, lam_layers=[2.0], rho_layers=[1.2], pol=(0.6, 0.2, -0.7),
             plane=[0.05, 0.1, 0.5, 1.2, 0.1, 0.55, 0.05, 0.95, 0.6, 5, 4],
             plane_times=[0.0, 0.36, 0.12], line=[0.1, 1.3, 0.2, 0.7, 0.3, 0.6, 7],
             line_times=[0.0, 0.36, 0.18])


@case
def ref2d_el_q4_plane():
    """Elastic 8x8 Q4 perturbed quadrilaterals with a SismoPlane box reaching
    outside the square every 50 steps and a SismoLine every 40 steps: one
    file per displacement component and per stress component."""
    rec = seeded_points(2, 3, 95)
    run_case("ref2d_el_q4_plane", 2, 8, 4, 0.1, 23, [1.5, 2.5], 10.0, 1e-3, 200, (0.47, 0.52),
             rec, elastic=True, lam_layers=[2.0, 1.2], rho_layers=[1.2, 1.0], pol=(0.6, -0.8),
             plane=[-0.1, 0.9, 0.2, 1.15, 7, 6], plane_times=[0.0, 0.2, 0.05],
             line=[0.1, 1.3, 0.2, 0.7, 7], line_times=[0.0, 0.2, 0.04])


@case
def ref3d_q7_pert():
    """3^3 Q7 perturbed hexahedra (the highest order), two layers."""
    rec = seeded_points(3, 6, 89)
    run_case("ref3d_q7_pert", 3, 3, 7, 0.1, 31, [1.0, 2.0], 6.0, 4e-3, 150, (0.48, 0.51, 0.46),
             rec)


@case
def ref3d_q2_pert():
    """5^3 Q2 perturbed hexahedra, three layers."""
    rec = seeded_points(3, 6, 90)
    run_case("ref3d_q2_pert", 3, 5, 2, 0.1, 33, [1.0, 2.0, 1.5], 4.0, 1.5e-2, 120,
             (0.52, 0.47, 0.5), rec)


@case
def ref2d_q1_pert():
    """16x16 Q1 perturbed quadrilaterals (the lowest order), two layers."""
    circ = O.circle_points(0.5, 0.5, 0.3, 6)
    run_case("ref2d_q1_pert", 2, 16, 1, 0.1, 32, [1.0, 3.0], 4.0, 1e-2, 150, (0.47, 0.52), circ)


@case
def ref2d_q6_pert():
    """6x6 Q6 perturbed quadrilaterals, two layers."""
    circ = O.circle_points(0.5, 0.5, 0.3, 6)
    run_case("ref2d_q6_pert", 2, 6, 6, 0.1, 34, [1.0, 2.0], 8.0, 2.5e-3, 200, (0.47, 0.52), circ)


@case
def ref2d_q4_damped():
    """8x8 Q4 perturbed quadrilaterals, three layers with damping sigma 0, 3, 12
    (MateriauDielec's third value)."""
    circ = O.circle_points(0.5, 0.5, 0.3, 6)
    run_case("ref2d_q4_damped", 2, 8, 4, 0.1, 43, [1.0, 2.0, 1.5], 10.0, 1e-3, 300, (0.47, 0.52),
             circ, sigma_layers=[0.0, 3.0, 12.0])


@case
def ref3d_q3_damped():
    """4^3 Q3 perturbed hexahedra, two layers with damping sigma 5 and 20."""
    rec = seeded_points(3, 6, 91)
    run_case("ref3d_q3_damped", 3, 4, 3, 0.1, 99, [1.0, 2.5], 8.0, 5e-3, 150, (0.43, 0.52, 0.47),
             rec, sigma_layers=[5.0, 20.0])


@case
def ref2d_q4_sinus_sigma():
    """8x8 Q4 perturbed quadrilaterals, two layers: the damping of the first
    is a PhysicalMedia sigma SINUS field (offset 4, amplitude 3, cut-off box),
    the second keeps MateriauDielec's sigma = 2."""
    circ = O.circle_points(0.5, 0.5, 0.3, 6)
    run_case("ref2d_q4_sinus_sigma", 2, 8, 4, 0.1, 44, [1.0, 2.0], 10.0, 1e-3, 300, (0.47, 0.52),
             circ, sigma_layers=[1.0, 2.0],
             sinus_sigma=[0.05, 0.1, 0.95, 0.9, 0.01, 0.02, 4.0, 3.0, 0.5, 0.7])


@case
def fmt3d_q3_plane_float():
    """The ref3d_q3_plane deck with OutputFormat = BINARY FLOAT: float snapshot
    files (data type 0) and 8-digit seismograms.  Named outside the ref*
    family: its 7-digit traces are not a 1e-10 parity fixture."""
    rec = seeded_points(3, 4, 81)
    run_case("fmt3d_q3_plane_float", 3, 4, 3, 0.1, 99, [1.0, 2.5], 8.0, 5e-3, 150,
             (0.43, 0.52, 0.47), rec, plane=[0.05, 0.1, 0.5, 1.2, 0.1, 0.55, 0.05, 0.95, 0.6, 6, 5],
             plane_times=[0.0, 0.6, 0.15], line=[0.1, 1.3, 0.2, 0.7, 0.3, 0.6, 9],
             line_times=[0.0, 0.75, 0.25], output_format="BINARY FLOAT")


@case
def fmt2d_q4_plane_ascii():
    """The ref2d_q4_plane deck with OutputFormat = ASCII DOUBLE (text snapshot
    files)."""
    circ = O.circle_points(0.5, 0.5, 0.3, 4)
    run_case("fmt2d_q4_plane_ascii", 2, 8, 4, 0.0, 0, [1.0], 10.0, 1e-3, 200, (0.47, 0.52), circ,
             regular=True, plane=[-0.1, 0.9, 0.2, 1.15, 7, 6], plane_times=[0.0, 0.2, 0.05],
             line=[0.1, 1.3, 0.2, 0.7, 7], line_times=[0.0, 0.2, 0.04],
             output_format="ASCII DOUBLE")


@case
def fmt3d_q3_plane_ascii_float():
    """The ref3d_q3_plane deck with OutputFormat = ASCII FLOAT."""
    rec = seeded_points(3, 4, 81)
    run_case("fmt3d_q3_plane_ascii_float", 3, 4, 3, 0.1, 99, [1.0, 2.5], 8.0, 5e-3, 150,
             (0.43, 0.52, 0.47), rec, plane=[0.05, 0.1, 0.5, 1.2, 0.1, 0.55, 0.05, 0.95, 0.6, 6, 5],
             plane_times=[0.0, 0.6, 0.15], line=[0.1, 1.3, 0.2, 0.7, 0.3, 0.6, 9],
             line_times=[0.0, 0.75, 0.25], output_format="ASCII FLOAT")


if __name__ == "__main__":
    if not os.path.exists(os.path.join(REF, "mjref_acous.x")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    names = sys.argv[1:] or list(CASES)
    for n in names:
        CASES[n]()
