"""Montjoie data files (.ini) and Medit meshes for the synthetic test
problems: written by make_ref_fixtures.py for Montjoie's own solvers and by
the deck-driver tests for montjoie_amd/mjwave_run, from the same parameters
(stored in tests/golden/<case>.npz)."""
from __future__ import annotations

import os

import numpy as np

from oracle import oracle as O


def _f(x: float) -> str:
    return repr(float(x))


def medit_topology(dim, nx):
    """Elements (Medit corner order) and boundary faces (edges in 2-D) of the
    structured mesh, 0-based vertex ids, in the order write_medit lists them."""
    nv1 = nx + 1
    vid = lambda a, b, c=0: a + nv1 * (b + nv1 * c)  # noqa: E731
    if dim == 3:
        corners = [[vid(a, b, c), vid(a + 1, b, c), vid(a + 1, b + 1, c), vid(a, b + 1, c),
                    vid(a, b, c + 1), vid(a + 1, b, c + 1), vid(a + 1, b + 1, c + 1),
                    vid(a, b + 1, c + 1)]
                   for c in range(nx) for b in range(nx) for a in range(nx)]
        faces = []
        for i in range(nx):
            for j in range(nx):
                faces.append([vid(i, j, 0), vid(i, j + 1, 0), vid(i + 1, j + 1, 0), vid(i + 1, j, 0)])
                faces.append([vid(i, j, nx), vid(i + 1, j, nx), vid(i + 1, j + 1, nx), vid(i, j + 1, nx)])
                faces.append([vid(i, 0, j), vid(i + 1, 0, j), vid(i + 1, 0, j + 1), vid(i, 0, j + 1)])
                faces.append([vid(i, nx, j), vid(i, nx, j + 1), vid(i + 1, nx, j + 1), vid(i + 1, nx, j)])
                faces.append([vid(0, i, j), vid(0, i, j + 1), vid(0, i + 1, j + 1), vid(0, i + 1, j)])
                faces.append([vid(nx, i, j), vid(nx, i + 1, j), vid(nx, i + 1, j + 1), vid(nx, i, j + 1)])
    else:
        corners = [[vid(a, b), vid(a + 1, b), vid(a + 1, b + 1), vid(a, b + 1)]
                   for b in range(nx) for a in range(nx)]
        faces = []
        for i in range(nx):
            faces += [[vid(i, 0), vid(i + 1, 0)], [vid(nx, i), vid(nx, i + 1)],
                      [vid(i + 1, nx), vid(i, nx)], [vid(0, i + 1), vid(0, i)]]
    return corners, faces


def write_medit(path, m: O.Mesh, elem_ref, bnd_ref=1):
    """Medit mesh of the structured (perturbed) mesh: vertices, hexahedra
    (quadrilaterals in 2-D) with their media reference, boundary faces."""
    dim, nx = m.dim, m.nx
    V = m.vert.reshape(-1, dim)
    lines = ["MeshVersionFormatted 1", "Dimension", str(dim), "Vertices", str(len(V))]
    for v in V:
        lines.append(" ".join(_f(c) for c in v) + " 0")
    corners, faces = medit_topology(dim, nx)
    lines += ["Hexahedra" if dim == 3 else "Quadrilaterals", str(len(corners))]
    lines += [" ".join(str(v + 1) for v in cs) + f" {int(elem_ref[e])}"
              for e, cs in enumerate(corners)]
    lines += ["Quadrilaterals" if dim == 3 else "Edges", str(len(faces))]
    lines += [" ".join(str(v + 1) for v in f) + f" {bnd_ref}" for f in faces]
    lines.append("End")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def sinus_line(name, ref, p):
    """PhysicalMedia line of a SINUS field; p = [xmin(dim), xmax(dim), coef(dim),
    offset, amplitude, period(dim)] (the layout stored in the fixtures)."""
    p = [float(v) for v in p]
    dim = (len(p) - 2) // 4
    xmin, xmax, coef = p[:dim], p[dim:2 * dim], p[2 * dim:3 * dim]
    off, amp, per = p[3 * dim], p[3 * dim + 1], p[3 * dim + 2:]
    box = []
    for a in range(dim):
        box += [xmin[a], xmax[a]]
    vals = box + coef + [off, amp] + per
    return f"PhysicalMedia = {name} {ref} ISOTROPE SINUS " + " ".join(_f(v) for v in vals)


def write_don(path, nu):
    """Binary data file of a PhysicalMedia MESH field (ReadMeshData,
    Output/CommonInputOutput.cxx:325-352): int 0 (real), int E, int offsets[E+1],
    then every element's values (double)."""
    nu = np.asarray(nu, np.float64)
    E, npn = nu.shape
    with open(path, "wb") as fh:
        fh.write(np.array([0, E], np.int32).tobytes())
        fh.write((np.arange(E + 1, dtype=np.int32) * npn).tobytes())
        fh.write(nu.tobytes())


def pml_spec(pml, dim):
    """PML parameters as a dict (axes, thickness, layers int or "AUTO",
    vsigma, profile 0 parabola / 1 constant / 2 linear / 3 shifted parabola,
    offset, transverse); a tuple (ignored, thickness, layers, vsigma) means
    every side with the parabolic profile."""
    if isinstance(pml, dict):
        return pml
    return dict(axes={2: "XY", 3: "XYZ"}[dim], thickness=float(pml[1]), layers=int(pml[2]),
                vsigma=float(pml[3]), profile=0, offset=0.0, transverse=0.0)


def write_deck(d, dim, nx, r, pert, seed, layers, f0, dt, nsteps, src, rec, elastic=False,
               lam_layers=None, rho_layers=None, pol=None, regular=False, sinus_rho=None,
               sinus_mu=None, plane=None, plane_times=None, line=None, line_times=None,
               mesh_media=None, save_reprise=0, sigma_layers=None, sinus_sigma=None,
               output_format="BINARY DOUBLE", circle=None, pml=None, young=None,
               temporal="RICKER", display_stress=True, time_order=2, volume=None):
    """case.ini (+ mesh.mesh unless REGULAR) in directory d.  circle = (x0, y0,
    radius, n): a SismoCircle line after the SismoPoint lines.  pml (pml_spec):
    AddPML = YES axes thickness layers, DampingPML (and TransverseDampingPML)
    and the first-order scheme OrderTimeScheme = 2 LEAP_FROG_PML.  sinus_rho /
    sinus_mu: SINUS fields (sinus_line layout) overriding rho / mu of medium 1.
    save_reprise > 0: checkpoint files every save_reprise steps.  sigma_layers:
    damping of each layer (MateriauDielec's value after the stiffness).
    young = (E_layers, nu_layers): elastic media written as YOUNG_POISSON
    rho E nu instead of ISOTROPE rho lambda mu.  temporal: the TemporalSource
    line's value (a kind and its parameter).  display_stress = False writes
    DisplayStress = NO (elastic outputs hold grad u instead of the stress).
    time_order = 4: OrderTimeScheme = 4 LEAP_FROG (the modified equation).
    volume = (refs, centre, radius, cut-off): TypeSource = SRC_VOLUME [refs]
    GAUSSIAN centre radius cut-off instead of the Dirac at src; an empty
    centre writes SRC_VOLUME [refs] UNIFORM."""
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
    if pml is not None and len(pml):
        spec = pml_spec(pml, dim)
        lay = spec["layers"] if spec["layers"] == "AUTO" else int(spec["layers"])
        ini.append(f"AddPML = YES {spec['axes']} {_f(spec['thickness'])} {lay}")
        prof, vs = spec.get("profile", 0), _f(spec["vsigma"])
        if prof == 0:
            ini.append(f"DampingPML = {vs}")
        elif prof == 3:
            ini.append(f"DampingPML = SHIFTED_PARABOLE {vs} {_f(spec['offset'])}")
        else:
            ini.append(f"DampingPML = {'CONSTANT' if prof == 1 else 'LINEAR'} {vs}")
        if spec.get("transverse"):
            ini.append(f"TransverseDampingPML = {_f(spec['transverse'])}")
    for k in range(nl):
        if elastic:
            if young is not None:
                mat = (f"MateriauDielec = {k + 1} YOUNG_POISSON {_f(rho_layers[k])} "
                       f"{_f(young[0][k])} {_f(young[1][k])}")
            else:
                mat = (f"MateriauDielec = {k + 1} ISOTROPE {_f(rho_layers[k])} "
                       f"{_f(lam_layers[k])} {_f(layers[k])}")
            if sigma_layers is not None and len(sigma_layers):
                mat += f" {_f(sigma_layers[k])}"
            ini.append(mat)
        else:
            sg = 0.0 if sigma_layers is None or not len(sigma_layers) else sigma_layers[k]
            ini.append(f"MateriauDielec = {k + 1} ISOTROPE 1.0 {_f(layers[k])} {_f(sg)}")
    for name, rm, off, amp, nu in (mesh_media or []):
        # PhysicalMedia ... SAME_MESH on the deck's Medit mesh (medium 1)
        write_don(os.path.join(d, name + ".don"), nu)
        ini.append(f"PhysicalMedia = {name} 1 ISOTROPE SAME_MESH mesh.mesh {_f(off)} {_f(amp)} "
                   f"{name}.don {int(rm)}")
    if sinus_rho is not None and len(sinus_rho):
        ini.append(sinus_line("rho", 1, sinus_rho))
    if sinus_mu is not None and len(sinus_mu):
        ini.append(sinus_line("mu", 1, sinus_mu))
    if sinus_sigma is not None and len(sinus_sigma):
        ini.append(sinus_line("sigma", 1, sinus_sigma))
    if volume is not None and not len(volume[1]):
        ini.append("TypeSource = SRC_VOLUME " + "".join(f"{int(k)} " for k in volume[0])
                   + "UNIFORM")
    elif volume is not None:
        refs, centre, rad, cut = volume
        ini.append("TypeSource = SRC_VOLUME " + "".join(f"{int(k)} " for k in refs) + "GAUSSIAN "
                   + " ".join(_f(c) for c in centre) + f" {_f(rad)} {_f(cut)}")
    else:
        ini.append("TypeSource = SRC_DIRAC")
    ini.append("OriginePhase = " + " ".join(_f(c) for c in src))
    if elastic:
        ini.append("Polarization = " + " ".join(_f(c) for c in pol))
    T = (nsteps - 0.5) * dt
    ini.append(f"ParametersOutputPoint = 0.0 {_f(nsteps * dt + 0.5 * dt)} {_f(dt)}")
    ini.append("FileOutputPoint = diffracPoint.dat totalPoint.dat")
    for p in rec:
        ini.append("SismoPoint = " + " ".join(_f(c) for c in p))
    if circle is not None and len(circle):
        ini.append("SismoCircle = " + " ".join(_f(v) for v in circle[:3]) + f" {int(circle[3])}")
        ini.append("FileOutputCircle = diffracCircle.dat totalCircle.dat")
        ini.append(f"ParametersOutputCircle = 0.0 {_f(nsteps * dt + 0.5 * dt)} {_f(dt)}")
    if plane is not None and len(plane):
        ini.append("SismoPlane = " + " ".join(_f(v) if i < len(plane) - 2 else str(int(v))
                                              for i, v in enumerate(plane)))
        ini.append("FileOutputPlane = diffracPlane totalPlane")
        ini.append("ParametersOutputPlane = " + " ".join(_f(v) for v in plane_times))
    if line is not None and len(line):
        ini.append("SismoLine = " + " ".join(_f(v) for v in line[:-1]) + f" {int(line[-1])}")
        ini.append("FileOutputLine = diffracLine totalLine")
        ini.append("ParametersOutputLine = " + " ".join(_f(v) for v in line_times))
    ini.append(f"OutputFormat = {output_format}")
    if save_reprise:
        ini.append(f"SaveReprise = {int(save_reprise)}")
    ini.append("ElectricOrMagnetic = " + ("-1" if elastic else "0"))
    ini.append(f"TimeStep = {_f(dt)}")
    ini.append(f"TimeInterval = 0.0 {_f(T)}")
    ini.append(f"OrderTimeScheme = {int(time_order)} " + ("LEAP_FROG_PML" if pml is not None and len(pml)
                                           else "LEAP_FROG"))
    ini.append(f"TemporalSource = {temporal or 'RICKER'}")
    if elastic and not display_stress:
        ini.append("DisplayStress = NO")
    ini.append("PrintLevel = 0")
    with open(os.path.join(d, "case.ini"), "w") as fh:
        fh.write("\n".join(ini) + "\n")
