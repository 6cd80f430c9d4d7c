"""ctypes front-end of the C oracle (oracle/mjw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s cpu_baseline leg, as the checker.  The product package
``montjoie_amd`` never imports it.

The oracle restates Montjoie's explicit leapfrog path in plain FP64 C; each C
routine cites the reference file:line it follows.  This module only marshals
numpy arrays and assembles the structured synthetic problems used by the
tests (regular / perturbed meshes of the unit square and cube).
"""
from __future__ import annotations

import ctypes as C
import itertools
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libmjw_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle shared library with the committed Makefile."""
    subprocess.run(["make", "-C", _HERE, "oracle"] + (["-B"] if force else []),
                   check=True, stdout=subprocess.DEVNULL)
    return _SO


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_SO):
        build()
    L = C.CDLL(_SO)
    L.or_gl.argtypes = [C.c_int, _dp, _dp]
    L.or_gl.restype = C.c_int
    L.or_gl_deriv.argtypes = [C.c_int, _dp, _dp]
    L.or_phi.argtypes = [C.c_int, _dp, C.c_int, C.c_double]
    L.or_phi.restype = C.c_double
    L.or_uniform.argtypes = [C.c_uint64, C.c_uint64]
    L.or_uniform.restype = C.c_double
    L.or_mesh_vertices.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint64, _dp]
    L.or_mesh_nodes.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, _bp]
    L.or_mesh_nodes.restype = C.c_int
    L.or_geometry.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_geometry.restype = C.c_double
    L.or_mass.argtypes = [C.c_long, C.c_int, C.c_long, _ip, _dp, _bp, _dp, _dp]
    L.or_color.argtypes = [C.c_long, C.c_int, C.c_long, _ip, _ip]
    L.or_color.restype = C.c_int
    L.or_stiffness.argtypes = [C.c_int, C.c_int, C.c_long, C.c_long, _ip, _dp, C.c_void_p,
                               C.c_int, C.c_double, _dp, C.c_double, _dp]
    L.or_stiffness_elastic.argtypes = [C.c_int, C.c_int, C.c_long, C.c_long, _ip, _dp, _dp, _dp, _dp,
                                       _ip, C.c_int, C.c_double, _dp, C.c_double, _dp]
    L.or_ricker.argtypes = [C.c_double, C.c_double]
    L.or_ricker.restype = C.c_double
    L.or_pulse_table.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    L.or_source_table.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_int, C.c_double, _dp]
    L.or_source_limit.argtypes = [C.c_int, C.c_double, C.c_double]
    L.or_source_limit.restype = C.c_double
    L.or_dirac.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _ip, C.c_long, _bp, _dp, C.c_int,
                           _dp, _dp]
    L.or_dirac.restype = C.c_int
    L.or_receiver.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    L.or_receiver.restype = C.c_long
    L.or_receiver_grad.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    L.or_receiver_grad.restype = C.c_long
    L.or_circle_points.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    L.or_leapfrog.argtypes = [C.c_int, C.c_int, C.c_int, C.c_long, C.c_long, _ip,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              _ip, C.c_int, _dp, C.c_void_p, _bp, C.c_void_p, C.c_void_p,
                              C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_int, _lp, _dp,
                              C.c_void_p, _dp, _dp, _dp]
    L.or_leapfrog4.argtypes = [C.c_int, C.c_int, C.c_int, C.c_long, C.c_long, _ip,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               _ip, C.c_int, _dp, C.c_void_p, _bp, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p,
                               C.c_void_p, C.c_int, _lp, _dp, C.c_void_p, _dp, _dp, _dp]
    L.or_lambda_max.argtypes = [C.c_int, C.c_int, C.c_int, C.c_long, C.c_long, _ip,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                _ip, C.c_int, _dp, _bp, C.c_int, C.c_uint64]
    L.or_lambda_max.restype = C.c_double
    L.or_adjoint_acoustic.argtypes = [C.c_int, C.c_int, C.c_long, C.c_long, _ip, _dp, _dp, _ip,
                                      C.c_int, _dp, _bp, C.c_void_p, C.c_void_p, C.c_double,
                                      C.c_int, C.c_int, _lp, _dp, _dp, _dp]
    L.or_time_table.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                C.c_double, _dp]
    L.or_leapfrog_pml.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_long, _ip, _dp, _dp,
                                  C.c_void_p, _bp, _dp, _bp, C.c_void_p, _dp, C.c_double,
                                  C.c_int, C.c_int, _lp, _dp, _dp, _dp]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------
# 1-D rule
# --------------------------------------------------------------------------

def gl(r: int):
    x = np.zeros(r + 1)
    w = np.zeros(r + 1)
    if lib().or_gl(r, x, w):
        raise ValueError("bad order")
    return x, w


def gl_deriv(r: int):
    x, _ = gl(r)
    D = np.zeros((r + 1) * (r + 1))
    lib().or_gl_deriv(r, x, D)
    return D.reshape(r + 1, r + 1)


def phi(r: int, i: int, t: float) -> float:
    x, _ = gl(r)
    return lib().or_phi(r, x, i, t)


def ricker(f0: float, t: float) -> float:
    return lib().or_ricker(f0, t)


SOURCE_KINDS = {"RICKER": 1, "DERIVATIVE_RICKER": 2, "MODIFIED_RICKER": 3, "GAUSSIAN": 4,
                "HARMONIC": 5, "SINUS_GAUSSIAN": 6}


def source_table(kind, f0: float, par: float, t_begin: float, dt: float, nsteps: int,
                 nb_deriv: int = 1, offset: float | None = None) -> np.ndarray:
    """or_source_table: the TemporalSource `kind` (keyword or number) with its
    value `par`; nb_deriv = 1 and offset = dt (default) is the leapfrog's
    pulse table, mjw_oracle.c:or_source_table."""
    k = SOURCE_KINDS[kind] if isinstance(kind, str) else int(kind)
    out = np.zeros(nsteps)
    lib().or_source_table(k, f0, par, t_begin, dt, nsteps, nb_deriv,
                          dt if offset is None else offset, out)
    return out


def source_limit(kind, f0: float, par: float) -> float:
    k = SOURCE_KINDS[kind] if isinstance(kind, str) else int(kind)
    return lib().or_source_limit(k, f0, par)


def pulse_table(f0: float, t_begin: float, dt: float, nsteps: int) -> np.ndarray:
    p = np.zeros(nsteps)
    lib().or_pulse_table(f0, t_begin, dt, nsteps, p)
    return p


THRESHOLD = 1e-10  # R2/R3::threshold as Mesh/NumberMesh.cxx:1621,2727 set it


def sinus_media(pts, xmin, xmax, coef, offset, amplitude, period):
    """PhysicalMedia ... SINUS: the sinusoidal field of
    Elliptic/PhysicalProperty.cxx:218-245 at the points pts[..., dim].  Inside
    the box [xmin, xmax] (PointInsideBoundingBox, SELDON
    vector/TinyVectorInline.cxx:1517-1540, with THRESHOLD):
    offset + amplitude * prod_i sin(2 pi x_i / period_i) * cutoff(x), the cutoff
    prod_i exp(-c_i / (x_i - xmin_i)^2) exp(-c_i / (x_i - xmax_i)^2) when some
    c_i != 0 (RectangleCutOff, PhysicalProperty.cxx:28-95; a factor is 0 within
    THRESHOLD of a face with c_i != 0).  Outside the box: offset."""
    pts = np.asarray(pts, float)
    dim = pts.shape[-1]
    xmin, xmax, coef = (np.asarray(v, float) for v in (xmin, xmax, coef))
    k = 2.0 * np.pi / np.asarray(period, float)
    inside = np.all((pts >= xmin - THRESHOLD) & (pts <= xmax + THRESHOLD), axis=-1)
    val = np.ones(pts.shape[:-1])
    for i in range(dim):
        val = val * np.sin(pts[..., i] * k[i])
    if np.any(coef != 0):
        for i in range(dim):
            for b in (xmin[i], xmax[i]):
                d = pts[..., i] - b
                near = np.abs(d) <= THRESHOLD
                f = np.exp(-coef[i] / np.where(near, 1.0, d * d))
                val = val * np.where(near, 0.0 if coef[i] != 0 else 1.0, f)
    return np.where(inside, offset + amplitude * val, offset)


def nodal_numbering(dim: int, r: int) -> np.ndarray:
    """num[i, j(, k)]: Montjoie's local number of the nodal point (x_i, x_j(,
    x_k)) of a Q_r quadrilateral / hexahedron (Quadrature/
    QuadrangleQuadrature.cxx:42-100, HexahedronQuadrature.cxx:43-182):
    vertices, then edges (square: bottom, right, top reversed, left reversed;
    cube: the 12 edges of GetEdgeCube from their first to their second
    vertex), then faces x=0, y=0, z=0, z=r, y=r, x=r (i outer, j inner), then
    the interior (i outer, k inner)."""
    if dim == 2:
        num = -np.ones((r + 1, r + 1), int)
        num[0, 0], num[r, 0], num[r, r], num[0, r] = 0, 1, 2, 3
        nb = 4
        for i in range(1, r):
            num[i, 0] = nb; nb += 1
        for i in range(1, r):
            num[r, i] = nb; nb += 1
        for i in range(1, r):
            num[r - i, r] = nb; nb += 1
        for i in range(1, r):
            num[0, r - i] = nb; nb += 1
        for i in range(1, r):
            for j in range(1, r):
                num[i, j] = nb; nb += 1
        return num
    num = -np.ones((r + 1,) * 3, int)
    corners = [(0, 0, 0), (r, 0, 0), (r, r, 0), (0, r, 0), (0, 0, r), (r, 0, r), (r, r, r),
               (0, r, r)]
    for n, c in enumerate(corners):
        num[c] = n
    nb = 8
    edges = [(0, 1), (1, 2), (3, 2), (0, 3), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6),
             (7, 6), (4, 7)]
    for a, b in edges:
        p1, p2 = np.array(corners[a]), np.array(corners[b])
        for l in range(1, r):
            num[tuple((l * p2 + (r - l) * p1) // r)] = nb
            nb += 1
    faces = [lambda i, j: (0, i, j), lambda i, j: (i, 0, j), lambda i, j: (i, j, 0),
             lambda i, j: (i, j, r), lambda i, j: (i, r, j), lambda i, j: (r, i, j)]
    for f in faces:
        for i in range(1, r):
            for j in range(1, r):
                num[f(i, j)] = nb
                nb += 1
    for i in range(1, r):
        for j in range(1, r):
            for k in range(1, r):
                num[i, j, k] = nb
                nb += 1
    return num


def mesh_media_points(m: "Mesh", rm: int) -> np.ndarray:
    """(E, (rm+1)^dim, dim) positions of the order-rm nodal points of every
    element in Montjoie's local numbering (GL points of order rm mapped by the
    element's multilinear map, HexahedronGeomReference.cxx:76-112), the points
    a PhysicalMedia MESH data file gives values at."""
    x, _ = gl(rm)
    num = nodal_numbering(m.dim, rm)
    V = m.vert.reshape(-1, m.dim)
    nv1 = m.nx + 1
    npn = (rm + 1) ** m.dim
    out = np.zeros((m.E, npn, m.dim))
    for e in range(m.E):
        if m.dim == 2:
            a, b = e % m.nx, e // m.nx
            cv = [a + nv1 * b, a + 1 + nv1 * b, a + 1 + nv1 * (b + 1), a + nv1 * (b + 1)]
            for i in range(rm + 1):
                for j in range(rm + 1):
                    s, t = x[i], x[j]
                    w = [(1 - s) * (1 - t), s * (1 - t), s * t, (1 - s) * t]
                    out[e, num[i, j]] = sum(w[c] * V[cv[c]] for c in range(4))
        else:
            a, b, c0 = e % m.nx, (e // m.nx) % m.nx, e // (m.nx * m.nx)
            vid = lambda p, q, z: p + nv1 * (q + nv1 * z)  # noqa: E731
            cv = [vid(a, b, c0), vid(a + 1, b, c0), vid(a + 1, b + 1, c0), vid(a, b + 1, c0),
                  vid(a, b, c0 + 1), vid(a + 1, b, c0 + 1), vid(a + 1, b + 1, c0 + 1),
                  vid(a, b + 1, c0 + 1)]
            ca = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1),
                  (0, 1, 1)]
            for i in range(rm + 1):
                for j in range(rm + 1):
                    for k in range(rm + 1):
                        xi = (x[i], x[j], x[k])
                        p = np.zeros(3)
                        for c in range(8):
                            w = 1.0
                            for d in range(3):
                                w *= xi[d] if ca[c][d] else 1 - xi[d]
                            p += w * V[cv[c]]
                        out[e, num[i, j, k]] = p
    return out


def mesh_media(m: "Mesh", nu, rm: int, offset: float = 0.0, amplitude: float = 1.0):
    """PhysicalMedia ... MESH|SAME_MESH on the computational mesh: nu[e] holds
    the element's values at its order-rm nodal points (Montjoie's local
    numbering); the field offset + amplitude nu (PhysicalMeshMedia::Init,
    Elliptic/PhysicalProperty.cxx:990-1000) is interpolated with the
    tensor Lagrange basis on the GL points of order rm at the quadrature
    points of order r (ProjectScalar, Harmonic/VarGeometryProblem.cxx:
    1428-1447).  Returns (E, nd) in the local order i + n j (+ n^2 k)."""
    xm, _ = gl(rm)
    xq, _ = gl(m.r)
    num = nodal_numbering(m.dim, rm)
    # L[q, a] = phi_a(xq_q) of the order-rm 1-D Lagrange basis
    L = np.array([[phi(rm, a, t) for a in range(rm + 1)] for t in xq])
    n = m.r + 1
    out = np.zeros((m.E, m.nd))
    nu = np.asarray(nu, float).reshape(m.E, -1)
    for e in range(m.E):
        v = offset + amplitude * nu[e]
        if m.dim == 2:
            T = v[num]                                    # T[a, b]
            out[e] = np.einsum("ia,jb,ab->ji", L, L, T).reshape(-1)
        else:
            T = v[num]                                    # T[a, b, c]
            out[e] = np.einsum("ia,jb,kc,abc->kji", L, L, L, T).reshape(-1)
    return out


def montjoie_dofs(dim: int, r: int, corners, boundary, elem_nodes, n_nodes: int) -> np.ndarray:
    """Montjoie's global DOF number of every node of a continuous Q_r mesh.

    corners[e]: the element's vertices (Medit order), boundary: the referenced
    boundary edges (2-D) or faces (3-D) in file order, elem_nodes[e][i + n j
    (+ n^2 k)]: the caller's node numbers.  Restates
      - referenced boundaries sorted: 2-D edges (min, max)
        (Mesh2DBoundaries.cxx:885-910); 3-D faces rotated to their lowest
        vertex, reversed when the next vertex exceeds the previous one
        (Mesh3DBoundaries.cxx:650-700);
      - edges numbered as first met on the referenced boundaries (local edges
        (v_k, v_k+1)), then on the elements (square (0,1) (1,2) (2,3) (3,0);
        cube: Quadrature/HexahedronQuadrature.cxx:41-47) (Mesh2D.cxx:2435-2454,
        Mesh3D.cxx:5201-5300,
        MeshBase.cxx:844-881); faces (3-D) from the referenced faces, then the
        elements' local faces (NumberMesh.cxx:2609) rotated to their lowest
        vertex (Mesh3D.cxx:4928-5170);
      - DOFs: vertices, then r-1 per edge from its lower to its higher vertex,
        then (r-1)^2 per face, index (s-1)(r-1) + (t-1) with s along
        v0 -> v1 and t along v0 -> v3 of the stored face, then (r-1)^d per
        element, interior nodes with the first index outermost
        (NumberMesh.cxx:695-900, face table NumberMesh.cxx:2609)."""
    n = r + 1
    nv = int(max(max(c) for c in corners)) + 1
    edges, eorder = {}, []

    def add_edge(a, b):
        k = (min(a, b), max(a, b))
        if k not in edges:
            edges[k] = len(eorder)
            eorder.append(k)
    if dim == 2:
        for a, b in boundary:
            add_edge(a, b)
        for c in corners:
            for a, b in ((0, 1), (1, 2), (2, 3), (3, 0)):
                add_edge(c[a], c[b])
        faces, nf = {}, 0
    else:
        def canon(f):
            f = list(f)
            k = int(np.argmin(f))
            g = f[k:] + f[:k]
            return [g[0], g[3], g[2], g[1]] if g[1] > g[3] else g
        ref = [canon(f) for f in boundary]
        for f in ref:
            for t in range(4):
                add_edge(f[t], f[(t + 1) % 4])
        cube_edges = [(0, 1), (1, 2), (3, 2), (0, 3), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5),
                      (5, 6), (7, 6), (4, 7)]
        for c in corners:
            for a, b in cube_edges:
                add_edge(c[a], c[b])
        faces = {}
        for f in ref:
            faces.setdefault(frozenset(f), (len(faces), f))
        cube_faces = [(0, 3, 7, 4), (0, 1, 5, 4), (0, 1, 2, 3), (4, 5, 6, 7), (3, 2, 6, 7),
                      (1, 2, 6, 5)]
        for c in corners:
            for fl in cube_faces:
                f = [c[x] for x in fl]
                k = int(np.argmin(f))
                f = f[k:] + f[:k]
                faces.setdefault(frozenset(f), (len(faces), f))
        nf = len(faces)
    ne = len(eorder)
    off_e, off_f = nv, nv + ne * (r - 1)
    off_i = off_f + nf * (r - 1) ** 2
    cc = ([(0, 0), (r, 0), (r, r), (0, r)] if dim == 2 else
          [(0, 0, 0), (r, 0, 0), (r, r, 0), (0, r, 0), (0, 0, r), (r, 0, r), (r, r, r), (0, r, r)])
    dof = -np.ones(n_nodes, np.int64)
    for e, c in enumerate(corners):
        ranges = [range(n)] * dim
        for idx in itertools.product(*ranges):
            q = sum(idx[a] * n ** a for a in range(dim))
            bnd = [x in (0, r) for x in idx]
            nb = sum(bnd)
            if nb == dim:
                d = c[cc.index(idx)]
            elif nb == dim - 1:
                ax = bnd.index(False)
                s0 = list(idx); s0[ax] = 0
                s1 = list(idx); s1[ax] = r
                a, b, t = c[cc.index(tuple(s0))], c[cc.index(tuple(s1))], idx[ax]
                if a > b:
                    a, b, t = b, a, r - t
                d = off_e + edges[(a, b)] * (r - 1) + t - 1
            elif dim == 3 and nb == 1:
                ax = bnd.index(True)
                pos = {c[v]: cc[v] for v in range(8) if cc[v][ax] == idx[ax]}
                fnum, f = faces[frozenset(pos)]
                p0 = np.array(pos[f[0]])
                s = int((np.array(idx) - p0) @ ((np.array(pos[f[1]]) - p0) // r))
                t = int((np.array(idx) - p0) @ ((np.array(pos[f[3]]) - p0) // r))
                d = off_f + fnum * (r - 1) ** 2 + (s - 1) * (r - 1) + t - 1
            else:
                li = 0
                for a in range(dim):
                    li = li * (r - 1) + idx[a] - 1
                d = off_i + e * (r - 1) ** dim + li
            dof[elem_nodes[e][q]] = d
    return dof


def regular_boundary(dim: int, nx: int):
    """Boundary edges (2-D) or faces (3-D) of FileMesh = REGULAR on an nx^dim
    grid of elements, 0-based vertex ids (vertex a + (nx+1)(b + (nx+1)c)), in
    the order the reference creates them: 2-D y = ymin and y = ymax edge pairs,
    then x = xmax and x = xmin (Mesh2D.cxx:4856-4873); 3-D x = xmin / x = xmax
    and y = ymin / y = ymax pairs row by row, then z = zmin / z = zmax
    (Mesh3D.cxx:8700-8840)."""
    v1 = nx + 1
    vid = lambda a, b, c=0: a + v1 * (b + v1 * c)  # noqa: E731
    out = []
    if dim == 2:
        for i in range(nx):
            out += [[vid(i, 0), vid(i + 1, 0)], [vid(i, nx), vid(i + 1, nx)]]
        for j in range(nx):
            out += [[vid(nx, j), vid(nx, j + 1)], [vid(0, j), vid(0, j + 1)]]
        return out
    for b in range(nx):
        for a in range(nx):
            out += [[vid(0, a, b), vid(0, a + 1, b), vid(0, a + 1, b + 1), vid(0, a, b + 1)],
                    [vid(nx, a, b), vid(nx, a + 1, b), vid(nx, a + 1, b + 1), vid(nx, a, b + 1)],
                    [vid(a, 0, b), vid(a + 1, 0, b), vid(a + 1, 0, b + 1), vid(a, 0, b + 1)],
                    [vid(a, nx, b), vid(a + 1, nx, b), vid(a + 1, nx, b + 1), vid(a, nx, b + 1)]]
    for b in range(nx):
        for a in range(nx):
            out += [[vid(a, b, 0), vid(a + 1, b, 0), vid(a + 1, b + 1, 0), vid(a, b + 1, 0)],
                    [vid(a, b, nx), vid(a + 1, b, nx), vid(a + 1, b + 1, nx), vid(a, b + 1, nx)]]
    return out


def reprise_saves(nsteps: int, every: int):
    """Checkpoints of SaveReprise = every over a run of nsteps steps: (nt, b)
    for each nt = 0, every, 2 every, ... <= nsteps (GiveIterate and
    GiveFinalIterate both reach SaveDataReprise), b alternating 0, 1, ...
    (VarInstationary.cxx:917-972).  Each writes SolAcous<b> and
    IterationAcousSaveNew = [nt, -1, b] after moving the previous New file to
    IterationAcousSave."""
    return [(nt, k % 2) for k, nt in enumerate(range(0, nsteps + 1, every))]


def quadrature_points(m: "Mesh"):
    """(E, nd, dim) positions of the element-local GL nodes (the quadrature
    points of the Lobatto elements)."""
    return m.xyz[m.elem_nodes]


def circle_points(cx, cy, radius, npts):
    p = np.zeros(2 * npts)
    lib().or_circle_points(cx, cy, radius, npts, p)
    return p.reshape(npts, 2)


def uniform(seed: int, idx: int) -> float:
    return lib().or_uniform(seed, idx)


# --------------------------------------------------------------------------
# Structured problems
# --------------------------------------------------------------------------

@dataclass
class Mesh:
    dim: int
    nx: int
    r: int
    pert: float
    seed: int
    vert: np.ndarray = field(repr=False)
    xyz: np.ndarray = field(repr=False)
    elem_nodes: np.ndarray = field(repr=False)
    dirichlet: np.ndarray = field(repr=False)

    @property
    def n(self):
        return self.r + 1

    @property
    def nd(self):
        return self.n ** self.dim

    @property
    def E(self):
        return self.nx ** self.dim

    @property
    def N(self):
        return (self.nx * self.r + 1) ** self.dim


def mesh(dim: int, nx: int, r: int, pert: float = 0.0, seed: int = 0) -> Mesh:
    L = lib()
    nv = (nx + 1) ** dim
    vert = np.zeros(nv * dim)
    L.or_mesh_vertices(dim, nx, pert, seed, vert)
    N = (nx * r + 1) ** dim
    E = nx ** dim
    nd = (r + 1) ** dim
    xyz = np.zeros(N * dim)
    en = np.zeros(E * nd, dtype=np.int32)
    dr = np.zeros(N, dtype=np.uint8)
    if L.or_mesh_nodes(dim, nx, r, vert, xyz, en, dr):
        raise ValueError("mesh")
    return Mesh(dim, nx, r, pert, seed, vert, xyz.reshape(N, dim), en.reshape(E, nd), dr)


def layered(m: Mesh, values) -> np.ndarray:
    """Per-element value of horizontal layers: layer = ez*len(values)//nx
    (y in 2-D, z in 3-D)."""
    E = m.E
    e = np.arange(E)
    ez = e // (m.nx ** (m.dim - 1))
    lay = (ez * len(values)) // m.nx
    return np.asarray(values, dtype=np.float64)[lay]


def _media(m: Mesh, v):
    """Per-element (scalar or (E,)) or per element-local node ((E, nd))."""
    v = np.asarray(v, np.float64)
    if v.ndim == 2:
        return np.ascontiguousarray(v.reshape(m.E * m.nd)), 1
    return np.ascontiguousarray(np.broadcast_to(v, (m.E,))), 0


def geometry(m: Mesh, mu, rho, elastic: bool = False):
    """Returns (G, Mloc) for the scalar case or (jw, dfinv) for elasticity.
    mu, rho: per element, or (E, nd) arrays of values at the quadrature points
    (both of the same kind)."""
    L = lib()
    mu, nod_m = _media(m, mu)
    rho, nod_r = _media(m, rho)
    if nod_m != nod_r:
        raise ValueError("mu and rho must both be per element or both per node")
    nc = 3 if m.dim == 2 else 6
    Mloc = np.zeros(m.E * m.nd)
    if not elastic:
        G = np.zeros(m.E * m.nd * nc)
        jmin = L.or_geometry(m.dim, m.nx, m.r, m.vert, mu, rho, nod_m, _ptr(G), _ptr(Mloc), None,
                             None)
        if jmin <= 0:
            raise ValueError("degenerate element")
        return G, Mloc
    jw = np.zeros(m.E * m.nd)
    dfi = np.zeros(m.E * m.nd * m.dim * m.dim)
    jmin = L.or_geometry(m.dim, m.nx, m.r, m.vert, mu, rho, nod_m, None, _ptr(Mloc), _ptr(jw),
                         _ptr(dfi))
    if jmin <= 0:
        raise ValueError("degenerate element")
    return jw, dfi, Mloc


def mass(m: Mesh, Mloc, dirichlet=True):
    M = np.zeros(m.N)
    Mi = np.zeros(m.N)
    dr = m.dirichlet if dirichlet else np.zeros(m.N, np.uint8)
    lib().or_mass(m.E, m.nd, m.N, m.elem_nodes.reshape(-1), Mloc, dr, M, Mi)
    return M, Mi


def color(m: Mesh):
    c = np.zeros(m.E, dtype=np.int32)
    nc = lib().or_color(m.E, m.nd, m.N, m.elem_nodes.reshape(-1), c)
    return c, nc


def stiffness(m: Mesh, G, u, colors=None, alpha=1.0):
    y = np.zeros(m.N)
    if colors is None:
        lib().or_stiffness(m.dim, m.r, m.E, m.N, m.elem_nodes.reshape(-1), G, None, 0, alpha,
                           np.ascontiguousarray(u, dtype=np.float64), 0.0, y)
    else:
        c, nc = colors
        lib().or_stiffness(m.dim, m.r, m.E, m.N, m.elem_nodes.reshape(-1), G, _ptr(c), nc,
                           alpha, np.ascontiguousarray(u, dtype=np.float64), 0.0, y)
    return y


def stiffness_elastic(m: Mesh, jw, dfi, lam, mu, u, colors, alpha=1.0):
    c, nc = colors
    y = np.zeros(m.dim * m.N)
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float64), (m.E,)))
    mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, np.float64), (m.E,)))
    lib().or_stiffness_elastic(m.dim, m.r, m.E, m.N, m.elem_nodes.reshape(-1), jw, dfi, lam, mu,
                               c, nc,
                               alpha, np.ascontiguousarray(u, dtype=np.float64), 0.0, y)
    return y


def dirac(m: Mesh, p, pol=(1.0,)):
    pol = np.asarray(pol, np.float64)
    b = np.zeros(m.N * len(pol))
    cnt = lib().or_dirac(m.dim, m.nx, m.r, m.vert, m.elem_nodes.reshape(-1), m.N, m.dirichlet,
                         np.asarray(p, np.float64), len(pol), pol, b)
    if cnt == 0:
        raise ValueError("source point outside the mesh")
    return b


def receiver_gradients(m: Mesh, pts):
    """(elements, value weights [npts, nd], gradient weights [npts, dim, nd]):
    u(p) = w . u_e and d u / d x_a (p) = g[a] . u_e on the first element
    containing each point (-1 outside)."""
    pts = np.asarray(pts, float).reshape(-1, m.dim)
    el = np.full(len(pts), -1, np.int64)
    W = np.zeros((len(pts), m.nd))
    Gw = np.zeros((len(pts), m.dim, m.nd))
    for k, p in enumerate(pts):
        w = np.zeros(m.nd)
        g = np.zeros(m.dim * m.nd)
        el[k] = lib().or_receiver_grad(m.dim, m.nx, m.r, m.vert, np.ascontiguousarray(p), w, g)
        W[k], Gw[k] = w, g.reshape(m.dim, m.nd)
    return el, W, Gw


def receivers(m: Mesh, pts):
    pts = np.asarray(pts, np.float64).reshape(-1, m.dim)
    elems = np.zeros(len(pts), dtype=np.int64)
    W = np.zeros((len(pts), m.nd))
    for i, p in enumerate(pts):
        w = np.zeros(m.nd)
        e = lib().or_receiver(m.dim, m.nx, m.r, m.vert, np.ascontiguousarray(p), w)
        if e < 0:
            raise ValueError(f"receiver {i} outside the mesh")
        elems[i] = e
        W[i] = w
    return elems, W


def plane_points(dim: int, plane):
    """SismoPlane grid points, index j*Nx + i.  2-D plane = (xmin, xmax, ymin,
    ymax, Nx, Ny): (xmin + i hx, ymin + j hy) (Output/GridInterpolation.cxx:
    2297-2345).  3-D plane = (A, B, C, Nx, Ny): A + i/(Nx-1) (B - A) +
    j/(Ny-1) (C - A) (GridInterpolation.cxx:5119-5166)."""
    plane = [float(v) for v in plane]
    nx, ny = int(plane[-2]), int(plane[-1])
    out = np.zeros((ny, nx, dim))
    if dim == 2:
        x0, x1, y0, y1 = plane[:4]
        hx = (x1 - x0) / (nx - 1) if nx > 1 else 0.0
        hy = (y1 - y0) / (ny - 1) if ny > 1 else 0.0
        for i in range(nx):
            for j in range(ny):
                out[j, i] = (x0 + i * hx, y0 + j * hy)
    else:
        a, b, c = (np.array(plane[3 * k:3 * k + 3]) for k in range(3))
        sx = 1.0 / (nx - 1) if nx > 1 else 0.0
        sy = 1.0 / (ny - 1) if ny > 1 else 0.0
        for i in range(nx):
            for j in range(ny):
                out[j, i] = a + i * sx * (b - a) + j * sy * (c - a)
    return out.reshape(-1, dim)


def line_points(dim: int, line):
    """SismoLine points: N points from (xmin, ymin[, zmin]) to (xmax, ymax[,
    zmax]); line = (xmin, xmax, ymin, ymax, [zmin, zmax,] N)
    (Output/GridInterpolation.cxx:1926-1943, :4782-4800)."""
    line = [float(v) for v in line]
    n = int(line[-1])
    out = np.zeros((n, dim))
    for a in range(dim):
        lo, hi = line[2 * a], line[2 * a + 1]
        h = (hi - lo) / (n - 1) if n > 1 else 0.0
        out[:, a] = lo + np.arange(n) * h
    return out


def snapshot(m: Mesh, u, pts):
    """u interpolated at pts (first element holding the point); 0 outside the
    mesh, as Montjoie leaves snapshot points it cannot place."""
    pts = np.asarray(pts, np.float64).reshape(-1, m.dim)
    out = np.zeros(len(pts))
    w = np.zeros(m.nd)
    for i, p in enumerate(pts):
        e = lib().or_receiver(m.dim, m.nx, m.r, m.vert, np.ascontiguousarray(p), w)
        if e >= 0:
            out[i] = w @ u[m.elem_nodes[e]]
    return out


def snapshot_elastic(m: Mesh, u, pts, lam_e, mu_e, stress: bool = True):
    """Elastic snapshot fields at pts: [d, npts] displacement and [d*d, npts]
    stress sigma_ij = lambda tr(grad u) delta_ij + mu (du_i/dx_j + du_j/dx_i)
    (row-major, k = d i + j) with the medium of the first element holding each
    point, or with stress = False (DisplayStress = NO, VarElastic.cxx:257-258)
    d u_i / d x_j itself; 0 outside the mesh."""
    d = m.dim
    el, W, Gw = receiver_gradients(m, pts)
    npts = len(el)
    disp = np.zeros((d, npts))
    sig = np.zeros((d * d, npts))
    en = m.elem_nodes.reshape(m.E, -1)
    for k in range(npts):
        if el[k] < 0:
            continue
        ue = np.stack([u[c * m.N + en[el[k]]] for c in range(d)])
        disp[:, k] = ue @ W[k]
        g = ue @ Gw[k].T  # g[i, a] = d u_i / d x_a
        if not stress:
            sig[:, k] = g.ravel()
            continue
        s = lam_e[el[k]] * np.trace(g) * np.eye(d) + mu_e[el[k]] * (g + g.T)
        sig[:, k] = s.ravel()
    return disp, sig


@dataclass
class Problem:
    """A structured acoustic (ncomp=1) or elastic (ncomp = dim) problem."""
    mesh: Mesh
    ncomp: int
    G: np.ndarray = None
    jw: np.ndarray = None
    dfi: np.ndarray = None
    lam: np.ndarray = None
    mu: np.ndarray = None
    rho: np.ndarray = None
    M: np.ndarray = None
    Minv: np.ndarray = None
    colors: tuple = None


def acoustic_problem(m: Mesh, mu_e, rho_e) -> Problem:
    """mu_e, rho_e per element, or both (E, nd) at the quadrature points."""
    G, Mloc = geometry(m, mu_e, rho_e)
    M, Mi = mass(m, Mloc)
    shp = (m.E, m.nd) if np.ndim(mu_e) == 2 else (m.E,)
    return Problem(m, 1, G=G, mu=np.broadcast_to(np.asarray(mu_e, float), shp).copy(),
                   rho=np.broadcast_to(np.asarray(rho_e, float), shp).copy(), M=M, Minv=Mi,
                   colors=color(m))


def elastic_problem(m: Mesh, lam_e, mu_e, rho_e) -> Problem:
    """lam_e, mu_e per element; rho_e per element or (E, nd) at the quadrature
    points (the mass rho J w only)."""
    rho_a = np.asarray(rho_e, float)
    if rho_a.ndim == 2:
        jw, dfi, Mloc = geometry(m, np.ones_like(rho_a), rho_a, elastic=True)
        M, Mi = mass(m, Mloc)
        return Problem(m, m.dim, jw=jw, dfi=dfi,
                       lam=np.ascontiguousarray(np.broadcast_to(np.asarray(lam_e, float), (m.E,))),
                       mu=np.ascontiguousarray(np.broadcast_to(np.asarray(mu_e, float), (m.E,))),
                       rho=rho_a.copy(), M=M, Minv=Mi, colors=color(m))
    jw, dfi, Mloc = geometry(m, mu_e, rho_e, elastic=True)
    M, Mi = mass(m, Mloc)
    return Problem(m, m.dim, jw=jw, dfi=dfi,
                   lam=np.ascontiguousarray(np.broadcast_to(np.asarray(lam_e, float), (m.E,))),
                   mu=np.ascontiguousarray(np.broadcast_to(np.asarray(mu_e, float), (m.E,))),
                   rho=np.broadcast_to(np.asarray(rho_e, float), (m.E,)).copy(), M=M, Minv=Mi,
                   colors=color(m))


def apply(P: Problem, u):
    if P.ncomp == 1:
        return stiffness(P.mesh, P.G, u, P.colors)
    return stiffness_elastic(P.mesh, P.jw, P.dfi, P.lam, P.mu, u, P.colors)


def lambda_max(P: Problem, m_iter: int = 120, seed: int = 7) -> float:
    m = P.mesh
    c, nc = P.colors
    return lib().or_lambda_max(m.dim, m.r, P.ncomp, m.E, m.N, m.elem_nodes.reshape(-1), _ptr(P.G),
                               _ptr(P.jw), _ptr(P.dfi), _ptr(P.lam), _ptr(P.mu), c, nc, P.M,
                               m.dirichlet, m_iter, seed)


def leapfrog(P: Problem, b, pulse, dt, nsteps, rec_pts=None, rec_comp=None, u0=None, du0=None,
             rec_weights=None, S=None):
    """Returns (U^N, U^{N+1}, traces[(N+1), nrec]).  rec_weights = (elements,
    weights [nrec, nd]) replaces the value weights of rec_pts (e.g. gradient
    weights from receiver_gradients)."""
    m = P.mesh
    c, nc = P.colors
    NT = m.N * P.ncomp
    if rec_weights is not None:
        re = np.ascontiguousarray(rec_weights[0], np.int64)
        rw = np.ascontiguousarray(rec_weights[1], np.float64)
    elif rec_pts is not None and len(rec_pts):
        re, rw = receivers(m, rec_pts)
    else:
        re, rw = np.zeros(0, np.int64), np.zeros(0)
    nrec = len(re)
    rc = None if rec_comp is None else np.ascontiguousarray(rec_comp, dtype=np.int32)
    uN = np.zeros(NT)
    uN1 = np.zeros(NT)
    tr = np.zeros(((nsteps + 1) * max(nrec, 1)))
    pulse = None if pulse is None else np.ascontiguousarray(pulse, np.float64)
    lib().or_leapfrog(m.dim, m.r, P.ncomp, m.E, m.N, m.elem_nodes.reshape(-1), _ptr(P.G),
                      _ptr(P.jw), _ptr(P.dfi), _ptr(P.lam), _ptr(P.mu), c, nc, P.M,
                      _ptr(None if S is None else np.ascontiguousarray(S, np.float64)),
                      m.dirichlet, _ptr(b), _ptr(pulse), dt, nsteps, _ptr(u0), _ptr(du0), nrec,
                      np.ascontiguousarray(re), np.ascontiguousarray(rw.reshape(-1)), _ptr(rc),
                      uN, uN1, tr)
    return uN, uN1, tr[: (nsteps + 1) * nrec].reshape(nsteps + 1, nrec)


def leapfrog4(P: Problem, b, tables, dt, nsteps, rec_pts=None, rec_comp=None, u0=None, du0=None,
              S=None, rec_weights=None):
    """Fourth-order modified-equation leapfrog (or_leapfrog4): tables = (f, f',
    f'') at t_n + dt (any may be None).  Returns (U^N, U^{N+1}, traces)."""
    m = P.mesh
    c, nc = P.colors
    NT = m.N * P.ncomp
    if rec_weights is not None:
        re = np.ascontiguousarray(rec_weights[0], np.int64)
        rw = np.ascontiguousarray(rec_weights[1], np.float64)
    elif rec_pts is not None and len(rec_pts):
        re, rw = receivers(m, rec_pts)
    else:
        re, rw = np.zeros(0, np.int64), np.zeros(0)
    nrec = len(re)
    rc = None if rec_comp is None else np.ascontiguousarray(rec_comp, dtype=np.int32)
    uN, uN1 = np.zeros(NT), np.zeros(NT)
    tr = np.zeros(((nsteps + 1) * max(nrec, 1)))
    tabs = [None if t is None else np.ascontiguousarray(t, np.float64) for t in tables]
    lib().or_leapfrog4(m.dim, m.r, P.ncomp, m.E, m.N, m.elem_nodes.reshape(-1), _ptr(P.G),
                       _ptr(P.jw), _ptr(P.dfi), _ptr(P.lam), _ptr(P.mu), c, nc, P.M,
                       _ptr(None if S is None else np.ascontiguousarray(S, np.float64)),
                       m.dirichlet, _ptr(b), _ptr(tabs[0]), _ptr(tabs[1]), _ptr(tabs[2]), dt,
                       nsteps, _ptr(u0), _ptr(du0), nrec, np.ascontiguousarray(re),
                       np.ascontiguousarray(rw.reshape(-1)), _ptr(rc), uN, uN1, tr)
    return uN, uN1, tr[: (nsteps + 1) * nrec].reshape(nsteps + 1, nrec)


def damping(m: Mesh, sigma_e):
    """Damping diagonal S = sum_e sigma J w phi_i (lumped like the mass,
    VarHelmholtz.cxx:64-66,108-109), from the oracle's geometry with sigma in
    the place of rho; sigma per element or [E, nd] at the quadrature points."""
    sg = np.asarray(sigma_e, float)
    _, Mloc = geometry(m, np.ones(sg.shape), sg)
    return mass(m, Mloc, dirichlet=False)[0]


def adjoint_gradient(P: Problem, b, pulse, dt, nsteps, rec_pts, dJ):
    """FP64 reference gradient dJ/dmu_e (zero initial state), dJ[(nsteps+1), nrec]."""
    m = P.mesh
    c, nc = P.colors
    re, rw = receivers(m, rec_pts)
    grad = np.zeros(m.E)
    lib().or_adjoint_acoustic(m.dim, m.r, m.E, m.N, m.elem_nodes.reshape(-1), P.G,
                              np.ascontiguousarray(P.mu, np.float64), c, nc, P.M, m.dirichlet,
                              _ptr(b), _ptr(np.ascontiguousarray(pulse, np.float64)), dt, nsteps,
                              len(re), np.ascontiguousarray(re), np.ascontiguousarray(rw.reshape(-1)),
                              np.ascontiguousarray(dJ, np.float64).reshape(-1), grad)
    return grad


def forward_traces(P: Problem, b, pulse, dt, nsteps, rec_pts):
    """Traces rows 0..nsteps of the zero-initial-state leapfrog."""
    _, _, tr = leapfrog(P, b, pulse, dt, nsteps, rec_pts)
    return tr


# --------------------------------------------------------------------------
# PML layers and the first-order split scheme (LEAP_FROG_PML)
# --------------------------------------------------------------------------

def time_table(f0: float, t_begin: float, dt: float, nsteps: int, nb_deriv: int,
               offset: float) -> np.ndarray:
    """f^(nb_deriv)(t_begin + n dt + offset), n < nsteps, of the GL-24
    interpolated Ricker primitive (or_time_table)."""
    out = np.zeros(nsteps)
    lib().or_time_table(f0, t_begin, dt, nsteps, nb_deriv, offset, out)
    return out


@dataclass
class PmlMesh:
    """A structured mesh of [0,1]^dim with PML layers on every side, and the
    PML data of its elements."""
    mesh: Mesh
    nx: int                   # elements per direction of the physical mesh
    layers: int
    thickness: float
    pml: np.ndarray           # (E,) uint8: element inside the layers
    kind: np.ndarray          # (E, dim) int: -1 / +1 layer below / above the box, 0 none
    phys: np.ndarray          # (E,) physical element whose medium the element takes


def pml_mesh(dim: int, nx: int, r: int, pert: float, seed: int, thickness: float,
             layers: int, sides=None) -> PmlMesh:
    """The structured mesh of [0,1]^dim (vertices perturbed like mesh())
    extended by `layers` layers of elements of width thickness/layers beyond
    the box faces of `sides` (per axis (below, above), default all), the way
    AddPML = YES <axes> thickness layers extrudes a box mesh
    (PmlRegionParameter::AddPMLElements, Mesh/Mesh2DBoundaries.cxx:104-192,
    Mesh3DBoundaries.cxx:59-; Mesh<>::ExtrudeCoordinate, Mesh/Mesh2D.cxx:
    4898-4990: new vertex coordinate pos + (i+1) delta / layers, the other
    coordinates copied from the boundary vertex; axes in the order x, y, z).
    A PML element takes the medium of the physical element it was extruded
    from (its indices clamped into the physical block) and the layer kind of
    its centre relative to the box [0,1]^dim."""
    L = lib()
    if sides is None:
        sides = [(True, True)] * dim
    lo = [layers if sides[a][0] else 0 for a in range(dim)]
    hi = [layers if sides[a][1] else 0 for a in range(dim)]
    nv1 = nx + 1
    vphys = np.zeros(nv1 ** dim * dim)
    L.or_mesh_vertices(dim, nx, pert, seed, vphys)
    vphys = vphys.reshape(-1, dim)
    nt = [nx + lo[a] + hi[a] for a in range(dim)]
    nt1 = [n + 1 for n in nt]
    if len(set(nt)) != 1:
        raise ValueError("pml_mesh: the extended grid must stay cubic (or_mesh_nodes)")
    ntc, ntc1 = nt[0], nt1[0]

    def coord(a, i):
        # layer coordinates of axis a, or None inside the physical range
        if i < lo[a]:
            return 0.0 + ((lo[a] - i) * (-thickness)) / layers
        if i > lo[a] + nx:
            return 1.0 + ((i - lo[a] - nx) * thickness) / layers
        return None

    vert = np.zeros((ntc1 ** dim, dim))
    for v in range(ntc1 ** dim):
        I = [(v // ntc1 ** a) % ntc1 for a in range(dim)]
        inside = all(lo[a] <= I[a] <= lo[a] + nx for a in range(dim))
        if inside:
            pv = sum((I[a] - lo[a]) * nv1 ** a for a in range(dim))
            vert[v] = vphys[pv]
            continue
        for a in range(dim):
            c = coord(a, I[a])
            if c is None:
                # copy of a boundary vertex of the box: regular coordinate
                ia = I[a] - lo[a]
                c = 1.0 if ia == nx else ia * (1.0 / nx)
            vert[v, a] = c
    vert = np.ascontiguousarray(vert.reshape(-1))
    N = (ntc * r + 1) ** dim
    E = ntc ** dim
    nd = (r + 1) ** dim
    xyz = np.zeros(N * dim)
    en = np.zeros(E * nd, dtype=np.int32)
    dr = np.zeros(N, dtype=np.uint8)
    if L.or_mesh_nodes(dim, ntc, r, vert, xyz, en, dr):
        raise ValueError("mesh")
    m = Mesh(dim, ntc, r, pert, seed, vert, xyz.reshape(N, dim), en.reshape(E, nd), dr)
    e = np.arange(E)
    idx = np.stack([(e // ntc ** a) % ntc for a in range(dim)], axis=1)
    lo_a, hi_a = np.array(lo), np.array(lo) + nx
    kind = np.where(idx < lo_a, -1, np.where(idx >= hi_a, 1, 0))
    cl = np.clip(idx - lo_a, 0, nx - 1)
    phys = sum(cl[:, a] * nx ** a for a in range(dim))
    pml = (kind != 0).any(axis=1).astype(np.uint8)
    return PmlMesh(m, nx, layers, thickness, pml, kind, phys)


PML_PARABOLE, PML_CONSTANT, PML_LINEAR, PML_SHIFTED_PARABOLE = 0, 1, 2, 3


def pml_tau(pm: PmlMesh, vsigma: float, vmax: float, profile: int = PML_PARABOLE,
            offset: float = 0.0, transverse: float = 0.0) -> np.ndarray:
    """Damping tau[e, q, a] at the quadrature points (GL nodes) of every
    element: the DampingPML profile of the distance dx = x_a - box_a beyond
    the box face of the element's kind, dsig = log(1000) 3 vsigma vmax /
    (2 a^3), a = thickness (VarBoundaryCondition_Base::GetDampingTauPML and
    EvaluateFunctionTauPML, Harmonic/BoundaryConditionHarmonic.cxx:1548-1570,
    1639-1760): parabola dsig dx^2, constant dsig a^2, linear dsig |dx| a,
    shifted parabola dsig (dx^2 + offset a^2); TransverseDampingPML = c mixes
    the axes (tau_a = max(tau_a, c max of the others)); the first-order
    formulation keeps tau itself (GetDampingFactorPML :1588-1636); vmax =
    max sqrt(mu / rho) over the media of the PML elements (:1486-1522)."""
    m = pm.mesh
    a = pm.thickness
    dsig = np.log(1000e0) * 3e0 * vsigma * vmax / (2 * a ** 3)

    def f(dx):
        if profile == PML_CONSTANT:
            return np.full_like(dx, dsig * a * a)
        if profile == PML_LINEAR:
            return dsig * np.abs(dx) * a
        if profile == PML_SHIFTED_PARABOLE:
            return dsig * (dx * dx + offset * a * a)
        return dsig * dx * dx

    X = m.xyz[m.elem_nodes]                       # (E, nd, dim)
    tau = np.zeros(X.shape)
    for ax in range(m.dim):
        k = pm.kind[:, ax][:, None]
        tau[:, :, ax] = np.where(k == -1, f(X[:, :, ax] - 0.0),
                                 np.where(k == 1, f(X[:, :, ax] - 1.0), 0.0))
    if transverse != 0:
        c = transverse
        z = tau.copy()
        for ax in range(m.dim):
            others = np.max(np.delete(z, ax, axis=2), axis=2)
            tau[:, :, ax] = np.maximum(z[:, :, ax], c * others)
    return tau


def pml_sides(axes: str, dim: int):
    """Sides per axis of AddPML's axis string (PmlRegionParameter_Base::
    SetParameters, Mesh/MeshBoundaries.cxx:44-70): 'X-' below only, 'X+'
    above only, 'X' both."""
    out = []
    for a in range(dim):
        c = "XYZ"[a]
        if c not in axes:
            out.append((False, False))
        elif c + "-" in axes:
            out.append((True, False))
        elif c + "+" in axes:
            out.append((False, True))
        else:
            out.append((True, True))
    return out


def leapfrog_pml(pm: PmlMesh, rho_e, mu_e, tau, b, table, dt, nsteps, rec_pts, sigma_e=None):
    """First-order split scheme with PML (or_leapfrog_pml): returns
    (u^nsteps, traces[nsteps+1, nrec]); rho_e, mu_e per element of the
    extended mesh, table = time_table(..., 0, dt/2)."""
    m = pm.mesh
    re, rw = receivers(m, rec_pts)
    uN = np.zeros(m.N)
    tr = np.zeros((nsteps + 1) * max(len(re), 1))
    sg = None if sigma_e is None else np.ascontiguousarray(sigma_e, np.float64)
    lib().or_leapfrog_pml(m.dim, m.nx, m.r, m.vert, m.N, m.elem_nodes.reshape(-1),
                          np.ascontiguousarray(rho_e, np.float64),
                          np.ascontiguousarray(mu_e, np.float64), _ptr(sg), pm.pml,
                          np.ascontiguousarray(tau, np.float64).reshape(-1), m.dirichlet,
                          _ptr(np.ascontiguousarray(b, np.float64)),
                          np.ascontiguousarray(table, np.float64), dt, nsteps, len(re),
                          np.ascontiguousarray(re), np.ascontiguousarray(rw.reshape(-1)), uN, tr)
    return uN, tr[: (nsteps + 1) * len(re)].reshape(nsteps + 1, len(re))
