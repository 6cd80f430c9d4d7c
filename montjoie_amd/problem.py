"""Host-side mirror of Montjoie's explicit wave-solver interface on top of the
C ABI (include/mjwave.h).

``WaveProblem`` plays the role of ``HyperbolicProblem<AcousticEquation<Dim>>``
/ ``HyperbolicProblem<TimeElasticEquation<Dimension3>>`` for the explicit
path (Instationary/VarInstationary.cxx:1629-1831 RunAll): it owns the mesh,
media, Dirichlet nodes, Dirac sources with a Ricker pulse, receivers and the
leapfrog state, and exposes the VirtualOdeSystem operators
(Instationary/TimeSchemes.hxx:35-151) and the leapfrog scheme
(ModifiedEquationIterator, Instationary/ModifiedEquation.hxx:78-114).

Every computation runs in the HIP library; this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr, u8ptr


@dataclass
class Mesh:
    """Q_r mesh: global GL node positions and element -> node map."""
    dim: int
    order: int
    xyz: np.ndarray          # (N, dim) float64
    elem_nodes: np.ndarray   # (E, (order+1)^dim) int32, local q = i + n j (+ n^2 k)
    dirichlet: np.ndarray    # (N,) uint8
    nx: int = 0              # elements per direction for regular meshes

    @property
    def N(self) -> int:
        return self.xyz.shape[0]

    @property
    def E(self) -> int:
        return self.elem_nodes.shape[0]

    @property
    def nd(self) -> int:
        return (self.order + 1) ** self.dim


def regular_mesh(dim: int, nx: int, order: int, pert: float = 0.0, seed: int = 0) -> Mesh:
    """REGULAR mesh of [0,1]^dim (Mesh/Mesh3D.cxx:8406-8486) with nx elements
    per direction; interior vertices displaced by <= pert*h (seeded)."""
    L = _lib.lib()
    N = C.c_int64()
    E = C.c_int64()
    check(L.mjw_mesh_regular_size(dim, nx, order, C.byref(N), C.byref(E)), None, "mesh size")
    nd = (order + 1) ** dim
    xyz = np.zeros((N.value, dim))
    en = np.zeros((E.value, nd), dtype=np.int32)
    dr = np.zeros(N.value, dtype=np.uint8)
    check(L.mjw_mesh_regular(dim, nx, order, float(pert), int(seed), dptr(xyz), iptr(en),
                             u8ptr(dr)), None, "mesh")
    return Mesh(dim, order, xyz, en, dr, nx)


def layered_values(mesh: Mesh, values) -> np.ndarray:
    """Per-element values of horizontal layers of a regular mesh: layer =
    ez*len(values)//nx (ez = element index along the last axis)."""
    e = np.arange(mesh.E)
    ez = e // (mesh.nx ** (mesh.dim - 1))
    return np.asarray(values, np.float64)[(ez * len(values)) // mesh.nx]


def ricker_pulse_table(f0: float, t_begin: float, dt: float, nsteps: int) -> np.ndarray:
    """f_n = pulse(t_begin + n dt + dt) through Montjoie's GL-24 interpolator."""
    out = np.zeros(nsteps)
    check(_lib.lib().mjw_ricker_pulse_table(f0, t_begin, dt, nsteps, dptr(out)), None, "pulse")
    return out


def time_source_table(f0: float, t_begin: float, dt: float, nsteps: int, nb_deriv: int = 1,
                      offset: float | None = None) -> np.ndarray:
    """f^(nb_deriv)(t_begin + n dt + offset) of the Ricker source (offset
    defaults to dt: the second-order leapfrog's table; the first-order PML
    scheme uses nb_deriv = 0, offset = dt/2)."""
    out = np.zeros(nsteps)
    off = dt if offset is None else offset
    check(_lib.lib().mjw_time_source_table(f0, t_begin, dt, nsteps, nb_deriv, off, dptr(out)),
          None, "time_source_table")
    return out


# TemporalSource kinds (MJW_TIME_* of mjwave.h) by Montjoie keyword
TEMPORAL_SOURCES = {"RICKER": 1, "DERIVATIVE_RICKER": 2, "MODIFIED_RICKER": 3, "GAUSSIAN": 4,
                    "HARMONIC": 5, "SINUS_GAUSSIAN": 6}


def temporal_source_table(kind, f0: float, param: float, t_begin: float, dt: float, nsteps: int,
                          nb_deriv: int = 1, offset: float | None = None) -> np.ndarray:
    """f^(nb_deriv)(t_begin + n dt + offset) of a TemporalSource (kind: a
    keyword of TEMPORAL_SOURCES or its number; param: the keyword's value,
    mjw_temporal_source_table); offset defaults to dt, the leapfrog's table."""
    k = TEMPORAL_SOURCES[kind] if isinstance(kind, str) else int(kind)
    out = np.zeros(nsteps)
    off = dt if offset is None else offset
    check(_lib.lib().mjw_temporal_source_table(k, f0, param, t_begin, dt, nsteps, nb_deriv, off,
                                               dptr(out)), None, "temporal_source_table")
    return out


class WaveProblem:
    """One acoustic (ncomp=1) or elastic (ncomp = dimension) problem on one
    device."""

    def __init__(self, mesh: Mesh, rho, mu, lam=None, ncomp: int = 1, nshots: int = 1,
                 precision: str = "fp64", device: int = 0, dirichlet=True):
        self.mesh = mesh
        self.ncomp = ncomp
        self.nshots = nshots
        self.precision = precision
        E = mesh.E
        rho = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, np.float64), (E,)))
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, np.float64), (E,)))
        lamv = None
        if ncomp > 1:
            lamv = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float64), (E,)))
        xyz = np.ascontiguousarray(mesh.xyz, np.float64)
        en = np.ascontiguousarray(mesh.elem_nodes, np.int32)
        dr = np.ascontiguousarray(mesh.dirichlet, np.uint8) if dirichlet is True else (
            None if dirichlet is False or dirichlet is None else
            np.ascontiguousarray(dirichlet, np.uint8))
        self._h = C.c_void_p()
        L = _lib.lib()
        check(L.mjw_create(C.byref(self._h), device, mesh.dim, mesh.order, ncomp, mesh.N,
                           dptr(xyz), E, iptr(en), dptr(rho), dptr(mu), dptr(lamv), u8ptr(dr),
                           nshots, _lib.PRECISIONS[precision]), None, "mjw_create")
        self.dt = None
        self.nrec = 0

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if self._h:
            _lib.lib().mjw_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        check(rc, self._h, what)

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        a = (C.c_int64 * 9)()
        self._chk(_lib.lib().mjw_info(self._h, a), "info")
        k = ["dim", "order", "ncomp", "n_nodes", "n_elem", "ncolors", "nshots", "precision",
             "device_bytes"]
        return dict(zip(k, list(a)))

    def set_mu(self, mu):
        """Replace the per-element coefficient mu in place (mjw_set_mu)."""
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, np.float64), (self.mesh.E,)))
        self._chk(_lib.lib().mjw_set_mu(self._h, dptr(mu)), "set_mu")

    def set_media_nodal(self, rho_q, mu_q=None):
        """rho and mu at every quadrature point, shape (n_elem, (r+1)^dim) in
        the local node order (mjw_set_media_nodal); elastic problems take rho
        only (mu_q None).  Call leapfrog_init afterwards."""
        shape = (self.mesh.E, self.mesh.elem_nodes.shape[1])
        rq = np.ascontiguousarray(np.broadcast_to(np.asarray(rho_q, np.float64), shape))
        mq = None if mu_q is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(mu_q, np.float64), shape))
        self._chk(_lib.lib().mjw_set_media_nodal(self._h, dptr(rq), dptr(mq)), "set_media_nodal")

    def set_geometry(self, mode: str):
        """'trilinear' (per-element vertex map, default when the mesh allows
        it) or 'stored' (per-node tensor read every step)."""
        self._chk(_lib.lib().mjw_set_geometry_mode(self._h, _lib.GEOMETRY_MODES[mode]),
                  "set_geometry")

    @property
    def geometry(self) -> str:
        m = C.c_int32()
        a = C.c_int32()
        self._chk(_lib.lib().mjw_get_geometry_mode(self._h, C.byref(m), C.byref(a)),
                  "get_geometry")
        return "trilinear" if m.value == _lib.MJW_GEOM_TRILINEAR else "stored"

    @property
    def trilinear_available(self) -> bool:
        m = C.c_int32()
        a = C.c_int32()
        self._chk(_lib.lib().mjw_get_geometry_mode(self._h, C.byref(m), C.byref(a)),
                  "get_geometry")
        return bool(a.value)

    def set_tile(self, epb: int):
        self._chk(_lib.lib().mjw_set_tile(self._h, int(epb)), "set_tile")

    def set_xcd_order(self, nxcd: int):
        self._chk(_lib.lib().mjw_set_xcd_order(self._h, int(nxcd)), "set_xcd_order")

    def tile(self) -> int:
        v = C.c_int()
        self._chk(_lib.lib().mjw_get_tile(self._h, C.byref(v)), "get_tile")
        return v.value

    # ------------------------------------------------------------ operators
    @property
    def ndof(self) -> int:
        return self.mesh.N * self.ncomp

    def apply_kh(self, U, alpha=1.0, beta=0.0, Y=None) -> np.ndarray:
        """Y = beta Y + alpha K U, K = -A (ApplyOperatorKh)."""
        U = np.ascontiguousarray(U, np.float64)
        Y = np.zeros(self.ndof) if Y is None else np.ascontiguousarray(Y, np.float64).copy()
        self._chk(_lib.lib().mjw_apply_kh(self._h, alpha, dptr(U), beta, dptr(Y)), "apply_kh")
        return Y

    def apply_kh_dev(self, U_ptr: int, Y_ptr: int, alpha=1.0, beta=0.0):
        self._chk(_lib.lib().mjw_apply_kh_dev(self._h, alpha, C.c_void_p(U_ptr), beta,
                                              C.c_void_p(Y_ptr)), "apply_kh_dev")

    def _vec(self, Y):
        return np.zeros(self.ndof) if Y is None else np.ascontiguousarray(Y, np.float64).copy()

    def apply_dh(self, U, alpha=1.0, beta=0.0, Y=None) -> np.ndarray:
        """Y = beta Y + alpha D U (ApplyOperatorDh)."""
        U, Y = np.ascontiguousarray(U, np.float64), self._vec(Y)
        self._chk(_lib.lib().mjw_apply_dh(self._h, alpha, dptr(U), beta, dptr(Y)), "apply_dh")
        return Y

    def apply_sh(self, U, alpha=1.0, beta=0.0, Y=None) -> np.ndarray:
        """Y = beta Y + alpha S U, S = lumped damping (ApplyOperatorSh)."""
        U, Y = np.ascontiguousarray(U, np.float64), self._vec(Y)
        self._chk(_lib.lib().mjw_apply_sh(self._h, alpha, dptr(U), beta, dptr(Y)), "apply_sh")
        return Y

    def apply_dh_minus_dt_sh(self, dt, U, alpha=1.0, beta=0.0, Y=None) -> np.ndarray:
        """Y = beta Y + alpha (D - dt/2 S) U (ApplyOperatorDhMinusdtSh)."""
        U, Y = np.ascontiguousarray(U, np.float64), self._vec(Y)
        self._chk(_lib.lib().mjw_apply_dh_minus_dt_sh(self._h, dt, alpha, dptr(U), beta, dptr(Y)),
                  "apply_dh_minus_dt_sh")
        return Y

    def solve_dh(self, Y) -> np.ndarray:
        """D^{-1} Y (SolveOperatorDh)."""
        Y = self._vec(Y)
        self._chk(_lib.lib().mjw_solve_dh(self._h, dptr(Y)), "solve_dh")
        return Y

    def solve_dh_plus_dt_sh(self, dt, Y) -> np.ndarray:
        """(D + dt/2 S)^{-1} Y (SolveOperatorDhPlusdtSh)."""
        Y = self._vec(Y)
        self._chk(_lib.lib().mjw_solve_dh_plus_dt_sh(self._h, dt, dptr(Y)), "solve_dh_plus_dt_sh")
        return Y

    def set_dirichlet_condition(self, Y) -> np.ndarray:
        """Y with its Dirichlet rows set to 0 (SetDirichletCondition)."""
        Y = self._vec(Y)
        self._chk(_lib.lib().mjw_set_dirichlet_condition(self._h, dptr(Y)), "set_dirichlet")
        return Y

    def set_time_source(self, f0: float, t_init: float = 0.0, eps_time: float = 0.0,
                        kind="RICKER", param: float = 0.0):
        """Time function of add_scalar_time_source: a TemporalSource keyword
        (or MJW_TIME_* number) and its value (mjw_set_temporal_source)."""
        k = TEMPORAL_SOURCES[kind] if isinstance(kind, str) else int(kind)
        self._chk(_lib.lib().mjw_set_temporal_source(self._h, k, f0, param, t_init, eps_time),
                  "set_time_source")

    def add_scalar_time_source(self, t: float, Y, alpha=1.0, nb_deriv=1, shot=0) -> np.ndarray:
        """Y + alpha f^(nb_deriv)(t) F_shot (AddScalarTimeSource)."""
        Y = self._vec(Y)
        self._chk(_lib.lib().mjw_add_scalar_time_source(self._h, shot, alpha, t, nb_deriv,
                                                        dptr(Y)), "add_scalar_time_source")
        return Y

    def mass(self) -> np.ndarray:
        D = np.zeros(self.mesh.N)
        self._chk(_lib.lib().mjw_get_mass(self._h, dptr(D)), "get_mass")
        return D

    # ---------------------------------------------------- sources/receivers
    def set_source_vectors(self, F):
        """Whole source vectors [nshots, ncomp*N] in place of Dirac points
        (mjw_set_source_vectors)."""
        F = np.ascontiguousarray(np.broadcast_to(np.asarray(F, np.float64),
                                                 (self.nshots, self.ndof)))
        self._chk(_lib.lib().mjw_set_source_vectors(self._h, dptr(F)), "set_source_vectors")

    def lumped_integral(self, f_q) -> np.ndarray:
        """integral of f phi_g with the Gauss-Lobatto quadrature, f at the
        quadrature points [E, nd] (mjw_lumped_integral)."""
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(f_q, np.float64),
                                                 (self.mesh.E, self.mesh.nd)))
        out = np.zeros(self.mesh.N)
        self._chk(_lib.lib().mjw_lumped_integral(self._h, dptr(f), dptr(out)), "lumped_integral")
        return out

    def set_sources(self, pts, pol=None, shot=None):
        pts = np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1, self.mesh.dim))
        k = pts.shape[0]
        polv = None if pol is None else np.ascontiguousarray(
            np.asarray(pol, np.float64).reshape(k, self.ncomp))
        sh = None if shot is None else np.ascontiguousarray(np.asarray(shot, np.int32))
        self._chk(_lib.lib().mjw_set_sources(self._h, k, dptr(pts), dptr(polv), iptr(sh)),
                  "set_sources")

    def set_pulse(self, table, d1=None, d2=None):
        """Source time factors f(t_n + dt); the fourth-order scheme also takes
        f' and f'' at the same times (d1, d2; mjw_set_pulse_derivatives)."""
        t = np.ascontiguousarray(table, np.float64)
        self._chk(_lib.lib().mjw_set_pulse(self._h, len(t), dptr(t)), "set_pulse")
        if d1 is not None or d2 is not None:
            a = np.ascontiguousarray(np.zeros(len(t)) if d1 is None else d1, np.float64)
            b = np.ascontiguousarray(np.zeros(len(t)) if d2 is None else d2, np.float64)
            if len(a) != len(t) or len(b) != len(t):
                raise ValueError("set_pulse: the derivative tables must match the pulse table")
            self._chk(_lib.lib().mjw_set_pulse_derivatives(self._h, len(t), dptr(a), dptr(b)),
                      "set_pulse_derivatives")

    def set_time_order(self, order: int):
        """2 (default) or 4: the modified-equation leapfrog of that order in
        time (mjw_set_time_order); call leapfrog_init afterwards."""
        self._chk(_lib.lib().mjw_set_time_order(self._h, int(order)), "set_time_order")

    def set_damping(self, sigma):
        """Damping sigma per element, or [E, nd] at the quadrature points
        (rho u'' + sigma u' - div(mu grad u) = f, elastic: rho u'' + sigma u'
        - div(C : grad u) = f; mjw_set_damping[_nodal]); call leapfrog_init
        afterwards."""
        sg = np.asarray(sigma, np.float64)
        if sg.ndim == 2:
            sg = np.ascontiguousarray(sg)
            self._chk(_lib.lib().mjw_set_damping_nodal(self._h, dptr(sg)), "set_damping")
            return
        sg = np.ascontiguousarray(np.broadcast_to(sg, (self.mesh.E,)))
        self._chk(_lib.lib().mjw_set_damping(self._h, dptr(sg)), "set_damping")

    def set_pml(self, pml, tau):
        """PML layers (mjw_set_pml): pml[e] marks the elements of the layers,
        tau[e, q, a] the damping at their quadrature points; the problem then
        runs Montjoie's first-order LEAP_FROG_PML scheme (pulse table:
        time_source_table(..., nb_deriv=0, offset=dt/2))."""
        fl = np.ascontiguousarray(np.asarray(pml, np.uint8).reshape(self.mesh.E))
        tv = np.ascontiguousarray(np.asarray(tau, np.float64).reshape(-1))
        if tv.size != self.mesh.E * self.mesh.nd * self.mesh.dim:
            raise ValueError("tau must hold E * nd * dim values")
        self._chk(_lib.lib().mjw_set_pml(self._h, u8ptr(fl), dptr(tv)), "set_pml")
        self.pml = True

    def set_receivers(self, pts, comp=None, deriv=None):
        """Receivers of u (deriv None or -1) or of d u / d x_deriv at pts on
        component comp; returns the element holding each point."""
        pts = np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1, self.mesh.dim))
        cm = None if comp is None else np.ascontiguousarray(np.asarray(comp, np.int32))
        dv = None if deriv is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(deriv, np.int32), (pts.shape[0],)))
        elem = np.zeros(pts.shape[0], np.int64)
        self._chk(_lib.lib().mjw_set_receivers_grad(
            self._h, pts.shape[0], dptr(pts), iptr(cm), iptr(dv),
            elem.ctypes.data_as(_lib.C.POINTER(_lib.C.c_int64))), "set_receivers")
        self.nrec = pts.shape[0]
        return elem

    # ------------------------------------------------------------ leapfrog
    def leapfrog_init(self, dt: float, t_begin: float = 0.0, u0=None, du0=None):
        u0a = None if u0 is None else np.ascontiguousarray(u0, np.float64)
        du0a = None if du0 is None else np.ascontiguousarray(du0, np.float64)
        self._chk(_lib.lib().mjw_leapfrog_init(self._h, t_begin, dt, dptr(u0a), dptr(du0a)),
                  "leapfrog_init")
        self.dt = dt

    def advance(self, nsteps: int, traces: bool = True, graph: bool = True):
        """Returns traces [nsteps, nshots, nrec] of U^n before each step."""
        tr = np.zeros((nsteps, self.nshots, max(self.nrec, 0))) if (traces and self.nrec) else None
        self._chk(_lib.lib().mjw_leapfrog_advance(self._h, nsteps, dptr(tr), int(bool(graph))),
                  "advance")
        return tr

    def record(self) -> np.ndarray:
        tr = np.zeros((self.nshots, self.nrec))
        if self.nrec:
            self._chk(_lib.lib().mjw_record(self._h, dptr(tr)), "record")
        return tr

    def set_probes(self, pts, deriv=None) -> np.ndarray:
        """Points where probe() samples the current iterate (deriv None or -1)
        or its derivative d / d x_deriv; returns the inside-the-mesh flags
        (mjw_set_probes, mjw_set_probes_grad)."""
        pts = np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1, self.mesh.dim))
        inside = np.zeros(pts.shape[0], np.uint8)
        if deriv is None:
            self._chk(_lib.lib().mjw_set_probes(self._h, pts.shape[0], dptr(pts),
                                                inside.ctypes.data_as(C.c_void_p)), "set_probes")
        else:
            dv = np.ascontiguousarray(np.broadcast_to(np.asarray(deriv, np.int32), (pts.shape[0],)))
            self._chk(_lib.lib().mjw_set_probes_grad(
                self._h, pts.shape[0], dptr(pts), iptr(dv),
                inside.ctypes.data_as(C.POINTER(C.c_uint8)), None), "set_probes")
        self.nprobe = pts.shape[0]
        return inside.astype(bool)

    def probe(self) -> np.ndarray:
        """[nshots, npts, ncomp] values of the current iterate (0 outside)."""
        out = np.zeros((self.nshots, getattr(self, "nprobe", 0), self.ncomp))
        if out.size:
            self._chk(_lib.lib().mjw_probe(self._h, dptr(out)), "probe")
        return out

    def run(self, nsteps: int, graph: bool = True):
        """Receiver traces for rows n = 0..nsteps (RunSecondOrderScheme)."""
        tr = self.advance(nsteps, traces=True, graph=graph)
        last = self.record()
        if tr is None:
            return None
        return np.concatenate([tr, last[None]], axis=0)

    def state(self):
        """(U^n, U^{n+1}) [nshots, ndof]; with PML layers (u^n, None)."""
        n = self.nshots * self.ndof
        a = np.zeros(n)
        if getattr(self, "pml", False):
            self._chk(_lib.lib().mjw_get_state(self._h, dptr(a), None), "get_state")
            return a.reshape(self.nshots, self.ndof), None
        b = np.zeros(n)
        self._chk(_lib.lib().mjw_get_state(self._h, dptr(a), dptr(b)), "get_state")
        return a.reshape(self.nshots, self.ndof), b.reshape(self.nshots, self.ndof)

    def node_order(self) -> np.ndarray:
        """solver_index[g]: position of node g in the device vectors."""
        out = np.zeros(self.mesh.N, dtype=np.int32)
        self._chk(_lib.lib().mjw_node_order(self._h, iptr(out)), "node_order")
        return out

    def state_dev(self):
        a = C.c_void_p()
        b = C.c_void_p()
        self._chk(_lib.lib().mjw_state_dev(self._h, C.byref(a), C.byref(b)), "state_dev")
        return a.value, b.value

    def step(self) -> int:
        v = C.c_int64()
        self._chk(_lib.lib().mjw_get_step(self._h, C.byref(v)), "get_step")
        return v.value

    def energy(self) -> np.ndarray:
        e = np.zeros(self.nshots)
        self._chk(_lib.lib().mjw_energy(self._h, dptr(e)), "energy")
        return e

    def lambda_max(self, iters: int = 80, seed: int = 1) -> float:
        v = C.c_double()
        self._chk(_lib.lib().mjw_lambda_max(self._h, iters, seed, C.byref(v)), "lambda_max")
        return v.value

    def profile(self, nsteps: int = 10) -> dict:
        """Per-launch device times (HIP events on the problem's stream)."""
        a = (C.c_double * 4)()
        self._chk(_lib.lib().mjw_profile(self._h, nsteps, a), "profile")
        return {"stiffness_launch_ms": a[0], "tail_ms": a[1], "step_ms": a[2],
                "launches_per_step": int(a[3])}

    def last_timing(self):
        a = (C.c_double * 2)()
        self._chk(_lib.lib().mjw_last_timing(self._h, a), "last_timing")
        return a[0], a[1]
