"""Shared definitions of the synthetic test problems and helpers to rebuild
the problems stored in tests/golden/*.npz (made by
tests/golden/make_ref_fixtures.py from Montjoie's own solvers)."""
from __future__ import annotations

import os
from dataclasses import dataclass, replace

import numpy as np

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@dataclass
class Case:
    name: str
    dim: int
    nx: int
    r: int
    pert: float
    seed: int
    elastic: bool
    mu_e: np.ndarray
    rho_e: np.ndarray
    lam_e: np.ndarray | None
    f0: float
    dt: float
    nsteps: int
    src: np.ndarray
    pol: np.ndarray | None
    rec: np.ndarray
    times: np.ndarray | None = None
    traces: np.ndarray | None = None
    sigma_e: np.ndarray | None = None
    # AddPML + LEAP_FROG_PML: (thickness, layers, vsigma, axes, profile, offset,
    # transverse); layers = the count used (AUTO decks: the resolved count)
    pml: tuple | None = None
    # the deck's TemporalSource line: a kind and its value
    temporal: str = "RICKER"
    # elastic outputs: the stress (DisplayStress = YES, the default) or grad u
    display_stress: bool = True
    # OrderTimeScheme = r LEAP_FROG: 2 or 4 (the modified equation)
    time_order: int = 2
    # TypeSource = SRC_VOLUME [refs] GAUSSIAN centre radius cut-off, or None
    volume: tuple | None = None

    def source_kind(self) -> tuple:
        w = self.temporal.split()
        return w[0], float(w[1]) if len(w) > 1 else 0.0


def golden_names():
    return sorted(f[:-4] for f in os.listdir(GOLD) if f.startswith("ref") and f.endswith(".npz"))


def load(name: str) -> Case:
    z = np.load(os.path.join(GOLD, name + ".npz"))
    dim, nx = int(z["dim"]), int(z["nx"])
    E = nx ** dim
    e = np.arange(E)
    layers = z["layers"]
    lay = ((e // nx ** (dim - 1)) * len(layers)) // nx
    elastic = bool(z["elastic"])
    mu_e = layers[lay]
    rho_e = z["rho_layers"][lay] if elastic else np.ones(E)
    lam_e = z["lam_layers"][lay] if elastic else None
    mm = {k.split("_")[1]: k for k in z.files if k.startswith("meshmedia_") and k.endswith("_nu")}
    if mm:
        # PhysicalMedia ... SAME_MESH data files of medium 1 (every element)
        m = O.mesh(dim, nx, int(z["r"]), float(z["pert"]), int(z["seed"]))
        fields = {}
        for nm in mm:
            g = lambda k: z[f"meshmedia_{nm}_{k}"]  # noqa: E731
            fields[nm] = O.mesh_media(m, g("nu"), int(g("order")), float(g("offset")),
                                      float(g("amplitude")))
        nd = m.nd
        rho_e = fields.get("rho", np.broadcast_to(rho_e[:, None], (E, nd)).copy())
        mu_e = fields.get("mu", np.broadcast_to(mu_e[:, None], (E, nd)).copy())
    srho = z["sinus_rho"] if "sinus_rho" in z.files else np.zeros(0)
    smu = z["sinus_mu"] if "sinus_mu" in z.files else np.zeros(0)
    if len(srho) or len(smu):
        # PhysicalMedia SINUS of medium 1 (every element): rho and mu at the
        # quadrature points, shape (E, nd)
        m = O.mesh(dim, nx, int(z["r"]), float(z["pert"]), int(z["seed"]))
        pts = O.quadrature_points(m)
        rho_e = sinus_field(srho, pts) if len(srho) else np.broadcast_to(rho_e[:, None],
                                                                         pts.shape[:2]).copy()
        if len(smu):
            mu_e = sinus_field(smu, pts)
        elif not elastic:  # elastic decks keep mu (and lambda) per element
            mu_e = np.broadcast_to(mu_e[:, None], pts.shape[:2]).copy()
    sg = z["sigma_layers"] if "sigma_layers" in z.files else np.zeros(0)
    sigma_e = sg[lay] if len(sg) else None
    ssg = z["sinus_sigma"] if "sinus_sigma" in z.files else np.zeros(0)
    if len(ssg):
        # PhysicalMedia sigma SINUS of medium 1 (the first layer): damping at
        # the quadrature points; the other layers keep MateriauDielec's sigma
        m = O.mesh(dim, nx, int(z["r"]), float(z["pert"]), int(z["seed"]))
        f = sinus_field(ssg, O.quadrature_points(m))
        base = sg[lay] if len(sg) else np.zeros(E)
        sigma_e = np.where((lay == 0)[:, None], f, base[:, None])
    pml = None
    if "pml" in z.files and len(z["pml"]):
        p = [float(v) for v in z["pml"]] + [0.0, 0.0, 0.0]
        axes = str(z["pml_axes"]) if "pml_axes" in z.files else {2: "XY", 3: "XYZ"}[dim]
        pml = (p[0], int(p[1]), p[2], axes, int(p[3]), p[4], p[5])
    return Case(name, dim, nx, int(z["r"]), float(z["pert"]), int(z["seed"]), elastic, mu_e,
                rho_e, lam_e, float(z["f0"]), float(z["dt"]), int(z["nsteps"]), z["src"],
                z["pol"] if elastic else None, z["rec"], z["times"], z["traces"], sigma_e, pml,
                str(z["temporal"]) if "temporal" in z.files else "RICKER",
                bool(z["display_stress"]) if "display_stress" in z.files else True,
                int(z["time_order"]) if "time_order" in z.files else 2,
                # the elements of the references (the layers, 1-based; all
                # when none is given), centre, radius, cut-off
                (np.isin(lay + 1, z["volume_refs"]) if len(z["volume_refs"]) else
                 np.ones(E, bool), z["volume_centre"],
                 *(list(z["volume_radius"]) if len(z["volume_radius"]) else [0.0, 0.0]))
                if "volume_kind" in z.files and str(z["volume_kind"]) else None)


def gaussian_volume(pts, centre, radius, cut):
    """The GAUSSIAN volume source's amplitude beta exp(-alpha r^2), alpha =
    -ln(1e-6)/radius^2, beta = sqrt(alpha/pi), 0 from r = cut on
    (GaussianSource::Init / GetAmplitude, Source/SourceSpatiale.cxx:72-97)."""
    alpha = -np.log(1e-6) / radius ** 2
    beta = np.sqrt(alpha / np.pi)
    r = np.sqrt(((np.asarray(pts) - np.asarray(centre)) ** 2).sum(-1))
    return np.where(r < cut, beta * np.exp(-alpha * r * r), 0.0)


def volume_field(c: Case, pts) -> np.ndarray:
    """f at the quadrature points pts [E, nd, dim] of the case's volume
    source, zero on the elements outside its references."""
    inside, centre, radius, cut = c.volume
    if not len(centre):  # UNIFORM
        return np.where(inside[:, None], 1.0, np.zeros(pts.shape[:2]))
    return np.where(inside[:, None], gaussian_volume(pts, centre, radius, cut), 0.0)


def oracle_source(c: Case, m) -> np.ndarray:
    """The case's source vector from the oracle: the Dirac at src, or the
    volume source's integral of f phi_i (component-blocked, times the
    polarisation for elastic)."""
    if c.volume is None:
        return O.dirac(m, c.src, c.pol) if c.elastic else O.dirac(m, c.src)
    b = O.damping(m, volume_field(c, O.quadrature_points(m)))
    if c.elastic:
        return np.concatenate([p * b for p in c.pol])
    return b


def oracle_pulse(c: Case, nsteps: int) -> np.ndarray:
    """The oracle's pulse table of the case's TemporalSource."""
    kind, par = c.source_kind()
    if kind == "RICKER":
        return O.pulse_table(c.f0, 0.0, c.dt, nsteps)
    return O.source_table(kind, c.f0, par, 0.0, c.dt, nsteps)


def oracle_pulse_derivatives(c: Case, nsteps: int):
    """f' and f'' of the case's TemporalSource at t_n + dt (the fourth-order
    scheme's source terms)."""
    kind, par = c.source_kind()
    return tuple(O.source_table(kind, c.f0, par, 0.0, c.dt, nsteps, k) for k in (2, 3))


def oracle_leapfrog(c: Case, P, b, nsteps: int, **kw):
    """O.leapfrog, or O.leapfrog4 for OrderTimeScheme = 4 cases."""
    pulse = oracle_pulse(c, nsteps)
    if c.time_order == 4:
        d1, d2 = oracle_pulse_derivatives(c, nsteps)
        return O.leapfrog4(P, b, (pulse, d1, d2), c.dt, nsteps, **kw)
    return O.leapfrog(P, b, pulse, c.dt, nsteps, **kw)


def gpu_pulse(c: Case, nsteps: int) -> np.ndarray:
    """The library's pulse table of the case's TemporalSource."""
    import montjoie_amd as M
    kind, par = c.source_kind()
    if kind == "RICKER":
        return M.ricker_pulse_table(c.f0, 0.0, c.dt, nsteps)
    return M.temporal_source_table(kind, c.f0, par, 0.0, c.dt, nsteps)


def gpu_set_pulse(c: Case, P, nsteps: int):
    """The case's source tables on a library problem (f' and f'' too for the
    fourth-order scheme)."""
    import montjoie_amd as M
    if c.time_order == 4:
        kind, par = c.source_kind()
        P.set_pulse(gpu_pulse(c, nsteps),
                    *(M.temporal_source_table(kind, c.f0, par, 0.0, c.dt, nsteps, k)
                      for k in (2, 3)))
    else:
        P.set_pulse(gpu_pulse(c, nsteps))


def pml_setup(c: Case):
    """(extended mesh, rho, mu per extended element, tau) of a PML case: the
    layers extruded as AddPML does, media of the element each was extruded
    from, parabolic damping with vmax over the PML media."""
    thick, layers, vsigma, axes, profile, offset, transverse = c.pml
    pm = O.pml_mesh(c.dim, c.nx, c.r, c.pert, c.seed, thick, layers,
                    O.pml_sides(axes, c.dim))
    mu, rho = c.mu_e[pm.phys], c.rho_e[pm.phys]
    vmax = float(np.sqrt(mu[pm.pml == 1] / rho[pm.pml == 1]).max())
    return pm, rho, mu, O.pml_tau(pm, vsigma, vmax, profile, offset, transverse)


def pml_deck(c: Case):
    """The deck-writer PML spec of a PML case (AUTO when the fixture's deck
    let Montjoie count the layers)."""
    z = np.load(os.path.join(GOLD, c.name + ".npz"))
    thick, layers, vsigma, axes, profile, offset, transverse = c.pml
    auto = "pml_auto" in z.files and int(z["pml_auto"]) == 1
    return dict(axes=axes, thickness=thick, layers="AUTO" if auto else layers, vsigma=vsigma,
                profile=profile, offset=offset, transverse=transverse)


def sinus_field(p, pts):
    """Oracle evaluation of a stored SINUS field (layout: xmin(dim) xmax(dim)
    cut-off(dim) offset amplitude period(dim)) at pts[..., dim]."""
    dim = pts.shape[-1]
    p = np.asarray(p, float)
    return O.sinus_media(pts, p[:dim], p[dim:2 * dim], p[2 * dim:3 * dim], p[3 * dim],
                         p[3 * dim + 1], p[3 * dim + 2:])


def receivers_inside(c: Case) -> np.ndarray:
    """Mask of the case's receivers that lie in the mesh (the oracle's point
    location); Montjoie writes 0 for a SismoPoint outside the mesh and keeps
    its columns."""
    if c.pml is not None:
        return np.ones(len(c.rec), bool)
    m = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    return O.receiver_gradients(m, c.rec)[0] >= 0


def expand_outside(c: Case, tr: np.ndarray, inside: np.ndarray) -> np.ndarray:
    """Traces over the receivers inside (last axis: component-major, the
    points inside per component) widened to every receiver, 0 for the points
    outside."""
    if inside.all():
        return tr
    ncomp = c.dim if c.elastic else 1
    npt = len(inside)
    out = np.zeros(tr.shape[:-1] + (ncomp, npt))
    out[..., inside] = tr.reshape(tr.shape[:-1] + (ncomp, int(inside.sum())))
    return out.reshape(tr.shape[:-1] + (ncomp * npt,))


def oracle_run(c: Case, nsteps: int | None = None):
    """Run the case through the C oracle: returns (uN, uN1, traces[n, rec]);
    receivers outside the mesh give 0 columns."""
    inside = receivers_inside(c)
    if not inside.all():
        uN, uN1, tr = oracle_run(replace(c, rec=c.rec[inside]), nsteps)
        return uN, uN1, expand_outside(c, tr, inside)
    nsteps = c.nsteps if nsteps is None else nsteps
    if c.pml is not None:
        pm, rho, mu, tau = pml_setup(c)
        tab = O.time_table(c.f0, 0.0, c.dt, nsteps, 0, 0.5 * c.dt)
        uN, tr = O.leapfrog_pml(pm, rho, mu, tau, O.dirac(pm.mesh, c.src), tab, c.dt, nsteps,
                                c.rec)
        return uN, None, tr
    m = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    if c.elastic:
        P = O.elastic_problem(m, c.lam_e, c.mu_e, c.rho_e)
        rc = np.zeros(0, np.int32)
    else:
        P = O.acoustic_problem(m, c.mu_e, c.rho_e)
        rc = None
    b = oracle_source(c, m)
    S = O.damping(m, c.sigma_e) if c.sigma_e is not None and np.any(c.sigma_e) else None
    if c.elastic:
        # every receiver on the three components, component-major
        pts = np.vstack([c.rec] * c.dim)
        comp = np.repeat(np.arange(c.dim, dtype=np.int32), len(c.rec))
        return oracle_leapfrog(c, P, b, nsteps, rec_pts=pts, rec_comp=comp, S=S)
    return oracle_leapfrog(c, P, b, nsteps, rec_pts=c.rec, rec_comp=rc, S=S)


def oracle_stress(c: Case, nsteps: int | None = None) -> np.ndarray:
    """Stress traces of an elastic case from the oracle: d u_i / d x_a at each
    receiver recorded by the leapfrog (gradient weights), then
    sigma = lambda tr(grad u) I + mu (grad u + grad u^T) with the medium of
    the element holding the point; rows n = 0..nsteps, d*d values per point
    (row-major), point-major.  DisplayStress = NO decks: grad u itself,
    d u_i / d x_a at i d + a."""
    inside = receivers_inside(c)
    if not inside.all():
        # a point outside the mesh: d*d zero columns
        g = oracle_stress(replace(c, rec=c.rec[inside]), nsteps)
        out = np.zeros((g.shape[0], len(c.rec), c.dim * c.dim))
        out[:, inside] = g.reshape(g.shape[0], int(inside.sum()), c.dim * c.dim)
        return out.reshape(g.shape[0], -1)
    nsteps = c.nsteps if nsteps is None else nsteps
    d = c.dim
    m = O.mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    P = O.elastic_problem(m, c.lam_e, c.mu_e, c.rho_e)
    b = oracle_source(c, m)
    el, _, Gw = O.receiver_gradients(m, c.rec)
    assert (el >= 0).all()
    npt = len(c.rec)
    # receiver (point r, component i, direction a), point-major
    re = np.repeat(el, d * d)
    rw = np.stack([Gw[r, a] for r in range(npt) for i in range(d) for a in range(d)])
    comp = np.array([i for r in range(npt) for i in range(d) for a in range(d)], np.int32)
    S = O.damping(m, c.sigma_e) if c.sigma_e is not None and np.any(c.sigma_e) else None
    _, _, g = oracle_leapfrog(c, P, b, nsteps, rec_comp=comp, rec_weights=(re, rw), S=S)
    g = g.reshape(nsteps + 1, npt, d, d)
    if not c.display_stress:
        return g.reshape(nsteps + 1, npt * d * d)
    lam = c.lam_e[el][None, :, None, None]
    mu = c.mu_e[el][None, :, None, None]
    div = np.trace(g, axis1=2, axis2=3)[:, :, None, None]
    sig = lam * div * np.eye(d) + mu * (g + g.transpose(0, 1, 3, 2))
    return sig.reshape(nsteps + 1, npt * d * d)


def reference_traces(c: Case) -> np.ndarray:
    """Montjoie's traces re-ordered like oracle_run's (rows n = 0..nsteps)."""
    tr = c.traces
    if c.elastic:
        nr, d = len(c.rec), c.dim
        # elastic SismoPoint rows start with the displacement of each point
        # (d components per point), followed by d*d further fields per point
        tr = tr[:, :d * nr].reshape(tr.shape[0], nr, d).transpose(0, 2, 1)
        tr = tr.reshape(tr.shape[0], d * nr)
    return tr


def plane_file(raw: bytes, npts: int):
    """(header bytes, values) of a SismoPlane binary double file."""
    raw = bytes(raw)
    return raw[:len(raw) - 8 * npts], np.frombuffer(raw[len(raw) - 8 * npts:], np.float64)


def plane_steps(c: Case, times) -> list:
    """Steps of the ParametersOutputPlane snapshots t0, t0 + dt_out, ... <= t1."""
    t0, t1, dto = (float(v) for v in times)
    every = max(1, int(round(dto / c.dt)))
    return [s for s in range(c.nsteps + 1)
            if t0 - 1e-9 * c.dt <= s * c.dt <= t1 + 1e-9 * c.dt
            and int(round((s * c.dt - t0) / c.dt)) % every == 0]


def rel_l2(a, b) -> float:
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_problem(c: Case, precision: str = "fp64", nshots: int = 1, srcs=None):
    """The same case on the HIP path (montjoie_amd); sources default to the
    case's single source replicated for every shot.  PML cases: the extended
    mesh handed over as a caller's mesh, with its layers (mjw_set_pml)."""
    import montjoie_amd as M
    if c.pml is not None:
        pm, rho, mu, tau = pml_setup(c)
        o = pm.mesh
        m = M.Mesh(c.dim, c.r, np.ascontiguousarray(o.xyz), np.ascontiguousarray(o.elem_nodes),
                   np.ascontiguousarray(o.dirichlet), o.nx)
        P = M.WaveProblem(m, rho, mu, precision=precision)
        P.set_pml(pm.pml, tau)
        P.set_sources(c.src)
        P.set_receivers(c.rec)
        return P
    m = M.regular_mesh(c.dim, c.nx, c.r, c.pert, c.seed)
    if c.elastic:
        nodal_rho = c.rho_e.ndim == 2
        P = M.WaveProblem(m, 1.0 if nodal_rho else c.rho_e, c.mu_e, lam=c.lam_e, ncomp=c.dim,
                          nshots=nshots, precision=precision)
        if nodal_rho:  # PhysicalMedia rho: the mass only
            P.set_media_nodal(c.rho_e)
    elif c.mu_e.ndim == 2:
        # media varying inside the elements
        P = M.WaveProblem(m, 1.0, 1.0, nshots=nshots, precision=precision)
        P.set_media_nodal(c.rho_e, c.mu_e)
    else:
        P = M.WaveProblem(m, c.rho_e, c.mu_e, nshots=nshots, precision=precision)
    if c.sigma_e is not None and np.any(c.sigma_e):
        P.set_damping(c.sigma_e)
    if c.time_order != 2:
        P.set_time_order(c.time_order)
    if c.volume is not None and srcs is None:
        # the volume source through the library's own quadrature sum
        b = P.lumped_integral(volume_field(c, m.xyz[m.elem_nodes]))
        F = np.concatenate([p * b for p in c.pol]) if c.elastic else b
        P.set_source_vectors(np.tile(F, (nshots, 1)))
    else:
        if srcs is None:
            srcs = [c.src] * nshots
        pol = None if not c.elastic else np.vstack([c.pol] * len(srcs))
        P.set_sources(np.asarray(srcs), pol=pol, shot=np.arange(len(srcs)) % nshots)
    rec = c.rec[receivers_inside(c)]  # the points outside give 0 columns (gpu_run)
    if c.elastic:
        P.set_receivers(np.vstack([rec] * c.dim),
                        comp=np.repeat(np.arange(c.dim), len(rec)))
    else:
        P.set_receivers(rec)
    return P


def gpu_run(c: Case, precision="fp64", nsteps=None, graph=True, nshots=1, srcs=None):
    import montjoie_amd as M
    nsteps = c.nsteps if nsteps is None else nsteps
    P = gpu_problem(c, precision, nshots, srcs)
    if c.pml is not None:
        # first-order system: the primitive at t_n + dt/2
        P.set_pulse(M.time_source_table(c.f0, 0.0, c.dt, nsteps, 0, 0.5 * c.dt))
    else:
        gpu_set_pulse(c, P, nsteps)
    P.leapfrog_init(c.dt)
    tr = P.run(nsteps, graph=graph)
    return P, expand_outside(c, tr, receivers_inside(c))
