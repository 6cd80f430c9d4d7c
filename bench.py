"""Benchmark: FP32 vs FP64 leapfrog on the Q5 3-D acoustic workload.

Workload (one GPU): unit cube, 48^3 Q5 hexahedra with interior vertices
displaced by <= 0.1 h (seed 2024), velocity in three horizontal layers
c_k ~ U[1,3] (seed 2025), rho = 1, mu = c^2, Dirichlet boundary, Dirac at
the centre with a Ricker pulse f0 = 8, dt = 0.5 x CFL limit, 16 receivers.
A step is one fused leapfrog step over the whole mesh (8 colour passes of
the stiffness kernel + source/receiver tail), inputs resident in HBM.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20] [--no-cpu]

Prints one JSON line.  `value` is the GDoF-steps/s of the FP32 state path
with FP64 geometric factors (the workload's precision); the FP64 path and the
all-FP32 path are reported beside it, with the errors of the FP32 paths
against FP64 over the workload's 500 steps (untimed runs).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

# torch ships its own HIP runtime under the same soname as ROCm's; load it
# before libmjwave.so so both use one runtime in this process
import torch  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def workload(nx=48, order=5):
    import montjoie_amd as M
    m = M.regular_mesh(3, nx, order, 0.1, 2024)
    c = np.array([1.0 + 2.0 * _unif(2025, k) for k in range(3)])
    mu = M.layered_values(m, c * c)
    return m, mu, c


def _unif(seed, idx):
    # same splitmix64 stream as the mesh perturbation (montjoie_amd/csrc/mjw_host.cpp)
    M64 = (1 << 64) - 1

    def mix(z):
        z = (z + 0x9E3779B97F4A7C15) & M64
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)
    h = mix(mix(seed) ^ ((idx * 0xD1B54A32D192ED03) & M64))
    return (h >> 11) * (1.0 / 9007199254740992.0)


def receivers(n=16, seed=77):
    return np.array([[0.1 + 0.8 * _unif(seed, 3 * i + a) for a in range(3)] for i in range(n)])


def algorithmic_bytes_per_step(E, nd, N, state_bytes, geom_bytes, ngeom=6):
    """Stored per-node geometry + element map + (read U, read/write
    U^{n-1 -> n+1}, read dt^2/M): the per-step byte count of the stored-
    geometry algorithm, no cache reuse credited."""
    return E * nd * (ngeom * geom_bytes + 4) + N * 4 * state_bytes


def trilinear_bytes_per_step(E, nd, N, state_bytes, geom_bytes, ncoef=24):
    """Same with the per-element trilinear coefficients in place of the
    per-node geometry: what the trilinear kernels have to move."""
    return E * (ncoef * geom_bytes + nd * 4) + N * 4 * state_bytes


def kernel_name(tile, tri, n=6, t="float"):
    """rocprofv3 name of the FP32 stiffness kernel a tile code selects."""
    tr = "true" if tri else "false"
    if tile > 1000:
        return f"mjw::k_pipe3d<{n}, {tile - 1000}, {t}, {t}, 1, {tr}>"
    if tile < 0:
        return f"mjw::k_node3d<{n}, {-tile}, {t}, {t}, 1, false, {tr}>"
    return f"mjw::k_scalar3d<{n}, {tile}, {t}, {t}, 1, false, false, {tr}>"


def measured_traffic(name="stiffness_q5_48cube_counters.md"):
    """HBM bytes per colour pass of the headline kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (line `traffic_bytes_per_launch:`)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def flops_per_step(E, n):
    # gradient + divergence 2*3*2n per node, tensor 2*9, update ~6 per node
    nd = n ** 3
    return E * nd * (12 * n + 18 + 6)


def timed_advance(P, dt, steps, warmup):
    """Untimed warm-up (graph capture and instantiation, traces buffer sized
    for `steps`), then `steps` steps from a fresh zero state timed with the
    device synchronised on both sides.  Returns (wall seconds, traces)."""
    P.leapfrog_init(dt)
    P.advance(max(warmup, steps, 16), traces=True, graph=True)
    P.leapfrog_init(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tr = P.advance(steps, traces=True, graph=True)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, tr


def run_gpu(m, mu, dt, prec, steps, warmup, f0=8.0, nshots=1, profile_steps=10, tile=None,
            nxcd=None, acc_steps=500):
    """One precision on the headline workload: `steps` timed steps (HIP graph,
    traces recorded), after an untimed warm-up of at least as many steps that
    captures and instantiates the graph and sizes the traces buffer; then, from
    a fresh start and untimed, the `acc_steps` run whose traces, state and FP64
    energy give the accuracy figures."""
    import montjoie_amd as M
    P = M.WaveProblem(m, 1.0, mu, precision=prec, nshots=nshots)
    srcs = [(0.5, 0.5, 0.5)] * nshots
    P.set_sources(srcs, shot=np.arange(nshots))
    P.set_receivers(receivers())
    total = max(warmup, steps, acc_steps, 16) + profile_steps
    P.set_pulse(M.ricker_pulse_table(f0, 0.0, dt, total))
    if tile is None:
        from montjoie_amd import tuner
        tuned = tuner.tune_problem(P, steps=8, dt=dt)  # launch shape only: result-neutral
    else:
        P.set_tile(tile)
        P.set_xcd_order(nxcd or 8)
        tuned = {"tile": tile, "nxcd": nxcd or 8}
    wall, tr = timed_advance(P, dt, steps, warmup)
    dev_ms = P.last_timing()[0]
    if steps != acc_steps:
        P.leapfrog_init(dt)
        tr = P.advance(acc_steps, traces=True, graph=True)
    uN, _ = P.state()
    energy = P.energy()  # FP64 accumulation, also for the FP32 state
    prof = P.profile(profile_steps)
    info = P.info()
    geo = P.geometry
    P.close()
    return dict(wall=wall, dev_ms_step=dev_ms, traces=tr, uN=uN, prof=prof, info=info, geometry=geo,
                tile=tuned["tile"], nxcd=tuned["nxcd"], energy=energy)


def shot_sources(nshots, seed=2027):
    return np.array([[0.25 + 0.5 * _unif(seed, 3 * s + a) for a in range(3)]
                     for s in range(nshots)])


def bench_shots(m, mu, dt, steps, warmup, nshots=16, acc_steps=500):
    """FP32 batch of `nshots` Dirac shots in one problem (geometry read once per
    element for all shots) against the FP64 path, run shot by shot and as the
    same interleaved batch (time-to-solution ratios, and the FP32 error against
    the FP64 batch)."""
    import montjoie_amd as M
    srcs = shot_sources(nshots)
    out = {}
    from montjoie_amd import tuner
    for key, prec, ns in (("fp32", "fp32", nshots), ("fp32_g64", "fp32_g64", nshots),
                          ("fp64_batch", "fp64", nshots), ("fp64", "fp64", 1)):
        P = M.WaveProblem(m, 1.0, mu, precision=prec, nshots=ns)
        P.set_sources(srcs[:ns], shot=np.arange(ns))
        P.set_receivers(receivers())
        P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, max(warmup, steps, acc_steps, 16) + 64))
        # launch shape only; every candidate but the matrix-core tile (FP32
        # rounding) gives the same bits
        tuner.tune_problem(P, steps=4, dt=dt)
        wall, tr = timed_advance(P, dt, steps, warmup)
        dev_ms = P.last_timing()[0]
        if key != "fp64" and steps != acc_steps:
            P.leapfrog_init(dt)
            tr = P.advance(acc_steps, traces=True)
        uN = P.state()[0] if key != "fp64" else None  # before mjw_profile advances it
        prof = P.profile(4) if key in ("fp32_g64", "fp32") else None
        out[key] = dict(wall=wall, shots=ns, dev_ms=dev_ms, traces=tr, prof=prof, uN=uN)
        P.close()
    t64_batch = out["fp64"]["wall"] * nshots           # FP64 path, one shot at a time
    # stored-geometry bytes per step: FP64 tensor and map per element node, three
    # FP32 state streams per node and shot, FP32 dt^2/M per node
    sb = m.E * m.nd * (6 * 8 + 4) + m.N * (3 * nshots * 4 + 4)
    g64_ms = out["fp32_g64"]["prof"]["stiffness_launch_ms"]
    f32_ms = out["fp32"]["prof"]["stiffness_launch_ms"]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    ref = out["fp64_batch"]
    return {
        "shots": nshots,
        "fp32_batch_s": out["fp32"]["wall"],
        "fp32_g64_batch_s": out["fp32_g64"]["wall"],
        "fp64_one_shot_s": out["fp64"]["wall"],
        "fp64_path_16_shots_s": t64_batch,
        "fp32_batch_gdof_steps_s": m.N * steps * nshots / out["fp32"]["wall"] / 1e9,
        "fp32_g64_batch_gdof_steps_s": m.N * steps * nshots / out["fp32_g64"]["wall"] / 1e9,
        "time_to_solution_ratio_fp64_over_fp32_batch": t64_batch / out["fp32"]["wall"],
        "time_to_solution_ratio_fp64_over_fp32_g64_batch": t64_batch / out["fp32_g64"]["wall"],
        "fp64_batch_s": ref["wall"],
        "time_to_solution_ratio_fp64_batch_over_fp32_batch": ref["wall"] / out["fp32"]["wall"],
        "accuracy_steps": acc_steps,
        "fp32_batch_rel_l2_state_vs_fp64_batch": rel(out["fp32"]["uN"], ref["uN"]),
        "fp32_batch_rel_l2_traces_vs_fp64_batch": rel(out["fp32"]["traces"], ref["traces"]),
        "fp32_g64_batch_rel_l2_state_vs_fp64_batch": rel(out["fp32_g64"]["uN"], ref["uN"]),
        "fp32_g64_batch_rel_l2_traces_vs_fp64_batch": rel(out["fp32_g64"]["traces"], ref["traces"]),
        "time_to_solution_ratio_fp64_batch_over_fp32_g64_batch":
            ref["wall"] / out["fp32_g64"]["wall"],
        "roofline": {"bound": "hbm", "achieved": sb / 8 / (g64_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": sb / 8 / (g64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_shots3d, FP32 state, FP64 geometry, one colour pass of "
                               f"{nshots} interleaved shots",
                     "algorithmic_bytes_per_launch": sb / 8, "launch_ms": g64_ms},
        "roofline_fp32": {"bound": "hbm", "achieved": sb / 8 / (f32_ms * 1e-3) / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": sb / 8 / (f32_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "kernel": "k_shots3d, FP32 state and geometry, one colour pass",
                          "launch_ms": f32_ms},
    }


def smooth_field(m, seed, ng=4):
    """Per-element sum of ng seeded Gaussians, scaled to [0, 1]."""
    X = m.xyz[m.elem_nodes].mean(axis=1)
    f = np.zeros(m.E)
    for g in range(ng):
        a = 2 * _unif(seed, 10 * g) - 1
        c = np.array([0.2 + 0.6 * _unif(seed, 10 * g + 1 + k) for k in range(3)])
        s = 0.1 + 0.15 * _unif(seed, 10 * g + 5)
        f += a * np.exp(-np.sum((X - c) ** 2, axis=1) / (2 * s * s))
    return (f - f.min()) / (f.max() - f.min())


def bench_adjoint(nx=40, order=4, steps=2000, nsnap=20, ndir=8):
    """Gradient of an L2 receiver misfit w.r.t. per-element c^2 on a 40^3 Q4
    cube: cost of forward + adjoint over a plain forward solve (FP32 state with
    FP64 geometric factors and FP64 gradient accumulation), central
    finite-difference check of the FP64 gradient on 8 random directions, and
    the FP32 vs FP64 gradient error."""
    import montjoie_amd as M
    from montjoie_amd.adjoint import AdjointSolver
    m = M.regular_mesh(3, nx, order, 0.0, 0)
    mtrue = 1.0 + 8.0 * smooth_field(m, 2029)
    m0 = mtrue * (1.0 + 0.1 * (2 * smooth_field(m, 2030) - 1))
    g1 = np.linspace(0.1, 0.9, 8)
    rec = np.array([[x, y, 0.95] for y in g1 for x in g1])
    P64 = M.WaveProblem(m, 1.0, m0, precision="fp64")
    dt = 0.5 * 2.0 / np.sqrt(P64.lambda_max(60, 2))
    res = {"workload": f"{nx}^3 Q{order} cube, mu = c^2 smooth in [1,9] (seed 2029), model "
                       f"perturbed 10% (seed 2030), {steps} steps, 64 receivers (8x8 at z=0.95), "
                       f"{nsnap} checkpoints", "dt": dt, "n_dof": m.N}
    solvers = {}
    for prec in ("fp32_g64", "fp64"):
        P = P64 if prec == "fp64" else M.WaveProblem(m, 1.0, m0, precision=prec)
        P.set_sources([(0.5, 0.5, 0.5)])
        P.set_receivers(rec)
        P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, steps + 4))
        solvers[prec] = (P, AdjointSolver(P, dt, steps, nsnap))
    grads = {}
    for prec in ("fp32_g64", "fp64"):
        P, S = solvers[prec]
        dobs = S.forward(mtrue)
        S.set_mu(m0)
        # plain forward solve (same steps, traces on)
        P.leapfrog_init(dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        P.advance(steps, traces=True)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        tr = S.forward()
        g = S.backward(tr - dobs)
        torch.cuda.synchronize()
        t_grad = time.perf_counter() - t0
        dev = S.timing()
        grads[prec] = g
        res[prec] = {"forward_s": t_fwd, "forward_plus_adjoint_s": t_grad,
                     "gradient_over_forward_cost": t_grad / t_fwd,
                     "device_ms_adjoint_forward": dev[0], "device_ms_adjoint_backward": dev[1]}
        if prec == "fp64":
            P, S = solvers["fp64"]
            J = lambda mu: 0.5 * np.sum((S.forward(mu) - dobs) ** 2)  # noqa: E731
            rng = np.random.default_rng(2031)
            errs2, errs4 = [], []
            for _ in range(ndir):
                d = rng.standard_normal(m.E)
                gd = g @ d
                # central difference, step small enough that its O(h^2)
                # truncation stays below the 1e-6 target
                h = 1e-5 * np.linalg.norm(m0) / np.linalg.norm(d)
                fd2 = (J(m0 + h * d) - J(m0 - h * d)) / (2 * h)
                errs2.append(abs(fd2 - gd) / abs(gd))
                # fourth-order stencil at a 10x larger step, as a cross-check
                h4 = 1e-4 * np.linalg.norm(m0) / np.linalg.norm(d)
                jp, jm = J(m0 + h4 * d), J(m0 - h4 * d)
                jp2, jm2 = J(m0 + 2 * h4 * d), J(m0 - 2 * h4 * d)
                fd4 = (-jp2 + 8 * jp - 8 * jm + jm2) / (12 * h4)
                errs4.append(abs(fd4 - gd) / abs(gd))
            res["fd_check_fp64_max_rel_err"] = float(max(errs2))
            res["fd_stencil"] = "central (J(m+hd) - J(m-hd)) / 2h, h = 1e-5 ||m|| / ||d||"
            res["fd4_check_fp64_max_rel_err"] = float(max(errs4))
            res["fd4_stencil"] = "4th-order central, h = 1e-4 ||m|| / ||d||"
            res["fd_directions"] = ndir
    res["fp32_vs_fp64_gradient_rel_l2"] = float(np.linalg.norm(grads["fp32_g64"] - grads["fp64"])
                                               / np.linalg.norm(grads["fp64"]))
    res["precision"] = ("fp32_g64: FP32 state, FP64 geometric factors, FP64 gradient "
                        "accumulation; fp64: all FP64")
    for P, _ in solvers.values():
        P.close()
    return res


def bench_batch(nprob=64, kmin=8, kmax=40, omin=3, omax=7, steps=300, seed=2028, tune=True,
                prec="fp32_g64"):
    """64 independent 3-D acoustic problems of uneven size and order (edge
    k ~ U{kmin..kmax}, order ~ U{omin..omax}, smooth velocity 2 + sum of 4
    Gaussians), FP32 state with FP64 geometric factors, one batch: per-colour
    launches over all problems, step loop in a HIP graph, launch order picked
    by the autotuner."""
    import montjoie_amd as M
    from montjoie_amd import tuner
    from montjoie_amd.batch import Batch
    t_setup = time.perf_counter()
    probs, dts, ndof, nbytes, obytes = [], [], 0, 0, 0
    for i in range(nprob):
        k = kmin + int(_unif(seed, 2 * i) * (kmax - kmin + 1))
        r = omin + int(_unif(seed, 2 * i + 1) * (omax - omin + 1))
        m = M.regular_mesh(3, k, r, 0.0, 0)
        c = 2.0 + 1.5 * (2 * smooth_field(m, seed * 1000 + i) - 1) * 0.5
        P = M.WaveProblem(m, 1.0, c * c, precision=prec)
        P.set_sources([(0.5, 0.5, 0.5)])
        P.set_receivers(receivers(4, seed + i))
        dt = 0.5 * 2.0 / np.sqrt(P.lambda_max(30, 3))
        P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, steps + 64))
        probs.append(P)
        dts.append(dt)
        ndof += m.N
        nbytes += algorithmic_bytes_per_step(m.E, m.nd, m.N, 4, 8)
        obytes += trilinear_bytes_per_step(m.E, m.nd, m.N, 4, 8)
        del m
        if i % 16 == 15:
            print(f"# batch setup: {i + 1}/{nprob} problems, {ndof} dof", file=sys.stderr, flush=True)
    B = Batch(probs)
    t_setup = time.perf_counter() - t_setup
    for P, dt in zip(probs, dts):
        P.leapfrog_init(dt)
    tuned = tuner.tune_batch(B, steps=4) if tune else None
    for P, dt in zip(probs, dts):
        P.leapfrog_init(dt)
    B.advance(steps)  # warm-up: graph capture and instantiation, traces buffers
    for P, dt in zip(probs, dts):
        P.leapfrog_init(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.advance(steps, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = B.ms_per_step()
    # the same problems one after the other (per-problem graphs)
    t1 = time.perf_counter()
    for P, dt in zip(probs, dts):
        P.leapfrog_init(dt)
        P.advance(steps, traces=True, graph=True)
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - t1
    orders = sorted({P.mesh.order for P in probs})
    B.close()
    for P in probs:
        P.close()
    achieved = nbytes / (dev_ms * 1e-3) / 1e9
    own = obytes / (dev_ms * 1e-3) / 1e9
    return {"problems": nprob, "total_dof": ndof, "steps": steps, "orders": orders,
            "precision": "FP32 state, FP64 geometric factors",
            "batch_s": wall, "batch_gdof_steps_s": ndof * steps / wall / 1e9,
            "one_by_one_s": wall_seq, "one_by_one_gdof_steps_s": ndof * steps / wall_seq / 1e9,
            "setup_s": t_setup,
            "autotuned": tuned and {"order_mode": tuned["order_mode"],
                                    "candidates_ms_per_step": {str(m): ms for ms, m in
                                                               tuned["candidates"]},
                                    "modes": "0/1: one launch per colour over every order; "
                                             "2/3/4: one launch per (colour, order), blocks of "
                                             "about 256/512/128 threads"},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "note": "whole step (8 colour launches over all problems + tails), "
                                 "stored-geometry algorithmic bytes summed over the problems",
                         "algorithmic_bytes_per_step": nbytes, "device_ms_per_step": dev_ms,
                         "own_bytes_per_step": obytes, "own_bytes_frac": own / HBM_PEAK_GBS,
                         "own_bytes_note": "trilinear kernels: 24 coefficients per element, "
                                           "map per element node, 4 state streams per node"}}


def bench_host_apply(m, mu, reps=5):
    """ApplyOperatorKh through host FP64 vectors (mjw_apply_kh: upload U,
    apply, download Y), the PCIe-inclusive rate of the operator boundary,
    against the same apply on device-resident vectors (mjw_apply_kh_dev)."""
    import montjoie_amd as M
    P = M.WaveProblem(m, 1.0, mu, precision="fp64")
    u = np.random.default_rng(0).standard_normal(m.N)
    P.apply_kh(u)
    t0 = time.perf_counter()
    for _ in range(reps):
        P.apply_kh(u)
    host = (time.perf_counter() - t0) / reps
    U = torch.tensor(u, device="cuda")
    Y = torch.zeros_like(U)
    P.apply_kh_dev(U.data_ptr(), Y.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        P.apply_kh_dev(U.data_ptr(), Y.data_ptr())
    torch.cuda.synchronize()
    dev = (time.perf_counter() - t0) / reps
    # the diagonal operators of Montjoie's Advance on host vectors, through the
    # C ABI on preallocated arrays: (D - dt/2 S) U and (D + dt/2 S)^{-1} Y
    from montjoie_amd import _lib
    L, h, dt = _lib.lib(), P.handle, 1e-3
    y = np.zeros(m.N)
    diag = {}
    for name, call in (("apply_dh_minus_dt_sh",
                        lambda: L.mjw_apply_dh_minus_dt_sh(h, dt, 1.0, _lib.dptr(u), 0.0,
                                                           _lib.dptr(y))),
                       ("solve_dh_plus_dt_sh", lambda: L.mjw_solve_dh_plus_dt_sh(h, dt,
                                                                                 _lib.dptr(y)))):
        call()
        t0 = time.perf_counter()
        for _ in range(reps):
            assert call() == 0
        diag[name + "_ms"] = (time.perf_counter() - t0) / reps * 1e3
    P.close()
    return {"precision": "fp64", "host_vectors_ms": host * 1e3, "device_vectors_ms": dev * 1e3,
            "host_gdof_s": m.N / host / 1e9, "device_gdof_s": m.N / dev / 1e9,
            "host_diagonal_operators": diag}


def sphere_dirs(n, seed=2027):
    """n unit vectors, uniform on the sphere (z ~ U[-1,1], phi ~ U[0,2pi))."""
    out = []
    for s in range(n):
        z = 2 * _unif(seed, 1000 + 2 * s) - 1
        ph = 2 * np.pi * _unif(seed, 1001 + 2 * s)
        r = np.sqrt(max(0.0, 1 - z * z))
        out.append([r * np.cos(ph), r * np.sin(ph), z])
    return np.array(out)


def elastic_bytes_per_step(E, nd, N, nshots, state_bytes, geom_bytes=8):
    """Bytes per step of the elastic 16-shot pass, no cache reuse credited: per
    element node the stored geometry (J w and J DF^{-1}, 10 values) and the map;
    per node and shot three state streams of 3 components, and M^{-1}."""
    return E * nd * (10 * geom_bytes + 4) + N * (3 * 3 * nshots * state_bytes + 8)


def elastic_flops_per_step(E, n, nshots):
    """Algorithmic FLOPs per step of the elastic pass: per element and shot,
    three gradients and three divergences of 3 directions (2 n flops per node
    each) and, per node, M^T g (54), the isotropic stress (about 30) and M
    sigma (54)."""
    nd = n ** 3
    return E * nshots * nd * (2 * 2 * 9 * n + 54 + 30 + 54)


def bench_elastic(nx=32, order=4, steps=800, nshots=16, warmup=10):
    """Isotropic elastodynamics, Q4 hexahedra nx^3 (vertices perturbed <= 0.1h,
    seed 2024), 3 layers lambda, mu ~ U[1,4], rho ~ U[1,2] (seed 2026);
    16 Dirac shots at U[0.25,0.75]^3 with directions uniform on the sphere
    (seed 2027), Ricker f0 = 8, dt = 0.5 CFL, 800 steps.  FP32 16-shot batch
    (FP32 state, FP64 geometric factors; all-FP32 beside it) against the FP64
    path (batch for the error, one shot for the shot-by-shot time)."""
    import montjoie_amd as M
    from montjoie_amd import tuner
    m = M.regular_mesh(3, nx, order, 0.1, 2024)
    lam = M.layered_values(m, [1.0 + 3.0 * _unif(2026, k) for k in range(3)])
    mu = M.layered_values(m, [1.0 + 3.0 * _unif(2026, 3 + k) for k in range(3)])
    rho = M.layered_values(m, [1.0 + _unif(2026, 6 + k) for k in range(3)])
    srcs, pol = shot_sources(nshots), sphere_dirs(nshots)
    rec = np.repeat(receivers(), 3, axis=0)
    comp = np.tile(np.arange(3, dtype=np.int32), len(rec) // 3)
    P = M.WaveProblem(m, rho, mu, lam=lam, ncomp=3, precision="fp64")
    dt = 0.5 * 2.0 / np.sqrt(P.lambda_max(60, 1))
    P.close()
    res = {}
    for key, prec, ns in (("fp32_g64", "fp32_g64", nshots), ("fp32", "fp32", nshots),
                          ("fp64", "fp64", nshots), ("fp64_1", "fp64", 1)):
        P = M.WaveProblem(m, rho, mu, lam=lam, ncomp=3, nshots=ns, precision=prec)
        P.set_sources(srcs[:ns], pol=pol[:ns], shot=np.arange(ns))
        P.set_receivers(rec, comp)
        P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, max(warmup, steps, 16) + 64))
        tuned = tuner.tune_problem(P, steps=4, dt=dt)  # launch shape only: result-neutral
        wall, tr = timed_advance(P, dt, steps, warmup)
        # FP64 energy of the state after the timed steps (the pulse ends at
        # t = 2.5/f0, long before), and 100 untimed steps later: conserved
        uN = P.state()[1]
        e1 = P.energy()
        P.advance(100, traces=False)
        e2 = P.energy()
        prof = P.profile(4) if key in ("fp32_g64", "fp32") else None
        res[key] = dict(wall=wall, traces=tr, uN=uN, e1=e1, e2=e2, prof=prof, tile=tuned["tile"],
                        ncolors=P.info()["ncolors"])
        P.close()
        print(f"# elastic {key}: {res[key]['wall']:.3f} s", file=sys.stderr, flush=True)
    ndof = 3 * m.N
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    t64_path = res["fp64_1"]["wall"] * nshots
    head = res["fp32_g64"]
    launch_ms = head["prof"]["stiffness_launch_ms"]
    launch_bytes = elastic_bytes_per_step(m.E, m.nd, m.N, nshots, 4) / head["ncolors"]
    achieved = launch_bytes / (launch_ms * 1e-3) / 1e9
    f32_ms = res["fp32"]["prof"]["stiffness_launch_ms"]
    launch_flops = elastic_flops_per_step(m.E, order + 1, nshots) / head["ncolors"]
    return {
        "workload": f"elastic Q{order} {nx}^3, 3 layers, {nshots} shots, {steps} steps",
        "n_dof": ndof, "dt": dt,
        "precision": "FP32 state, FP64 geometric factors; energy in FP64",
        "fp32_g64_batch_s": head["wall"],
        "fp32_g64_batch_gdof_steps_s": ndof * steps * nshots / head["wall"] / 1e9,
        "fp32_batch_gdof_steps_s": ndof * steps * nshots / res["fp32"]["wall"] / 1e9,
        "fp64_batch_gdof_steps_s": ndof * steps * nshots / res["fp64"]["wall"] / 1e9,
        "fp64_one_shot_gdof_steps_s": ndof * steps / res["fp64_1"]["wall"] / 1e9,
        "time_to_solution_ratio_fp64_shot_by_shot_over_fp32_g64_batch": t64_path / head["wall"],
        "time_to_solution_ratio_fp64_shot_by_shot_over_fp32_batch":
            t64_path / res["fp32"]["wall"],
        "fp32_g64_rel_l2_state_vs_fp64": rel(head["uN"], res["fp64"]["uN"]),
        "fp32_g64_rel_l2_traces_vs_fp64": rel(head["traces"], res["fp64"]["traces"]),
        "fp32_rel_l2_traces_vs_fp64": rel(res["fp32"]["traces"], res["fp64"]["traces"]),
        "energy_fp64_check": {
            "fp32_g64_vs_fp64_max_rel": float(np.max(np.abs(head["e1"] - res["fp64"]["e1"])
                                                     / res["fp64"]["e1"])),
            "fp32_g64_drift_100_steps_max_rel": float(np.max(np.abs(head["e2"] - head["e1"])
                                                             / head["e1"])),
            "fp64_drift_100_steps_max_rel": float(np.max(np.abs(res["fp64"]["e2"] -
                                                                res["fp64"]["e1"])
                                                         / res["fp64"]["e1"])),
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"k_shots_el (tile {head['tile']}, one colour pass, FP32 state, "
                               "FP64 geometry)",
                     "algorithmic_bytes_per_launch": launch_bytes, "launch_ms": launch_ms,
                     "flops_per_launch": launch_flops,
                     "achieved_TFLOPs": launch_flops / (launch_ms * 1e-3) / 1e12},
        "roofline_fp32": {"bound": "hbm", "achieved": launch_bytes / (f32_ms * 1e-3) / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": launch_bytes / (f32_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "kernel": f"k_shots_el (tile {res['fp32']['tile']}, one colour pass, "
                                    "FP32 state and geometry: the workload's stated precision)",
                          "launch_ms": f32_ms,
                          "achieved_TFLOPs": launch_flops / (f32_ms * 1e-3) / 1e12,
                          "peak_TFLOPs_fp32_non_packed": 78.6,
                          "note": "packed FP32 instructions are disabled in this build, so the "
                                  "vector FP32 rate is half the 157.3 TFLOP/s packed peak"},
    }


def bench_2d(nx=64, order=4, steps=1000):
    """2-D acoustic Q4, nx^2 unit square, Dirichlet, Dirac at the centre (a
    vertex), Ricker f0 = 10, dt = 0.5 CFL, 16 receivers on the circle of
    radius 0.3; FP64 and FP32, graph-captured step loop against one
    launch sequence per step."""
    import montjoie_amd as M
    m = M.regular_mesh(2, nx, order)
    ang = 2 * np.pi * np.arange(16) / 16
    rec = np.stack([0.5 + 0.3 * np.cos(ang), 0.5 + 0.3 * np.sin(ang)], axis=1)
    P = M.WaveProblem(m, 1.0, 1.0, precision="fp64")
    dt = 0.5 * 2.0 / np.sqrt(P.lambda_max(60, 1))
    P.close()
    out = {"workload": f"2-D acoustic Q{order} {nx}^2, {steps} steps", "n_dof": m.N, "dt": dt}
    for prec in ("fp64", "fp32"):
        for graph in (True, False):
            P = M.WaveProblem(m, 1.0, 1.0, precision=prec)
            P.set_sources([(0.5, 0.5)])
            P.set_receivers(rec)
            P.set_pulse(M.ricker_pulse_table(10.0, 0.0, dt, steps + 32))
            P.leapfrog_init(dt)
            P.advance(32, traces=False, graph=graph)
            P.leapfrog_init(dt)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            P.advance(steps, traces=True, graph=graph)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            P.close()
            key = f"{prec}_{'graph' if graph else 'eager'}"
            out[key] = {"us_per_step": wall / steps * 1e6,
                        "gdof_steps_s": m.N * steps / wall / 1e9}
    return out


def _ref_acoustic_deck(d, nx, order, steps, dt):
    """Montjoie data file + Medit mesh of the CPU sample (acoustic, the
    headline's layered perturbed cube at a bounded size); returns the DoF count."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from oracle import oracle as O
    import decks as F
    os.makedirs(d, exist_ok=True)
    m = O.mesh(3, nx, order, 0.1, 2024)
    c = [1.0 + 2.0 * _unif(2025, k) for k in range(3)]
    lay = ((np.arange(m.E) // nx ** 2) * 3) // nx
    F.write_medit(os.path.join(d, "mesh.mesh"), m, lay + 1, bnd_ref=4)
    ini = ["TypeElement = HEXAHEDRON_LOBATTO", "TypeEquation = ACOUSTIC", "MeshPath = ./",
           "Frequency = 8.0 0.0", "FileMesh = mesh.mesh", "TypeMesh = HEXA",
           f"OrderDiscretization = {order}", "ConditionReference = 4 DIRICHLET"]
    ini += [f"MateriauDielec = {k + 1} ISOTROPE 1.0 {c[k] ** 2!r} 0.0" for k in range(3)]
    ini += ["TypeSource = SRC_DIRAC", "OriginePhase = 0.5 0.5 0.5",
            f"ParametersOutputPoint = 0.0 {steps * dt + dt / 2!r} {dt!r}",
            "FileOutputPoint = diffracPoint.dat totalPoint.dat", "SismoPoint = 0.3 0.4 0.6",
            "OutputFormat = BINARY DOUBLE", "ElectricOrMagnetic = 0", f"TimeStep = {dt!r}",
            f"TimeInterval = 0.0 {(steps - 0.5) * dt!r}", "OrderTimeScheme = 2 LEAP_FROG",
            "TemporalSource = RICKER", "PrintLevel = 1"]
    open(os.path.join(d, "case.ini"), "w").write("\n".join(ini) + "\n")
    return m.N


def _ref_elastic_deck(d, nx, order, steps, dt, shot=0):
    """Elastic sample (the C3 family: layered lambda, mu, rho, perturbed cube,
    one of the 16 Dirac shots with its direction); returns the DoF count."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from oracle import oracle as O
    import decks as F
    lam = [1.0 + 3.0 * _unif(2026, k) for k in range(3)]
    mu = [1.0 + 3.0 * _unif(2026, 3 + k) for k in range(3)]
    rho = [1.0 + _unif(2026, 6 + k) for k in range(3)]
    src = shot_sources(16)[shot]
    pol = sphere_dirs(16)[shot]
    F.write_deck(d, 3, nx, order, 0.1, 2024, mu, 8.0, dt, steps, tuple(src),
                 np.array([[0.3, 0.4, 0.6]]), elastic=True, lam_layers=lam, rho_layers=rho,
                 pol=tuple(pol))
    ini = os.path.join(d, "case.ini")
    txt = open(ini).read().replace("PrintLevel = 0", "PrintLevel = 1")  # prints the loop time
    open(ini, "w").write(txt)
    return 3 * O.mesh(3, nx, order, 0.1, 2024).N


def _ref_run(d, elastic, timeout=900):
    exe = os.path.join(ROOT, "oracle", "_ref", "mjref_elastic.x" if elastic else "mjref_acous.x")
    args = [exe, "case.ini"] + ([] if elastic else ["3"])
    return subprocess.Popen(args, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                            text=True)


def _ref_loop_seconds(proc, timeout=900):
    out, _ = proc.communicate(timeout=timeout)
    mt = re.search(r"Time to complete time iterations : ([0-9.eE+-]+)", out)
    if proc.returncode != 0 or not mt:
        return None
    return float(mt.group(1))


def cpu_reference(nx=24, order=5, steps=150, dt=1e-3):
    """Montjoie's own solver (oracle/_ref, built from the reference sources)
    on a bounded sample of the same workload family, one core."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "mjref_acous.x")):
        return None
    with tempfile.TemporaryDirectory(dir=ROOT, prefix=".mjref_") as d:
        ndof = _ref_acoustic_deck(d, nx, order, steps, dt)
        secs = _ref_loop_seconds(_ref_run(d, False))
        if secs is None:
            return None
        return dict(value=ndof * steps / secs / 1e9, unit="GDoF-steps/s", cores=1,
                    kind="reference",
                    sample=f"Montjoie's own leapfrog (oracle/_ref, g++ -O2), {nx}^3 Q{order} "
                           f"perturbed layered cube ({ndof} DoF), {steps} steps, time loop "
                           f"{secs:.2f} s, one core")


def cpu_reference_procs(elastic, nproc=8, nx=None, order=None, steps=None, dt=None):
    """Batch baseline of the reference: `nproc` concurrent single-core
    processes of Montjoie's own solver (the path has no threads), each on a
    bounded sample of the batch's workload family (acoustic: the 24^3 Q5
    layered cube; elastic: 12^3 Q4 layered cube, shots of the 16-shot set);
    aggregate throughput = nproc * DoF * steps / slowest time loop."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref",
                                       "mjref_elastic.x" if elastic else "mjref_acous.x")):
        return None
    nx = nx or (12 if elastic else 24)
    order = order or (4 if elastic else 5)
    steps = steps or (100 if elastic else 150)
    dt = dt or (5e-4 if elastic else 1e-3)
    with tempfile.TemporaryDirectory(dir=ROOT, prefix=".mjref_") as root:
        dirs = [os.path.join(root, f"p{k}") for k in range(nproc)]
        ndof = 0
        for k, d in enumerate(dirs):
            ndof = (_ref_elastic_deck(d, nx, order, steps, dt, shot=k) if elastic
                    else _ref_acoustic_deck(d, nx, order, steps, dt))
        t0 = time.perf_counter()
        procs = [_ref_run(d, elastic) for d in dirs]
        secs = [_ref_loop_seconds(p) for p in procs]
        wall = time.perf_counter() - t0
        if any(x is None for x in secs):
            return None
        slow = max(secs)
        kind = "elastic" if elastic else "acoustic"
        return dict(value=nproc * ndof * steps / slow / 1e9, unit="GDoF-steps/s", cores=nproc,
                    kind="reference",
                    sample=f"{nproc} concurrent processes of Montjoie's own {kind} leapfrog "
                           f"(oracle/_ref, g++ -O2, one core each), {nx}^3 Q{order} layered "
                           f"perturbed cube ({ndof} DoF each), {steps} steps; time loops "
                           f"{min(secs):.2f}-{slow:.2f} s, wall {wall:.2f} s",
                    one_process_value=ndof * steps / min(secs) / 1e9)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=48)
    ap.add_argument("--acc-steps", type=int, default=500,
                    help="step count of the untimed accuracy runs (the workload's 500)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32_g64", "fp32"], default=None,
                    help="run one precision only (profiling runs)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--quick", action="store_true",
                    help="skip the 16-shot batch, the adjoint and the many-problem batch")
    ap.add_argument("--tile", type=int, default=None,
                    help="stiffness tile code (default: autotuned per precision)")
    ap.add_argument("--xcd", type=int, default=None, help="XCD block order with --tile")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 64-problem uneven batch (prints its JSON)")
    ap.add_argument("--extra-only", choices=["batch", "elastic", "2d", "shots", "adjoint"],
                    default=None, help="only one secondary workload (prints its JSON)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {a.gpus}): the solver runs each "
              "problem on one device; partitioning a mesh over several GPUs is not built",
              file=sys.stderr)
        sys.exit(2)
    import montjoie_amd as M
    if a.batch_only:
        a.extra_only = "batch"
    if a.extra_only:
        fn = {"batch": bench_batch, "elastic": bench_elastic, "2d": bench_2d,
              "adjoint": bench_adjoint}.get(a.extra_only)
        if fn is None:
            m, mu, _ = workload(a.nx)
            lam = M.WaveProblem(m, 1.0, mu, precision="fp64")
            dt = 0.5 * 2.0 / np.sqrt(lam.lambda_max(60, 1))
            lam.close()
            print(json.dumps(bench_shots(m, mu, dt, a.steps, a.warmup)))
        else:
            print(json.dumps(fn()))
        return
    t_setup = time.perf_counter()
    m, mu, c = workload(a.nx)
    # dt = 0.5 x CFL limit from FP64 Lanczos on D^{-1} A
    P = M.WaveProblem(m, 1.0, mu, precision="fp64")
    lam = P.lambda_max(60, 1)
    P.close()
    dt = 0.5 * 2.0 / np.sqrt(lam)
    print(f"# setup: N={m.N} E={m.E} lambda_max={lam:.6e} dt={dt:.6e} "
          f"({time.perf_counter() - t_setup:.1f} s)", file=sys.stderr, flush=True)
    if a.calib_copy:
        x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("# calibration: 3 x 1 GiB device copies", file=sys.stderr, flush=True)
    res = {}
    for prec in ((a.only,) if a.only else ("fp64", "fp32_g64", "fp32")):
        res[prec] = run_gpu(m, mu, dt, prec, a.steps, a.warmup, tile=a.tile, nxcd=a.xcd,
                            acc_steps=a.acc_steps)
        r = res[prec]
        print(f"# {prec}: {r['wall'] / a.steps * 1e3:.4f} ms/step wall, "
              f"{r['dev_ms_step']:.4f} ms/step device, stiffness launch "
              f"{r['prof']['stiffness_launch_ms']:.4f} ms, tile {r['tile']} xcd {r['nxcd']}",
              file=sys.stderr, flush=True)
    if a.only:
        return
    ref = res["fp64"]
    for prec in ("fp32_g64", "fp32"):
        r = res[prec]
        r["err_state"] = float(np.linalg.norm(r["uN"] - ref["uN"]) / np.linalg.norm(ref["uN"]))
        r["err_traces"] = float(np.linalg.norm(r["traces"] - ref["traces"]) /
                                np.linalg.norm(ref["traces"]))
        r["err_energy"] = float(abs(r["energy"][0] - ref["energy"][0]) / abs(ref["energy"][0]))
    gdof = lambda r: m.N * a.steps / r["wall"] / 1e9  # noqa: E731
    ms = lambda r: r["wall"] / a.steps * 1e3  # noqa: E731
    head = res["fp32_g64"]
    nd = 216
    ncol = head["info"]["ncolors"]
    tri = head["geometry"] == "trilinear"
    launch_ms = head["prof"]["stiffness_launch_ms"]
    # bytes per step, no cache reuse credited: the stored-geometry algorithm
    # (per-node FP64 tensor + map + 4 FP32 state streams) and this kernel's own
    # (24 FP64 trilinear coefficients per element in place of the tensor)
    bytes_step = algorithmic_bytes_per_step(m.E, nd, m.N, 4, 8)
    launch_bytes = bytes_step / ncol
    achieved = launch_bytes / (launch_ms * 1e-3) / 1e9
    kernel_bytes = (trilinear_bytes_per_step(m.E, nd, m.N, 4, 8) if tri else bytes_step) / ncol
    node_slot = m.E / ncol * (nd * (4 * 4 + 4) + (24 * 8 if tri else 6 * 8 * nd))
    traffic = (measured_traffic("stiffness_q5_48cube_g64_counters.md")
               if (a.nx == 48 and tri) else None)
    flops = flops_per_step(m.E, 6)
    tflops = lambda r: flops / (r["wall"] / a.steps) / 1e12  # noqa: E731
    out = {
        "metric": "GDoF-steps/s, Q5 hex acoustic, FP32 vs FP64 at rel-L2 err; gradient/forward cost",
        "value": gdof(head),
        "unit": "GDoF-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms(head),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 state, fp64 geometric factors",
        "data": "synthetic",
        "config": {"workload": f"3-D acoustic Q5 hexahedra {a.nx}^3 unit cube, vertices "
                               "perturbed <=0.1h (seed 2024), 3 layers c~U[1,3] (seed 2025), "
                               "Dirac centre, Ricker f0=8, dt=0.5 CFL, 16 receivers",
                   "n_dof": m.N, "n_elem": m.E, "order": 5, "dt": dt, "shots": 1,
                   "precision": "FP32 state with FP64 geometric factors (MJW_FP32_G64); "
                                "energy and traces accumulated in FP64",
                   "parallelism": "single GPU"},
        "accuracy": {"steps": a.acc_steps,
                     "note": "untimed runs of the full step count from a zero state, each "
                             "precision against the FP64 GPU path (which matches Montjoie at "
                             "<= 1e-10 on the golden fixtures)",
                     "rel_l2_traces_vs_fp64": head["err_traces"],
                     "rel_l2_state_vs_fp64": head["err_state"],
                     "energy_rel_diff_vs_fp64": head["err_energy"],
                     "fp64_energy": float(ref["energy"][0])},
        "flops": {"per_step": flops, "note": "algorithmic count: 2*3*n per node and "
                  "direction for the gradient and the divergence, 18 for the tensor, 6 for "
                  "the update (no geometry rebuild counted)",
                  "achieved_TFLOPs": tflops(head), "peak_TFLOPs_fp32": 157.3,
                  "frac_fp32": tflops(head) / 157.3,
                  "fp64_achieved_TFLOPs": tflops(ref), "peak_TFLOPs_fp64": 78.6,
                  "fp64_frac": tflops(ref) / 78.6},
        "fp64": {"value": gdof(ref), "ms_per_step": ms(ref)},
        "fp32": {"value": gdof(res["fp32"]), "ms_per_step": ms(res["fp32"]),
                 "note": "FP32 geometric factors too (narrower than the workload asks)",
                 "rel_l2_state_vs_fp64": res["fp32"]["err_state"],
                 "rel_l2_traces_vs_fp64": res["fp32"]["err_traces"]},
        "fp32_g64_over_fp64_speedup": gdof(head) / gdof(ref),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "geometry": head["geometry"],
                     "algorithmic_bytes_per_launch": launch_bytes,
                     "algorithmic_bytes_note": "per colour pass: E/8 elements x 216 nodes x "
                                               "(6 x 8 B tensor + 4 B map) + N/8 nodes x 4 x 4 B "
                                               "state (stored-geometry algorithm)",
                     "own_bytes_per_launch": kernel_bytes,
                     "own_bytes_frac": kernel_bytes / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     # what a colour pass must move when no node survives in cache between
                     # passes (the 224 MB state exceeds the 32 MB of L2): every node slot
                     # of its elements (4 state streams + map) plus the coefficients
                     "node_slot_bytes_per_launch": node_slot,
                     "traffic_over_node_slot_bytes": None if traffic is None else
                     traffic / node_slot,
                     "traffic_GBs": None if traffic is None else
                     traffic / (launch_ms * 1e-3) / 1e9,
                     "kernel": kernel_name(head["tile"], tri, t="float").replace(
                         "float, float", "float, double") + " (one colour pass)",
                     "tile": head["tile"], "xcd_order": head["nxcd"],
                     "launch_ms": launch_ms,
                     "fp64_launch_ms": ref["prof"]["stiffness_launch_ms"],
                     "fp64_achieved_GBs": algorithmic_bytes_per_step(m.E, nd, m.N, 8, 8) / ncol /
                     (ref["prof"]["stiffness_launch_ms"] * 1e-3) / 1e9},
        "cpu_baseline": None,
    }
    out["host_operator"] = bench_host_apply(m, mu)
    print(f"# host operator: {out['host_operator']}", file=sys.stderr, flush=True)
    if not a.quick:
        out["batch_16_shots"] = bench_shots(m, mu, dt, a.steps, a.warmup)
        print(f"# 16 shots: {out['batch_16_shots']}", file=sys.stderr, flush=True)
        out["adjoint"] = bench_adjoint()
        print(f"# adjoint: {out['adjoint']}", file=sys.stderr, flush=True)
        out["batch_64_problems"] = bench_batch()
        print(f"# batch: {out['batch_64_problems']}", file=sys.stderr, flush=True)
        out["elastic_16_shots"] = bench_elastic()
        print(f"# elastic: {out['elastic_16_shots']}", file=sys.stderr, flush=True)
        out["acoustic_2d"] = bench_2d()
        print(f"# 2d: {out['acoustic_2d']}", file=sys.stderr, flush=True)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_reference()
        if not a.quick:
            # batch baselines: 8 concurrent single-core reference processes
            out["batch_64_problems"]["cpu_baseline_8proc"] = cpu_reference_procs(False)
            out["elastic_16_shots"]["cpu_baseline_8proc"] = cpu_reference_procs(True)
            print(f"# cpu 8 processes: {out['batch_64_problems']['cpu_baseline_8proc']}, "
                  f"{out['elastic_16_shots']['cpu_baseline_8proc']}", file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
