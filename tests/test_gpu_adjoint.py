"""Adjoint gradient on the GPU against the oracle's FP64 reference gradient
and central finite differences (FP64), and its torch.autograd.Function."""
import numpy as np
import pytest

import montjoie_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(dim, nx, r, pert, seed, prec="fp64", nsteps=90, dt=4e-3):
    from montjoie_amd.adjoint import AdjointSolver
    m = M.regular_mesh(dim, nx, r, pert, seed)
    o = O.mesh(dim, nx, r, pert, seed)
    mu0 = 1.0 + 0.5 * (np.arange(m.E) % 3)
    src = (0.43, 0.52) if dim == 2 else (0.43, 0.52, 0.47)
    rec = np.array([[0.3, 0.3], [0.7, 0.6], [0.55, 0.2]]) if dim == 2 else \
        np.array([[0.3, 0.3, 0.6], [0.7, 0.6, 0.4], [0.55, 0.2, 0.5]])
    P = M.WaveProblem(m, 1.0, mu0, precision=prec)
    P.set_sources(src)
    P.set_receivers(rec)
    pulse = M.ricker_pulse_table(10.0, 0.0, dt, nsteps)
    P.set_pulse(pulse)
    S = AdjointSolver(P, dt, nsteps, nsnap=7)
    return S, m, o, mu0, src, rec, pulse


@pytest.mark.parametrize("dim,nx,r", [(2, 4, 3), (3, 3, 3), (3, 3, 4)])
def test_adjoint_matches_oracle_gradient(dim, nx, r):
    S, m, o, mu0, src, rec, pulse = _setup(dim, nx, r, 0.1, 3)
    tr = S.forward()
    Po = O.acoustic_problem(o, mu0, 1.0)
    b = O.dirac(o, src)
    tro = O.forward_traces(Po, b, pulse, S.dt, S.nsteps, rec)
    assert np.linalg.norm(tr - tro) / np.linalg.norm(tro) < 1e-12
    dJ = np.sin(np.arange(tr.size)).reshape(tr.shape)
    g = S.backward(dJ)
    go = O.adjoint_gradient(Po, b, pulse, S.dt, S.nsteps, rec, dJ)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"dim {dim}: gradient rel-L2 vs oracle {err:.2e}")
    assert err < 1e-10


@pytest.mark.parametrize("r", [3, 4])
def test_adjoint_finite_differences_fp64(r):
    """Central differences (J(m + h d) - J(m - h d)) / 2h on 4 random
    directions within 1e-6 of the adjoint gradient (Q3 and Q4, the order of
    the 40^3 gradient workload)."""
    S, m, o, mu0, *_ = _setup(3, 3, r, 0.1, 5, nsteps=120)
    dobs = S.forward(mu0 * 1.1)

    def J(mu):
        tr = S.forward(mu)
        return 0.5 * np.sum((tr - dobs) ** 2), tr

    J0, tr0 = J(mu0)
    g = S.backward(tr0 - dobs)
    rng = np.random.default_rng(0)
    for _ in range(4):
        dm = rng.standard_normal(m.E)
        eps = 1e-4
        fd = (J(mu0 + eps * dm)[0] - J(mu0 - eps * dm)[0]) / (2 * eps)
        assert abs(fd - g @ dm) < 1e-6 * abs(fd)


def test_checkpoint_count_does_not_change_gradient():
    S, m, o, mu0, *_ = _setup(3, 3, 3, 0.1, 6)
    tr = S.forward()
    dJ = np.cos(np.arange(tr.size)).reshape(tr.shape)
    g1 = S.backward(dJ)
    S.nsnap = 2
    S.forward()
    g2 = S.backward(dJ)
    assert np.linalg.norm(g1 - g2) < 1e-11 * np.linalg.norm(g1)


@pytest.mark.parametrize("prec", ["fp32", "fp32_g64"])
def test_adjoint_fp32_close_to_fp64(prec):
    S64, m, o, mu0, *_ = _setup(3, 3, 4, 0.1, 7)
    S32, *_ = _setup(3, 3, 4, 0.1, 7, prec=prec)
    tr = S64.forward()
    dJ = np.sin(0.3 * np.arange(tr.size)).reshape(tr.shape)
    g64 = S64.backward(dJ)
    S32.forward()
    g32 = S32.backward(dJ)
    err = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"fp32 gradient rel-L2 vs fp64: {err:.2e}")
    assert err < 1e-4


def test_torch_autograd_function():
    import torch
    from montjoie_amd.adjoint import WaveTraces
    S, m, o, mu0, *_ = _setup(2, 4, 3, 0.1, 8)
    dobs = torch.from_numpy(S.forward(mu0 * 0.9))
    mu = torch.tensor(mu0, dtype=torch.float64, requires_grad=True)
    loss = 0.5 * ((WaveTraces.apply(mu, S) - dobs) ** 2).sum()
    loss.backward()
    tr = S.forward(mu0)
    g = S.backward(tr - dobs.numpy())
    assert np.allclose(mu.grad.numpy(), g, rtol=1e-12, atol=1e-14 * np.abs(g).max())


def test_adjoint_shot_loop_rebuilds_backward_graph():
    """One problem, shots after one another (sources and pulse reset between
    adjoint runs, as a shot-by-shot gradient does): each gradient equals that
    of a fresh problem with the same shot, bit for bit.  The backward step
    graph holds the source arrays and the pulse table, so resetting either
    replaces it."""
    from montjoie_amd.adjoint import AdjointSolver
    dim, nx, r, pert, seed, dt, nsteps = 3, 3, 3, 0.1, 4, 4e-3, 90
    rec = np.array([[0.3, 0.3, 0.6], [0.7, 0.6, 0.4], [0.55, 0.2, 0.5]])
    # a point inside an element (n^3 source entries) and one on a vertex (one entry)
    shots = [((0.43, 0.52, 0.47), 10.0), ((1.0 / 3.0, 2.0 / 3.0, 1.0 / 3.0), 12.0)]

    def problem():
        m = M.regular_mesh(dim, nx, r, pert, seed)
        P = M.WaveProblem(m, 1.0, 1.0 + 0.5 * (np.arange(m.E) % 3))
        P.set_receivers(rec)
        return P

    shared = problem()
    S = AdjointSolver(shared, dt, nsteps, nsnap=7)
    for src, f0 in shots:
        pulse = M.ricker_pulse_table(f0, 0.0, dt, nsteps)
        shared.set_sources(src)
        shared.set_pulse(pulse)
        tr = S.forward()
        dJ = np.sin(np.arange(tr.size)).reshape(tr.shape)
        g = S.backward(dJ)
        fresh = problem()
        fresh.set_sources(src)
        fresh.set_pulse(pulse)
        F = AdjointSolver(fresh, dt, nsteps, nsnap=7)
        assert np.array_equal(tr, F.forward())
        assert np.array_equal(g, F.backward(dJ))
        fresh.close()
