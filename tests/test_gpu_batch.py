"""Batched problems of uneven size and order: every problem against the
oracle (orders 3-7), bitwise equal to solving each problem alone; graph ==
eager; autotuning is result-neutral."""
import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _spec(nx, r, seed):
    """(mu per element, source, receivers) of a batch test problem."""
    c = 2.0 + 0.5 * np.sin(np.arange(nx ** 3) * 0.37 + seed)
    rng = np.random.default_rng(seed)
    return c * c, 0.3 + 0.4 * rng.random(3), 0.2 + 0.6 * rng.random((4, 3))


def _problem(nx, r, seed, prec="fp32"):
    m = M.regular_mesh(3, nx, r, 0.1, seed)
    mu, src, rec = _spec(nx, r, seed)
    P = M.WaveProblem(m, 1.0, mu, precision=prec)
    P.set_sources(src)
    P.set_receivers(rec)
    dt = 0.5 * 2 / np.sqrt(P.lambda_max(40))
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, dt, 200))
    return P, dt


SPECS = [(3, 3, 1), (2, 5, 2), (4, 2, 3), (2, 7, 4), (3, 4, 5)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_equals_individual_runs(prec):
    from montjoie_amd.batch import Batch
    probs = [_problem(*s, prec=prec) for s in SPECS]
    alone = []
    for P, dt in probs:
        P.leapfrog_init(dt)
        alone.append(P.advance(40, graph=False)[:, 0, :])
        P.leapfrog_init(dt)
    B = Batch([P for P, _ in probs])
    # every work order: one launch per colour (0, 1) or per (colour, order)
    # with blocks of about 256, 512, 128 threads (2, 3, 4)
    for mode in (0, 1, 2, 3, 4):
        print(f"batch order_mode {mode}", flush=True)
        B.config(mode)
        for P, dt in probs:
            P.leapfrog_init(dt)
        B.advance(40, graph=True)
        for i, (P, dt) in enumerate(probs):
            assert np.array_equal(B.traces(i), alone[i]), (mode, i)
    B.close()


# orders 3..7, as the 64-problem workload draws them
ORACLE_SPECS = [(3, 3, 1), (2, 4, 2), (3, 5, 3), (2, 6, 4), (2, 7, 5)]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32_g64", 1e-5), ("fp32", 1e-5)])
def test_batch_matches_oracle(prec, tol):
    """Each problem of one batch (orders 3-7, perturbed meshes, graph-captured
    step loop) against the oracle's FP64 leapfrog of the same problem."""
    from montjoie_amd.batch import Batch
    probs = [_problem(*s, prec=prec) for s in ORACLE_SPECS]
    for P, dt in probs:
        P.leapfrog_init(dt)
    B = Batch([P for P, _ in probs])
    n = 64
    B.advance(n, graph=True)
    for i, ((nx, r, seed), (P, dt)) in enumerate(zip(ORACLE_SPECS, probs)):
        mu, src, rec = _spec(nx, r, seed)
        o = O.mesh(3, nx, r, 0.1, seed)
        _, _, otr = O.leapfrog(O.acoustic_problem(o, mu, 1.0), O.dirac(o, src),
                               O.pulse_table(8.0, 0.0, dt, n), dt, n, rec)
        err = CS.rel_l2(B.traces(i), otr[:n])
        print(f"{prec} problem {i} (Q{r}, {nx}^3): rel-L2 vs oracle {err:.2e}")
        assert err < tol, (i, err)
    B.close()


def test_batch_graph_equals_eager():
    from montjoie_amd.batch import Batch
    probs = [_problem(*s) for s in SPECS]
    B = Batch([P for P, _ in probs])
    for P, dt in probs:
        P.leapfrog_init(dt)
    B.advance(37, graph=True)
    a = [B.traces(i) for i in range(len(probs))]
    for P, dt in probs:
        P.leapfrog_init(dt)
    B.advance(37, graph=False)
    for i in range(len(probs)):
        assert np.array_equal(a[i], B.traces(i))
    B.close()


def test_tuners_are_result_neutral():
    from montjoie_amd import tuner
    from montjoie_amd.batch import Batch
    P, dt = _problem(3, 5, 9)
    P.leapfrog_init(dt)
    ref = P.advance(30, graph=False)
    best = tuner.tune_problem(P, steps=3, dt=dt, use_cache=False)
    assert best["tile"] in tuner.tile_choices(P)
    got = P.advance(30, graph=False)
    assert np.array_equal(ref, got)
    probs = [_problem(*s) for s in SPECS[:3]]
    B = Batch([Q for Q, _ in probs])
    for Q, d in probs:
        Q.leapfrog_init(d)
    B.advance(20)
    a = [B.traces(i) for i in range(3)]
    res = tuner.tune_batch(B, steps=2)
    for Q, d in probs:
        Q.leapfrog_init(d)
    B.advance(20)
    for i in range(3):
        assert np.array_equal(a[i], B.traces(i))
    assert res["order_mode"] in (0, 1, 2, 3, 4)
    B.close()
