"""Edge cases of the HIP path: empty source / receiver sets, ragged element
counts for every launch shape, meshes without Dirichlet nodes, shots
without sources, pulse tables shorter than the run, state errors."""
import numpy as np
import pytest

import cases as CS
import montjoie_amd as M
from montjoie_amd._lib import MjwError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_no_receivers_no_sources():
    m = M.regular_mesh(3, 2, 3, 0.1, seed=1)
    P = M.WaveProblem(m, 1.0, 1.0)
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, 1e-3, 20))
    P.leapfrog_init(1e-3)
    assert P.advance(20) is None                  # no receivers: no traces
    un, un1 = P.state()
    assert not un.any() and not un1.any()         # no source, zero start: stays zero
    assert P.energy()[0] == 0.0


@pytest.mark.parametrize("nx,r", [(3, 5), (5, 3), (1, 6)])
def test_ragged_element_counts_every_tile(nx, r):
    """E per colour is not a multiple of the elements per block."""
    from montjoie_amd import tuner
    m = M.regular_mesh(3, nx, r, 0.1, seed=2)
    o = O.mesh(3, nx, r, 0.1, 2)
    P = M.WaveProblem(m, 1.0, 1.0 + (np.arange(m.E) % 3))
    G, _ = O.geometry(o, 1.0 + (np.arange(m.E) % 3), 1.0)
    u = np.random.default_rng(nx).standard_normal(m.N)
    ref = -O.stiffness(o, G, u, O.color(o))
    outs = []
    for t in tuner.tile_choices(P):
        P.set_tile(t)
        outs.append(P.apply_kh(u))
        assert CS.rel_l2(outs[-1], ref) < 1e-12, t
    assert all(np.array_equal(outs[0], x) for x in outs[1:])


@pytest.mark.parametrize("dim", [2, 3])
def test_no_dirichlet_constant_in_kernel(dim):
    m = M.regular_mesh(dim, 3, 4, 0.1, seed=3)
    P = M.WaveProblem(m, 1.0, 2.0, dirichlet=False)
    assert np.abs(P.apply_kh(np.ones(m.N))).max() < 1e-12


def test_shot_without_source_stays_zero():
    m = M.regular_mesh(3, 3, 3, 0.1, seed=4)
    P = M.WaveProblem(m, 1.0, 1.0, nshots=3)
    P.set_sources([(0.5, 0.5, 0.5)], shot=[1])
    P.set_receivers([(0.4, 0.4, 0.6)])
    P.set_pulse(M.ricker_pulse_table(8.0, 0.0, 1e-3, 60))
    P.leapfrog_init(1e-3)
    tr = P.run(60)
    assert not tr[:, 0].any() and not tr[:, 2].any() and np.abs(tr[:, 1]).max() > 0


def test_pulse_shorter_than_run_is_zero_afterwards():
    """Steps past the pulse table see no source: the run equals one with the
    table padded by zeros."""
    m = M.regular_mesh(3, 3, 3, 0.1, seed=5)
    tab = M.ricker_pulse_table(8.0, 0.0, 1e-3, 30)
    out = []
    for t in (tab, np.concatenate([tab, np.zeros(50)])):
        P = M.WaveProblem(m, 1.0, 1.0)
        P.set_sources([(0.5, 0.5, 0.5)])
        P.set_receivers([(0.3, 0.6, 0.5)])
        P.set_pulse(t)
        P.leapfrog_init(1e-3)
        out.append(P.run(80))
    assert np.array_equal(out[0], out[1])


def test_errors_are_reported_not_crashes():
    m = M.regular_mesh(3, 2, 2, 0.0)
    P = M.WaveProblem(m, 1.0, 1.0, nshots=2)
    with pytest.raises(MjwError):
        P.advance(5)                                   # before leapfrog_init
    with pytest.raises(MjwError):
        P.set_sources([(1.5, 0.5, 0.5)])               # outside the mesh
    with pytest.raises(MjwError):
        P.set_sources([(0.5, 0.5, 0.5)], shot=[2])     # shot index out of range
    with pytest.raises(MjwError):
        P.set_receivers([(0.5, -0.2, 0.5)])            # outside the mesh
    with pytest.raises(MjwError):
        P.set_receivers([(0.5, 0.5, 0.5)], comp=[1])   # acoustic: component 0 only
    with pytest.raises(MjwError):
        P.set_tile(3)                                  # not a tile of order 2
    with pytest.raises(MjwError):
        M.WaveProblem(m, 1.0, -1.0)                    # mu <= 0
