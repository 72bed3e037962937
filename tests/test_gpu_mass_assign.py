"""GPU MassAssignor (cc/particle_mesh/mass_assignor.h, distribution_functions.h)
against the oracle: grids to 1e-13 relative (summation order differs),
stats exact for counts, bitwise-repeatable runs."""
import numpy as np
import pytest

from fftcorr_amd.grid import ConfigSpaceGrid
from fftcorr_amd.particle_mesh import MassAssignor
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu

CASES = [((2, 2, 2), 0), ((2, 2, 2), 1), ((2, 2, 2), 3), ((5, 4, 6), 1), ((5, 4, 6), 3),
         ((16, 16, 16), 0), ((16, 16, 16), 1), ((16, 16, 16), 3), ((33, 17, 70), 1),
         ((64, 64, 64), 1), ((64, 64, 64), 3), ((1, 3, 40), 1)]


def _grid(shape, posmin, cell):
    posmin = np.asarray(posmin, float)
    return ConfigSpaceGrid(shape, posmin, posmin + cell * np.array(shape), cell_size=cell)


@pytest.mark.parametrize("shape,scheme", CASES)
@pytest.mark.parametrize("wrap", [False, True])
def test_against_oracle(shape, scheme, wrap):
    rng = np.random.default_rng(sum(shape) * 7 + scheme)
    n = 4000
    posmin = np.array([-10.0, 3.0, 100.0])
    cell = 7.0
    margin = 2 * cell if wrap else 0.0
    pos = rng.uniform(posmin - margin, posmin + cell * np.array(shape) + margin, size=(n, 3))
    # Cluster a third of the particles into a few cells.
    pos[: n // 3] = pos[0] + rng.normal(0, 0.3 * cell, size=(n // 3, 3))
    w = rng.uniform(0.5, 1.5, n)
    g = _grid(shape, posmin, cell)
    ma = MassAssignor(g, scheme, periodic_wrap=wrap, buffer_size=1000)
    ma.add_particles(pos, w)
    x, ok = O.grid_coords(pos, posmin, cell, shape, wrap)
    ref = O.assign_mass(np.zeros(shape), x[ok], w[ok], scheme)
    np.testing.assert_allclose(g.data, ref, rtol=1e-13, atol=1e-12)
    assert ma.num_added == ok.sum() and ma.num_skipped == (~ok).sum()
    assert ma.totw == pytest.approx(w[ok].sum(), rel=1e-13)
    assert ma.totwsq == pytest.approx((w[ok] ** 2).sum(), rel=1e-13)


def test_input_forms_and_known_answer():
    # test_mass_assignor.py:15-96: 2^3 grid, TSC, (n,4) / (n,3)+array / (n,3)+scalar.
    g = ConfigSpaceGrid([2, 2, 2], [2, -5, 77], [50, 45, 120], cell_size=25)
    ma = MassAssignor(g, distribution_scheme=1, buffer_size=100)
    ma.add_particles(np.array([[10, 10, 100, 12]], dtype=np.float64))
    assert ma.num_added == 1 and ma.totw == 12 and ma.totwsq == 144
    assert np.sum(g.data) == pytest.approx(12.0, rel=1e-15)
    rng = np.random.default_rng(0)
    pos = rng.uniform([2, -5, 77], [50, 45, 120], size=(10, 3))
    weight = np.arange(10, dtype=np.float64)
    posw = np.concatenate([pos, weight[:, None]], axis=1)
    sums = []
    for args in [(posw,), (pos, weight), (pos, 5.0)]:
        g.clear()
        ma.clear()
        ma.add_particles(*args)
        sums.append(np.sum(g.data))
        assert ma.num_added == 10
    assert sums[0] == pytest.approx(45.0) and sums[1] == pytest.approx(45.0)
    assert sums[2] == pytest.approx(50.0)


def test_input_errors():
    g = ConfigSpaceGrid([2, 2, 2], [2, -5, 77], [50, 45, 120], cell_size=25)
    ma = MassAssignor(g, distribution_scheme=1, buffer_size=100)
    with pytest.raises(RuntimeError):
        ma.add_particles(np.zeros((2, 3, 4)), 5.0)
    with pytest.raises(RuntimeError):
        ma.add_particles(np.zeros((10, 3), np.float32), 5.0)
    with pytest.raises(RuntimeError):
        ma.add_particles(np.zeros((10, 4))[:, 0:3], 5.0)
    ro = np.zeros((10, 4))
    ro.setflags(write=False)
    with pytest.raises(RuntimeError):
        ma.add_particles(ro)
    with pytest.raises(ValueError):
        ma.add_particles(np.zeros((10, 4)), 1.0)
    with pytest.raises(ValueError):
        ma.add_particles(np.zeros((10, 5)))
    with pytest.raises(ValueError):
        ma.add_particles(np.zeros((10, 3)), np.ones(9))


def test_bitwise_repeatable_and_fp32_storage():
    rng = np.random.default_rng(5)
    shape = (32, 32, 32)
    pos = rng.uniform(0, 320, size=(200_000, 3))
    pos[:50000] = 160 + rng.normal(0, 1.0, size=(50000, 3))
    grids = []
    for _ in range(2):
        g = _grid(shape, [0, 0, 0], 10.0)
        MassAssignor(g, 1, periodic_wrap=True).add_particles(pos, 1.0)
        grids.append(g.data.copy())
    assert np.array_equal(grids[0], grids[1])
    g32 = ConfigSpaceGrid(shape, [0, 0, 0], [320] * 3, cell_size=10.0, precision="fp32")
    MassAssignor(g32, 1, periodic_wrap=True).add_particles(pos, 1.0)
    np.testing.assert_allclose(g32.data, grids[0], rtol=1e-7, atol=1e-6)


def test_interlaced_shift_matches_shifted_positions():
    rng = np.random.default_rng(8)
    shape = (16, 16, 16)
    cell = 5.0
    pos = rng.uniform(0, 80, size=(3000, 3))
    g1 = _grid(shape, [0, 0, 0], cell)
    MassAssignor(g1, 3, periodic_wrap=True, shift=0.5).add_particles(pos, 1.0)
    g2 = _grid(shape, [0, 0, 0], cell)
    MassAssignor(g2, 3, periodic_wrap=True).add_particles(np.mod(pos + 0.5 * cell, 80.0), 1.0)
    np.testing.assert_allclose(g1.data, g2.data, rtol=1e-12, atol=1e-12)
