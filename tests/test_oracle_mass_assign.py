"""Pins grid coordinates (cc/grid/config_space_grid.h:31-46) and mass
assignment (cc/particle_mesh/distribution_functions.h:25-160,
cc/particle_mesh/mass_assignor.h:105-191)."""
import numpy as np
import pytest

from oracle import fftcorr_oracle as O

SHAPE = (2, 2, 2)
POSMIN = (2.0, -5.0, 77.0)
CELL = 25.0


def test_grid_coords_known_answers():
    # test_grid.py:112-119 inputs with the grid of test_grid.py:7-16.
    x, ok = O.grid_coords([[2, -4, 79]], POSMIN, CELL, SHAPE, False)
    assert ok[0]
    np.testing.assert_allclose(x[0], [0.0, 0.04, 0.08], rtol=1e-15)
    _, ok = O.grid_coords([[1, 0, 77]], POSMIN, CELL, SHAPE, False)
    assert not ok[0]
    x, ok = O.grid_coords([[1, 0, 77]], POSMIN, CELL, SHAPE, True)
    assert ok[0]
    np.testing.assert_allclose(x[0], [1.96, 0.2, 0.0], rtol=1e-15)


def test_grid_coords_against_reference(ref):
    rng = np.random.default_rng(1)
    pos = rng.uniform(-60, 140, size=(300, 3))
    for wrap in (False, True):
        x, ok = O.grid_coords(pos, POSMIN, CELL, SHAPE, wrap)
        for p, xo, oko in zip(pos, x, ok):
            okr, xr = ref.grid_coords(SHAPE, POSMIN, CELL, wrap, p)
            assert okr == oko
            if okr:
                np.testing.assert_array_equal(xo, xr)


def test_tsc_single_particle_known_answer():
    # The particle of test_mass_assignor.py:40-41 sits at grid coords
    # (0.32, 0.6, 0.92); the 27 TSC weights wrap onto the 8 cells of a 2^3 grid
    # and must sum to its weight, 12.
    x, ok = O.grid_coords([[10, 10, 100]], POSMIN, CELL, SHAPE, False)
    np.testing.assert_allclose(x[0], [0.32, 0.6, 0.92], rtol=1e-15)
    g = O.assign_mass(np.zeros(SHAPE), x, [12.0], O.TRIANGULAR_SHAPED_CLOUD)
    assert g.sum() == pytest.approx(12.0, rel=1e-15)
    # Along x: cell 0 gets x0 = 0.5+r-r^2, cell 1 gets xm+xp.
    r = 0.32
    fx = np.array([0.5 + r - r * r, 0.5 * (1 - r) ** 2 + 0.5 * r * r])
    np.testing.assert_allclose(g.sum(axis=(1, 2)), 12 * fx, rtol=1e-14)


@pytest.mark.parametrize("scheme", [O.NEAREST_CELL, O.TRIANGULAR_SHAPED_CLOUD])
@pytest.mark.parametrize("shape", [(2, 2, 2), (5, 4, 6), (16, 16, 16)])
@pytest.mark.parametrize("wrap", [False, True])
def test_against_reference(ref, scheme, shape, wrap):
    rng = np.random.default_rng(sum(shape) + scheme)
    n = 500
    posmin = np.array([-10.0, 3.0, 100.0])
    cell = 7.0
    lo = posmin - (cell * 2 if wrap else 0)
    hi = posmin + cell * np.array(shape) + (cell * 2 if wrap else 0)
    pos = rng.uniform(lo, hi, size=(n, 3))
    posw = np.concatenate([pos, rng.uniform(0.5, 1.5, (n, 1))], axis=1)
    gref, st = ref.mass_assign(shape, posmin, cell, scheme, wrap, posw, buffer_size=64)
    x, ok = O.grid_coords(pos, posmin, cell, shape, wrap)
    g = O.assign_mass(np.zeros(shape), x[ok], posw[ok, 3], scheme)
    np.testing.assert_allclose(g, gref, rtol=1e-13, atol=1e-13)
    assert st["num_added"] == ok.sum() and st["num_skipped"] == (~ok).sum()
    assert st["totw"] == pytest.approx(posw[ok, 3].sum(), rel=1e-13)


def test_cic_closed_form_and_invariants():
    rng = np.random.default_rng(5)
    shape = (8, 9, 10)
    x = rng.uniform(0, 1, (400, 3)) * np.array(shape)
    w = rng.uniform(0.5, 2, 400)
    g = O.assign_mass(np.zeros(shape), x, w, O.CLOUD_IN_CELL)
    assert g.sum() == pytest.approx(w.sum(), rel=1e-13)
    # Brute force: weight of cell c from a particle at x is prod(max(0, 1-|x - (c+1/2)|))
    # taken over periodic images.
    gb = np.zeros(shape)
    for p, wp in zip(x, w):
        f = []
        for d in range(3):
            c = np.arange(shape[d]) + 0.5
            dist = np.abs(((p[d] - c) + shape[d] / 2) % shape[d] - shape[d] / 2)
            f.append(np.maximum(0.0, 1 - dist))
        gb += wp * f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :]
    np.testing.assert_allclose(g, gb, rtol=1e-12, atol=1e-13)
    # CIC reproduces a linear field exactly away from the wrap: the weighted
    # mean cell centre equals the particle position.
    p = np.array([[3.3, 4.7, 5.5]])
    g1 = O.assign_mass(np.zeros(shape), p, [1.0], O.CLOUD_IN_CELL)
    centres = [np.arange(s) + 0.5 for s in shape]
    for d in range(3):
        marg = g1.sum(axis=tuple(i for i in range(3) if i != d))
        assert np.dot(marg, centres[d]) == pytest.approx(p[0, d], rel=1e-14)


def test_d12_table_invariants(ref):
    # Each sub-cell sample's 8 cell weights sum to one (mass conservation),
    # and the weights' first moment over the offsets -1..6 moves by one
    # sample (1/64 cell) per sample step, to within 1%.
    tab = ref.wavelet_table().reshape(8, 64)
    np.testing.assert_allclose(tab.sum(axis=0), 1.0, atol=1e-9)
    step = np.diff(np.arange(-1, 7) @ tab)
    np.testing.assert_allclose(np.abs(step), 1 / 64, rtol=0.01)


@pytest.mark.parametrize("shape", [(8, 8, 8), (9, 12, 10), (16, 16, 16)])
@pytest.mark.parametrize("wrap", [False, True])
def test_d12_against_reference(ref, shape, wrap):
    rng = np.random.default_rng(sum(shape) * 3 + wrap)
    n = 400
    posmin = np.array([-10.0, 3.0, 100.0])
    cell = 7.0
    lo = posmin - (cell * 2 if wrap else 0)
    hi = posmin + cell * np.array(shape) + (cell * 2 if wrap else 0)
    pos = rng.uniform(lo, hi, size=(n, 3))
    posw = np.concatenate([pos, rng.uniform(-1.5, 1.5, (n, 1))], axis=1)
    gref, st = ref.mass_assign(shape, posmin, cell, O.WAVELET, wrap, posw, buffer_size=64)
    x, ok = O.grid_coords(pos, posmin, cell, shape, wrap)
    tab = ref.wavelet_table()
    g = O.assign_mass(np.zeros(shape), x[ok], posw[ok, 3], O.WAVELET, tab)
    np.testing.assert_allclose(g, gref, rtol=1e-12, atol=1e-13)
    assert g.sum() == pytest.approx(posw[ok, 3].sum(), abs=1e-9)
    assert st["num_added"] == ok.sum()
