"""Device random stream and displacement field: Philox4x32-10 known answers,
positions equal to the numpy restatement bit for bit, apply_displacement_field
equal to the oracle, and add_random_particles (host or device stream) equal to
assigning the same positions with MassAssignor."""
import ctypes

import numpy as np
import pytest

from fftcorr_amd import _lib
from fftcorr_amd.catalog import add_random_particles
from fftcorr_amd.grid import ConfigSpaceGrid, apply_displacement_field
from fftcorr_amd.particle_mesh import MassAssignor
from oracle import fftcorr_oracle as O
from philox_ref import uniform_positions
from test_randoms_cpu import KAT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_device_philox_known_answers(ctr, key, want):
    out = (ctypes.c_uint32 * 4)()
    first = ctr[0] | (ctr[1] << 32)
    seed = key[0] | (key[1] << 32)
    _lib.check(_lib.lib().fc_random_raw(_lib.context().handle, out, 1, seed, first, ctr[2],
                                        ctr[3]))
    assert list(out) == list(want)


def test_device_uniform_positions_match_restatement():
    n = 10000
    lo, hi = np.array([1.0, -2.0, 3.0]), np.array([5.0, 2.0, 11.0])
    got = np.zeros((n, 3))
    _lib.check(_lib.lib().fc_random_uniform(_lib.context().handle, _lib.dptr(got), n,
                                            _lib.dptr(lo), _lib.dptr(hi), 12345, 77))
    np.testing.assert_array_equal(got, uniform_positions(n, lo, hi, 12345, first=77))


def _grid(shape=(16, 12, 20), cell=2.5, posmin=(10.0, -20.0, 5.0)):
    posmax = [posmin[i] + cell * shape[i] for i in range(3)]
    return ConfigSpaceGrid(shape, posmin, posmax)


def test_apply_displacement_matches_oracle():
    g = _grid()
    rng = np.random.default_rng(4)
    disp = rng.normal(size=g.shape + (3,))
    pos = rng.uniform(g.posmin, g.posmax, size=(5000, 3))
    got = apply_displacement_field(g, pos, disp)
    want = O.apply_displacement_field(pos, disp, g.posmin, g.cell_size, g.shape, False)
    np.testing.assert_array_equal(got, want)
    outside = pos.copy()
    outside[7, 1] = g.posmax[1] + 3.0
    with pytest.raises(ValueError, match="periodic_wrap"):
        apply_displacement_field(g, outside, disp)
    wrapped = apply_displacement_field(g, outside, disp, periodic_wrap=True)
    np.testing.assert_array_equal(
        wrapped, O.apply_displacement_field(outside, disp, g.posmin, g.cell_size, g.shape, True))


def _assign(pos, scheme, shape=(16, 12, 20), weight=1.0, wrap=True):
    g = _grid(shape)
    ma = MassAssignor(g, scheme, wrap)
    ma.add_particles(np.ascontiguousarray(pos), weight)
    return g.data.copy(), ma


@pytest.mark.parametrize("use_disp", [False, True])
def test_device_randoms_equal_assigning_the_same_positions(use_disp):
    g = _grid()
    rng = np.random.default_rng(8)
    disp = rng.normal(scale=0.7, size=g.shape + (3,)) if use_disp else None
    n = 30000
    st = add_random_particles(n, g, 1, particle_weight=0.5, periodic_wrap=True, disp=disp,
                              seed=99, batch_size=7000, device_rng=True)
    assert st.num_added == n
    pos = uniform_positions(n, g.posmin, g.posmax, 99)
    if use_disp:
        pos = O.apply_displacement_field(pos, disp, g.posmin, g.cell_size, g.shape, True)
    want, ma = _assign(pos, 1, weight=0.5)
    # Batches are assigned in turn, so sums agree to rounding, not bitwise.
    np.testing.assert_allclose(g.data, want, rtol=1e-12, atol=1e-12)
    assert st.totw == pytest.approx(ma.totw, rel=1e-12)


def test_host_rng_matches_reference_procedure():
    g = _grid()
    ma = add_random_particles(20000, g, 3, total_weight=100.0, periodic_wrap=True, seed=5,
                              batch_size=6000)
    rng = np.random.default_rng(5)
    pos = np.concatenate([rng.uniform(g.posmin, g.posmax, (m, 3)) for m in (6000, 6000, 6000,
                                                                            2000)])
    want, ref_ma = _assign(pos, 3, weight=100.0 / 20000)
    # The buffered batches are assigned in turn: sums agree to rounding.
    np.testing.assert_allclose(g.data, want, rtol=1e-12, atol=1e-15)
    assert ma.num_added == 20000 and ma.totw == pytest.approx(100.0)
