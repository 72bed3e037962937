"""Interlaced correlations with FP32 spectra run the z and y forward passes of
both grids as one launch each (the two spectra back to back in one buffer);
grids of two storage precisions take the separate launches.  Both routes
compute every element with the same arithmetic, so a mixed-form (and an
FP32-form) correlation of an FP32 grid pair and of the same values with grid
a stored in FP64 agree bitwise, and the correlator's footprint covers the
buffers the paired route holds."""
import ctypes

import numpy as np
import pytest

from fftcorr_amd import _lib
from fftcorr_amd.correlate import PeriodicCorrelator

pytestmark = pytest.mark.gpu

N, BOX = 64, 500.0


def _grid(values, precision):
    g = _lib.DeviceGrid((N, N, N), (0, 0, 0), BOX / N, precision)
    g.upload(values)
    return g


def _correlate(c, ga, gb):
    _lib.check(_lib.lib().fc_corr_autocorrelate_interlaced(c._handle, ga.handle, gb.handle))
    c._fetch()
    return {k: np.array(v, copy=True) for k, v in c.results().items()
            if isinstance(v, np.ndarray)}


def _footprint(c):
    held, full = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(_lib.lib().fc_corr_footprint(c._handle, 1, ctypes.byref(held),
                                            ctypes.byref(full)))
    return held.value, full.value


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_paired_forward_matches_separate_launches(precision):
    rng = np.random.default_rng(31)
    # FP32-representable values, so the FP64 copy of grid a holds the same.
    a = rng.standard_normal((N, N, N)).astype(np.float32).astype(np.float64)
    b = rng.standard_normal((N, N, N)).astype(np.float32).astype(np.float64)
    c = PeriodicCorrelator((N, N, N), BOX / N, 3, 60.0, 5.0, 0.3, 0.02, 4,
                           precision=precision)
    ga, gb = _grid(a, "fp32"), _grid(b, "fp32")
    paired = _correlate(c, ga, gb)
    held, full = _footprint(c)
    assert held <= full  # the footprint covers what the paired route holds
    separate = _correlate(c, _grid(a, "fp64"), gb)
    assert set(paired) == set(separate)
    for k in paired:
        np.testing.assert_array_equal(paired[k], separate[k], err_msg=k)
    # Calls alternate between the routes on one correlator.
    again = _correlate(c, ga, gb)
    for k in paired:
        np.testing.assert_array_equal(again[k], paired[k], err_msg=k)
