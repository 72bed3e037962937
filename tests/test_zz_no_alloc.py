"""fc_hist_accumulate_device_noalloc (the entry point that allocates nothing)
against fc_hist_accumulate on the same inputs: counts and FP64 sums equal
bit for bit, on two calls in a row with the second input smaller, rows
accumulated in place; nbins beyond the scratch limit is rejected.

The caller-owned device inputs and rows are allocated here with the HIP
runtime the library itself runs on (libamdhip64, already loaded with it)."""
import ctypes

import numpy as np
import pytest

from fftcorr_amd import _lib

pytestmark = pytest.mark.gpu

MAX_BINS = 682
H2D, D2H = 1, 2


def _hip():
    _lib.lib()  # loads libamdhip64 as the library's dependency
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return h


class DevArray:
    """Device copy of a numpy array (hipMalloc'd by the caller)."""

    def __init__(self, host):
        self.hip = _hip()
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 1)) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def __del__(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)


def _host(minval, binsize, nbins, x, y, counts, hist):
    _lib.check(_lib.lib().fc_hist_accumulate(
        _lib.context().handle, minval, binsize, nbins, _lib.dptr(x), _lib.dptr(y), len(x),
        _lib.iptr(counts), _lib.dptr(hist)))


def _device(minval, binsize, nbins, x, y, dcounts, dhist):
    dx, dy = DevArray(x), DevArray(y)
    ctx = _lib.context()
    _lib.check(_lib.lib().fc_hist_accumulate_device_noalloc(
        ctx.handle, minval, binsize, nbins, dx.ptr, dy.ptr, len(x), dcounts.ptr, dhist.ptr))
    ctx.synchronize()


@pytest.mark.parametrize("nbins", [1, 40, MAX_BINS])
def test_noalloc_equals_allocating_entry_point(nbins):
    rng = np.random.default_rng(nbins)
    minval, binsize = -3.0, 0.017
    counts = np.zeros(nbins, np.intc)
    hist = np.zeros(nbins)
    dcounts = DevArray(counts)
    dhist = DevArray(hist)
    for n in (300_000, 4_321):  # the second call smaller than the first
        x = rng.uniform(minval - 1.0, minval + binsize * nbins + 1.0, n)
        x[:50] = minval + binsize * rng.integers(0, nbins + 1, 50)  # bin edges, x == max
        y = rng.normal(size=n)
        _host(minval, binsize, nbins, x, y, counts, hist)
        _device(minval, binsize, nbins, x, y, dcounts, dhist)
        np.testing.assert_array_equal(dcounts.get(), counts)
        np.testing.assert_array_equal(dhist.get(), hist)
    assert counts.sum() > 0


def test_noalloc_rejects_oversized_and_bad_arguments():
    d = DevArray(np.zeros(MAX_BINS + 1))
    c = DevArray(np.zeros(MAX_BINS + 1, np.intc))
    ctx = _lib.context()
    L = _lib.lib()
    p = d.ptr
    pc = c.ptr
    rc = L.fc_hist_accumulate_device_noalloc(ctx.handle, 0.0, 1.0, MAX_BINS + 1, p, p, 10, pc, p)
    assert rc == _lib.FC_ERR_INVALID and b"no-allocation limit" in L.fc_last_error()
    rc = L.fc_hist_accumulate_device_noalloc(ctx.handle, 0.0, 0.0, 10, p, p, 10, pc, p)
    assert rc == _lib.FC_ERR_INVALID
    rc = L.fc_hist_accumulate_device_noalloc(ctx.handle, 0.0, 1.0, 10, None, None, 10, pc, p)
    assert rc == _lib.FC_ERR_INVALID
    # n == 0 is a no-op that leaves the rows alone.
    assert L.fc_hist_accumulate_device_noalloc(ctx.handle, 0.0, 1.0, 10, None, None, 0, pc,
                                               p) == 0
    assert not d.get().any()
