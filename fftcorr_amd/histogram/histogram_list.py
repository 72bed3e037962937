"""HistogramList (fftcorr/histogram/histogram_list.pyx:9-44,
cc/histogram/histogram_list.h:10-67) with accumulate() on the GPU.

bins, counts (int32 [n, nbins]) and hist_values (float64 [n, nbins]) are
read-only views, as in the reference; accumulate(ih, x, y) bins x with
floor((x - minval) / binsize), drops bins outside [0, nbins) and adds y in
FP64 with a deterministic order.
"""

import math

import numpy as np

from fftcorr_amd import _lib


class HistogramList:
    def __init__(self, n, minval, maxval, binsize, device=None):
        n = int(n)
        minval = float(minval)
        maxval = float(maxval)
        binsize = float(binsize)
        if n <= 0:
            raise ValueError("n must be positive.")
        if minval >= maxval:
            raise ValueError("maxval must be greater than minval.")
        if binsize <= 0:
            raise ValueError("binsize must be positive.")
        nbins = int(math.floor((maxval - minval) / binsize))
        if nbins <= 0:
            raise ValueError("Shape must be nonnegative.")
        self._minval = minval
        self._binsize = binsize
        self._nbins = nbins
        self._device = device
        self._bins_rw = minval + (np.arange(nbins) + 0.5) * binsize
        self._counts_rw = np.zeros((n, nbins), np.intc)
        self._hist_rw = np.zeros((n, nbins), np.float64)
        self._bins = self._bins_rw.view()
        self._counts = self._counts_rw.view()
        self._hist_values = self._hist_rw.view()
        for a in (self._bins, self._counts, self._hist_values):
            a.setflags(write=False)

    @property
    def nbins(self):
        return self._nbins

    @property
    def bins(self):
        return self._bins

    @property
    def counts(self):
        return self._counts

    @property
    def hist_values(self):
        return self._hist_values

    def accumulate(self, ih, x, y):
        x = np.asarray(x)
        y = np.asarray(y)
        if x.shape != y.shape:
            raise ValueError("x and y should have the same dimensions")
        x = x.flatten()
        y = y.flatten()
        for a in (x, y):
            if a.dtype != np.float64:
                raise RuntimeError("Numpy array does not have the expected type.")
        ih = int(ih)
        if ih < 0 or ih >= self._counts_rw.shape[0]:
            raise IndexError("ih out of range.")
        ctx = _lib.context(self._device)
        crow = np.ascontiguousarray(self._counts_rw[ih])
        hrow = np.ascontiguousarray(self._hist_rw[ih])
        _lib.check(_lib.lib().fc_hist_accumulate(
            ctx.handle, self._minval, self._binsize, self._nbins, _lib.dptr(x), _lib.dptr(y),
            x.size, _lib.iptr(crow), _lib.dptr(hrow)))
        self._counts_rw[ih] = crow
        self._hist_rw[ih] = hrow

    def reset(self):
        self._counts_rw[...] = 0
        self._hist_rw[...] = 0.0
