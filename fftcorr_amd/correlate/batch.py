"""Batches of same-geometry periodic autocorrelations on one GPU.

``PeriodicBatch`` holds ``nbatch`` density grids back to back in HBM and one
PeriodicCorrelator geometry (cc/correlate/correlator.h:255-357), and
correlates members 0..n-1 with one launch per pass (z, y, fused x, pruned
inverse, histograms) over all of them.  With ``use_graph=True`` the launch
sequence is captured once as a HIP graph per member count and replayed.
Each member's results equal those of a PeriodicCorrelator on the same grid
(same kernels, same reduction order).
"""

import ctypes

import numpy as np

from fftcorr_amd import _lib
from fftcorr_amd._table import make_table


class _MemberGrid(_lib.DeviceGrid):
    """A batch member's grid: an fc_grid owned by the batch.  It holds no
    reference to the batch (so ``del batch`` frees the batch's HBM at once);
    when the batch closes, the member's handle is cleared and later calls
    through it fail with an error instead of touching freed memory."""

    def __init__(self, ctx, handle, shape, precision):
        self.ctx = ctx
        self.shape = shape
        self.precision = precision
        self.handle = handle

    def close(self):
        self.handle = None


class PeriodicBatch:
    def __init__(self, shape, cell_size, window_correct, rmax, dr, kmax, dk, maxell, nbatch,
                 precision="mixed", device=None, use_graph=True):
        self._shape = tuple(int(s) for s in shape)
        if len(self._shape) != 3:
            raise ValueError("shape must have 3 entries")
        self._ctx = _lib.context(device)
        self._prec = _lib.precision_code(precision)
        self.nbatch = int(nbatch)
        s = np.ascontiguousarray(self._shape, np.intc)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().fc_batch_create(
            self._ctx.handle, _lib.iptr(s), float(cell_size), int(window_correct), float(rmax),
            float(dr), float(kmax), float(dk), int(maxell), self._prec, self.nbatch,
            ctypes.byref(h)))
        self._handle = h
        _lib.check(_lib.lib().fc_batch_set_graph(h, int(bool(use_graph))))
        corr = ctypes.c_void_p()
        _lib.check(_lib.lib().fc_batch_correlator(h, ctypes.byref(corr)))
        nr, nk, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rs = np.zeros(3, np.intc)
        ks = np.zeros(3, np.intc)
        _lib.check(_lib.lib().fc_corr_geometry(corr, ctypes.byref(nr), ctypes.byref(nk),
                                               ctypes.byref(ne), _lib.iptr(rs), _lib.iptr(ks)))
        self._nr, self._nk, self._ne = nr.value, nk.value, ne.value
        self.rshape = tuple(int(v) for v in rs)
        self.kshape = tuple(int(v) for v in ks)
        self._r = np.zeros(self._nr)
        self._k = np.zeros(self._nk)
        _lib.check(_lib.lib().fc_corr_results(corr, _lib.dptr(self._r), None, None,
                                              _lib.dptr(self._k), None, None, None))
        gprec = _lib.grid_precision_code(precision)
        self._grids = []
        for i in range(self.nbatch):
            g = ctypes.c_void_p()
            _lib.check(_lib.lib().fc_batch_grid(h, i, ctypes.byref(g)))
            self._grids.append(_MemberGrid(self._ctx, g, self._shape, gprec))

    def close(self):
        """Destroys the batch and frees its device memory now; the member
        grids become invalid."""
        for g in getattr(self, "_grids", ()):
            g.close()
        h = getattr(self, "_handle", None)
        self._handle = None
        if h:
            _lib.check(_lib.lib().fc_batch_destroy(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _live(self):
        if not self._handle:
            raise RuntimeError("the batch has been closed")
        return self._handle

    def grid(self, i):
        """Member i's device grid (fill with fc_mass_assign / upload, then normalize)."""
        self._live()
        return self._grids[i]

    def assign(self, i, positions, weight=None, scheme=1, periodic_wrap=True, shift=0.0):
        """Mass-assigns an (n, 3) or (n, 4) float64 catalogue (survey
        coordinates; the member grids start at the origin with the batch's
        cell size) to member i, then turns it into delta = rho / mean - 1."""
        pos = np.ascontiguousarray(positions, np.float64)
        if pos.ndim != 2 or pos.shape[1] not in (3, 4):
            raise ValueError("positions must be (n, 3) or (n, 4)")
        g = self.grid(i)
        g.clear()
        st = _lib.MaStats()
        w = None if weight is None or np.isscalar(weight) else np.ascontiguousarray(weight,
                                                                                     np.float64)
        ws = 1.0 if weight is None else (float(weight) if np.isscalar(weight) else 1.0)
        _lib.check(_lib.lib().fc_mass_assign(g.handle, _lib.dptr(pos), pos.shape[0], pos.shape[1],
                                             _lib.dptr(w) if w is not None else None, ws,
                                             int(scheme), int(bool(periodic_wrap)), float(shift),
                                             ctypes.byref(st)))
        g.normalize(st.totw / float(np.prod(self._shape)))
        return st

    def autocorrelate(self, n=None):
        """Correlates members 0..n-1 (all by default)."""
        n = self.nbatch if n is None else int(n)
        _lib.check(_lib.lib().fc_batch_autocorrelate(self._live(), n))

    @property
    def last_time_ms(self):
        t = ctypes.c_double()
        _lib.check(_lib.lib().fc_batch_last_time_ms(self._live(), ctypes.byref(t)))
        return t.value

    def results(self, i):
        """Member i: dict of bin centres, int32 counts, FP64 histograms and zero lag."""
        rh = np.zeros((self._ne, self._nr))
        rc = np.zeros((self._ne, self._nr), np.intc)
        kh = np.zeros((self._ne, self._nk))
        kc = np.zeros((self._ne, self._nk), np.intc)
        z = ctypes.c_double()
        _lib.check(_lib.lib().fc_batch_results(self._live(), int(i), _lib.dptr(rh), _lib.iptr(rc),
                                               _lib.dptr(kh), _lib.iptr(kc), ctypes.byref(z)))
        return {"r": self._r.copy(), "r_counts": rc, "r_hist": rh, "k": self._k.copy(),
                "k_counts": kc, "k_hist": kh, "zerolag": z.value}

    def tables(self, i, squeeze=True):
        """Member i as (power_spectrum_table, correlation_table), like
        PeriodicCorrelator.autocorrelate."""
        res = self.results(i)
        out = []
        for x, h, c, names in ((res["k"], res["k_hist"], res["k_counts"],
                                ("k", "ps", "histogram", "count")),
                               (res["r"], res["r_hist"], res["r_counts"],
                                ("r", "xi", "histogram", "count"))):
            with np.errstate(invalid="ignore", divide="ignore"):
                ratio = h / c
            if squeeze:
                ratio = np.squeeze(ratio)
                h = np.squeeze(h)
            out.append(make_table((x, np.transpose(ratio), np.transpose(h), c[0, :]), names))
        return tuple(out)
