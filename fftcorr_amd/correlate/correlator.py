"""PeriodicCorrelator and Correlator (fftcorr/correlate/correlator.pyx:12-184,
cc/correlate/correlator.h:18-505) running on the GPU.

Same constructors, ``from_grid_spec``, ``autocorrelate``, ``cross_correlate``,
``set_grid2`` and ``set_grid2_fft`` as the reference, returning
(power_spectrum_table, correlation_table).  Additions:
  * ``precision``: "fp64" (default; every stage in FP64 like the reference),
    "fp32" (FP32 grids/FFTs/products, FP64 sums and histograms) or "mixed"
    (FP32 storage as "fp32"; the z and y forward FFT passes compute in FP64
    and round to FP32 only where values go to LDS or HBM; the pruned
    inverse FFT in FP64);
  * ``algorithm``: "fused" (default; fused x pass, pruned inverse) or
    "plain" (the reference's stage sequence on the device, fp64/fp32 only);
  * ``device``: the GPU to run on; ``context``: an explicit fftcorr_amd
    context (its own HIP stream), used by the serving layer;
  * ``PeriodicCorrelator.autocorrelate_interlaced(grid_a, grid_b)`` for a
    pair of grids assigned with shift 0 and 0.5 cells;
  * ``zerolag`` and ``last_time_ms`` properties.
Every failure raises a Python exception (the reference aborts on C++ throws
outside its constructors).
"""

import ctypes

import numpy as np

from fftcorr_amd import _lib
from fftcorr_amd._table import make_table
from fftcorr_amd.grid import ConfigSpaceGrid


class BaseCorrelator:
    _kind = None

    def _init(self, shape, cell_size, posmin, window_correct, rmax, dr, kmax, dk, maxell,
              fftw_flags, precision, device, context=None):
        self._shape = tuple(int(s) for s in np.ascontiguousarray(shape, dtype=np.intc))
        if len(self._shape) != 3:
            raise ValueError("shape must have 3 entries")
        self._ctx = context if context is not None else _lib.context(device)
        self._precision = _lib.precision_code(precision)
        self._fftw_flags = int(fftw_flags)
        s = np.ascontiguousarray(self._shape, np.intc)
        pm = np.ascontiguousarray(posmin if posmin is not None else (0.0, 0.0, 0.0), np.float64)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().fc_corr_create(
            self._ctx.handle, self._kind, _lib.iptr(s), float(cell_size), _lib.dptr(pm),
            int(window_correct), float(rmax), float(dr), float(kmax), float(dk), int(maxell),
            self._precision, ctypes.byref(h)))
        self._handle = h
        nr, nk, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rs = np.zeros(3, np.intc)
        ks = np.zeros(3, np.intc)
        _lib.check(_lib.lib().fc_corr_geometry(h, ctypes.byref(nr), ctypes.byref(nk),
                                               ctypes.byref(ne), _lib.iptr(rs), _lib.iptr(ks)))
        self._nrbins, self._nkbins, self._nell = nr.value, nk.value, ne.value
        self.rshape = tuple(int(v) for v in rs)
        self.kshape = tuple(int(v) for v in ks)
        self._scratch = None
        self._have_results = False
        self._algorithm = "fused"

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h:
            try:
                _lib.lib().fc_corr_destroy(h)
            except Exception:
                pass
            self._handle = None

    # ---- inputs --------------------------------------------------------------
    def _device_grid(self, grid):
        if isinstance(grid, ConfigSpaceGrid):
            if grid.shape != self._shape:
                raise ValueError("Incompatible shapes for copy.")
            return grid.device_grid()
        arr = np.ascontiguousarray(grid, dtype=np.float64)
        if arr.shape != self._shape:
            raise ValueError("Incompatible shapes for copy.")
        if self._scratch is None:
            self._scratch = _lib.DeviceGrid(self._shape, precision=self._precision,
                                            ctx=self._ctx)
        self._scratch.upload(arr)
        return self._scratch

    def set_grid2(self, grid2):
        _lib.check(_lib.lib().fc_corr_set_grid2(self._handle, self._device_grid(grid2).handle))

    def set_grid2_fft(self, grid2_fft):
        f = np.ascontiguousarray(grid2_fft, dtype=np.complex128)
        want = (self._shape[0], self._shape[1], self._shape[2] // 2 + 1)
        if f.shape != want:
            raise ValueError("Incompatible shapes for copy: expected %s, got %s"
                             % (want, f.shape))
        _lib.check(_lib.lib().fc_corr_set_grid2_fft(
            self._handle, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def autocorrelate(self, grid, squeeze=True):
        _lib.check(_lib.lib().fc_corr_autocorrelate(self._handle, self._device_grid(grid).handle))
        self._fetch()
        return self._power_spectrum(squeeze), self._correlation(squeeze)

    def cross_correlate(self, grid1, grid2=None, squeeze=True):
        if grid2 is not None:
            self.set_grid2(grid2)
        _lib.check(_lib.lib().fc_corr_cross_correlate(self._handle,
                                                      self._device_grid(grid1).handle))
        self._fetch()
        return self._power_spectrum(squeeze), self._correlation(squeeze)

    # ---- outputs -------------------------------------------------------------
    def _fetch(self):
        nr, nk, ne = self._nrbins, self._nkbins, self._nell
        self._r = np.zeros(nr)
        self._rc = np.zeros((ne, nr), np.intc)
        self._rh = np.zeros((ne, nr))
        self._k = np.zeros(nk)
        self._kc = np.zeros((ne, nk), np.intc)
        self._kh = np.zeros((ne, nk))
        z = ctypes.c_double()
        _lib.check(_lib.lib().fc_corr_results(
            self._handle, _lib.dptr(self._r), _lib.iptr(self._rc), _lib.dptr(self._rh),
            _lib.dptr(self._k), _lib.iptr(self._kc), _lib.dptr(self._kh), ctypes.byref(z)))
        self._zerolag = z.value
        self._have_results = True

    def _table(self, x, hist, counts, names, squeeze):
        with np.errstate(invalid="ignore", divide="ignore"):
            ratio = hist / counts
        if squeeze:
            ratio = np.squeeze(ratio)
            hist = np.squeeze(hist)
        return make_table((x, np.transpose(ratio), np.transpose(hist), counts[0, :]), names)

    def _correlation(self, squeeze=True):
        return self._table(self._r, self._rh, self._rc, ("r", "xi", "histogram", "count"),
                           squeeze)

    def _power_spectrum(self, squeeze=True):
        return self._table(self._k, self._kh, self._kc, ("k", "ps", "histogram", "count"),
                           squeeze)

    @property
    def zerolag(self):
        return self._zerolag if self._have_results else None

    @property
    def last_time_ms(self):
        t = ctypes.c_double()
        _lib.check(_lib.lib().fc_corr_last_time_ms(self._handle, ctypes.byref(t)))
        return t.value

    @property
    def algorithm(self):
        return self._algorithm

    @algorithm.setter
    def algorithm(self, name):
        codes = {"fused": _lib.FC_ALGO_FUSED, "plain": _lib.FC_ALGO_PLAIN}
        if name not in codes:
            raise ValueError("algorithm must be 'fused' or 'plain', got %r" % (name,))
        _lib.check(_lib.lib().fc_corr_set_algorithm(self._handle, codes[name]))
        self._algorithm = name

    def set_profiling(self, on=True):
        """Per-stage HIP-event timing of later correlate calls (resets totals)."""
        _lib.check(_lib.lib().fc_corr_set_profiling(self._handle, int(bool(on))))

    def stage_times(self):
        """{stage: (total ms, launches)} accumulated since set_profiling(True)."""
        ms = np.zeros(len(_lib.STAGES))
        n = np.zeros(len(_lib.STAGES), np.int64)
        _lib.check(_lib.lib().fc_corr_stage_times(
            self._handle, _lib.dptr(ms), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(_lib.STAGES)}

    def results(self):
        """Raw outputs: dict of bin centres, int32 counts and FP64 histograms
        ([n_ell, nbins]) plus the zero-lag value."""
        return {"r": self._r, "r_counts": self._rc, "r_hist": self._rh, "k": self._k,
                "k_counts": self._kc, "k_hist": self._kh, "zerolag": self._zerolag}


class PeriodicCorrelator(BaseCorrelator):
    _kind = _lib.FC_PERIODIC

    def __init__(self, shape, cell_size, window_correct, rmax, dr, kmax, dk, maxell,
                 fftw_flags=0, precision="fp64", device=None, context=None):
        self._init(shape, cell_size, None, window_correct, rmax, dr, kmax, dk, maxell,
                   fftw_flags, precision, device, context)

    @classmethod
    def from_grid_spec(cls, grid, rmax, dr, kmax, dk, maxell, window_correct=0, fftw_flags=0,
                       precision="fp64", device=None):
        return cls(shape=grid.shape, cell_size=grid.cell_size, window_correct=window_correct,
                   rmax=rmax, dr=dr, kmax=kmax, dk=dk, maxell=maxell, fftw_flags=fftw_flags,
                   precision=precision, device=device)

    def autocorrelate_interlaced(self, grid_a, grid_b, squeeze=True):
        ga = self._device_grid(grid_a)
        if not isinstance(grid_b, ConfigSpaceGrid):
            raise TypeError("the interlaced grid must be a ConfigSpaceGrid")
        _lib.check(_lib.lib().fc_corr_autocorrelate_interlaced(
            self._handle, ga.handle, self._device_grid(grid_b).handle))
        self._fetch()
        return self._power_spectrum(squeeze), self._correlation(squeeze)


class Correlator(BaseCorrelator):
    _kind = _lib.FC_YAMAMOTO

    def __init__(self, shape, cell_size, posmin, window_correct, rmax, dr, kmax, dk, maxell,
                 fftw_flags=0, precision="fp64", device=None):
        posmin = np.ascontiguousarray(posmin, dtype=np.float64)
        self._init(shape, cell_size, posmin, window_correct, rmax, dr, kmax, dk, maxell,
                   fftw_flags, precision, device)

    @classmethod
    def from_grid_spec(cls, grid, rmax, dr, kmax, dk, maxell, window_correct=0, fftw_flags=0,
                       precision="fp64", device=None):
        return cls(shape=grid.shape, cell_size=grid.cell_size, posmin=grid.posmin,
                   window_correct=window_correct, rmax=rmax, dr=dr, kmax=kmax, dk=dk,
                   maxell=maxell, fftw_flags=fftw_flags, precision=precision, device=device)
