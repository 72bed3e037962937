"""ctypes loader for oracle/_ref/libfftcorr_ref.so (the reference's own
FFTW-free code built by oracle/Makefile).  TEST INFRASTRUCTURE ONLY.

``load()`` returns None when the library has not been built (it is built only
where the reference sources are present), and tests skip in that case.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libfftcorr_ref.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        return None
    lib = ctypes.CDLL(LIB_PATH)
    lib.ref_sequence.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
    lib.ref_make_ylm.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                 ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, _dp, _dp]
    lib.ref_window.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.ref_window.restype = ctypes.c_double
    lib.ref_histogram.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, _dp, _dp,
                                  ctypes.c_int64, _dp, _ip, _dp]
    lib.ref_grid_coords.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp]
    lib.ref_mass_assign.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int64, _dp, ctypes.c_int64, _dp, _u64p,
                                    _u64p, _dp, _dp]
    lib.ref_cli_grid.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                 _dp, _u64p, _dp, _dp]
    _lib = lib
    return lib


def sequence(start, step, n):
    out = np.empty(n)
    load().ref_sequence(start, step, n, _ptr(out))
    return out


def make_ylm(ell, m, x, y, z, coeff, exponent=0, mult=None):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    out = np.zeros((len(x), len(y), len(z)))
    mu = None if mult is None else np.ascontiguousarray(mult, np.float64)
    rc = load().ref_make_ylm(ell, m, _ptr(x), len(x), _ptr(y), len(y), _ptr(z), len(z),
                             coeff, exponent, _ptr(mu), _ptr(out))
    if rc:
        raise ValueError("make_ylm rejected its arguments")
    return out


def window(kind, k, cell):
    return load().ref_window(kind, k, cell)


def histogram(n, minval, maxval, binsize, ih, x, y):
    x = np.ascontiguousarray(np.ravel(x), np.float64)
    y = np.ascontiguousarray(np.ravel(y), np.float64)
    nb_max = int(np.floor((maxval - minval) / binsize)) + 2
    bins = np.zeros(nb_max)
    counts = np.zeros(n * nb_max, np.int32)
    hist = np.zeros(n * nb_max)
    nb = load().ref_histogram(n, minval, maxval, binsize, ih, _ptr(x), _ptr(y), len(x),
                              _ptr(bins), _ptr(counts, _ip), _ptr(hist))
    if nb < 0:
        raise ValueError("HistogramList rejected its arguments")
    return bins[:nb], counts[: n * nb].reshape(n, nb), hist[: n * nb].reshape(n, nb)


def grid_coords(shape, posmin, cell, wrap, pos):
    s = np.array(shape, np.intc)
    pm = np.array(posmin, np.float64)
    p = np.array(pos, np.float64)
    out = np.zeros(3)
    ok = load().ref_grid_coords(_ptr(s, _ip), _ptr(pm), cell, int(wrap), _ptr(p), _ptr(out))
    return bool(ok), out


def wavelet_table():
    """The reference's D12 sample table (512 values, 8 cells x 64 samples)."""
    out = np.zeros(512)
    lib = load()
    lib.ref_wavelet_table.argtypes = [_dp]
    lib.ref_wavelet_table.restype = ctypes.c_int
    if lib.ref_wavelet_table(_ptr(out)) != 512:
        raise RuntimeError("unexpected wavelet table size")
    return out


def mass_assign(shape, posmin, cell, scheme, wrap, posw, buffer_size=0):
    s = np.array(shape, np.intc)
    pm = np.array(posmin, np.float64)
    posw = np.ascontiguousarray(posw, np.float64)
    grid = np.zeros(tuple(shape))
    na = ctypes.c_uint64()
    ns = ctypes.c_uint64()
    tw = ctypes.c_double()
    tw2 = ctypes.c_double()
    rc = load().ref_mass_assign(_ptr(s, _ip), _ptr(pm), cell, scheme, int(wrap),
                                buffer_size, _ptr(posw), len(posw), _ptr(grid),
                                ctypes.byref(na), ctypes.byref(ns), ctypes.byref(tw),
                                ctypes.byref(tw2))
    if rc:
        raise RuntimeError("reference MassAssignor failed")
    return grid, dict(num_added=na.value, num_skipped=ns.value, totw=tw.value,
                      totwsq=tw2.value)


def cli_grid(file1, file2, ngrid, sep, periodic, scheme, geometry_only=False):
    posmin = np.zeros(3)
    cell = ctypes.c_double()
    grid = None if geometry_only else np.zeros((ngrid,) * 3)
    na = ctypes.c_uint64()
    tw = ctypes.c_double()
    tw2 = ctypes.c_double()
    rc = load().ref_cli_grid(file1.encode(), (file2 or "").encode(), ngrid, sep,
                             int(periodic), scheme, _ptr(posmin), ctypes.byref(cell),
                             _ptr(grid), ctypes.byref(na), ctypes.byref(tw),
                             ctypes.byref(tw2))
    if rc:
        raise RuntimeError("reference CLI gridding failed")
    return grid, posmin, cell.value, dict(num_added=na.value, totw=tw.value,
                                          totwsq=tw2.value)
