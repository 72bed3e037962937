"""Shared helpers of the GPU parity tests (numbers compared with the oracle)."""
import ctypes

import numpy as np

from fftcorr_amd import _lib


def fft_r2c(real, precision):
    real = np.ascontiguousarray(real, np.float64)
    shape = np.array(real.shape, np.intc)
    out = np.zeros(real.shape[:2] + (real.shape[2] // 2 + 1,), np.complex128)
    _lib.check(_lib.lib().fc_fft_r2c(_lib.context().handle, _lib.iptr(shape),
                                     _lib.precision_code(precision), _lib.dptr(real),
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def fft_c2r(cplx, shape, precision):
    cplx = np.ascontiguousarray(cplx, np.complex128)
    s = np.array(shape, np.intc)
    out = np.zeros(tuple(shape), np.float64)
    _lib.check(_lib.lib().fc_fft_c2r(_lib.context().handle, _lib.iptr(s),
                                     _lib.precision_code(precision),
                                     cplx.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     _lib.dptr(out)))
    return out


def floor_rel_err(x, ref, floor_frac=1e-3):
    """|x - ref| / max(|ref|, floor_frac * max|ref_row0|) per element, the
    error metric for multipoles that cross zero."""
    ref = np.atleast_2d(ref)
    x = np.atleast_2d(x)
    floor = floor_frac * np.max(np.abs(ref[0])) if ref.size else 0.0
    den = np.maximum(np.abs(ref), floor)
    den = np.where(den == 0, 1.0, den)
    return np.abs(x - ref) / den
