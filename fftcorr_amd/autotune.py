"""FFT autotuner: the GPU counterpart of FFTW's measuring planner
(FftGrid::plan_fft with FFTW_MEASURE, cc/grid/fft_grid.cc:69-115).

``autotune(shape, precision)`` times, on the device, the candidate radix
plans (first radix 8 or 16, where the register-chained kernels have both)
together with the tile shapes of the z (rows per workgroup), y and x
(columns per workgroup) passes, and the z+y chunk size; it keeps the
fastest in the plan cache and returns them.  Plans created afterwards (new
correlators, ``fc_fft_*``) use them; ``fftcorr_amd.fftw.export_wisdom_to_file``
saves them for later runs.
"""

import ctypes

import numpy as np

from fftcorr_amd import _lib


def plan_tiles(shape, precision="fp32", device=None):
    """Tiles a new plan of `shape` would use now."""
    s = np.ascontiguousarray(shape, np.intc)
    out = np.zeros(4, np.intc)
    dev = 0 if device is None else int(device)
    _lib.check(_lib.lib().fc_plan_tiles(dev, _lib.iptr(s), _lib.precision_code(precision),
                                        _lib.iptr(out)))
    return {"x_cols": int(out[0]), "y_cols": int(out[1]), "z_rows": int(out[2]),
            "zy_chunk_kib": int(out[3])}


def fft_radices(length, variant):
    """Radices (first to last) of plan `variant` for an FFT of `length`, or
    None where the register-chained kernels have no such variant."""
    out = np.zeros(8, np.intc)
    n = ctypes.c_int(0)
    _lib.check(_lib.lib().fc_fft_radices(int(length), int(variant), _lib.iptr(out),
                                         ctypes.byref(n)))
    return [int(r) for r in out[:n.value]] if n.value else None


def plan_variants(shape, precision="fp32", device=None):
    """Radix plans a new plan of `shape` would use now: variant numbers and
    radices of the x and y column passes and of the z row FFT (n2/2)."""
    s = np.ascontiguousarray(shape, np.intc)
    out = np.zeros(3, np.intc)
    dev = 0 if device is None else int(device)
    _lib.check(_lib.lib().fc_plan_variants(dev, _lib.iptr(s), _lib.precision_code(precision),
                                           _lib.iptr(out)))
    lengths = (int(shape[0]), int(shape[1]), int(shape[2]) // 2)
    res = {}
    for name, v, n in zip(("x", "y", "z"), out, lengths):
        res[name + "_variant"] = int(v)
        res[name + "_radices"] = fft_radices(n, int(v))
    return res


def autotune(shape, precision="fp32", device=None, reps=5):
    """Tunes the plan for `shape` at `precision`; returns the chosen tiles
    and the best times (ms) of each pass."""
    if len(shape) != 3:
        raise ValueError("shape must have 3 entries")
    s = np.ascontiguousarray(shape, np.intc)
    ms = np.zeros(4)
    ctx = _lib.context(device)
    _lib.check(_lib.lib().fc_autotune(ctx.handle, _lib.iptr(s), _lib.precision_code(precision),
                                      int(reps), _lib.dptr(ms)))
    out = plan_tiles(shape, precision, ctx.device)
    out.update(plan_variants(shape, precision, ctx.device))
    out["ms"] = {"z": float(ms[0]), "y": float(ms[1]), "x": float(ms[2]), "zy": float(ms[3])}
    return out
