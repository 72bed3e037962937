"""FFT autotuner: the GPU counterpart of FFTW's measuring planner
(FftGrid::plan_fft with FFTW_MEASURE, cc/grid/fft_grid.cc:69-115).

``autotune(shape, precision)`` times the candidate tile shapes of the z
(rows per workgroup), y and x (columns per workgroup) passes and the z+y
chunk size on the device, keeps the fastest in the plan cache, and returns
them.  Plans created afterwards (new correlators, ``fc_fft_*``) use them;
``fftcorr_amd.fftw.export_wisdom_to_file`` saves them for later runs.
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
    out["ms"] = {"z": float(ms[0]), "y": float(ms[1]), "x": float(ms[2]), "zy": float(ms[3])}
    return out
