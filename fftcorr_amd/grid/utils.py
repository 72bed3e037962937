"""apply_displacement_field (fftcorr/grid/utils.pyx:19-55) on the GPU."""

import numpy as np

from fftcorr_amd import _lib


def apply_displacement_field(grid, pos, disp, periodic_wrap=False, out=None):
    """out[i] = pos[i] + disp[cell of pos[i]]: the cell is the truncated grid
    coordinate of pos[i] (with the periodic wrap when asked).  Raises
    ValueError if a position is outside the grid and periodic_wrap is off."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("pos must have shape (n, 3)")
    disp = np.ascontiguousarray(disp, dtype=np.float64)
    if disp.shape != tuple(grid.shape) + (3,):
        raise ValueError("disp must have shape %s" % (tuple(grid.shape) + (3,),))
    if out is None:
        out = np.empty_like(pos)
    res = np.empty_like(pos)
    _lib.check(_lib.lib().fc_apply_displacement(grid.device_grid().handle, _lib.dptr(pos),
                                                pos.shape[0], _lib.dptr(disp),
                                                int(bool(periodic_wrap)), _lib.dptr(res)))
    out[...] = res
    return out
