"""add_random_particles (fftcorr/catalog/random_catalog.py:8-87).

With ``device_rng=False`` (default) the randoms are drawn exactly as the
reference draws them -- numpy's default_rng(seed).uniform(posmin, posmax) in
batches -- then displaced and assigned on the GPU.  With ``device_rng=True``
they are drawn on the device from the counter-based Philox4x32-10 stream
``seed`` (particle i depends only on (seed, i)), displaced and assigned
there: nothing crosses PCIe but the displacement field.
"""

import numpy as np

from fftcorr_amd import _lib
from fftcorr_amd.grid.utils import apply_displacement_field
from fftcorr_amd.particle_mesh import MassAssignor


def add_random_particles(n, grid, window_type, particle_weight=None, total_weight=None,
                         periodic_wrap=False, disp=None, seed=None, batch_size=int(1e8),
                         buffer_size=int(1e7), device_rng=False):
    """Adds n uniform random particles (optionally displaced) to grid.
    Returns the MassAssignor (host RNG) or the assignment statistics
    (device RNG)."""
    if total_weight is not None:
        if particle_weight is not None:
            raise ValueError("particle_weight and total_weight cannot both be set")
        particle_weight = total_weight / n
    if particle_weight is None:
        particle_weight = 1.0
    n = int(n)
    if disp is not None:
        disp = np.ascontiguousarray(disp, dtype=np.float64)
    if device_rng:
        if seed is None:
            seed = int(np.random.default_rng().integers(0, 2 ** 63))
        st = _lib.MaStats()
        g = grid.device_grid()
        _lib.check(_lib.lib().fc_add_random_particles(
            g.handle, n, float(particle_weight), int(window_type), int(bool(periodic_wrap)),
            _lib.dptr(disp) if disp is not None else None, int(seed) & (2 ** 64 - 1),
            int(batch_size), st))
        grid._device_modified()
        return st
    gridmin, gridmax = grid.posmin, grid.posmax
    rng = np.random.default_rng(seed)
    ma = MassAssignor(grid, window_type, periodic_wrap, buffer_size)
    added = 0
    while added < n:
        nbatch = int(min(batch_size, n - added))
        pos = rng.uniform(gridmin, gridmax, (nbatch, 3))
        if disp is not None:
            apply_displacement_field(grid, pos, disp, periodic_wrap=periodic_wrap, out=pos)
        ma.add_particles_to_buffer(pos, particle_weight)
        added += nbatch
        if added == n:
            ma.flush()
    assert ma.num_added + ma.num_skipped == n
    return ma
