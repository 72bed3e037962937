"""Seeded synthetic catalogues (numpy PCG64) for benchmarks and tests:
uniform points in a box and a Thomas cluster process (uniform parents with
Poisson-distributed numbers of Gaussian-displaced children)."""

import numpy as np


def uniform_catalog(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    lo = np.broadcast_to(np.asarray(lo, np.float64), (3,))
    hi = np.broadcast_to(np.asarray(hi, np.float64), (3,))
    return rng.uniform(lo, hi, size=(int(n), 3))


def thomas_catalog(n_parents, mean_children, sigma, lo, hi, seed, wrap=True):
    """Positions [n, 3]; wrap=True folds points back into [lo, hi)."""
    rng = np.random.default_rng(seed)
    lo = np.broadcast_to(np.asarray(lo, np.float64), (3,))
    hi = np.broadcast_to(np.asarray(hi, np.float64), (3,))
    parents = rng.uniform(lo, hi, size=(int(n_parents), 3))
    counts = rng.poisson(mean_children, size=int(n_parents))
    pos = np.repeat(parents, counts, axis=0)
    pos += rng.normal(0.0, sigma, size=pos.shape)
    if wrap:
        box = hi - lo
        pos = lo + np.mod(pos - lo, box)
        pos = np.where(pos >= hi, lo, pos)
    return pos
