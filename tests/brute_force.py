"""Direct-definition multipole estimators on tiny grids, written from the
textbook definitions and not from the reference's code path.

* Power-spectrum multipoles (Yamamoto et al. 2006, periodic, line of sight
  z): for every Fourier mode m (signed integer indices, |m_i| <= K_i) of the
  grid, k = 2 pi m / (n cell), mu = k_z / |k|,

      P(k) = (V / N^2) |sum_x delta(x) exp(-i k.x)|^2,   V = N cell^3,
      P_l(bin) = sum_{modes in bin} (2l+1) L_l(mu) P(k) W^-2(k),

  with the DFT evaluated as a plain sum over cells (three DFT matrices), the
  Legendre polynomial from scipy and the window 1 / prod_i |W(k_i)|^2 from
  the closed form of the assignment kernel's transform.
* Correlation-function multipoles: for every cell lag s (|s_i| <= R_i),

      xi(s) = (1/N) sum_x delta(x) delta(x + s)      (periodic pair product),
      xi_l(bin) = sum_{lags in bin} (2l+1) L_l(s_z / |s|) xi(s).

The l > 0 terms at k = 0 and s = 0 are 0 (the direction is undefined and the
harmonic's numerator vanishes there).  Bins are floor(|x| / width); values at
or beyond the last edge are dropped.  The mode and lag ranges follow the
reference's subgrid extents (cc/correlate/correlator.h:109-164): R =
ceil(rmax / cell) on every axis; K_i = ceil(kmax / (2 pi / (n_i cell))),
capped at n_i // 2 (so an even axis counts its Nyquist index at both +n/2 and
-n/2, as the reference's modulo indexing does, cc/grid/fft_grid.cc:229-234).
Pure numpy loops over lags: use grids of at most ~12^3.
"""
import math

import numpy as np
from scipy.special import eval_legendre


def _bin(x, width, nbins):
    b = np.floor(x / width).astype(np.int64)
    ok = (b >= 0) & (b < nbins)
    return b, ok


def _sinc_half(k, cell):
    x = k * cell / 2.0
    return np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))


# |W(k)|^2 along one axis of the assignment kernel's transform: the NGP
# kernel (a top hat of one cell) transforms to sinc(k cell / 2); CIC is two
# top hats convolved (sinc^2, squared norm sinc^4), TSC three (sinc^6).
_WINDOW_SQ = {"cic": lambda k, h: _sinc_half(k, h) ** 4,
              "tsc": lambda k, h: _sinc_half(k, h) ** 6}


def _dft3(delta, m):
    """F(m) = sum_x delta(x) exp(-2 pi i m.x / n) for the signed index lists
    m[0], m[1], m[2] (an outer product of the three index ranges)."""
    mats = []
    for ax in range(3):
        n = delta.shape[ax]
        x = np.arange(n)
        mats.append(np.exp(-2j * np.pi * np.outer(m[ax], x) / n))
    return np.einsum("ai,bj,ck,ijk->abc", mats[0], mats[1], mats[2], delta, optimize=True)


def power_multipoles(delta, cell, kmax, dk, maxell, window=None, delta2=None):
    """Returns (counts [nell, nbins] int, sums [nell, nbins] float64).
    window: None, "cic" or "tsc".  delta2: cross power Re F1 F2* when given."""
    shape = delta.shape
    N = delta.size
    V = N * cell ** 3
    nbins = int(math.floor(kmax / dk))
    m = []
    for n in shape:
        K = int(math.ceil(kmax / (2 * math.pi / (n * cell))))
        K = min(K, n // 2)
        m.append(np.arange(-K, K + 1))
    F1 = _dft3(np.asarray(delta, np.float64), m)
    F2 = F1 if delta2 is None else _dft3(np.asarray(delta2, np.float64), m)
    P = (V / N ** 2) * np.real(F1 * np.conj(F2))
    k = [2 * np.pi * m[ax] / (shape[ax] * cell) for ax in range(3)]
    kx, ky, kz = np.meshgrid(k[0], k[1], k[2], indexing="ij")
    kn = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    if window is not None:
        wsq = _WINDOW_SQ[window]
        P = P / (wsq(kx, cell) * wsq(ky, cell) * wsq(kz, cell))
    mu = np.where(kn > 0, kz / np.where(kn > 0, kn, 1.0), 0.0)
    b, ok = _bin(kn, dk, nbins)
    nell = maxell // 2 + 1
    counts = np.zeros((nell, nbins), np.int64)
    sums = np.zeros((nell, nbins))
    for q in range(nell):
        ell = 2 * q
        leg = (2 * ell + 1) * eval_legendre(ell, mu)
        if ell > 0:
            leg = np.where(kn > 0, leg, 0.0)
        np.add.at(sums[q], b[ok], (leg * P)[ok])
        np.add.at(counts[q], b[ok], 1)
    return counts, sums


def correlation_multipoles(delta, cell, rmax, dr, maxell, delta2=None):
    """Returns (counts [nell, nbins] int, sums [nell, nbins] float64) of the
    periodic pair-product estimator over cell lags."""
    delta = np.asarray(delta, np.float64)
    d2 = delta if delta2 is None else np.asarray(delta2, np.float64)
    R = int(math.ceil(rmax / cell))
    nbins = int(math.floor(rmax / dr))
    nell = maxell // 2 + 1
    counts = np.zeros((nell, nbins), np.int64)
    sums = np.zeros((nell, nbins))
    for sx in range(-R, R + 1):
        for sy in range(-R, R + 1):
            for sz in range(-R, R + 1):
                r = cell * math.sqrt(sx * sx + sy * sy + sz * sz)
                b = int(math.floor(r / dr))
                if b < 0 or b >= nbins:
                    continue
                xi = np.mean(delta * np.roll(d2, shift=(-sx, -sy, -sz), axis=(0, 1, 2)))
                mu = (cell * sz / r) if r > 0 else 0.0
                for q in range(nell):
                    ell = 2 * q
                    if ell > 0 and r == 0:
                        w = 0.0
                    else:
                        w = (2 * ell + 1) * float(eval_legendre(ell, mu))
                    sums[q, b] += w * xi
                    counts[q, b] += 1
    return counts, sums
