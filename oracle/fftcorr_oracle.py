"""FP64 CPU oracle for the fftcorr correlation / power-spectrum path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or as the timed CPU baseline).

Every function restates the reference algorithm in numpy, op for op where the
rounding matters (bin indices, Y_lm products, window tables), and cites the
reference file:line it follows.  Paths are relative to the reference root.

Pins (what checks each function; see tests/test_oracle_*.py):
  HistogramList        closed form + test_histogram.py known answers + oracle/_ref
  sequence / subgrids  golden geometry (north/south .out files: +-19 r, +-14 k)
  make_ylm             scipy.special.sph_harm_y up to sign, addition theorem,
                       bit-exact against oracle/_ref (spherical_harmonics.h)
  window functions     closed forms + bit-exact against oracle/_ref
  grid coords / skip   test_grid.py known answers + oracle/_ref
  NGP / TSC            oracle/_ref (distribution_functions.h) + sum(grid)=sum(w)
  D12 wavelet          oracle/_ref (its own table and stencil) + partition of
                       unity of the table columns + sum(grid)=sum(w)
  CIC (new, no ref)    closed form (trilinear about cell centres) + mass
                       conservation + linear-field reproduction
  FFT conventions      brute-force DFT on tiny grids, Parseval
  PeriodicCorrelator   brute-force periodic pair sum over cell lags (xi_0),
                       zero-lag = sum(delta^2)/N, P_0 integral identity,
                       xi_2 = xi_4 = 0 for isotropic-by-construction fields
  Correlator           north/south golden .out files (D12 wavelet grids built
                       by oracle/_ref), far-observer limit -> periodic with LOS x
"""

import math

import numpy as np

# Enum codes kept as in the reference (ints), plus the new CIC codes.
# cc/particle_mesh/distribution_functions.h:11-15
NEAREST_CELL = 0
TRIANGULAR_SHAPED_CLOUD = 1
WAVELET = 2
CLOUD_IN_CELL = 3  # new: not in the reference
# cc/correlate/window_correct.h:8-12
NO_CORRECTION = 0
TSC_CORRECTION = 1
TSC_ALIASED_CORRECTION = 2
CIC_CORRECTION = 3  # new
CIC_ALIASED_CORRECTION = 4  # new


# ---------------------------------------------------------------------------
# cc/array/array_ops.cc:7-13
def sequence(start, step, size):
    """start + i*step in FP64 (one multiply and one add per element)."""
    i = np.arange(size, dtype=np.float64)
    return start + i * step


# ---------------------------------------------------------------------------
# cc/histogram/histogram_list.h:10-67
class HistogramList:
    def __init__(self, n, minval, maxval, binsize):
        n = int(n)
        minval = float(minval)
        maxval = float(maxval)
        binsize = float(binsize)
        if n <= 0:
            raise ValueError("n must be positive.")
        if minval >= maxval:
            raise ValueError("maxval must be greater than minval.")
        if binsize <= 0:
            raise ValueError("binsize must be positive.")
        self.minval = minval
        self.binsize = binsize
        self.nbins = int(math.floor((maxval - minval) / binsize))  # :15
        self.bins = minval + (np.arange(self.nbins) + 0.5) * binsize  # :24
        self.counts = np.zeros((n, self.nbins), dtype=np.int32)
        self.hist_values = np.zeros((n, self.nbins), dtype=np.float64)

    def bin_index(self, x):
        # :57  floor((val - minval) / binsize)
        return np.floor((np.asarray(x, np.float64) - self.minval) / self.binsize)

    def accumulate(self, ih, x, y):
        # :33-45 ; out-of-range bins (b < 0 or b >= nbins) are dropped.
        if ih >= self.counts.shape[0]:
            raise IndexError("ih out of range.")
        x = np.asarray(x, np.float64).ravel()
        y = np.asarray(y, np.float64).ravel()
        b = self.bin_index(x)
        ok = (b >= 0) & (b < self.nbins)
        bi = b[ok].astype(np.int64)
        yv = y[ok]
        self.counts[ih] += np.bincount(bi, minlength=self.nbins).astype(np.int32)
        # Sequential sum per bin in element order (matches the reference's
        # single loop up to the per-bin order, which is the same).
        hv = self.hist_values[ih]
        order = np.argsort(bi, kind="stable")
        bs = bi[order]
        ys = yv[order]
        if bs.size:
            starts = np.flatnonzero(np.r_[True, bs[1:] != bs[:-1]])
            ends = np.r_[starts[1:], bs.size]
            for s, e in zip(starts, ends):
                acc = hv[bs[s]]
                for v in ys[s:e]:
                    acc += v
                hv[bs[s]] = acc

    def accumulate_fast(self, ih, x, y):
        """Same as accumulate but with numpy's pairwise sums (large inputs)."""
        x = np.asarray(x, np.float64).ravel()
        y = np.asarray(y, np.float64).ravel()
        b = self.bin_index(x)
        ok = (b >= 0) & (b < self.nbins)
        bi = b[ok].astype(np.int64)
        self.counts[ih] += np.bincount(bi, minlength=self.nbins).astype(np.int32)
        self.hist_values[ih] += np.bincount(bi, weights=y[ok], minlength=self.nbins)

    def reset(self):
        self.counts[:] = 0
        self.hist_values[:] = 0.0


# ---------------------------------------------------------------------------
# cc/correlate/spherical_harmonics.h:32-179
def make_ylm(ell, m, xcell, ycell, zcell, coeff, exponent=0, mult=None):
    """Real Y_lm on the grid spanned by the axes, times coeff * r^exp * mult.

    m > 0 -> sqrt(2) Re Y_lm, m < 0 -> sqrt(2) Im Y_l|m|, sign per the
    reference's table.  Returns an array of shape (len(x), len(y), len(z)).
    Op order follows the reference line by line so FP64 results agree
    bit for bit with it.
    """
    if ell not in (0, 2, 4):
        raise ValueError("Unsupported ell")
    if m < -ell or m > ell:
        raise ValueError("Illegal (ell, m)")
    if exponent % 2 != 0:
        raise ValueError("Exponent must be even.")
    x = np.asarray(xcell, np.float64)[:, None, None]
    y = np.asarray(ycell, np.float64)[None, :, None]
    z = np.asarray(zcell, np.float64)[None, None, :]
    shape = (x.shape[0], y.shape[1], z.shape[2])
    if mult is None:
        D = np.ones(shape)
    else:
        D = np.asarray(mult, np.float64)[: shape[0], : shape[1], : shape[2]]
    if ell == 0 and m == 0 and exponent == 0:
        value = coeff / math.sqrt(4.0 * math.pi)  # :51
        return D * value
    z2 = z * z
    z3 = z2 * z
    z4 = z3 * z
    isqpi = math.sqrt(1.0 / math.pi)
    isqpi *= coeff
    if m != 0:
        isqpi *= math.sqrt(2.0)
    tiny = 1e-20
    x2 = x * x
    y2 = y * y
    y3 = y2 * y
    y4 = y3 * y
    ir2 = 1.0 / (x2 + y2 + z2 + tiny)  # :104
    if exponent == 0:
        R = np.ones(shape)
    else:
        R = np.ones(shape)
        rmult = (x2 + y2 + z2) if exponent > 0 else ir2
        for _ in range(abs(exponent) // 2):
            R = R * rmult
    DR = D * R
    s = math.sqrt
    if ell == 2:
        if m == 2:
            Y = DR * isqpi * s(15.0 / 32.0) * (x2 - y2) * ir2
        elif m == 1:
            Y = DR * isqpi * s(15.0 / 8.0) * x * z * ir2
        elif m == 0:
            Y = DR * isqpi * s(5.0 / 16.0) * (2.0 * z2 - x2 - y2) * ir2
        elif m == -1:
            Y = DR * isqpi * s(15.0 / 8.0) * y * z * ir2
        else:
            Y = DR * isqpi * s(15.0 / 8.0) * x * y * ir2
    else:  # ell == 4, :139-177
        if m == 4:
            Y = (DR * isqpi * 3.0 / 16.0 * s(35.0 / 2.0)
                 * (x2 * x2 - 6.0 * x2 * y2 + y4) * ir2 * ir2)
        elif m == 3:
            Y = (DR * isqpi * 3.0 / 8.0 * s(35.0)
                 * (x2 - 3.0 * y2) * z * x * ir2 * ir2)
        elif m == 2:
            Y = (DR * isqpi * 3.0 / 8.0 * s(5.0 / 2.0)
                 * (6.0 * z2 * (x2 - y2) - x2 * x2 + y4) * ir2 * ir2)
        elif m == 1:
            Y = (DR * isqpi * 3.0 / 8.0 * s(5.0) * 3.0
                 * (4.0 / 3.0 * z2 - x2 - y2) * x * z * ir2 * ir2)
        elif m == 0:
            Y = (DR * isqpi * 3.0 / 16.0 * 8.0
                 * (z4 - 3.0 * z2 * (x2 + y2) + 3.0 / 8.0 * (x2 * x2 + 2.0 * x2 * y2 + y4))
                 * ir2 * ir2)
        elif m == -1:
            Y = (DR * isqpi * 3.0 / 8.0 * s(5.0) * 3.0
                 * (4.0 / 3.0 * z2 - x2 - y2) * y * z * ir2 * ir2)
        elif m == -2:
            Y = (DR * isqpi * 3.0 / 8.0 * s(5.0 / 2.0)
                 * (2.0 * x * y * (6.0 * z2 - x2 - y2)) * ir2 * ir2)
        elif m == -3:
            Y = (DR * isqpi * 3.0 / 8.0 * s(35.0)
                 * (3.0 * x2 * y - y3) * z * ir2 * ir2)
        else:
            Y = (DR * isqpi * 3.0 / 16.0 * s(35.0 / 2.0)
                 * (4.0 * x * (x2 * y - y3)) * ir2 * ir2)
    return np.broadcast_to(Y, shape).copy()


# ---------------------------------------------------------------------------
# cc/correlate/window_correct.h:14-60 (1D squared-norm factor per axis value)
def sinc(x):
    if x == 0:
        return 1.0
    return math.sin(x) / x


def window_squared_norm(kind, k, cell_size):
    """W^2 along one axis.  Scalar libm calls (math.sin/math.pow) so the
    values agree bit for bit with the reference's C++ on the same host."""
    if kind == NO_CORRECTION:
        return 1.0
    if kind == TSC_CORRECTION:  # sinc^6
        return math.pow(sinc(k * cell_size / 2.0), 6)
    if kind == TSC_ALIASED_CORRECTION:  # 1 - s^2 + 2/15 s^4
        sinsq = math.sin(k * cell_size / 2.0)
        sinsq *= sinsq
        return 1 - sinsq + 2.0 / 15.0 * sinsq * sinsq
    if kind == CIC_CORRECTION:  # new: sinc^4
        return math.pow(sinc(k * cell_size / 2.0), 4)
    if kind == CIC_ALIASED_CORRECTION:  # new: 1 - 2/3 s^2
        sinsq = math.sin(k * cell_size / 2.0)
        sinsq *= sinsq
        return 1 - 2.0 / 3.0 * sinsq
    raise ValueError("Unknown window correction %r" % (kind,))


def window_axis(kind, kaxis, cell_size):
    return np.array([window_squared_norm(kind, float(k), cell_size) for k in kaxis])


def inv_window(kind, kx, ky, kz, cell_size):
    # cc/correlate/correlator.h:191-204: 1/(Wx*Wy*Wz)
    wx = window_axis(kind, kx, cell_size)[:, None, None]
    wy = window_axis(kind, ky, cell_size)[None, :, None]
    wz = window_axis(kind, kz, cell_size)[None, None, :]
    return 1.0 / (wx * wy * wz)


# ---------------------------------------------------------------------------
# cc/grid/config_space_grid.h:31-46 (vectorised over particles)
def grid_coords(pos, posmin, cell_size, shape, periodic_wrap):
    """Returns (coords [n,3], ok [n]).  Wrapped coords that round up to the
    grid size are set to 0 (the periodic image), where the reference would
    index one past the end."""
    pos = np.asarray(pos, np.float64)
    x = (pos[:, :3] - np.asarray(posmin, np.float64)[None, :]) / cell_size
    xmax = np.asarray(shape, np.float64)[None, :]
    out = (x < 0) | (x >= xmax)
    ok = np.ones(len(x), bool)
    if not periodic_wrap:
        ok = ~out.any(axis=1)
        return x, ok
    xw = np.fmod(x, xmax)
    xw = np.where(xw < 0, xw + xmax, xw)
    xw = np.where(xw >= xmax, 0.0, xw)
    x = np.where(out, xw, x)
    return x, ok


# ---------------------------------------------------------------------------
# fftcorr/grid/utils.pyx:19-55 (apply_displacement_field), vectorised.
def apply_displacement_field(pos, disp, posmin, cell_size, shape, periodic_wrap):
    """out[i] = pos[i] + disp[int(gx), int(gy), int(gz)] with (gx, gy, gz)
    the grid coordinates of pos[i]; ValueError if any is outside the grid
    without wrap (the reference stops at the first such particle)."""
    x, ok = grid_coords(pos, posmin, cell_size, shape, periodic_wrap)
    if not ok.all():
        raise ValueError("Failed to apply displacement field: coordinates out of bounds? "
                         "Try setting periodic_wrap=True.")
    ic = x.astype(np.int64)
    return np.asarray(pos, np.float64) + disp[ic[:, 0], ic[:, 1], ic[:, 2]]


# ---------------------------------------------------------------------------
# Mass assignment.  cc/particle_mesh/distribution_functions.h:25-160 and
# cc/particle_mesh/mass_assignor.h:105-191.
def _stencil_1d(scheme, r, table=None):
    """(offsets, weights[n, len(offsets)]) along one axis, r = x - floor(x).
    ``table`` is the D12 sample table (8 x 64, d12.h order) for WAVELET."""
    if scheme == WAVELET:  # distribution_functions.h:176-216, d12.h:14-16
        if table is None:
            raise ValueError("the D12 wavelet needs its sample table")
        tab = np.asarray(table, np.float64).reshape(8, 64)
        sub = np.floor(r * 64).astype(np.int64)
        return list(range(-1, 7)), tab[:, sub].T
    if scheme == NEAREST_CELL:
        return [0], np.ones((len(r), 1))
    if scheme == TRIANGULAR_SHAPED_CLOUD:  # :59-67
        wm = 0.5 * (1 - r) * (1 - r)
        w0 = 0.5 + r - r * r
        wp = 0.5 * r * r
        return [-1, 0, 1], np.stack([wm, w0, wp], axis=1)
    if scheme == CLOUD_IN_CELL:
        # Trilinear about cell centres (cell i is centred at i + 1/2):
        # r >= 1/2 -> cells (i, i+1) with (3/2 - r, r - 1/2);
        # r <  1/2 -> cells (i-1, i) with (1/2 - r, r + 1/2).
        hi = r >= 0.5
        wm = np.where(hi, 0.0, 0.5 - r)
        w0 = np.where(hi, 1.5 - r, r + 0.5)
        wp = np.where(hi, r - 0.5, 0.0)
        return [-1, 0, 1], np.stack([wm, w0, wp], axis=1)
    raise ValueError("Unsupported distribution scheme %r" % (scheme,))


def assign_mass(grid, coords, weights, scheme, table=None):
    """Adds the particles (grid coords, already wrapped/filtered) to grid
    (float64 [n0,n1,n2]) in place.  The stencil always wraps periodically, as
    the reference's TSC and D12 do (distribution_functions.h:69-119, 189-214)."""
    n0, n1, n2 = grid.shape
    coords = np.asarray(coords, np.float64)
    w = np.asarray(weights, np.float64)
    ix = np.floor(coords).astype(np.int64)
    r = coords - ix
    offs = []
    wts = []
    for d in range(3):
        o, wd = _stencil_1d(scheme, r[:, d], table)
        offs.append(o)
        wts.append(wd)
    flat = grid.reshape(-1)
    for a, oa in enumerate(offs[0]):
        xa = (ix[:, 0] + oa) % n0
        # The reference folds the weight into the x factor:  xm = 0.5(1-r)^2 w.
        fx = wts[0][:, a] * w
        for b, ob in enumerate(offs[1]):
            yb = (ix[:, 1] + ob) % n1
            fxy = fx * wts[1][:, b]
            for c, oc in enumerate(offs[2]):
                zc = (ix[:, 2] + oc) % n2
                val = fxy * wts[2][:, c]
                np.add.at(flat, (xa * n1 + yb) * n2 + zc, val)
    return grid


# ---------------------------------------------------------------------------
# cc/grid/fft_grid.cc: unnormalised R2C forward (sign -1) and C2R inverse
# without 1/N (fftw_plan_many_dft_r2c/c2r, :81-84).
# FFT_WORKERS > 1 switches to scipy.fft (multithreaded pocketfft) for the
# CPU-baseline timing; the default is numpy's pocketfft.
FFT_WORKERS = 1


def rfft3(real):
    real = np.asarray(real, np.float64)
    if FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.rfftn(real, workers=FFT_WORKERS)
    return np.fft.rfftn(real)


def irfft3(cplx, shape):
    if FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.irfftn(cplx, s=shape, axes=(0, 1, 2), norm="forward",
                                workers=FFT_WORKERS)
    return np.fft.irfftn(cplx, s=shape, axes=(0, 1, 2), norm="forward")


def interlace(fa, fb, shape):
    """Interlaced density (new; no reference symbol): grid b holds the same
    particles displaced by +1/2 cell along every axis, so its transform is
    F(k) exp(-i k.h/2 (1,1,1)); averaging F_a with F_b exp(+i k.h/2 (1,1,1))
    cancels the aliased images with odd sum of image indices.  Signed
    frequencies follow numpy.fft.fftfreq (Nyquist negative)."""
    n0, n1, n2 = shape
    f0 = np.fft.fftfreq(n0) * n0
    f1 = np.fft.fftfreq(n1) * n1
    f2 = np.where(np.arange(fa.shape[2]) < (n2 + 1) // 2, np.arange(fa.shape[2]),
                  np.arange(fa.shape[2]) - n2)
    ph = np.pi * (f0[:, None, None] / n0 + f1[None, :, None] / n1 + f2[None, None, :] / n2)
    return 0.5 * (fa + fb * np.exp(1j * ph))


# cc/grid/fft_grid.cc:182-210
def extract_submatrix(real, oshape, mult=None):
    n0, n1, n2 = real.shape
    ox, oy, oz = (s // 2 for s in oshape)
    ii = (n0 - ox + np.arange(oshape[0])) % n0
    jj = (n1 - oy + np.arange(oshape[1])) % n1
    kk = (n2 - oz + np.arange(oshape[2])) % n2
    sub = real[np.ix_(ii, jj, kk)]
    return sub * mult if mult is not None else sub.copy()


# cc/grid/fft_grid.cc:215-257
def extract_fft_submatrix_c2r(cplx, oshape, mult=None):
    c0, c1, _ = cplx.shape
    ox, oy, oz = (s // 2 for s in oshape)
    i = np.arange(oshape[0])
    j = np.arange(oshape[1])
    ii = (c0 - ox + i) % c0
    iin = (c0 - ii) % c0
    jj = (c1 - oy + j) % c1
    jjn = (c1 - jj) % c1
    re = np.real(cplx)
    out = np.empty(tuple(oshape))
    kneg = np.arange(oz)
    out[:, :, :oz] = re[np.ix_(iin, jjn, oz - kneg)]
    kpos = np.arange(oz, oshape[2])
    out[:, :, oz:] = re[np.ix_(ii, jj, kpos - oz)]
    return out * mult if mult is not None else out


# ---------------------------------------------------------------------------
# cc/correlate/correlator.h:109-205
class SubgridGeometry:
    def __init__(self, shape, cell_size, rmax, kmax, window_correct):
        shape = [int(s) for s in shape]
        self.shape = shape
        cell = float(cell_size)
        rmax_cells = int(math.ceil(rmax / cell))  # :111
        sizex = 2 * rmax_cells + 1
        self.rshape = (sizex, sizex, sizex)
        minsep = -cell * rmax_cells
        self.rx = sequence(minsep, cell, sizex)
        self.ry = sequence(minsep, cell, sizex)
        self.rz = sequence(minsep, cell, sizex)
        self.rnorm = np.sqrt(self.rx[:, None, None] * self.rx[:, None, None]
                             + self.ry[None, :, None] * self.ry[None, :, None]
                             + self.rz[None, None, :] * self.rz[None, None, :])
        self.rzero = (rmax_cells,) * 3
        k_nyq = math.pi / cell  # :146
        kshape = []
        for i in range(3):
            ks = int(2 * math.ceil(kmax / (2.0 * k_nyq / shape[i])) + 1)
            if ks > shape[i]:
                ks = 2 * (shape[i] // 2) + 1  # :154-164
            kshape.append(ks)
        self.kshape = tuple(kshape)
        axes = []
        for i in range(3):
            start = (-(kshape[i] // 2)) * 2.0 * k_nyq / shape[i]
            step = 2.0 * k_nyq / shape[i]
            axes.append(sequence(start, step, kshape[i]))
        self.kx, self.ky, self.kz = axes
        self.knorm = np.sqrt(self.kx[:, None, None] * self.kx[:, None, None]
                             + self.ky[None, :, None] * self.ky[None, :, None]
                             + self.kz[None, None, :] * self.kz[None, None, :])
        self.inv_window = inv_window(window_correct, self.kx, self.ky, self.kz, cell)


def _results(rhist, khist, zerolag):
    return {
        "r": rhist.bins.copy(), "r_counts": rhist.counts.copy(),
        "r_hist": rhist.hist_values.copy(),
        "k": khist.bins.copy(), "k_counts": khist.counts.copy(),
        "k_hist": khist.hist_values.copy(), "zerolag": zerolag,
    }


class PeriodicCorrelator:
    """cc/correlate/correlator.h:255-357 (line of sight = z)."""

    def __init__(self, shape, cell_size, window_correct, rmax, dr, kmax, dk, maxell):
        self.shape = tuple(int(s) for s in shape)
        self.cell_size = float(cell_size)
        self.maxell = int(maxell)
        self.geom = SubgridGeometry(shape, cell_size, rmax, kmax, window_correct)
        self.nell = self.maxell // 2 + 1
        self.rmax, self.dr, self.kmax, self.dk = rmax, dr, kmax, dk
        self.grid2_fft = None
        self.fast_hist = False

    def _hists(self):
        return (HistogramList(self.nell, 0.0, self.rmax, self.dr),
                HistogramList(self.nell, 0.0, self.kmax, self.dk))

    def set_grid2(self, grid2):
        self.grid2_fft = rfft3(grid2)

    def set_grid2_fft(self, f):
        self.grid2_fft = np.array(f, np.complex128)

    def autocorrelate(self, grid):
        return self._correlate(grid, None)

    def autocorrelate_interlaced(self, grid_a, grid_b):
        F = interlace(rfft3(grid_a), rfft3(grid_b), self.shape)
        return self._correlate(None, None, fft1=F)

    def cross_correlate(self, grid1, grid2=None):
        if grid2 is not None:
            self.set_grid2(grid2)
        return self._correlate(grid1, self.grid2_fft)

    def _correlate(self, grid1, grid2_fft, fft1=None):
        g = self.geom
        rhist, khist = self._hists()
        acc = "accumulate_fast" if self.fast_hist else "accumulate"
        F = rfft3(grid1) if fft1 is None else fft1
        if grid2_fft is None:
            # (a+bi)(a-bi): real part a*a + b*b (array_ops.cc:139-163)
            P = F * np.conj(F)
        else:
            P = F * np.conj(grid2_fft)
        ncells = int(np.prod(self.shape))
        k_rescale = self.cell_size * self.cell_size * self.cell_size / ncells
        for ell in range(0, self.maxell + 1, 2):
            coeff = math.sqrt((4.0 * math.pi) * (2 * ell + 1)) * k_rescale
            kylm = make_ylm(ell, 0, g.kx, g.ky, g.kz, coeff, 0, g.inv_window)
            ksub = extract_fft_submatrix_c2r(P, g.kshape, kylm)
            getattr(khist, acc)(ell // 2, g.knorm, ksub)
        xi = irfft3(P, self.shape)
        r_rescale = 1.0 / ncells / ncells
        zerolag = None
        for ell in range(0, self.maxell + 1, 2):
            coeff = math.sqrt((4.0 * math.pi) * (2 * ell + 1)) * r_rescale
            rylm = make_ylm(ell, 0, g.rx, g.ry, g.rz, coeff, 0, None)
            rsub = extract_submatrix(xi, g.rshape, rylm)
            getattr(rhist, acc)(ell // 2, g.rnorm, rsub)
            if ell == 0:
                zerolag = rsub[g.rzero]
        return _results(rhist, khist, zerolag)


class Correlator(PeriodicCorrelator):
    """cc/correlate/correlator.h:360-505: Yamamoto-style, observer at the
    survey-coordinate origin."""

    def __init__(self, shape, cell_size, posmin, window_correct, rmax, dr, kmax, dk,
                 maxell, observer=None):
        super().__init__(shape, cell_size, window_correct, rmax, dr, kmax, dk, maxell)
        posmin = [float(p) for p in posmin]
        if observer is None:
            observer = [-posmin[i] / self.cell_size for i in range(3)]  # :384-386
        # :398-400
        self.xcell = sequence(0.5 - observer[0], 1.0, self.shape[0])
        self.ycell = sequence(0.5 - observer[1], 1.0, self.shape[1])
        self.zcell = sequence(0.5 - observer[2], 1.0, self.shape[2])

    def autocorrelate(self, grid):
        self.set_grid2(grid)
        return self._correlate(grid, self.grid2_fft)

    def cross_correlate(self, grid1, grid2=None):
        if grid2 is not None:
            self.set_grid2(grid2)
        return self._correlate(grid1, self.grid2_fft)

    def _correlate(self, grid1, grid2_fft, fft1=None):
        g = self.geom
        rhist, khist = self._hists()
        acc = "accumulate_fast" if self.fast_hist else "accumulate"
        grid1 = np.asarray(grid1, np.float64)
        ncells = int(np.prod(self.shape))
        zerolag = None
        G2c = np.conj(grid2_fft)
        for ell in range(0, self.maxell + 1, 2):
            rsub = np.zeros(g.rshape)
            ksub = np.zeros(g.kshape)
            for m in range(-ell, ell + 1):
                work = make_ylm(ell, m, self.xcell, self.ycell, self.zcell, 1.0, 0, grid1)
                P = rfft3(work) * G2c
                kylm = make_ylm(ell, m, g.kx, g.ky, g.kz, 4.0 * math.pi, 0, g.inv_window)
                ksub += extract_fft_submatrix_c2r(P, g.kshape, kylm)
                xi = irfft3(P, self.shape)
                rylm = make_ylm(ell, m, g.rx, g.ry, g.rz, 4.0 * math.pi / ncells, 0, None)
                rsub += extract_submatrix(xi, g.rshape, rylm)
            getattr(rhist, acc)(ell // 2, g.rnorm, rsub)
            getattr(khist, acc)(ell // 2, g.knorm, ksub)
            if ell == 0:
                zerolag = rsub[g.rzero]
        return _results(rhist, khist, zerolag)


# ---------------------------------------------------------------------------
# fftcorr/correlate/correlator.pyx:42-70 (table assembly, numpy only)
def correlation_table(res, squeeze=True):
    hist = res["r_hist"]
    counts = res["r_counts"]
    with np.errstate(invalid="ignore", divide="ignore"):
        xi = hist / counts
    if squeeze:
        xi = np.squeeze(xi)
        hist = np.squeeze(hist)
    return {"r": res["r"], "xi": np.transpose(xi), "histogram": np.transpose(hist),
            "count": counts[0, :]}


# cc/fftcorr.cc:213-228 print_hist, as rows of numbers (no text formatting)
def print_hist_rows(bins, counts, hist, prefix, normalize):
    rows = []
    for j in range(len(bins)):
        count = float(counts[0, j])
        denom = count if (normalize and count != 0) else 1.0
        rows.append([prefix, bins[j], count] + [hist[i, j] / denom for i in range(hist.shape[0])])
    return np.array(rows)
