"""Pins the oracle's FFT conventions, subgrid extraction and the periodic
correlator (cc/grid/fft_grid.cc, cc/correlate/correlator.h:109-357) against
brute force and closed forms."""
import math

import numpy as np
import pytest

from oracle import fftcorr_oracle as O


def test_fft_convention_brute_force():
    rng = np.random.default_rng(0)
    shape = (4, 3, 5)
    a = rng.normal(size=shape)
    F = O.rfft3(a)
    n = np.indices(shape).reshape(3, -1).T
    for k in [(0, 0, 0), (1, 2, 1), (3, 1, 2)]:
        ph = np.exp(-2j * np.pi * sum(n[:, d] * k[d] / shape[d] for d in range(3)))
        assert F[k] == pytest.approx(np.sum(a.reshape(-1) * ph), rel=1e-12, abs=1e-12)
    # C2R is unnormalised: irfft(rfft(a)) = N a.
    np.testing.assert_allclose(O.irfft3(F, shape), a * a.size, rtol=1e-12, atol=1e-12)


def test_extract_submatrix_index_map():
    rng = np.random.default_rng(1)
    g = rng.normal(size=(6, 7, 8))
    sub = O.extract_submatrix(g, (5, 5, 5))
    for i, j, k in [(0, 0, 0), (2, 2, 2), (4, 1, 3)]:
        assert sub[i, j, k] == g[(i - 2) % 6, (j - 2) % 7, (k - 2) % 8]


def test_extract_fft_submatrix_hermitian_reflection():
    rng = np.random.default_rng(2)
    shape = (8, 6, 10)
    a = rng.normal(size=shape)
    half = O.rfft3(a)
    full = np.fft.fftn(a)
    ks = (5, 5, 7)
    sub = O.extract_fft_submatrix_c2r(half, ks)
    o = [s // 2 for s in ks]
    for i in range(ks[0]):
        for j in range(ks[1]):
            for k in range(ks[2]):
                q = ((i - o[0]) % shape[0], (j - o[1]) % shape[1], (k - o[2]) % shape[2])
                assert sub[i, j, k] == pytest.approx(full[q].real, rel=1e-10, abs=1e-10)


def _brute_xi0(delta, geom, rmax, dr):
    """Periodic pair sum over cell lags: xi(r) = <d(x) d(x+r)> binned."""
    n = delta.shape
    h = O.HistogramList(1, 0.0, rmax, dr)
    acc = np.zeros(geom.rshape)
    o = geom.rzero
    for i in range(geom.rshape[0]):
        for j in range(geom.rshape[1]):
            for k in range(geom.rshape[2]):
                s = (i - o[0], j - o[1], k - o[2])
                acc[i, j, k] = np.mean(delta * np.roll(delta, shift=(-s[0], -s[1], -s[2]),
                                                       axis=(0, 1, 2)))
    h.accumulate_fast(0, geom.rnorm, acc)
    return h


def test_periodic_xi0_brute_force_pair_sum():
    rng = np.random.default_rng(3)
    shape = (8, 8, 8)
    cell = 10.0
    delta = rng.normal(size=shape)
    c = O.PeriodicCorrelator(shape, cell, 0, 30.0, 10.0, 0.3, 0.05, 0)
    res = c.autocorrelate(delta)
    hb = _brute_xi0(delta, c.geom, 30.0, 10.0)
    np.testing.assert_array_equal(res["r_counts"], hb.counts)
    np.testing.assert_allclose(res["r_hist"], hb.hist_values, rtol=1e-11, atol=1e-12)
    assert res["zerolag"] == pytest.approx(np.sum(delta ** 2) / delta.size, rel=1e-12)


def test_power_spectrum_integral_identity():
    # cc/fftcorr.cc:441-448: sum_k P_0 / V over all modes = <delta^2>
    # (Parseval).  Odd n: the truncated k-subgrid covers every mode once.
    rng = np.random.default_rng(4)
    for n, cell in [(7, 5.0), (8, 5.0)]:
        shape = (n, n, n)
        delta = rng.normal(size=shape)
        kmax = math.pi / cell * math.sqrt(3) * 1.01  # covers every mode
        c = O.PeriodicCorrelator(shape, cell, 0, 10.0, 5.0, kmax, kmax / 4, 0)
        res = c.autocorrelate(delta)
        V = cell ** 3 * delta.size
        total = res["k_hist"][0].sum() / V
        if n % 2:
            assert total == pytest.approx(np.mean(delta ** 2), rel=1e-12)
        else:
            # kshape truncates to 2*(n//2)+1 = n+1: the Nyquist index appears
            # twice per axis (modulo indexing, fft_grid.cc:229-234).
            full = np.abs(np.fft.fftn(delta)) ** 2
            idx = (np.arange(n + 1) - n // 2) % n
            dbl = full[np.ix_(idx, idx, idx)].sum()
            assert total == pytest.approx(dbl * cell ** 3 / delta.size / V, rel=1e-12)
        assert res["k_counts"][0].sum() == c.geom.kshape[0] ** 3


def test_isotropic_field_has_no_quadrupole():
    # A field symmetric under x<->z exchange gives xi_2 summed over a shell
    # ~0 only statistically; use an exactly isotropic-in-mean construction:
    # the lattice-symmetrised field (average over the 6 axis permutations)
    # has a correlation with zero l=2,4 sums over every cubic-symmetric shell.
    rng = np.random.default_rng(6)
    n = 8
    a = rng.normal(size=(n, n, n))
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    c = O.PeriodicCorrelator((n, n, n), 1.0, 0, 3.0, 1.0, 1.0, 0.25, 4)
    hists = [c.autocorrelate(np.transpose(a, p)) for p in perms]
    r2 = sum(h["r_hist"][1] for h in hists)
    r0 = sum(h["r_hist"][0] for h in hists)
    # Y_20 averaged over the permutations of (x,y,z) vanishes identically.
    assert np.max(np.abs(r2)) <= 1e-12 * np.max(np.abs(r0))


def test_interlace_recovers_band_limited_field():
    # Grid b samples the field at x_i - h/2 (particles displaced by +h/2).
    # For a band-limited field F_b = F_a exp(-i k h/2) exactly, so the
    # interlaced transform must equal F_a; a sign error would not.
    rng = np.random.default_rng(12)
    n = (8, 10, 12)
    modes = [(1, 2, 3), (-3, 1, 0), (2, -4, 5), (0, 0, 1)]
    amps = rng.normal(size=len(modes)) + 1j * rng.normal(size=len(modes))
    idx = np.indices(n).astype(float)

    def field(offset):
        f = np.zeros(n)
        for (a, b, c), amp in zip(modes, amps):
            ph = 2 * np.pi * ((idx[0] + offset) * a / n[0] + (idx[1] + offset) * b / n[1]
                              + (idx[2] + offset) * c / n[2])
            f += np.real(amp * np.exp(1j * ph))
        return f

    fa = O.rfft3(field(0.0))
    fb = O.rfft3(field(-0.5))
    np.testing.assert_allclose(O.interlace(fa, fb, n), fa, atol=1e-10 * np.abs(fa).max())
    # A flipped phase sign would not reproduce F_a.
    bad = 0.5 * (fa + fb * np.conj(O.interlace(np.zeros_like(fa), np.ones_like(fa), n) * 2))
    assert np.abs(bad - fa).max() > 1e-3 * np.abs(fa).max()
