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


# ---------------------------------------------------------------------------
# Periodic multipoles l = 0, 2, 4 against the direct-definition estimators of
# tests/brute_force.py (Legendre polynomials, plain DFT sums, pair products):
# pins the (2l+1) L_l normalisation, the line of sight (z), the cell^3/N and
# 1/N^2 factors and the window division independently of make_ylm and of the
# oracle's coefficients (cc/correlate/correlator.h:296-357).
import brute_force as BF  # noqa: E402

_PERIODIC_PIN_CASES = [
    # shape, cell, window (oracle code, brute-force name), rmax, dr, kmax, dk
    ((8, 8, 8), 10.0, (0, None), 30.0, 10.0, 0.33, 0.037),
    ((6, 7, 9), 7.0, (3, "cic"), 21.0, 7.0, 0.47, 0.061),
    ((8, 6, 10), 5.0, (1, "tsc"), 15.0, 5.0, 0.6, 0.087),
]


def _anisotropic_field(shape, seed):
    """A random field with different correlation lengths along x, y and z,
    so every multipole and every choice of line of sight differ."""
    rng = np.random.default_rng(seed)
    a = rng.normal(size=shape)
    F = np.fft.rfftn(a)
    k = [np.fft.fftfreq(n) for n in shape[:2]] + [np.fft.rfftfreq(shape[2])]
    kk = (0.6 * k[0][:, None, None] ** 2 + 1.0 * k[1][None, :, None] ** 2
          + 3.0 * k[2][None, None, :] ** 2)
    return np.fft.irfftn(F * np.exp(-kk / 0.05), s=shape, axes=(0, 1, 2)) + 0.1 * a


@pytest.mark.parametrize("case", _PERIODIC_PIN_CASES)
def test_periodic_multipoles_match_legendre_brute_force(case):
    shape, cell, (wcode, wname), rmax, dr, kmax, dk = case
    delta = _anisotropic_field(shape, 40 + shape[2])
    c = O.PeriodicCorrelator(shape, cell, wcode, rmax, dr, kmax, dk, 4)
    res = c.autocorrelate(delta)
    kc, ks = BF.power_multipoles(delta, cell, kmax, dk, 4, window=wname)
    rc, rsum = BF.correlation_multipoles(delta, cell, rmax, dr, 4)
    np.testing.assert_array_equal(res["k_counts"], kc)
    np.testing.assert_array_equal(res["r_counts"], rc)
    # Every multipole is far from zero somewhere, so a wrong coefficient or
    # axis cannot hide inside the tolerance.
    for q in range(3):
        assert np.abs(ks[q]).max() > 1e-3 * np.abs(ks[0]).max()
        assert np.abs(rsum[q]).max() > 1e-3 * np.abs(rsum[0]).max()
    np.testing.assert_allclose(res["k_hist"], ks, rtol=1e-10, atol=1e-12 * np.abs(ks).max())
    np.testing.assert_allclose(res["r_hist"], rsum, rtol=1e-10,
                               atol=1e-12 * np.abs(rsum).max())


def test_periodic_cross_multipoles_match_brute_force():
    shape, cell = (8, 8, 8), 10.0
    d1 = _anisotropic_field(shape, 50)
    d2 = 0.5 * d1 + _anisotropic_field(shape, 51)
    c = O.PeriodicCorrelator(shape, cell, 0, 30.0, 10.0, 0.33, 0.037, 4)
    res = c.cross_correlate(d1, d2)
    kc, ks = BF.power_multipoles(d1, cell, 0.33, 0.037, 4, delta2=d2)
    rc, rsum = BF.correlation_multipoles(d1, cell, 30.0, 10.0, 4, delta2=d2)
    np.testing.assert_array_equal(res["k_counts"], kc)
    np.testing.assert_allclose(res["k_hist"], ks, rtol=1e-10, atol=1e-12 * np.abs(ks).max())
    np.testing.assert_allclose(res["r_hist"], rsum, rtol=1e-10,
                               atol=1e-12 * np.abs(rsum).max())


def test_periodic_multipole_pin_rejects_plausible_slips():
    """The brute force tells apart the slips a restatement could make: a
    sqrt(4 pi / (2l+1)) coefficient and a line of sight along x."""
    shape, cell = (8, 8, 8), 10.0
    delta = _anisotropic_field(shape, 48)
    kc, ks = BF.power_multipoles(delta, cell, 0.33, 0.037, 4)
    c = O.PeriodicCorrelator(shape, cell, 0, 30.0, 10.0, 0.33, 0.037, 4)
    good = c.autocorrelate(delta)["k_hist"]
    # LOS along x: the field transposed so its z axis becomes x.
    swapped = c.autocorrelate(np.ascontiguousarray(np.transpose(delta, (2, 1, 0))))["k_hist"]
    for q in (1, 2):
        ell = 2 * q
        wrong_coeff = good[q] / (2 * ell + 1)  # sqrt(4pi/(2l+1)) Y_l0 = L_l
        assert not np.allclose(wrong_coeff, ks[q], rtol=1e-3)
        assert not np.allclose(swapped[q], ks[q], rtol=1e-3)


# ---------------------------------------------------------------------------
# The far-observer limit (cc/correlate/correlator.h:388-394): with the observer
# 1e9 cells away along -z, every cell's line of sight is +z to 1e-8, and the
# Yamamoto Correlator must reproduce the periodic multipoles after its own
# normalisations: k histograms carry 4 pi without cell^3/N (periodic / Yamamoto
# = cell^3/N), r histograms 4 pi/N instead of 1/N^2 (periodic / Yamamoto = 1/N).
@pytest.mark.parametrize("shape,cell,wcode", [((8, 8, 8), 10.0, 0), ((6, 8, 7), 6.0, 3)])
def test_far_observer_yamamoto_equals_periodic(shape, cell, wcode):
    delta = _anisotropic_field(shape, 60 + shape[1])
    rmax, dr, kmax, dk = 3 * cell, cell, 0.4 * 10.0 / cell, 0.05 * 10.0 / cell
    per = O.PeriodicCorrelator(shape, cell, wcode, rmax, dr, kmax, dk, 4).autocorrelate(delta)
    observer = [shape[0] / 2.0, shape[1] / 2.0, -1e9]
    yam = O.Correlator(shape, cell, [0, 0, 0], wcode, rmax, dr, kmax, dk, 4,
                       observer=observer).autocorrelate(delta)
    N = delta.size
    np.testing.assert_array_equal(per["k_counts"], yam["k_counts"])
    np.testing.assert_array_equal(per["r_counts"], yam["r_counts"])
    k_scale = cell ** 3 / N
    for q in range(3):
        ref_k = per["k_hist"][q]
        ref_r = per["r_hist"][q]
        np.testing.assert_allclose(yam["k_hist"][q] * k_scale, ref_k, rtol=1e-6,
                                   atol=1e-7 * np.abs(per["k_hist"][0]).max())
        np.testing.assert_allclose(yam["r_hist"][q] / N, ref_r, rtol=1e-6,
                                   atol=1e-7 * np.abs(per["r_hist"][0]).max())
    # An observer along -x instead gives a line of sight along x: the l = 2
    # multipoles then differ from the periodic (LOS z) ones.
    yx = O.Correlator(shape, cell, [0, 0, 0], wcode, rmax, dr, kmax, dk, 4,
                      observer=[-1e9, shape[1] / 2.0, shape[2] / 2.0]).autocorrelate(delta)
    assert not np.allclose(yx["k_hist"][1] * k_scale, per["k_hist"][1], rtol=1e-3)
