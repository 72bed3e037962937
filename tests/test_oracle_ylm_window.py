"""Pins make_ylm (cc/correlate/spherical_harmonics.h:32-179) and the window
functions (cc/correlate/window_correct.h:14-60)."""
import math

import numpy as np
import pytest
import scipy.special

from oracle import fftcorr_oracle as O

PAIRS = [(0, 0)] + [(2, m) for m in range(-2, 3)] + [(4, m) for m in range(-4, 5)]


def _axes(rng):
    x = np.concatenate([[0.0], rng.uniform(-3, 3, 6)])
    y = np.concatenate([[0.0], rng.uniform(-3, 3, 5)])
    z = np.concatenate([[0.0], rng.uniform(-3, 3, 4)])
    return x, y, z


@pytest.mark.parametrize("ell,m", PAIRS)
def test_bit_exact_against_reference(ref, ell, m):
    rng = np.random.default_rng(ell * 10 + m + 5)
    x, y, z = _axes(rng)
    mult = rng.normal(size=(len(x), len(y), len(z)))
    for coeff, mu in [(1.0, None), (4 * math.pi, mult), (0.37, mult)]:
        np.testing.assert_array_equal(O.make_ylm(ell, m, x, y, z, coeff, 0, mu),
                                      ref.make_ylm(ell, m, x, y, z, coeff, 0, mu))


def _textbook_real_ylm(ell, m, x, y, z):
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    theta = np.arccos(np.clip(Z / np.where(r > 0, r, 1), -1, 1))
    phi = np.arctan2(Y, X)
    c = scipy.special.sph_harm_y(ell, abs(m), theta, phi)
    if m == 0:
        return np.real(c), r
    return (math.sqrt(2) * (np.real(c) if m > 0 else np.imag(c))), r


@pytest.mark.parametrize("ell,m", PAIRS)
def test_textbook_up_to_sign(ell, m):
    rng = np.random.default_rng(3)
    x, y, z = (rng.uniform(-2, 2, 7) for _ in range(3))
    ours = O.make_ylm(ell, m, x, y, z, 1.0)
    tb, _ = _textbook_real_ylm(ell, m, x, y, z)
    sign = np.sign(np.sum(ours * tb))
    np.testing.assert_allclose(ours, sign * tb, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("ell", [0, 2, 4])
def test_addition_theorem_sign_free(ell):
    # sum_m Y_lm(a) Y_lm(b) = (2l+1)/(4 pi) P_l(a.b)
    rng = np.random.default_rng(11 + ell)
    a = rng.normal(size=3)
    b = rng.normal(size=3)
    s = 0.0
    for m in range(-ell, ell + 1):
        ya = O.make_ylm(ell, m, [a[0]], [a[1]], [a[2]], 1.0)[0, 0, 0]
        yb = O.make_ylm(ell, m, [b[0]], [b[1]], [b[2]], 1.0)[0, 0, 0]
        s += ya * yb
    mu = a @ b / np.linalg.norm(a) / np.linalg.norm(b)
    expect = (2 * ell + 1) / (4 * math.pi) * scipy.special.eval_legendre(ell, mu)
    assert s == pytest.approx(expect, rel=1e-12, abs=1e-14)


def test_ylm_argument_errors():
    for ell, m, e in [(1, 0, 0), (6, 0, 0), (2, 3, 0), (2, 0, 1)]:
        with pytest.raises(ValueError):
            O.make_ylm(ell, m, [1.0], [1.0], [1.0], 1.0, e)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_window_bit_exact_against_reference(ref, kind):
    for k in np.concatenate([[0.0], np.linspace(-0.8, 0.8, 41)]):
        for cell in (3.90625, 10.856001976670939):
            assert O.window_squared_norm(kind, float(k), cell) == ref.window(kind, float(k), cell)


def test_window_closed_forms():
    cell = 2.0
    for k in np.linspace(0.05, 1.5, 9):
        s = math.sin(k * cell / 2) / (k * cell / 2)
        assert O.window_squared_norm(O.TSC_CORRECTION, k, cell) == pytest.approx(s ** 6, rel=1e-14)
        assert O.window_squared_norm(O.CIC_CORRECTION, k, cell) == pytest.approx(s ** 4, rel=1e-14)
        sn = math.sin(k * cell / 2) ** 2
        # Aliased sums: sum_n sinc^{2p}(x + pi n) closed forms for p = 2, 3.
        n = np.arange(-4000, 4001)
        xs = k * cell / 2 + math.pi * n
        sinc2 = np.where(xs == 0, 1.0, np.sin(xs) / np.where(xs == 0, 1, xs)) ** 2
        assert O.window_squared_norm(O.CIC_ALIASED_CORRECTION, k, cell) == pytest.approx(
            np.sum(sinc2 ** 2), rel=1e-9)
        assert O.window_squared_norm(O.TSC_ALIASED_CORRECTION, k, cell) == pytest.approx(
            np.sum(sinc2 ** 3), rel=1e-9)
        assert 1 - sn + 2 / 15 * sn * sn == O.window_squared_norm(O.TSC_ALIASED_CORRECTION, k, cell)
