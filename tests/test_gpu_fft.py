"""GPU 3D R2C/C2R FFT (FftGrid, cc/grid/fft_grid.cc:69-178) against numpy's
pocketfft in FP64.  Tolerances: FP64 max|err| <= 1e-12 max|F|; FP32 <= 3e-6."""
import numpy as np
import pytest

from gpu_helpers import fft_c2r, fft_r2c

pytestmark = pytest.mark.gpu

# (4, 512, 64) and (32, 512, 512): 512-point y passes, which run as the
# persistent prefetching column kernel (the second with several tiles per
# workgroup).
SHAPES = [(8, 8, 8), (16, 16, 16), (4, 6, 10), (5, 6, 7), (12, 10, 9), (24, 48, 96),
          (64, 32, 128), (128, 128, 128), (8, 16, 384), (3, 5, 2), (1, 1, 4), (4, 512, 64),
          (32, 512, 512)]
TOL = {"fp64": 1e-12, "fp32": 3e-6}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_matches_numpy(shape, precision):
    rng = np.random.default_rng(sum(shape))
    a = rng.normal(size=shape)
    got = fft_r2c(a, precision)
    ref = np.fft.rfftn(a)
    assert got.shape == ref.shape
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err <= TOL[precision] * np.log2(max(a.size, 2)), err


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", SHAPES)
def test_c2r_matches_numpy_unnormalised(shape, precision):
    rng = np.random.default_rng(sum(shape) + 1)
    F = np.fft.rfftn(rng.normal(size=shape))
    got = fft_c2r(F, shape, precision)
    ref = np.fft.irfftn(F, s=shape, axes=(0, 1, 2), norm="forward")
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err <= TOL[precision] * np.log2(max(np.prod(shape), 2)), err


def test_c2r_ignores_imaginary_dc_and_nyquist():
    # FFTW's c2r treats X[0] and X[n/2] along the last axis as real.
    shape = (8, 8, 8)
    rng = np.random.default_rng(3)
    F = np.fft.rfftn(rng.normal(size=shape))
    G = F.copy()
    G[:, :, 0] += 1j * rng.normal(size=G[:, :, 0].shape)
    G[:, :, -1] += 1j * rng.normal(size=G[:, :, -1].shape)
    # x, y are full complex axes, so only the kz = 0 / Nyquist planes' own
    # Hermitian partners matter; compare against numpy on the same input.
    ref = np.fft.irfftn(G, s=shape, axes=(0, 1, 2), norm="forward")
    got = fft_c2r(G, shape, "fp64")
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.max(np.abs(ref)))


def test_linearity_and_roundtrip_at_256():
    shape = (256, 256, 256)
    rng = np.random.default_rng(9)
    a = rng.normal(size=shape).astype(np.float64)
    F = fft_r2c(a, "fp32")
    back = fft_c2r(F, shape, "fp32") / a.size
    assert np.max(np.abs(back - a)) < 2e-5 * np.max(np.abs(a))
    # Parseval: sum |a|^2 = (1/N) sum over the full spectrum.
    w = np.full(F.shape[2], 2.0)
    w[0] = 1.0
    w[-1] = 1.0  # even n2: Nyquist counted once
    p = np.sum(np.abs(F) ** 2 * w) / a.size
    assert p == pytest.approx(np.sum(a * a), rel=1e-5)
