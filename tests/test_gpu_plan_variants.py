"""Radix plan variants of the register-chained FFT kernels (first radix 8 or
16), forced through the plan cache: each variant transforms like numpy, a
correlation whose z and y passes run variant 1 equals the variant-0 result
to FP64 rounding and matches the oracle, and the autotuner's choice is one of
the variants and keeps the results."""
import numpy as np
import pytest

from fftcorr_amd import fftw
from fftcorr_amd.autotune import autotune, fft_radices, plan_variants
from fftcorr_amd.correlate import PeriodicCorrelator
from gpu_helpers import fft_c2r, fft_r2c, floor_rel_err
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_cache():
    fftw.forget_wisdom()
    yield
    fftw.forget_wisdom()


def _force_variant(tmp_path, var, lengths):
    """Plan-cache lines selecting variant `var` for the column (kind 3) and
    row (kind 4) passes of these lengths, at both plan precisions."""
    lines = ["fftcorr_hip_tiles v1"]
    for n in lengths:
        for prec in (0, 1):
            lines += ["%d %d 3 %d" % (n, prec, var), "%d %d 4 %d" % (n, prec, var)]
    path = tmp_path / ("variant%d.txt" % var)
    path.write_text("\n".join(lines) + "\n")
    fftw.import_wisdom_from_file(str(path))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(128, 256, 512), (512, 128, 256)])
def test_variant1_transforms_match_numpy(tmp_path, shape, precision):
    _force_variant(tmp_path, 1, (128, 256, 512))
    pv = plan_variants(shape, precision)
    assert (pv["x_variant"], pv["y_variant"], pv["z_variant"]) == (1, 1, 1)
    assert pv["y_radices"][0] == 16 and np.prod(pv["y_radices"]) == shape[1]
    x = np.random.default_rng(11).normal(size=shape)
    ref = np.fft.rfftn(x)
    tol = 1e-12 if precision == "fp64" else 3e-6
    got = fft_r2c(x, precision)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    back = fft_c2r(ref, shape, precision)
    want = np.fft.irfftn(ref, shape, axes=(0, 1, 2)) * x.size
    assert np.abs(back - want).max() <= tol * np.abs(want).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_variant1_correlation_equals_variant0_and_oracle(tmp_path, precision):
    # y = 128 and the z row FFT (n2 / 2 = 128) have variant 1; the fused x
    # pass keeps its own plan.
    shape = (64, 128, 256)
    rng = np.random.default_rng(21)
    a = rng.normal(size=shape)
    args = (shape, 4.0, 2, 60.0, 5.0, 0.6, 0.05, 4)
    c0 = PeriodicCorrelator(*args, precision=precision)
    c0.autocorrelate(a)
    r0 = c0.results()
    _force_variant(tmp_path, 1, (128,))
    assert plan_variants(shape, precision)["z_variant"] == 1
    c1 = PeriodicCorrelator(*args, precision=precision)
    c1.autocorrelate(a)
    r1 = c1.results()
    ref = O.PeriodicCorrelator(*args)
    ref.fast_hist = True
    want = ref.autocorrelate(a)
    np.testing.assert_array_equal(r1["k_counts"], want["k_counts"])
    np.testing.assert_array_equal(r1["r_counts"], want["r_counts"])
    same = 1e-12 if precision == "fp64" else 1e-5
    tol = 1e-9 if precision == "fp64" else 1e-5
    for key in ("k_hist", "r_hist"):
        assert floor_rel_err(r1[key], r0[key]).max() <= same, key
        assert floor_rel_err(r1[key], want[key]).max() <= tol, key


def test_autotune_mixed_chooses_a_variant_and_keeps_results():
    shape = (128, 128, 256)
    res = autotune(shape, "mixed", reps=2)
    for ax, n in (("x", 128), ("y", 128), ("z", 128)):
        assert res[ax + "_variant"] in (0, 1)
        assert res[ax + "_radices"] == fft_radices(n, res[ax + "_variant"])
    assert plan_variants(shape, "mixed") == {k: res[k] for k in res
                                             if k.endswith(("_variant", "_radices"))}
    a = np.random.default_rng(5).normal(size=shape)
    args = (shape, 4.0, 2, 60.0, 5.0, 0.6, 0.05, 2)
    cm = PeriodicCorrelator(*args, precision="mixed")
    cm.autocorrelate(a)
    c64 = PeriodicCorrelator(*args, precision="fp64")
    c64.autocorrelate(a)
    got, want = cm.results(), c64.results()
    for key in ("k_hist", "r_hist"):
        assert floor_rel_err(got[key], want[key]).max() <= 1e-5, key
