"""Autotuner and plan-cache tunables on the GPU: tuned plans still transform
exactly, and a correlation whose z+y passes run in several x-slab chunks
(forced through the plan cache) matches the oracle."""
import numpy as np
import pytest

from fftcorr_amd import fftw
from fftcorr_amd.autotune import autotune, plan_tiles
from fftcorr_amd.correlate import Correlator, PeriodicCorrelator
from gpu_helpers import fft_r2c, floor_rel_err
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh_cache():
    fftw.forget_wisdom()
    yield
    fftw.forget_wisdom()


def test_autotune_records_tiles_and_keeps_transforms_exact():
    shape = (32, 16, 64)
    res = autotune(shape, "fp32", reps=2)
    for k in ("x_cols", "y_cols", "z_rows"):
        assert res[k] >= 1
    assert res["zy_chunk_kib"] >= 0
    assert all(v > 0 for v in res["ms"].values())
    assert plan_tiles(shape, "fp32") == {k: res[k] for k in
                                         ("x_cols", "y_cols", "z_rows", "zy_chunk_kib")}
    x = np.random.default_rng(3).normal(size=shape)
    ref = np.fft.rfftn(x)
    got = fft_r2c(x, "fp32")
    assert np.abs(got - ref).max() <= 3e-6 * np.abs(ref).max()


def _force_chunk(tmp_path, kib):
    path = tmp_path / "wisdom.txt"
    path.write_text("fftcorr_hip_tiles v1\n0 0 2 %d\n" % kib)
    fftw.import_wisdom_from_file(str(path))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chunked_forward_matches_oracle(tmp_path, precision):
    _force_chunk(tmp_path, 16)  # 16 KiB chunks: a few x slabs each
    shape = (32, 24, 20)
    assert plan_tiles(shape, precision)["zy_chunk_kib"] == 16
    rng = np.random.default_rng(7)
    a = rng.normal(size=shape)
    args = (shape, 5.0, 2, 40.0, 5.0, 0.5, 0.05, 4)
    ref = O.PeriodicCorrelator(*args)
    ref.fast_hist = True
    want = ref.autocorrelate(a)
    c = PeriodicCorrelator(*args, precision=precision)
    c.autocorrelate(a)
    got = c.results()
    tol = 1e-9 if precision == "fp64" else 3e-5
    np.testing.assert_array_equal(got["k_counts"], want["k_counts"])
    assert floor_rel_err(got["k_hist"], want["k_hist"]).max() <= tol
    assert floor_rel_err(got["r_hist"], want["r_hist"]).max() <= tol
    # Yamamoto: the Y_lm(x) loader must see each chunk's x offset.
    yargs = (shape, 6.0, (100.0, -50.0, 400.0), 1, 40.0, 5.0, 0.4, 0.04, 4)
    yref = O.Correlator(*yargs)
    yref.fast_hist = True
    ywant = yref.autocorrelate(a)
    yc = Correlator(*yargs, precision=precision)
    yc.autocorrelate(a)
    ygot = yc.results()
    assert floor_rel_err(ygot["k_hist"], ywant["k_hist"]).max() <= tol
    assert floor_rel_err(ygot["r_hist"], ywant["r_hist"]).max() <= tol


def test_wisdom_roundtrip_keeps_chunk_setting(tmp_path):
    _force_chunk(tmp_path, 4096)
    out = tmp_path / "out.txt"
    fftw.export_wisdom_to_file(str(out))
    assert "0 0 2 4096" in out.read_text().splitlines()
    fftw.forget_wisdom()
    fftw.import_wisdom_from_file(str(out))
    assert plan_tiles((64, 64, 64), "fp32")["zy_chunk_kib"] == 4096
