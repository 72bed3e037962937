"""Accuracy of the reduced-precision forms on a clustered catalogue at a
realistic size (256^3, CIC with interlacing, l <= 4): the mixed form stays
within 1e-5 of the FP64 form on every multipole value, |x - ref| /
max(|ref|, 1e-3 max|ref_0|) per bin, and is clearly more accurate than the
all-FP32 form (the FP64 arithmetic of its z and y passes at work); counts
are identical."""
import numpy as np
import pytest

from fftcorr_amd import _lib
from fftcorr_amd.catalog import thomas_catalog
from fftcorr_amd.correlate import PeriodicCorrelator

pytestmark = pytest.mark.gpu

N, BOX = 256, 1000.0


def _grids(precision, pos):
    cell = BOX / N
    out = []
    for shift in (0.0, 0.5):
        g = _lib.DeviceGrid((N, N, N), (0, 0, 0), cell, precision)
        st = _lib.MaStats()
        _lib.check(_lib.lib().fc_mass_assign(g.handle, _lib.dptr(pos), pos.shape[0], 3, None,
                                             1.0, 3, 1, shift, st))
        g.normalize(st.totw / float(N) ** 3)
        out.append(g)
    return out


def _run(precision, pos):
    ga, gb = _grids(precision, pos)
    c = PeriodicCorrelator((N, N, N), BOX / N, 3, 100.0, 5.0, 0.5, 0.01, 4,
                           precision=precision)
    _lib.check(_lib.lib().fc_corr_autocorrelate_interlaced(c._handle, ga.handle, gb.handle))
    c._fetch()
    return {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
            for k, v in c.results().items()}


def _err(res, ref, key):
    cnt = key.replace("hist", "counts")
    ok = ref[cnt][0] > 0
    x = res[key][:, ok] / res[cnt][:, ok]
    r = ref[key][:, ok] / ref[cnt][:, ok]
    floor = 1e-3 * np.abs(r[0]).max()
    return (np.abs(x - r) / np.maximum(np.abs(r), floor)).max(axis=1)


@pytest.fixture(scope="module")
def results():
    pos = np.ascontiguousarray(thomas_catalog(10_000, 100, 3.0, 0.0, BOX, 11), np.float64)
    return {p: _run(p, pos) for p in ("fp64", "mixed", "fp32")}


def test_mixed_within_1e5_of_fp64(results):
    ref, mx = results["fp64"], results["mixed"]
    for key in ("k_hist", "r_hist"):
        cnt = key.replace("hist", "counts")
        np.testing.assert_array_equal(mx[cnt], ref[cnt])
        assert _err(mx, ref, key).max() < 1e-5, (key, _err(mx, ref, key))


def test_mixed_more_accurate_than_fp32(results):
    ref = results["fp64"]
    e_mx = _err(results["mixed"], ref, "k_hist")
    e_32 = _err(results["fp32"], ref, "k_hist")
    assert e_mx.max() < 0.5 * e_32.max(), (e_mx, e_32)
