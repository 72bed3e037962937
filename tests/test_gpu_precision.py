"""Accuracy of the reduced-precision forms at realistic sizes: the mixed
form stays within 1e-5 of the FP64 form on every multipole value, |x - ref| /
max(|ref|, 1e-3 max|ref_0|) per bin, for the periodic correlator on a
clustered catalogue (256^3, CIC with interlacing, l <= 4) -- where it is
also clearly more accurate than the all-FP32 form (the FP64 arithmetic of
its z and y passes at work) -- and for the anisotropic correlator on data
minus randoms (256^3, TSC); counts are identical."""
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


def test_mixed_yamamoto_data_minus_randoms():
    # The survey configuration of the anisotropic correlator: 2M clustered
    # points minus alpha x 20M randoms (FKP-like weights), 256^3 TSC grid,
    # observer at the origin, l <= 4.  Every multipole value of the mixed
    # form is within 1e-5 of the FP64 form on the same grid (worst: P_4).
    from fftcorr_amd.catalog import uniform_catalog
    from fftcorr_amd.correlate import Correlator
    d = thomas_catalog(20_000, 100, 5.0, 500.0, 1500.0, 2, wrap=False)
    r = uniform_catalog(10 * d.shape[0], 500.0, 1500.0, 3)
    wgt = 1.0 / (1 + d.shape[0] / 1e9 * 1e4)
    alpha = d.shape[0] / r.shape[0]
    pos = np.ascontiguousarray(np.concatenate([d, r]), np.float64)
    w = np.ascontiguousarray(np.concatenate([np.full(d.shape[0], wgt),
                                             np.full(r.shape[0], -alpha * wgt)]))
    pad = 200.0 / 1.5
    posmin = np.full(3, 500.0 - pad)
    cell = (1000.0 + 2 * pad) / 256
    res = {}
    for prec in ("fp64", "mixed"):
        g = _lib.DeviceGrid((256,) * 3, posmin, cell, prec)
        st = _lib.MaStats()
        _lib.check(_lib.lib().fc_mass_assign(g.handle, _lib.dptr(pos), pos.shape[0], 3,
                                             _lib.dptr(w), 1.0, 1, 0, 0.0, st))
        c = Correlator((256,) * 3, cell, posmin, 2, 200.0, 5.0, 0.1, 0.005, 4, precision=prec)
        _lib.check(_lib.lib().fc_corr_autocorrelate(c._handle, g.handle))
        c._fetch()
        res[prec] = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                     for k, v in c.results().items()}
    for key in ("k_hist", "r_hist"):
        cnt = key.replace("hist", "counts")
        np.testing.assert_array_equal(res["mixed"][cnt], res["fp64"][cnt])
        e = _err(res["mixed"], res["fp64"], key)
        assert e.max() < 1e-5, (key, e)
