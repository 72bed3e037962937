"""The GPU correlators against the direct-definition multipole estimators of
tests/brute_force.py (plain DFT sums, periodic pair products and Legendre
polynomials), with no oracle in between: pins the l = 2, 4 normalisation and
the line of sight of the HIP path itself (cc/correlate/correlator.h:296-357,
:405-491).  FP64 forms within 1e-9, FP32/mixed forms within 3e-5 of the floor
metric; counts bit-exact."""
import numpy as np
import pytest

import brute_force as BF
from fftcorr_amd.correlate import Correlator, PeriodicCorrelator
from gpu_helpers import floor_rel_err

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-9, "fp32": 3e-5, "mixed": 3e-5}
FORMS = [("fp64", "fused"), ("fp64", "plain"), ("mixed", "fused"), ("fp32", "fused")]
CASES = [
    ((8, 8, 8), 10.0, (0, None), 30.0, 10.0, 0.33, 0.037),
    ((6, 7, 9), 7.0, (3, "cic"), 21.0, 7.0, 0.47, 0.061),
    ((8, 6, 10), 5.0, (1, "tsc"), 15.0, 5.0, 0.6, 0.087),
]


def _anisotropic_field(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=shape)
    F = np.fft.rfftn(a)
    k = [np.fft.fftfreq(n) for n in shape[:2]] + [np.fft.rfftfreq(shape[2])]
    kk = (0.6 * k[0][:, None, None] ** 2 + 1.0 * k[1][None, :, None] ** 2
          + 3.0 * k[2][None, None, :] ** 2)
    return np.fft.irfftn(F * np.exp(-kk / 0.05), s=shape, axes=(0, 1, 2)) + 0.1 * a


@pytest.mark.parametrize("form", FORMS, ids=["-".join(f) for f in FORMS])
@pytest.mark.parametrize("case", CASES, ids=[str(c[0]) for c in CASES])
def test_periodic_gpu_matches_brute_force(case, form):
    precision, algorithm = form
    shape, cell, (wcode, wname), rmax, dr, kmax, dk = case
    delta = _anisotropic_field(shape, 40 + shape[2])
    c = PeriodicCorrelator(shape, cell, wcode, rmax, dr, kmax, dk, 4, precision=precision)
    c.algorithm = algorithm
    c.autocorrelate(delta)
    res = c.results()
    kc, ks = BF.power_multipoles(delta, cell, kmax, dk, 4, window=wname)
    rc, rs = BF.correlation_multipoles(delta, cell, rmax, dr, 4)
    np.testing.assert_array_equal(res["k_counts"], kc)
    np.testing.assert_array_equal(res["r_counts"], rc)
    assert floor_rel_err(res["k_hist"], ks).max() <= TOL[precision]
    assert floor_rel_err(res["r_hist"], rs).max() <= TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_far_observer_gpu_yamamoto_equals_brute_force(precision):
    """Observer 1e9 cells away along -z: every line of sight is +z, so the
    Yamamoto correlator's multipoles are the periodic ones times N / cell^3
    (k) and N (r)."""
    shape, cell, wcode, wname = (8, 8, 8), 10.0, 0, None
    rmax, dr, kmax, dk = 30.0, 10.0, 0.33, 0.037
    delta = _anisotropic_field(shape, 77)
    N = delta.size
    posmin = (-shape[0] / 2.0 * cell, -shape[1] / 2.0 * cell, 1e9 * cell)
    c = Correlator(shape, cell, posmin, wcode, rmax, dr, kmax, dk, 4, precision=precision)
    c.autocorrelate(delta)
    res = c.results()
    kc, ks = BF.power_multipoles(delta, cell, kmax, dk, 4, window=wname)
    rc, rs = BF.correlation_multipoles(delta, cell, rmax, dr, 4)
    np.testing.assert_array_equal(res["k_counts"], kc)
    np.testing.assert_array_equal(res["r_counts"], rc)
    tol = 1e-6 if precision == "fp64" else 3e-5
    assert floor_rel_err(res["k_hist"] * cell ** 3 / N, ks).max() <= tol
    assert floor_rel_err(res["r_hist"] / N, rs).max() <= tol
