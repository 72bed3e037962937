"""Output tables, Yamamoto set_grid2_fft and the NN/RR estimator flow on the GPU.

* Tables (fftcorr/correlate/correlator.pyx:42-70): the values of the
  (power_spectrum, correlation) tables the GPU correlators return -- x = bin
  centres, xi/ps = histogram / count per multipole with 0/0 -> nan (empty
  bins), np.squeeze then transpose to [nbins, nell], count = counts of the
  l = 0 row -- against the FP64 oracle's raw histograms, for l <= 4 (2-D
  columns), l = 0 (squeezed to 1-D) and squeeze=False.
* Correlator.set_grid2_fft (cc/correlate/correlator.h:69-77) for the
  Yamamoto kind: the cross-correlation against a caller-supplied complex
  grid equals the one against the real grid it transforms.
* NN/RR flow (fftcorr/fftcorr.py:434-452, test_fftcorr_aniso.py:115-134):
  the GPU CLI (D12 gridding) correlates D - R against R and R alone on the
  reference's north and south catalogues; compute_xi of the printed
  correlation rows reproduces the reference's xi.txt at rtol 1e-5.
"""
import os

import numpy as np
import pytest

from fftcorr_amd import cli
from fftcorr_amd.correlate import Correlator, PeriodicCorrelator
from fftcorr_amd.correlate.utils import compute_xi
from fftcorr_amd.particle_mesh import set_wavelet_table
from gpu_helpers import floor_rel_err
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=shape)
    F = np.fft.rfftn(a)
    k = [np.fft.fftfreq(n) for n in shape[:2]] + [np.fft.rfftfreq(shape[2])]
    kk = np.sqrt(k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2
                 + (1.5 * k[2][None, None, :]) ** 2)
    return np.fft.irfftn(F * np.exp(-(kk / 0.2) ** 2), s=shape, axes=(0, 1, 2))


def _check_table(table, xname, vname, x, hist, counts, nell, squeeze):
    """hist, counts: the oracle's [nell, nbins] rows."""
    nb = len(x)
    assert table.colnames == [xname, vname, "histogram", "count"]
    assert len(table) == nb
    np.testing.assert_array_equal(table[xname], x)
    np.testing.assert_array_equal(table["count"], counts[0])
    vals, hcol = np.asarray(table[vname]), np.asarray(table["histogram"])
    want_shape = (nb,) if (squeeze and nell == 1) else (nb, nell)
    assert vals.shape == want_shape and hcol.shape == want_shape
    vals = vals.reshape(nb, nell)
    hcol = hcol.reshape(nb, nell)
    empty = counts.T == 0
    assert empty.any() and (~empty).any()  # the case must exercise both
    # 0/0 -> nan exactly where a bin has no cells, finite elsewhere.
    np.testing.assert_array_equal(np.isnan(vals), empty)
    ref_vals = np.where(empty, 0.0, hist.T / np.where(empty, 1, counts.T))
    got_vals = np.where(empty, 0.0, vals)
    assert floor_rel_err(got_vals.T, ref_vals.T).max() < 1e-9
    assert floor_rel_err(hcol.T, hist).max() < 1e-9


@pytest.mark.parametrize("maxell,squeeze", [(4, True), (0, True), (2, False), (0, False)])
def test_periodic_table_values(maxell, squeeze):
    # dr below the cell size and dk below the fundamental mode leave empty bins.
    shape, cell = (16, 16, 16), 10.0
    args = (shape, cell, 1, 60.0, 4.0, 0.4, 0.01, maxell)
    delta = _field(shape, 7 + maxell)
    c = PeriodicCorrelator(*args)
    ps, xi = c.autocorrelate(delta, squeeze=squeeze)
    ref = O.PeriodicCorrelator(*args)
    ref.fast_hist = True
    want = ref.autocorrelate(delta)
    nell = maxell // 2 + 1
    _check_table(xi, "r", "xi", want["r"], want["r_hist"], want["r_counts"], nell, squeeze)
    _check_table(ps, "k", "ps", want["k"], want["k_hist"], want["k_counts"], nell, squeeze)


def test_yamamoto_table_values_and_grid2_fft():
    shape = (16, 20, 24)
    args = (shape, 6.0, (200.0, -40.0, 90.0), 2, 50.0, 3.0, 0.3, 0.01, 4)
    a, b = _field(shape, 21), _field(shape, 22)
    ref = O.Correlator(*args)
    ref.fast_hist = True
    want = ref.cross_correlate(a, b)
    # grid2 handed over as its FFT (complex128 [n0, n1, n2/2+1], FFTW's
    # unnormalised forward transform = numpy's rfftn).
    c = Correlator(*args)
    c.set_grid2_fft(np.fft.rfftn(b))
    ps, xi = c.cross_correlate(a)
    res = c.results()
    np.testing.assert_array_equal(res["r_counts"], want["r_counts"])
    np.testing.assert_array_equal(res["k_counts"], want["k_counts"])
    assert floor_rel_err(res["r_hist"], want["r_hist"]).max() < 1e-9
    assert floor_rel_err(res["k_hist"], want["k_hist"]).max() < 1e-9
    _check_table(xi, "r", "xi", want["r"], want["r_hist"], want["r_counts"], 3, True)
    _check_table(ps, "k", "ps", want["k"], want["k_hist"], want["k_counts"], 3, True)
    # The same correlator against the real grid gives the same numbers.
    c2 = Correlator(*args)
    c2.cross_correlate(a, b)
    r2 = c2.results()
    assert floor_rel_err(r2["r_hist"], res["r_hist"]).max() < 1e-12
    assert floor_rel_err(r2["k_hist"], res["k_hist"]).max() < 1e-12


def _corr_rows(path):
    data = np.loadtxt(path)
    data = data[data[:, 0] == 0]  # correlation rows carry 0 in the first column
    return data[:, 3:], data[:, 1]


@pytest.fixture
def d12_table(ref):
    set_wavelet_table(ref.wavelet_table())
    yield
    set_wavelet_table(None)


@pytest.mark.parametrize("hemi", ["north", "south"])
def test_nn_rr_flow_reproduces_reference_xi(d12_table, tmp_path, hemi):
    dd = os.path.join(GOLDEN, hemi, "corrDD.dat")
    rr = os.path.join(GOLDEN, hemi, "corrRR.dat")
    common = ["-n", "256", "-ell", "2", "-r", "200", "-dr", "10",
              "-wisdom", str(tmp_path / "w.txt")]
    out_nn, out_rr = tmp_path / "corrDD.dat.out", tmp_path / "corrRR.dat.out"
    cli.run(common + ["-in", dd, "-in2", rr, "-out", str(out_nn)])
    cli.run(common + ["-in", rr, "-out", str(out_rr)])
    nn, rcen = _corr_rows(out_nn)
    rrh, rcen2 = _corr_rows(out_rr)
    np.testing.assert_array_equal(rcen, rcen2)
    xi = compute_xi(nn, rrh)
    gold = np.loadtxt(os.path.join(GOLDEN, hemi, "nonperiodic", "xi.txt")).T
    np.testing.assert_allclose(rcen, gold[0])
    np.testing.assert_allclose(xi, gold[1:].T, rtol=1e-5)
