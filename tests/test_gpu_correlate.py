"""GPU PeriodicCorrelator / Correlator (cc/correlate/correlator.h) against the
FP64 oracle.  Counts bit-exact.  Histogram values relative to
max(|ref|, 1e-3 max|ref_l0|) per bin: FP64 within 1e-9; FP32 and mixed (FP32
forward, FP64 inverse) within 3e-5 -- the smooth test fields put most of
xi_0 into a few small-r cells, where the three FP32 roundings of the
forward passes alone (numpy float32 emulation of the same Stockham passes,
and scipy's float32 FFT, on the same inputs) reach 1e-5 of the floor.  Both
the fused algorithm and the reference's plain stage sequence."""
import numpy as np
import pytest

from fftcorr_amd.correlate import Correlator, PeriodicCorrelator
from fftcorr_amd.grid import ConfigSpaceGrid
from gpu_helpers import floor_rel_err
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-9, "fp32": 3e-5, "mixed": 3e-5}
FORMS = [("fp64", "fused"), ("fp32", "fused"), ("mixed", "fused"), ("fp64", "plain"),
         ("fp32", "plain")]


def _field(shape, seed, smooth=True):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=shape)
    if smooth:  # give the field power at low k so multipoles are not pure noise
        F = np.fft.rfftn(a)
        k = [np.fft.fftfreq(n) for n in shape[:2]] + [np.fft.rfftfreq(shape[2])]
        kk = np.sqrt(k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2
                     + (1.3 * k[2][None, None, :]) ** 2)
        a = np.fft.irfftn(F * np.exp(-(kk / 0.15) ** 2), s=shape, axes=(0, 1, 2))
    return a


def _compare(res, ref, precision):
    np.testing.assert_array_equal(res["r_counts"], ref["r_counts"])
    np.testing.assert_array_equal(res["k_counts"], ref["k_counts"])
    np.testing.assert_array_equal(res["r"], ref["r"])
    np.testing.assert_array_equal(res["k"], ref["k"])
    er = floor_rel_err(res["r_hist"], ref["r_hist"])
    ek = floor_rel_err(res["k_hist"], ref["k_hist"])
    assert er.max() <= TOL[precision], ("r", er.max())
    assert ek.max() <= TOL[precision], ("k", ek.max())
    assert res["zerolag"] == pytest.approx(ref["zerolag"], rel=TOL[precision])


PERIODIC = [((16, 16, 16), 10.0, 0, 40.0, 10.0, 0.3, 0.05, 0),
            ((32, 32, 32), 5.0, 2, 50.0, 5.0, 0.5, 0.02, 4),
            ((24, 16, 20), 4.0, 1, 30.0, 3.0, 0.9, 0.05, 2),
            ((16, 16, 16), 10.0, 3, 40.0, 10.0, 0.4, 0.05, 4),
            ((16, 16, 16), 10.0, 4, 300.0, 10.0, 0.4, 0.05, 4),
            ((64, 64, 64), 7.8125, 0, 200.0, 5.0, 0.03, 0.01, 0),
            ((9, 10, 11), 3.0, 2, 20.0, 2.0, 2.0, 0.1, 4),
            ((512, 8, 8), 4.0, 3, 60.0, 4.0, 0.5, 0.05, 4),
            ((256, 8, 6), 4.0, 1, 40.0, 4.0, 0.5, 0.05, 2),
            ((128, 12, 16), 5.0, 0, 40.0, 5.0, 0.4, 0.05, 4),
            ((8, 8, 8), 5.0, 2, 20.0, 5.0, 0.6, 0.05, 4)]


def _make(cls, args, precision, algorithm):
    c = cls(*args, precision=precision)
    c.algorithm = algorithm
    return c


@pytest.mark.parametrize("form", FORMS, ids=["-".join(f) for f in FORMS])
@pytest.mark.parametrize("case", PERIODIC)
def test_periodic_against_oracle(case, form):
    precision, algorithm = form
    shape, cell, win, rmax, dr, kmax, dk, maxell = case
    delta = _field(shape, sum(shape) + maxell)
    c = _make(PeriodicCorrelator, case, precision, algorithm)
    ps, xi = c.autocorrelate(delta)
    ref = O.PeriodicCorrelator(shape, cell, win, rmax, dr, kmax, dk, maxell)
    ref.fast_hist = True
    _compare(c.results(), ref.autocorrelate(delta), precision)
    assert set(xi.colnames) == {"r", "xi", "histogram", "count"}
    assert set(ps.colnames) == {"k", "ps", "histogram", "count"}


INTERLACED = [((32, 32, 32), 5.0, 3, 50.0, 5.0, 0.5, 0.02, 4),
              ((64, 16, 24), 4.0, 4, 40.0, 4.0, 0.6, 0.05, 4),
              ((24, 16, 20), 4.0, 3, 30.0, 3.0, 0.9, 0.05, 2),
              ((512, 8, 8), 4.0, 3, 60.0, 4.0, 0.5, 0.05, 4),
              ((16, 16, 16), 10.0, 3, 300.0, 10.0, 0.4, 0.05, 4)]


@pytest.mark.parametrize("form", FORMS, ids=["-".join(f) for f in FORMS])
@pytest.mark.parametrize("case", INTERLACED)
def test_periodic_interlaced_against_oracle(case, form):
    """F = (F_a + F_b exp(i pi (fx/n0 + fy/n1 + fz/n2))) / 2 from two grids."""
    precision, algorithm = form
    shape = case[0]
    a = _field(shape, 3 + sum(shape))
    b = _field(shape, 4 + sum(shape))
    ref = O.PeriodicCorrelator(*case)
    ref.fast_hist = True
    want = ref.autocorrelate_interlaced(a, b)
    c = _make(PeriodicCorrelator, case, precision, algorithm)
    ga = ConfigSpaceGrid(shape, [0, 0, 0], [case[1] * n for n in shape], precision=precision)
    gb = ConfigSpaceGrid(shape, [0, 0, 0], [case[1] * n for n in shape], precision=precision)
    np.copyto(ga.data, a)
    np.copyto(gb.data, b)
    c.autocorrelate_interlaced(ga, gb)
    _compare(c.results(), want, precision)


# (512, 8, 8): the cross-mode fused x pass over a 512-point x axis runs as
# the persistent prefetching kernel for FP32 spectra.
@pytest.mark.parametrize("shape", [(32, 16, 24), (512, 8, 8)])
@pytest.mark.parametrize("form", FORMS, ids=["-".join(f) for f in FORMS])
def test_periodic_cross_and_grid2_fft(form, shape):
    precision, algorithm = form
    a = _field(shape, 1)
    b = _field(shape, 2)
    args = (shape, 5.0, 0, 40.0, 5.0, 0.6, 0.05, 4)
    ref = O.PeriodicCorrelator(*args)
    ref.fast_hist = True
    want = ref.cross_correlate(a, b)
    c = _make(PeriodicCorrelator, args, precision, algorithm)
    c.cross_correlate(a, b)
    _compare(c.results(), want, precision)
    c2 = _make(PeriodicCorrelator, args, precision, algorithm)
    c2.set_grid2_fft(np.fft.rfftn(b))
    c2.cross_correlate(a)
    _compare(c2.results(), want, precision)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_periodic_cross_512_x_several_tiles_per_workgroup(precision):
    """576 column tiles of the cross-mode x pass (64 rows x 9 tiles of 8
    kz columns): more than one workgroup per CU can hold, so the persistent
    kernel prefetches and swaps its buffers between tiles."""
    shape = (512, 64, 128)
    a = _field(shape, 21)
    b = _field(shape, 22)
    args = (shape, 4.0, 1, 60.0, 4.0, 0.5, 0.05, 4)
    ref = O.PeriodicCorrelator(*args)
    ref.fast_hist = True
    want = ref.cross_correlate(a, b)
    c = _make(PeriodicCorrelator, args, precision, "fused")
    c.cross_correlate(a, b)
    _compare(c.results(), want, precision)


YAMAMOTO = [((16, 16, 16), 10.0, (-300.0, 50.0, 20.0), 0, 40.0, 10.0, 0.3, 0.05, 2),
            ((24, 20, 16), 6.0, (100.0, -50.0, 400.0), 2, 50.0, 5.0, 0.4, 0.04, 4),
            ((32, 32, 32), 5.0, (500.0, 500.0, 500.0), 1, 60.0, 5.0, 0.3, 0.03, 4),
            ((512, 8, 8), 4.0, (-1500.0, 40.0, -20.0), 3, 60.0, 4.0, 0.5, 0.05, 4)]


@pytest.mark.parametrize("form", FORMS, ids=["-".join(f) for f in FORMS])
@pytest.mark.parametrize("case", YAMAMOTO)
def test_yamamoto_against_oracle(case, form):
    precision, algorithm = form
    shape, cell, posmin, win, rmax, dr, kmax, dk, maxell = case
    delta = _field(shape, sum(shape), smooth=True)
    c = _make(Correlator, case, precision, algorithm)
    c.autocorrelate(delta)
    ref = O.Correlator(shape, cell, posmin, win, rmax, dr, kmax, dk, maxell)
    ref.fast_hist = True
    _compare(c.results(), ref.autocorrelate(delta), precision)


def test_yamamoto_cross():
    shape = (16, 16, 16)
    args = (shape, 8.0, (10.0, 20.0, 30.0), 0, 40.0, 8.0, 0.5, 0.05, 4)
    a, b = _field(shape, 5), _field(shape, 6)
    ref = O.Correlator(*args)
    ref.fast_hist = True
    want = ref.cross_correlate(a, b)
    c = Correlator(*args)
    c.cross_correlate(a, b)
    _compare(c.results(), want, "fp64")


def test_device_resident_grid_path_and_timing():
    shape = (32, 32, 32)
    delta = _field(shape, 11)
    g = ConfigSpaceGrid(shape, [0, 0, 0], [160] * 3)
    np.copyto(g.data, delta)
    c = PeriodicCorrelator.from_grid_spec(g, 50.0, 5.0, 0.5, 0.05, 4)
    c.autocorrelate(g)
    ref = O.PeriodicCorrelator(shape, g.cell_size, 0, 50.0, 5.0, 0.5, 0.05, 4)
    ref.fast_hist = True
    _compare(c.results(), ref.autocorrelate(delta), "fp64")
    assert c.last_time_ms > 0


def test_constructor_errors():
    with pytest.raises(ValueError):
        PeriodicCorrelator((8, 8, 8), 1.0, 0, 0.0, 1.0, 0.1, 0.01, 0)  # rmax <= min
    with pytest.raises(ValueError):
        PeriodicCorrelator((8, 8, 8), 1.0, 0, 5.0, -1.0, 0.1, 0.01, 0)
    with pytest.raises(ValueError):
        PeriodicCorrelator((8, 8, 8), 1.0, 0, 5.0, 1.0, 0.1, 0.01, 6)
    with pytest.raises(ValueError):
        PeriodicCorrelator((8, 8, 8), 1.0, 9, 5.0, 1.0, 0.1, 0.01, 0)
    c = PeriodicCorrelator((8, 8, 8), 1.0, 0, 5.0, 1.0, 0.5, 0.1, 0)
    with pytest.raises(ValueError):
        c.autocorrelate(np.zeros((8, 8, 9)))
    with pytest.raises(ValueError):
        c.algorithm = "fast"
    with pytest.raises(ValueError):
        PeriodicCorrelator((8, 8, 8), 1.0, 0, 5.0, 1.0, 0.5, 0.1, 0, precision="fp16")
    cm = PeriodicCorrelator((8, 8, 8), 1.0, 0, 5.0, 1.0, 0.5, 0.1, 0, precision="mixed")
    with pytest.raises(ValueError):
        cm.algorithm = "plain"
    with pytest.raises(RuntimeError):
        c.cross_correlate(np.zeros((8, 8, 8)))  # no grid2 yet
