"""Batched periodic autocorrelations (fc_batch_*): every member equals a
PeriodicCorrelator run on the same grid bit for bit (same kernels, same
reduction order), eagerly and through a replayed HIP graph, and a small
batch of TSC mocks matches the FP64 oracle."""
import numpy as np
import pytest

from fftcorr_amd.catalog import thomas_catalog
from fftcorr_amd.correlate import PeriodicBatch, PeriodicCorrelator
from fftcorr_amd.grid import ConfigSpaceGrid
from gpu_helpers import floor_rel_err
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu

ARGS = (5.0, 2, 50.0, 5.0, 0.5, 0.05, 4)


def _fields(shape, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.normal(size=shape)
        F = np.fft.rfftn(a)
        k = [np.fft.fftfreq(m) for m in shape[:2]] + [np.fft.rfftfreq(shape[2])]
        kk = np.sqrt(k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2
                     + (1.3 * k[2][None, None, :]) ** 2)
        out.append(np.fft.irfftn(F * np.exp(-(kk / 0.2) ** 2), s=shape, axes=(0, 1, 2)))
    return out


def _single(shape, field, precision):
    c = PeriodicCorrelator(shape, *ARGS, precision=precision)
    g = ConfigSpaceGrid(shape, [0, 0, 0], [ARGS[0] * n for n in shape], precision=precision)
    np.copyto(g.data, field)
    c.autocorrelate(g)
    return c.results()


def _same(a, b):
    for key in ("r_hist", "k_hist", "r_counts", "k_counts"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert a["zerolag"] == b["zerolag"]


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_members_equal_single_correlator(precision, use_graph):
    shape = (32, 16, 24)
    fields = _fields(shape, 3, 11)
    b = PeriodicBatch(shape, *ARGS, nbatch=3, precision=precision, use_graph=use_graph)
    for i, f in enumerate(fields):
        b.grid(i).upload(f)
    b.autocorrelate()
    for i, f in enumerate(fields):
        _same(b.results(i), _single(shape, f, precision))
    assert b.last_time_ms > 0


def test_graph_replays_match_eager_runs():
    shape = (32, 32, 32)
    sets = [_fields(shape, 4, s) for s in (1, 2, 3)]
    eager = PeriodicBatch(shape, *ARGS, nbatch=4, precision="mixed", use_graph=False)
    graph = PeriodicBatch(shape, *ARGS, nbatch=4, precision="mixed", use_graph=True)
    for fields in sets:  # new data between replays (eager uploads in between)
        for i, f in enumerate(fields):
            eager.grid(i).upload(f)
            graph.grid(i).upload(f)
        eager.autocorrelate()
        graph.autocorrelate()
        for i in range(4):
            _same(graph.results(i), eager.results(i))
    # a smaller member count captures a second graph; members beyond it keep
    # their previous results
    before = graph.results(3)
    graph.autocorrelate(2)
    eager.autocorrelate(2)
    for i in range(2):
        _same(graph.results(i), eager.results(i))
    _same(graph.results(3), before)


def test_tsc_mocks_match_oracle():
    shape = (64, 64, 64)
    box = 250.0
    cell = box / shape[0]
    args = (cell, 2, 60.0, 5.0, 0.3, 0.03, 4)
    b = PeriodicBatch(shape, *args, nbatch=2, precision="fp64")
    wants = []
    for i in range(2):
        pos = thomas_catalog(300, 30, 4.0, 0.0, box, 100 + i)
        b.assign(i, pos, scheme=1)
        x, ok = O.grid_coords(pos, [0, 0, 0], cell, shape, True)
        ref_grid = np.zeros(shape)
        O.assign_mass(ref_grid, x[ok], np.ones(int(ok.sum())), O.TRIANGULAR_SHAPED_CLOUD)
        delta = ref_grid / ref_grid.mean() - 1
        ref = O.PeriodicCorrelator(shape, *args)
        ref.fast_hist = True
        wants.append(ref.autocorrelate(delta))
    b.autocorrelate()
    for i, want in enumerate(wants):
        got = b.results(i)
        np.testing.assert_array_equal(got["k_counts"], want["k_counts"])
        np.testing.assert_array_equal(got["r_counts"], want["r_counts"])
        assert floor_rel_err(got["k_hist"], want["k_hist"]).max() <= 1e-9
        assert floor_rel_err(got["r_hist"], want["r_hist"]).max() <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(24, 16, 16), (48, 20, 24), (384, 8, 8)])
def test_members_equal_single_correlator_non_power_of_two_x(shape, precision):
    """x lengths without the fused kernel (3 x 2^k as in 384^3) run the x
    pass member by member; members still equal single correlations."""
    fields = _fields(shape, 2, 21)
    b = PeriodicBatch(shape, *ARGS, nbatch=2, precision=precision, use_graph=True)
    for i, f in enumerate(fields):
        b.grid(i).upload(f)
    b.autocorrelate()
    b.autocorrelate()  # graph replay
    for i, f in enumerate(fields):
        got = b.results(i)
        want = _single(shape, f, precision)
        for key in ("r_counts", "k_counts"):
            np.testing.assert_array_equal(got[key], want[key])
        # the single correlator may take the fused LDS-staged x kernel where
        # the batch runs separate passes: equal to rounding
        tol = 1e-12 if precision == "fp64" else 1e-5
        assert floor_rel_err(got["k_hist"], want["k_hist"]).max() <= tol
        assert floor_rel_err(got["r_hist"], want["r_hist"]).max() <= tol
    ref = O.PeriodicCorrelator(shape, *ARGS)
    ref.fast_hist = True
    want = ref.autocorrelate(fields[0])
    got = b.results(0)
    np.testing.assert_array_equal(got["k_counts"], want["k_counts"])
    tol = 1e-9 if precision == "fp64" else 3e-5
    assert floor_rel_err(got["k_hist"], want["k_hist"]).max() <= tol
    assert floor_rel_err(got["r_hist"], want["r_hist"]).max() <= tol


def test_batch_errors():
    b = PeriodicBatch((16, 16, 16), *ARGS, nbatch=2)
    with pytest.raises(IndexError):
        b.autocorrelate(3)
    with pytest.raises(IndexError):
        b.results(2)
