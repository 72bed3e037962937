"""The five benchmark workloads at their full sizes, on the GPU.

c1  periodic 128^3, 100,000 uniform points (seed 1) in a 1000 Mpc/h box,
    CIC, xi_0, r bins 0-200 by 5, k bins 0-0.03 by 0.01, FP64: grid and
    multipoles against the FP64 oracle; the reference-faithful variant (NGP,
    as the reference's isotropic test runs) against the reference's own NGP
    code.
c3  periodic 512^3 of a 2000 Mpc/h box, 8M Thomas points (80,000 parents,
    Poisson(100), sigma 3, seed 4), CIC with interlacing, sinc^4 window,
    l <= 4, k 0-0.5 by 0.005, r 0-200 by 5: FP64 against the oracle
    (multithreaded pocketfft) within 1e-9, mixed within 1e-5 of it, counts
    bit-exact; the 512^3 CIC grid itself against the oracle on a slab.
c4  64 mocks (seeds 1000..1063) of 1M Thomas points (10,000 parents,
    sigma 4) in a 1000 Mpc/h box, 256^3 TSC, TSC-aliased window, l <= 4,
    r 0-200 by 5, k 0-0.03 by 0.01, in one graph-captured batch (mixed):
    every member bitwise equal to a single correlator on its grid, two
    members (FP64 batch) against the oracle.
c5  a serving stream over the grid sizes {128, 256, 384, 512}^3 with
    0.1M-10M points (uniform or Thomas, float32 weights in [0.5, 1.5],
    xi or P(k) jobs): each job's result the same bits alone and in the
    4-slot mix, and one FP64 job per size against the oracle.

Error metric: |x - ref| / max(|ref|, 1e-3 max|ref_l0|) per bin on the
histogram values (floor_rel_err).
"""
import os

import numpy as np
import pytest

from fftcorr_amd import _lib
from fftcorr_amd.catalog import thomas_catalog, uniform_catalog
from fftcorr_amd.correlate import PeriodicBatch, PeriodicCorrelator
from fftcorr_amd.serve import CorrelationServer, Job
from gpu_helpers import floor_rel_err
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def threaded_oracle():
    """The oracle's FFTs on every host core (scipy pocketfft)."""
    old = O.FFT_WORKERS
    O.FFT_WORKERS = min(16, os.cpu_count() or 1)
    yield O
    O.FFT_WORKERS = old


def _assign(pos, n, box, scheme, precision, shift=0.0, weights=None):
    g = _lib.DeviceGrid((n, n, n), (0, 0, 0), box / n, precision)
    st = _lib.MaStats()
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    _lib.check(_lib.lib().fc_mass_assign(g.handle, _lib.dptr(pos), pos.shape[0], 3,
                                         _lib.dptr(w) if w is not None else None, 1.0, scheme,
                                         1, shift, st))
    return g, st


def _correlate(corr, *grids):
    L = _lib.lib()
    if len(grids) == 2:
        _lib.check(L.fc_corr_autocorrelate_interlaced(corr._handle, grids[0].handle,
                                                      grids[1].handle))
    else:
        _lib.check(L.fc_corr_autocorrelate(corr._handle, grids[0].handle))
    corr._fetch()
    return {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
            for k, v in corr.results().items()}


def _compare(res, want, tol):
    np.testing.assert_array_equal(res["r_counts"], want["r_counts"])
    np.testing.assert_array_equal(res["k_counts"], want["k_counts"])
    er = floor_rel_err(res["r_hist"], want["r_hist"]).max()
    ek = floor_rel_err(res["k_hist"], want["k_hist"]).max()
    assert er <= tol and ek <= tol, (er, ek)


def _same(a, b):
    for key in ("r_hist", "k_hist", "r_counts", "k_counts"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert a["zerolag"] == b["zerolag"]


# ---- c1 ------------------------------------------------------------------
C1 = dict(n=128, box=1000.0, npart=100_000, seed=1, rmax=200.0, dr=5.0, kmax=0.03, dk=0.01)


@pytest.mark.parametrize("algorithm", ["fused", "plain"])
def test_c1_isotropic_128_cic_fp64(algorithm):
    n, box = C1["n"], C1["box"]
    pos = np.ascontiguousarray(uniform_catalog(C1["npart"], 0.0, box, C1["seed"]))
    g, st = _assign(pos, n, box, 3, "fp64")
    dens = g.download()
    x, ok = O.grid_coords(pos, [0, 0, 0], box / n, (n, n, n), True)
    want_grid = O.assign_mass(np.zeros((n, n, n)), x[ok], np.ones(int(ok.sum())),
                              O.CLOUD_IN_CELL)
    np.testing.assert_allclose(dens, want_grid, rtol=1e-13, atol=1e-13)
    assert st.num_added == C1["npart"] and st.totw == C1["npart"]
    mean = st.totw / n ** 3
    g.normalize(mean)
    delta = (want_grid - mean) / mean
    c = PeriodicCorrelator((n, n, n), box / n, 0, C1["rmax"], C1["dr"], C1["kmax"], C1["dk"], 0)
    c.algorithm = algorithm
    res = _correlate(c, g)
    ref = O.PeriodicCorrelator((n, n, n), box / n, 0, C1["rmax"], C1["dr"], C1["kmax"],
                               C1["dk"], 0)
    ref.fast_hist = True
    _compare(res, ref.autocorrelate(delta), 1e-9)
    assert res["zerolag"] == pytest.approx(float(np.mean(delta ** 2)), rel=1e-9)


def test_c1_reference_variant_ngp(ref):
    """The reference has no CIC; its isotropic test grids with NGP (-d 0)."""
    n, box = C1["n"], C1["box"]
    pos = np.ascontiguousarray(uniform_catalog(C1["npart"], 0.0, box, C1["seed"]))
    g, st = _assign(pos, n, box, 0, "fp64")
    posw = np.column_stack([pos, np.ones(len(pos))])
    want_grid, rst = ref.mass_assign((n, n, n), [0, 0, 0], box / n, 0, True, posw)
    np.testing.assert_array_equal(g.download(), want_grid)  # integer counts: exact
    assert st.num_added == rst["num_added"]


# ---- c3 ------------------------------------------------------------------
C3 = dict(n=512, box=2000.0, parents=80_000, children=100, sigma=3.0, seed=4, rmax=200.0,
          dr=5.0, kmax=0.5, dk=0.005)


@pytest.fixture(scope="module")
def c3_catalogue():
    return np.ascontiguousarray(thomas_catalog(C3["parents"], C3["children"], C3["sigma"], 0.0,
                                               C3["box"], C3["seed"]))


def test_c3_cic_grid_512_against_oracle_slab(c3_catalogue):
    """The full 8M-point CIC assignment at 512^3, checked on the x slab of
    cells [200, 240): it receives contributions from source cells 199..240."""
    n, box = C3["n"], C3["box"]
    pos = c3_catalogue
    for shift in (0.0, 0.5):
        g, st = _assign(pos, n, box, 3, "fp64", shift=shift)
        assert st.num_added == pos.shape[0]
        x, ok = O.grid_coords(pos, [0, 0, 0], box / n, (n, n, n), True)
        x = np.mod(x + shift, n)
        s0, s1 = 199, 241
        sel = (np.floor(x[:, 0]) >= s0) & (np.floor(x[:, 0]) < s1)
        xs = x[sel] - np.array([s0, 0.0, 0.0])
        slab = O.assign_mass(np.zeros((s1 - s0, n, n)), xs, np.ones(int(sel.sum())),
                             O.CLOUD_IN_CELL)
        got = g.download()[s0 + 1:s1 - 1]
        np.testing.assert_allclose(got, slab[1:-1], rtol=1e-13, atol=1e-12)


def test_c3_periodic_512_interlaced(c3_catalogue, threaded_oracle):
    n, box = C3["n"], C3["box"]
    pos = c3_catalogue
    args = ((n, n, n), box / n, 3, C3["rmax"], C3["dr"], C3["kmax"], C3["dk"], 4)
    res = {}
    grids64 = None
    for prec in ("fp64", "mixed"):
        grids = []
        for shift in (0.0, 0.5):
            g, st = _assign(pos, n, box, 3, prec, shift=shift)
            g.normalize(st.totw / float(n) ** 3)
            grids.append(g)
        c = PeriodicCorrelator(*args, precision=prec)
        res[prec] = _correlate(c, *grids)
        if prec == "fp64":
            grids64 = [gg.download() for gg in grids]
        del c, grids
    ref = threaded_oracle.PeriodicCorrelator(*args)
    ref.fast_hist = True
    want = ref.autocorrelate_interlaced(grids64[0], grids64[1])
    del grids64
    _compare(res["fp64"], want, 1e-9)
    _compare(res["mixed"], want, 1e-5)


# ---- c4 ------------------------------------------------------------------
C4 = dict(n=256, box=1000.0, parents=10_000, children=100, sigma=4.0, rmax=200.0, dr=5.0,
          kmax=0.03, dk=0.01, nmocks=64)


def _c4_mock(i):
    return np.ascontiguousarray(thomas_catalog(C4["parents"], C4["children"], C4["sigma"], 0.0,
                                               C4["box"], 1000 + i))


def _c4_args():
    n, box = C4["n"], C4["box"]
    return ((n, n, n), box / n, 2, C4["rmax"], C4["dr"], C4["kmax"], C4["dk"], 4)


def test_c4_64_mocks_256_batched_members_equal_single():
    args = _c4_args()
    b = PeriodicBatch(*args, nbatch=C4["nmocks"], precision="mixed", use_graph=True)
    try:
        for i in range(C4["nmocks"]):
            b.assign(i, _c4_mock(i), scheme=1)
        b.autocorrelate()
        b.autocorrelate()  # graph replay
        single = PeriodicCorrelator(*args, precision="mixed")
        for i in range(C4["nmocks"]):
            _same(b.results(i), _correlate(single, b.grid(i)))
    finally:
        b.close()


def test_c4_two_mocks_against_oracle(threaded_oracle):
    args = _c4_args()
    n, box = C4["n"], C4["box"]
    b = PeriodicBatch(*args, nbatch=2, precision="fp64", use_graph=True)
    try:
        wants = []
        for i in range(2):
            pos = _c4_mock(i)
            b.assign(i, pos, scheme=1)
            x, ok = O.grid_coords(pos, [0, 0, 0], box / n, (n, n, n), True)
            grid = O.assign_mass(np.zeros((n, n, n)), x[ok], np.ones(int(ok.sum())),
                                 O.TRIANGULAR_SHAPED_CLOUD)
            delta = grid / grid.mean() - 1
            ref = threaded_oracle.PeriodicCorrelator(*args)
            ref.fast_hist = True
            wants.append(ref.autocorrelate(delta))
        b.autocorrelate()
        for i, want in enumerate(wants):
            _compare(b.results(i), want, 1e-9)
    finally:
        b.close()


# ---- c5 ------------------------------------------------------------------
def _c5_jobs(seed=5):
    rng = np.random.default_rng(seed)
    jobs = []
    sizes = [128, 256, 384, 512, 512, 384, 256, 128]
    for i, n in enumerate(sizes):
        npart = int(round(np.exp(rng.uniform(np.log(1e5), np.log(1e7)))))
        if i % 2 == 0:
            pos = uniform_catalog(npart, 0.0, 1000.0, 500 + i)
        else:
            pos = thomas_catalog(max(1, npart // 100), 100, 4.0, 0.0, 1000.0, 500 + i)
        w = rng.uniform(0.5, 1.5, size=pos.shape[0]).astype(np.float32)
        kw = (dict(rmax=200.0, dr=5.0, kmax=0.05, dk=0.01) if rng.uniform() < 0.5 else
              dict(rmax=20.0, dr=5.0, kmax=0.5, dk=0.01))
        jobs.append(Job(n=n, box=1000.0, positions=pos, weights=w, tag=i, **kw))
    return jobs


def test_c5_serving_stream_alone_equals_mix():
    jobs = _c5_jobs()
    assert sorted({j.n for j in jobs}) == [128, 256, 384, 512]
    alone = {}
    with CorrelationServer(slots=1, precision="mixed") as srv:
        for j in jobs:
            alone[j.tag] = srv.submit(j).result(timeout=240).results
    with CorrelationServer(slots=4, precision="mixed") as srv:
        futs = [srv.submit(j) for j in jobs]
        mixed = [f.result(timeout=240) for f in futs]
    for r in mixed:
        _same(r.results, alone[r.tag])


@pytest.mark.parametrize("n", [128, 256, 384, 512])
def test_c5_fp64_job_per_size_against_oracle(n, threaded_oracle):
    job = [j for j in _c5_jobs() if j.n == n][0]
    with CorrelationServer(slots=1, precision="fp64") as srv:
        got = srv.submit(job).result(timeout=240).results
    # The same grid through the device path, downloaded for the oracle.
    pos = np.ascontiguousarray(job.positions, np.float64)
    g, st = _assign(pos, n, job.box, job.scheme, "fp64", weights=job.weights)
    mean = st.totw / float(n) ** 3
    delta = (g.download() - mean) / mean
    del g
    ref = threaded_oracle.PeriodicCorrelator((n, n, n), job.box / n, job.window, job.rmax,
                                             job.dr, job.kmax, job.dk, job.maxell)
    ref.fast_hist = True
    _compare(got, ref.autocorrelate(delta), 1e-9)
