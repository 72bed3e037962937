"""Serving loop (fftcorr_amd.serve): jobs of mixed grid sizes give the same
bits alone and when they overlap on the device, admission respects the HBM
budget, and a job equals the direct ConfigSpaceGrid / MassAssignor /
PeriodicCorrelator path."""
import numpy as np
import pytest

from fftcorr_amd.catalog import thomas_catalog, uniform_catalog
from fftcorr_amd.correlate import PeriodicCorrelator
from fftcorr_amd.grid import ConfigSpaceGrid
from fftcorr_amd.particle_mesh import MassAssignor
from fftcorr_amd.serve import CorrelationServer, Job, job_bytes, poisson_stream

pytestmark = pytest.mark.gpu

BOX = 400.0


def _jobs(count, seed):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        n = int(rng.choice([16, 32, 48, 64]))
        npart = int(np.exp(rng.uniform(np.log(2e3), np.log(5e4))))
        if rng.uniform() < 0.5:
            pos = uniform_catalog(npart, 0.0, BOX, seed + 10 * i)
        else:
            pos = thomas_catalog(max(1, npart // 50), 50, 4.0, 0.0, BOX, seed + 10 * i)
        w = rng.uniform(0.5, 1.5, size=pos.shape[0]).astype(np.float32)
        if rng.uniform() < 0.5:   # xi job
            kw = dict(rmax=60.0, dr=5.0, kmax=0.05, dk=0.01)
        else:                     # P(k) job
            kw = dict(rmax=20.0, dr=5.0, kmax=0.5, dk=0.02)
        jobs.append(Job(n=n, box=BOX, positions=pos, weights=w, tag=i, **kw))
    return jobs


def _same(a, b):
    for k in ("r_hist", "k_hist", "r_counts", "k_counts"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert a["zerolag"] == b["zerolag"]


def test_alone_and_mixed_give_identical_results():
    jobs = _jobs(10, 5)
    alone = {}
    with CorrelationServer(slots=1) as srv:
        for j in jobs:
            alone[j.tag] = srv.submit(j).result(timeout=120).results
    with CorrelationServer(slots=4) as srv:
        futs = [srv.submit(j) for j in jobs]
        mixed = [f.result(timeout=120) for f in futs]
    for r in mixed:
        _same(r.results, alone[r.tag])
        assert r.latency_ms > 0 and r.finished >= r.started >= r.submitted


def test_job_equals_direct_path():
    job = _jobs(1, 9)[0]
    with CorrelationServer(slots=2, precision="mixed") as srv:
        got = srv.submit(job).result(timeout=120).results
    n = job.n
    g = ConfigSpaceGrid((n, n, n), [0, 0, 0], [BOX] * 3, precision="mixed")
    ma = MassAssignor(g, job.scheme, periodic_wrap=True)
    ma.add_particles(np.ascontiguousarray(job.positions),
                     np.ascontiguousarray(job.weights, np.float64))
    g.normalize(ma.totw / float(n) ** 3)
    c = PeriodicCorrelator((n, n, n), BOX / n, job.window, job.rmax, job.dr, job.kmax, job.dk,
                           job.maxell, precision="mixed")
    c.autocorrelate(g)
    _same(got, c.results())


def test_admission_keeps_measured_device_bytes_within_budget():
    """The library's own count of the device bytes it holds (every buffer of
    every handle and context) never rises above the budget while a stream
    of jobs of three geometries runs on four slots."""
    from fftcorr_amd import _lib
    jobs = [Job(n=n, box=BOX, positions=uniform_catalog(20000 + 5000 * i, 0.0, BOX, 70 + i),
                tag=i, rmax=40.0, dr=5.0, kmax=0.3, dk=0.02)
            for i, n in enumerate([48, 32, 64, 48, 64, 32, 48, 64])]
    need = max(job_bytes(j, "mixed") for j in jobs)
    budget = int(1.5 * need)
    base, _ = _lib.device_bytes()
    _lib.reset_device_peak()
    with CorrelationServer(slots=4, hbm_budget=budget) as srv:
        res = [f.result(timeout=120) for f in [srv.submit(j) for j in jobs]]
        _, peak = _lib.device_bytes()
    assert peak - base <= budget, (peak - base, budget)
    assert sorted(r.tag for r in res) == list(range(len(jobs)))
    with CorrelationServer(slots=1, hbm_budget=1000) as srv:
        with pytest.raises(MemoryError):
            srv.submit(jobs[0]).result(timeout=10)


@pytest.mark.parametrize("n", [16, 48, 64])
def test_footprint_estimates_bound_measured_bytes(n):
    """The admission estimates (state_bytes, assign_bytes) bound what a job's
    geometry and mass assignment then hold, as the library measures them."""
    from fftcorr_amd.serve import assign_bytes, state_bytes
    job = Job(n=n, box=BOX, positions=thomas_catalog(400, 50, 4.0, 0.0, BOX, n),
              rmax=60.0, dr=5.0, kmax=0.4, dk=0.02)
    for prec in ("mixed", "fp64"):
        with CorrelationServer(slots=1, precision=prec) as srv:
            srv.submit(job).result(timeout=120)
            slot = srv._slots[0]
            held = slot.state[slot.key(job)][2]
            assert 0 < held <= state_bytes(job, prec), (held, state_bytes(job, prec))
            assert 0 < slot.ctx_bytes <= assign_bytes(job)


def test_poisson_stream_is_seeded():
    a = poisson_stream(20.0, 50, 5)
    b = poisson_stream(20.0, 50, 5)
    np.testing.assert_array_equal(a, b)
    assert np.all(np.diff(a) > 0)
