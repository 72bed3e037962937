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


def test_budget_serialises_admission_and_rejects_oversized_jobs():
    jobs = [Job(n=48, box=BOX, positions=uniform_catalog(20000, 0.0, BOX, 70 + i), tag=i,
                rmax=40.0, dr=5.0, kmax=0.3, dk=0.02) for i in range(6)]
    need = max(job_bytes(j, "mixed") for j in jobs)
    with CorrelationServer(slots=4, hbm_budget=int(1.5 * need)) as srv:
        res = [f.result(timeout=120) for f in [srv.submit(j) for j in jobs]]
    spans = sorted((r.started, r.finished) for r in res)
    # Two of these jobs do not fit the budget together: one runs at a time
    # (a job starts only after the previous one released its bytes).
    for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
        assert s1 >= f0 - 1e-3
    with CorrelationServer(slots=1, hbm_budget=1000) as srv:
        with pytest.raises(MemoryError):
            srv.submit(jobs[0]).result(timeout=10)


def test_poisson_stream_is_seeded():
    a = poisson_stream(20.0, 50, 5)
    b = poisson_stream(20.0, 50, 5)
    np.testing.assert_array_equal(a, b)
    assert np.all(np.diff(a) > 0)
