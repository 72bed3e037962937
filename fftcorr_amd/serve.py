"""Serving loop: correlation jobs of mixed grid sizes on one GPU.

A ``CorrelationServer`` owns ``slots`` workers, each with its own HIP stream
(an fc_ctx).  Jobs wait in arrival order and are admitted as soon as a slot
is free and their HBM footprint fits the budget next to the jobs already
running; admitted jobs run concurrently on their slots' streams, so their
kernels overlap on the device.  Grids and correlators (FFT plans, bin-sorted
subgrids, weights) stay resident per slot and geometry and are reused by
later jobs of the same geometry; the least recently used ones are dropped
when a new geometry needs the memory.

Each job is one periodic autocorrelation (cc/correlate/correlator.h:255-357)
of a catalogue assigned onto its grid: mass assignment, delta = rho/mean - 1,
then xi_l(r) and P_l(k), l <= maxell.  Results do not depend on what else
runs: every kernel reduces in a fixed order, so a job gives the same bits
alone and in any mix.
"""

import collections
import concurrent.futures as cf
import ctypes
import dataclasses
import threading
import time

import numpy as np

from fftcorr_amd import _lib
from fftcorr_amd.correlate import PeriodicCorrelator


@dataclasses.dataclass
class Job:
    """One periodic correlation job.  ``positions`` are (n, 3) float64 survey
    coordinates in [0, box); ``weights`` (n,) (float32 or float64) or None."""
    n: int
    box: float
    positions: np.ndarray
    weights: np.ndarray = None
    scheme: int = 3
    window: int = 3
    rmax: float = 200.0
    dr: float = 5.0
    kmax: float = 0.5
    dk: float = 0.01
    maxell: int = 4
    tag: object = None


@dataclasses.dataclass
class Result:
    tag: object
    results: dict
    submitted: float
    started: float
    finished: float

    @property
    def latency_ms(self):
        return 1e3 * (self.finished - self.submitted)


def job_bytes(job, precision):
    """HBM footprint of a job: resident grid and correlator state plus the
    transient particle buffers of its mass assignment."""
    n = job.n
    s = 8 if precision == "fp64" else 4
    ldc = (n // 2 + 1 + 15) // 16 * 16
    work = n * n * ldc * 2 * s
    grid = n ** 3 * s
    rs = 2 * int(np.ceil(job.rmax / (job.box / n))) + 1
    inv = rs * n * ldc * 16 * 2 + rs ** 3 * 8 * 4
    ks = min(2 * int(np.ceil(job.kmax / (2 * np.pi / job.box))) + 1, n + 1)
    kside = ks ** 3 * 4 * 2
    particles = job.positions.shape[0] * 64
    return grid + work + inv + kside + particles


class _Slot:
    def __init__(self, device, precision):
        self.ctx = _lib.Context(device)
        self.precision = precision
        self.state = collections.OrderedDict()  # geometry -> (correlator, grid, bytes)

    def resident_bytes(self):
        return sum(v[2] for v in self.state.values())

    def get(self, job, budget_check, lock):
        """(correlator, grid) for the job's geometry on this slot; builds them
        (outside the lock) after evicting least recently used geometries
        until the new state fits the budget."""
        key = (job.n, job.box, job.window, job.rmax, job.dr, job.kmax, job.dk, job.maxell)
        nbytes = job_bytes(job, self.precision) - job.positions.shape[0] * 64
        with lock:
            if key in self.state:
                self.state.move_to_end(key)
                return self.state[key][:2]
            while self.state and not budget_check(nbytes):
                self.state.popitem(last=False)
        n = job.n
        cell = job.box / n
        corr = PeriodicCorrelator((n, n, n), cell, job.window, job.rmax, job.dr, job.kmax,
                                  job.dk, job.maxell, precision=self.precision,
                                  context=self.ctx)
        grid = _lib.DeviceGrid((n, n, n), (0, 0, 0), cell, self.precision, self.ctx)
        with lock:
            self.state[key] = (corr, grid, nbytes)
        return corr, grid


class CorrelationServer:
    """Continuous-batching correlation server on one GPU.

    slots: jobs that may run at once (one HIP stream each).
    hbm_budget: bytes the running jobs and resident state may use (default:
    half of the device's free memory at start-up).
    """

    def __init__(self, device=None, slots=4, precision="mixed", hbm_budget=None):
        self.device = _lib.default_device() if device is None else int(device)
        self.precision = precision
        if hbm_budget is None:
            free = ctypes.c_size_t()
            total = ctypes.c_size_t()
            _lib.check(_lib.lib().fc_device_memory(self.device, ctypes.byref(free),
                                                   ctypes.byref(total)))
            hbm_budget = free.value // 2
        self.hbm_budget = int(hbm_budget)
        self._slots = [_Slot(self.device, precision) for _ in range(int(slots))]
        self._free = list(range(len(self._slots)))
        self._queue = collections.deque()
        self._inflight_bytes = 0
        self._cv = threading.Condition()
        self._closed = False
        self._pool = cf.ThreadPoolExecutor(max_workers=len(self._slots))
        self._admitter = threading.Thread(target=self._admit_loop, daemon=True)
        self._admitter.start()

    # ---- public -------------------------------------------------------------
    def submit(self, job):
        """Queues a job; returns a Future of Result."""
        fut = cf.Future()
        need = job_bytes(job, self.precision)
        if need > self.hbm_budget:
            fut.set_exception(MemoryError("job needs %d bytes, budget is %d"
                                          % (need, self.hbm_budget)))
            return fut
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            self._queue.append((job, fut, time.perf_counter(), need))
            self._cv.notify_all()
        return fut

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        self._admitter.join()
        self._pool.shutdown(wait=True)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- scheduling ---------------------------------------------------------
    def _resident(self):
        return sum(s.resident_bytes() for s in self._slots)

    def _admit_loop(self):
        while True:
            with self._cv:
                while True:
                    if self._closed and not self._queue:
                        return
                    if self._queue and self._free:
                        job, fut, t_sub, need = self._queue[0]
                        if self._inflight_bytes + need + self._resident() <= self.hbm_budget \
                                or self._inflight_bytes == 0:
                            break
                    self._cv.wait()
                self._queue.popleft()
                slot_id = self._free.pop()
                self._inflight_bytes += need
            self._pool.submit(self._run, slot_id, job, fut, t_sub, need)

    def _budget_ok(self, extra):
        return self._inflight_bytes + self._resident() + extra <= self.hbm_budget

    def _run(self, slot_id, job, fut, t_sub, need):
        slot = self._slots[slot_id]
        t0 = time.perf_counter()
        try:
            corr, grid = slot.get(job, self._budget_ok, self._cv)
            pos = np.ascontiguousarray(job.positions, np.float64)
            w = None
            if job.weights is not None:
                w = np.ascontiguousarray(job.weights, np.float64)
            grid.clear()
            st = _lib.MaStats()
            _lib.check(_lib.lib().fc_mass_assign(
                grid.handle, _lib.dptr(pos), pos.shape[0], 3,
                _lib.dptr(w) if w is not None else None, 1.0, int(job.scheme), 1, 0.0,
                ctypes.byref(st)))
            grid.normalize(st.totw / float(job.n) ** 3)
            _lib.check(_lib.lib().fc_corr_autocorrelate(corr._handle, grid.handle))
            corr._fetch()
            res = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                   for k, v in corr.results().items()}
            fut.set_result(Result(job.tag, res, t_sub, t0, time.perf_counter()))
        except BaseException as e:  # noqa: BLE001 -- delivered to the caller's future
            fut.set_exception(e)
        finally:
            with self._cv:
                self._inflight_bytes -= need
                self._free.append(slot_id)
                self._cv.notify_all()


def poisson_stream(rate_hz, count, seed):
    """Seeded exponential inter-arrival times (s) of a Poisson stream."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_hz, size=count))
