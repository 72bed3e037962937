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


def _ldc(n, s):
    """Complex row stride of a work grid: n/2+1 rounded up to 128 bytes."""
    per = 128 // (2 * s)
    return (n // 2 + 1 + per - 1) // per * per


def _geometry(job):
    """(r subgrid, k subgrid) edge lengths (cc/correlate/correlator.h:109-164;
    the k spacing 2 k_Nyq / n is 2 pi / box)."""
    cell = job.box / job.n
    rs = 2 * int(np.ceil(job.rmax / cell)) + 1
    ks = min(2 * int(np.ceil(job.kmax / (2 * np.pi / job.box))) + 1, job.n + 1)
    return rs, ks


def state_bytes(job, precision):
    """Upper bound of the device bytes one geometry keeps resident on a slot:
    the density grid and every buffer of its periodic correlator once it has
    run (work grid, compact k values, pruned inverse arrays, bin-sorted
    subgrid cells and chunk tables, the half-subgrid k entries, weights,
    axes, outputs)."""
    n = job.n
    s = 8 if precision == "fp64" else 4         # grid, forward spectra, k values
    si = 4 if precision == "fp32" else 8        # pruned inverse
    rs, ks = _geometry(job)
    ldc, ldci = _ldc(n, s), _ldc(n, si)
    K, R = ks ** 3, rs ** 3
    nv = job.maxell // 2 + 1
    nbins = int(np.ceil(max(job.kmax, job.rmax) / min(job.dk, job.dr))) + 2
    grid = n ** 3 * s
    work = n * n * ldc * 2 * s
    pk = ks * ks * (ks // 2 + 1) * s
    inv = rs * n * ldci * 2 * si + rs * rs * ldci * 2 * si + R * si
    cells = 4 * (K + R) + 20 * ((K + R) // 4096 + 2 * nbins + 2) + 8 * (nbins + 1)
    weights = 8 * nv * R + 8 * nv * ((K + R) // 4096 + 2 * nbins + 2)
    small = 8 * 16 * (3 * n + ldc) + 64 * nv * nbins + (1 << 16)
    H = ks * ks * (ks // 2 + 1)  # compact k values: the half-subgrid k entries
    half = 4 * H + 20 * (H // 4096 + 2 * nbins + 2) + 8 * (nbins + 1)
    return grid + work + pk + inv + cells + weights + small + half


def assign_bytes(job):
    """Upper bound of the device bytes one mass assignment of the job's
    catalogue needs on a fresh slot: the context's reused work buffers
    (keys, indices, records, run tables, occupancy bits, sort space) plus
    the staging of the host catalogue."""
    npart = job.positions.shape[0]
    ncells = job.n ** 3
    has_w = job.weights is not None
    scratch = npart * (4 * 4 + 32 * 2 + 16) + 8 * ncells + ncells // 8 + (4 << 20)
    staging = 8 * npart * (3 + (1 if has_w else 0))
    return scratch + staging


def job_bytes(job, precision):
    """Upper bound of a job's device footprint on a slot that holds nothing
    yet (the admission estimate; admission then uses measured bytes)."""
    return state_bytes(job, precision) + assign_bytes(job)


class _Slot:
    """One worker: a context (HIP stream) and the geometries resident on it.
    Its state is changed only with the server's lock held, by its own job
    (while the slot is busy) or by admission (while it is idle)."""

    def __init__(self, device, precision):
        self.ctx = _lib.Context(device)
        self.precision = precision
        self.state = collections.OrderedDict()  # geometry -> (correlator, grid, bytes)
        self.ctx_bytes = 0  # the context's work buffers, measured when the slot goes idle

    @staticmethod
    def key(job):
        return (job.n, job.box, job.window, job.rmax, job.dr, job.kmax, job.dk, job.maxell)

    def resident_bytes(self):
        """Measured device bytes of this slot: its resident geometries
        (grid + correlator buffers) and its context's work buffers."""
        return self.ctx_bytes + sum(v[2] for v in self.state.values())

    def measure_ctx(self):
        b = ctypes.c_size_t()
        _lib.check(_lib.lib().fc_ctx_footprint(self.ctx.handle, ctypes.byref(b)))
        self.ctx_bytes = b.value

    def need(self, job):
        """Device bytes the job may add to this slot: a new geometry's state
        (upper bound) unless resident, and the mass assignment's growth of
        the context's work buffers plus its host staging (exact when the
        geometry's grid exists, else an upper bound)."""
        k = self.key(job)
        if k not in self.state:
            return state_bytes(job, self.precision) + assign_bytes(job)
        grid = self.state[k][1]
        grow, trans = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(_lib.lib().fc_mass_assign_footprint(
            grid.handle, job.positions.shape[0], 3, int(job.weights is not None), 1,
            ctypes.byref(grow), ctypes.byref(trans)))
        return grow.value + trans.value

    def get(self, job, lock):
        """(correlator, grid) for the job's geometry on this slot, built on
        first use (outside the lock; the slot is busy, so nothing else
        touches its state meanwhile)."""
        k = self.key(job)
        with lock:
            if k in self.state:
                self.state.move_to_end(k)
                return self.state[k][:2]
        n = job.n
        cell = job.box / n
        corr = PeriodicCorrelator((n, n, n), cell, job.window, job.rmax, job.dr, job.kmax,
                                  job.dk, job.maxell, precision=self.precision,
                                  context=self.ctx)
        grid = _lib.DeviceGrid((n, n, n), (0, 0, 0), cell, self.precision, self.ctx)
        with lock:
            self.state[k] = (corr, grid, 0)  # counted by the job's reservation until measured
        return corr, grid

    def measure(self, job):
        """Replaces the geometry's recorded bytes by its measured footprint."""
        k = self.key(job)
        corr, grid, _ = self.state[k]
        held, full = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(_lib.lib().fc_corr_footprint(corr._handle, 0, ctypes.byref(held),
                                                ctypes.byref(full)))
        gbytes = job.n ** 3 * (8 if self.precision == "fp64" else 4)
        self.state[k] = (corr, grid, full.value + gbytes)

    def evict_lru(self, keep=None):
        """Drops the least recently used geometry (not `keep`), or, with no
        geometry left to drop and keep None, the context's work buffers;
        False if there was nothing to free."""
        for k in list(self.state):
            if k != keep:
                del self.state[k]
                return True
        if keep is None and self.ctx_bytes:
            _lib.check(_lib.lib().fc_ctx_release_scratch(self.ctx.handle))
            self.ctx_bytes = 0
            return True
        return False

    def can_free(self):
        return bool(self.state) or self.ctx_bytes > 0


class CorrelationServer:
    """Continuous-batching correlation server on one GPU.

    slots: jobs that may run at once (one HIP stream each).
    hbm_budget: device bytes the server's resident state and running jobs
    may hold (default: half of the device's free memory at start-up).
    Admission counts the measured bytes each slot holds (fc_ctx_footprint,
    fc_corr_footprint, grids) plus, for every running job, an upper bound of
    what it may still allocate; idle slots' least recently used geometries
    are dropped when a job would not fit otherwise.
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
            fut.set_exception(MemoryError("job needs up to %d bytes, budget is %d"
                                          % (need, self.hbm_budget)))
            return fut
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            self._queue.append((job, fut, time.perf_counter()))
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

    def _try_admit(self, job):
        """(slot id, reserved bytes) if the job fits now (after dropping idle
        slots' least recently used geometries if needed), else None.  Called
        with the lock held."""
        free = sorted(self._free, key=lambda i: _Slot.key(job) not in self._slots[i].state)
        for sid in free:
            slot = self._slots[sid]
            need = slot.need(job)
            while self._inflight_bytes + self._resident() + need > self.hbm_budget:
                idle = [self._slots[i] for i in self._free]
                victim = next((s for s in idle if s is not slot and s.can_free()), None)
                if victim is not None:
                    victim.evict_lru()
                elif not slot.evict_lru(keep=_Slot.key(job)):
                    break
                need = slot.need(job)
            if self._inflight_bytes + self._resident() + need <= self.hbm_budget:
                return sid, need
        return None

    def _admit_loop(self):
        while True:
            with self._cv:
                while True:
                    if self._closed and not self._queue:
                        return
                    got = None
                    if self._queue and self._free:
                        got = self._try_admit(self._queue[0][0])
                        if got is None and len(self._free) == len(self._slots):
                            # Nothing runs and nothing more can be freed: the
                            # job cannot fit (submit() screens such jobs).
                            job, fut, _ = self._queue.popleft()
                            fut.set_exception(MemoryError(
                                "job does not fit the HBM budget of %d bytes"
                                % self.hbm_budget))
                            continue
                    if got is not None:
                        break
                    self._cv.wait()
                job, fut, t_sub = self._queue.popleft()
                slot_id, need = got
                self._free.remove(slot_id)
                self._inflight_bytes += need
            self._pool.submit(self._run, slot_id, job, fut, t_sub, need)

    def _run(self, slot_id, job, fut, t_sub, need):
        slot = self._slots[slot_id]
        t0 = time.perf_counter()
        outcome = None
        try:
            corr, grid = slot.get(job, self._cv)
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
            outcome = Result(job.tag, res, t_sub, t0, time.perf_counter())
        except BaseException as e:  # noqa: BLE001 -- delivered to the caller's future
            outcome = e
        finally:
            # The slot's measured bytes are recorded before the caller sees
            # the result.
            with self._cv:
                if _Slot.key(job) in slot.state:
                    slot.measure(job)
                slot.measure_ctx()
                self._inflight_bytes -= need
                self._free.append(slot_id)
                self._cv.notify_all()
        if isinstance(outcome, Result):
            fut.set_result(outcome)
        else:
            fut.set_exception(outcome)


def poisson_stream(rate_hz, count, seed):
    """Seeded exponential inter-arrival times (s) of a Poisson stream."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_hz, size=count))
