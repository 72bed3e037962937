"""Benchmark: periodic P_l(k) / xi_l(r) multipoles, l <= 4, on a 512^3 grid.

Workload (one "correlation"): the interlaced autocorrelation of a
device-resident density grid pair -- 8M Thomas-process points (80,000
parents, Poisson(100) children, sigma 3 Mpc/h) in a periodic 2000 Mpc/h box,
CIC assignment with and without a half-cell shift, delta = rho/mean - 1,
sinc^4 window correction, k bins 0-0.5 h/Mpc width 0.005 (k subgrid 321^3),
r bins 0-200 Mpc/h width 5 (r subgrid 105^3).  Each step runs both forward
FFTs, the interlacing combine, |F|^2, the k-space multipole histograms, the
inverse FFT and the r-space multipole histograms, and leaves FP64 histograms
on the device.

Forms timed on the same grids:
  * mixed, fused (the headline `value`): FP32 grids and spectra in HBM;
    the z and y forward FFT passes compute in FP64 (rounding to FP32 at
    each LDS exchange and store), the x pass and products in FP32; FP64
    pruned inverse FFT, Y_lm, windows, bins and sums;
  * fp32, fused: the inverse in FP32 as well;
  * fp64, fused: every stage in FP64;
  * fp64, plain: the reference's stage sequence (full forward FFT,
    conjugate product, per-multipole k extraction, full inverse FFT,
    per-multipole r extraction) in FP64 on the GPU -- the denominator of
    the mixed/plain ratio.
Reported beside them: the error of each reduced-precision form against the
FP64 fused result per multipole (floor 1e-3 max|l=0|), bit-exact counts,
p50/p99 latency, the dominant kernel's HBM roofline, and the FP64 CPU oracle
(scipy pocketfft, all host threads) timed on one correlation.

Timing: `value` is the throughput of the timed steps queued back to back on
the correlator's stream (synchronised before and after); p50/p99 come from
a second pass that times each correlation on its own (HIP events around
it, the host waiting for each), and the per-stage times from a third pass
with stage events.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
"""

import argparse
import contextlib
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

CFG = dict(n=512, box=2000.0, parents=80_000, children=100, sigma=3.0, seed=4,
           rmax=200.0, dr=5.0, kmax=0.5, dk=0.005, maxell=4, window=3, scheme=3)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_grids(lib, ctx, pos, cfg, precision):
    n = cfg["n"]
    cell = cfg["box"] / n
    grids = []
    ma_ms = 0.0
    for shift in (0.0, 0.5):
        g = lib.DeviceGrid((n, n, n), (0, 0, 0), cell, precision, ctx)
        st = lib.MaStats()
        t0 = time.perf_counter()
        lib.check(lib.lib().fc_mass_assign(g.handle, lib.dptr(pos), pos.shape[0], 3, None, 1.0,
                                           cfg["scheme"], 1, shift, st))
        ma_ms += (time.perf_counter() - t0) * 1e3
        g.normalize(st.totw / float(n) ** 3)
        grids.append(g)
    return grids, ma_ms, st


def stage_bytes(cfg, s_work, s_grid, kshape, rs, s_inv=None, k_entries=None, fwd_grids=1):
    """Algorithmic HBM bytes per launch of each stage: the data each stage
    must read and write once.  The forward z and y passes transform
    fwd_grids grids per launch (2 when an interlaced pair runs as one launch
    per pass).  Complex half grid n^2 (n/2+1) elements of
    2 s_work bytes; real grid n^3 of s_grid; compact k-subgrid values
    ks0 ks1 (ks2/2+1) of s_work; pruned inverse arrays rs n (n/2+1) and
    rs^2 (n/2+1) complex and the r subgrid rs^3 real, of s_inv bytes.  The k
    histogram reads, per compact k value inside kmax (the half subgrid
    kz >= 0, whose mirror cells share the values), its 4-byte index and the
    s_work value; the r histogram reads, per subgrid cell, its 4-byte cell
    index, its s_inv value and three FP64 multipole weights."""
    s_inv = s_inv or s_work
    n = cfg["n"]
    nh = n // 2 + 1
    half = n * n * nh * 2 * s_work
    pk = kshape[0] * kshape[1] * (kshape[2] // 2 + 1) * s_work
    q1 = rs * n * nh * 2 * s_inv
    q2 = rs * rs * nh * 2 * s_inv
    xs = rs ** 3 * s_inv
    return {
        "fft_z_fwd": fwd_grids * (n ** 3 * s_grid + half),
        "fft_y_fwd": fwd_grids * 2 * half,
        "combine": 2 * half + pk + q1,  # fused x pass: read a, b; write pk, pruned q1
        "hist_k": (k_entries if k_entries is not None else pk // s_work) * (4 + s_work),
        "fft_y_inv": q1 + q2,
        "fft_z_inv": q2 + xs,
        "hist_r": rs ** 3 * (4 + s_inv + 3 * 8),
    }


def k_half_entries(cfg, kshape):
    """Entries of the k histogram: compact k values (half subgrid kz >= 0)
    whose |k| falls in a bin, with the correlator's FP64 axes and bin rule."""
    n = cfg["n"]
    knyq = np.pi / (cfg["box"] / n)
    step = 2.0 * knyq / n
    ax = [(-(ks // 2)) * 2.0 * knyq / n + np.arange(ks) * step for ks in kshape]
    kz = ax[2][kshape[2] // 2:]
    nb = np.floor(cfg["kmax"] / cfg["dk"])
    xy = ax[0][:, None] * ax[0][:, None] + ax[1][None, :] * ax[1][None, :]
    count = 0
    for z in kz:
        f = np.floor(np.sqrt(xy + z * z) / cfg["dk"])
        count += int(np.count_nonzero((f >= 0) & (f < nb)))
    return count


def run_form(lib, corr_mod, ctx, pos, cfg, precision, steps, warmup, profile,
             algorithm="fused"):
    n = cfg["n"]
    cell = cfg["box"] / n
    grids, ma_ms, st = make_grids(lib, ctx, pos, cfg, precision)
    c = corr_mod.PeriodicCorrelator((n, n, n), cell, cfg["window"], cfg["rmax"], cfg["dr"],
                                    cfg["kmax"], cfg["dk"], cfg["maxell"], precision=precision)
    c.algorithm = algorithm
    L = lib.lib()
    h = c._handle
    ga, gb = grids[0].handle, grids[1].handle
    for _ in range(warmup):
        lib.check(L.fc_corr_autocorrelate_interlaced(h, ga, gb))
    # Throughput: the timed steps queued back to back on the correlator's
    # stream, synchronised before and after.
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.check(L.fc_corr_autocorrelate_interlaced(h, ga, gb))
    ctx.synchronize()
    wall = time.perf_counter() - t0
    # Latency: each correlation on its own (device time between the events
    # that bracket it), the host waiting for each before the next.
    lat = []
    t = ctypes.c_double()
    for _ in range(steps):
        lib.check(L.fc_corr_autocorrelate_interlaced(h, ga, gb))
        lib.check(L.fc_corr_last_time_ms(h, ctypes.byref(t)))
        lat.append(t.value)
    # Per-stage device times (HIP events between the stages).
    stages = None
    if profile:
        c.set_profiling(True)
        for _ in range(steps):
            lib.check(L.fc_corr_autocorrelate_interlaced(h, ga, gb))
        ctx.synchronize()
        stages = c.stage_times()
        c.set_profiling(False)
    c._fetch()
    res = {k: np.array(v, copy=True) if isinstance(v, np.ndarray) else v
           for k, v in c.results().items()}
    out = dict(wall_s=wall, lat_ms=np.array(lat), stages=stages, res=res, ma_ms=ma_ms,
               totw=st.totw, kshape=c.kshape, rshape=c.rshape, steps=steps)
    return out, grids, c


def host_grid_rate(lib, corr_mod, grids, cfg, steps):
    """Correlations/s when the caller hands over host (numpy) grids, as the
    reference's Python API does: each call uploads both grids over PCIe."""
    from fftcorr_amd.grid import ConfigSpaceGrid
    n = cfg["n"]
    box = cfg["box"]
    host = []
    for g in grids:
        cg = ConfigSpaceGrid((n, n, n), (0, 0, 0), (box, box, box), precision="mixed")
        g.download(cg.data)
        host.append(cg)
    c = corr_mod.PeriodicCorrelator((n, n, n), box / n, cfg["window"], cfg["rmax"], cfg["dr"],
                                    cfg["kmax"], cfg["dk"], cfg["maxell"], precision="mixed")
    c.autocorrelate_interlaced(host[0], host[1])
    t0 = time.perf_counter()
    for _ in range(steps):
        for h in host:
            h.data[0, 0, 0] += 0.0  # marks the host copy as written: forces the upload
        c.autocorrelate_interlaced(host[0], host[1])
    return steps / (time.perf_counter() - t0)


BATCH_CFG = dict(n=256, box=1000.0, parents=10_000, children=100, sigma=4.0, mocks=64,
                 rmax=200.0, dr=5.0, kmax=0.03, dk=0.01, maxell=4, window=2, scheme=1)


def batched_mocks(lib, corr_mod, ctx, reps):
    """64 seeded Thomas mocks (10,000 parents x Poisson(100), sigma 4 Mpc/h,
    1000 Mpc/h periodic box), TSC onto 256^3, TSC-aliased window, xi_l and
    P_l (kmax 0.03), l <= 4: mixed-precision batch (one HIP graph per batch)
    against the FP64 plain form run mock by mock on the same grids."""
    from fftcorr_amd.catalog import thomas_catalog
    b = BATCH_CFG
    n = b["n"]
    cell = b["box"] / n
    args = ((n, n, n), cell, b["window"], b["rmax"], b["dr"], b["kmax"], b["dk"], b["maxell"])
    batch = corr_mod.PeriodicBatch(*args, nbatch=b["mocks"], precision="mixed", use_graph=True)
    t_ma = 0.0
    for i in range(b["mocks"]):
        pos = thomas_catalog(b["parents"], b["children"], b["sigma"], 0.0, b["box"], 1000 + i)
        t0 = time.perf_counter()
        batch.assign(i, pos, scheme=b["scheme"])
        t_ma += time.perf_counter() - t0
    ctx.synchronize()
    batch.autocorrelate()  # captures the graph
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.autocorrelate()
    ctx.synchronize()
    rate = reps * b["mocks"] / (time.perf_counter() - t0)
    # FP64 plain, mock by mock, on the same grids (downloaded once).
    plain = corr_mod.PeriodicCorrelator(*args, precision="fp64")
    plain.algorithm = "plain"
    g = lib.DeviceGrid((n, n, n), (0, 0, 0), cell, "fp64", ctx)
    nplain = 8
    hosts = [batch.grid(i).download() for i in range(nplain)]
    err = 0.0
    t_plain = 0.0
    for i in range(nplain):
        g.upload(hosts[i])
        ctx.synchronize()
        t0 = time.perf_counter()
        lib.check(lib.lib().fc_corr_autocorrelate(plain._handle, g.handle))
        ctx.synchronize()
        t_plain += time.perf_counter() - t0
        plain._fetch()
        ref = plain.results()
        got = batch.results(i)
        err = max(err, float(multipole_err(got, ref, "k_hist").max()),
                  float(multipole_err(got, ref, "r_hist").max()))
    plain_rate = nplain / t_plain
    return {"value": rate, "unit": "mocks/s", "mocks": b["mocks"], "grid": n,
            "particles_per_mock": b["parents"] * b["children"],
            "fp64_plain_mock_by_mock": plain_rate, "batched_over_fp64_plain": rate / plain_rate,
            "max_floor_rel_err_vs_fp64_plain": err,
            "mass_assign_ms_per_mock": 1e3 * t_ma / b["mocks"]}


SERVE_CFG = dict(rate=20.0, jobs=200, seed=5, box=1000.0, sizes=(128, 256, 384, 512),
                 npart=(1e5, 1e7), slots=4)

YAMA_CFG = dict(n=512, posmin=500.0, box=2000.0, rmax=200.0, dr=5.0, kmax=0.5, dk=0.005,
                maxell=4, steps=2)


def yamamoto_512(lib, corr_mod, ctx, pos):
    """The anisotropic stress case: the Yamamoto Correlator (line of sight
    from an observer at the origin, 15 (l, m) transform pairs for l <= 4) on
    the headline's 8M points moved to [500, 2500)^3, 512^3 CIC grid, kmax 0.5
    dk 0.005, rmax 200 dr 5; mixed and FP64 fused forms on the same grid."""
    c = YAMA_CFG
    n = c["n"]
    cell = c["box"] / n
    posmin = np.full(3, c["posmin"])
    shifted = np.ascontiguousarray(pos + posmin, np.float64)
    out = {}
    for prec in ("mixed", "fp64"):
        g = lib.DeviceGrid((n, n, n), posmin, cell, prec, ctx)
        st = lib.MaStats()
        lib.check(lib.lib().fc_mass_assign(g.handle, lib.dptr(shifted), shifted.shape[0], 3,
                                           None, 1.0, 3, 0, 0.0, st))
        g.normalize(st.totw / float(n) ** 3)
        corr = corr_mod.Correlator((n, n, n), cell, posmin, 3, c["rmax"], c["dr"], c["kmax"],
                                   c["dk"], c["maxell"], precision=prec)
        lib.check(lib.lib().fc_corr_autocorrelate(corr._handle, g.handle))  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(c["steps"]):
            lib.check(lib.lib().fc_corr_autocorrelate(corr._handle, g.handle))
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / c["steps"]
        corr._fetch()
        res = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
               for k, v in corr.results().items()}
        out[prec] = (dt, res)
        del corr, g
    ref = out["fp64"][1]
    err = {}
    for key, name in (("k_hist", "P"), ("r_hist", "xi")):
        e = multipole_err(out["mixed"][1], ref, key)
        for q in range(e.shape[0]):
            err["%s_%d" % (name, 2 * q)] = float(e[q].max())
    return {"value": 1.0 / out["mixed"][0], "unit": "correlations/s",
            "ms_per_correlation": 1e3 * out["mixed"][0],
            "fp64_fused": 1.0 / out["fp64"][0], "max_floor_rel_err_vs_fp64": err,
            "config": "Correlator 512^3 l<=4 (15 Y_lm transforms), observer at the origin, "
                      "headline points in [500, 2500)^3, CIC, kmax 0.5 dk 0.005, rmax 200 dr 5"}


def yamamoto_stages(lib, corr_mod, ctx, pos, steps):
    """The mixed Yamamoto case of yamamoto_512 alone (for profiler runs):
    wall time per correlation and HIP-event times per stage, on stderr."""
    c = YAMA_CFG
    n = c["n"]
    cell = c["box"] / n
    posmin = np.full(3, c["posmin"])
    shifted = np.ascontiguousarray(pos + posmin, np.float64)
    g = lib.DeviceGrid((n, n, n), posmin, cell, "mixed", ctx)
    st = lib.MaStats()
    lib.check(lib.lib().fc_mass_assign(g.handle, lib.dptr(shifted), shifted.shape[0], 3, None,
                                       1.0, 3, 0, 0.0, st))
    g.normalize(st.totw / float(n) ** 3)
    corr = corr_mod.Correlator((n, n, n), cell, posmin, 3, c["rmax"], c["dr"], c["kmax"],
                               c["dk"], c["maxell"], precision="mixed")
    lib.check(lib.lib().fc_corr_autocorrelate(corr._handle, g.handle))  # warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.check(lib.lib().fc_corr_autocorrelate(corr._handle, g.handle))
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    corr.set_profiling(True)
    for _ in range(steps):
        lib.check(lib.lib().fc_corr_autocorrelate(corr._handle, g.handle))
    ctx.synchronize()
    stages = {k: {"ms_per_correlation": v[0] / steps, "launches_per_correlation": v[1] / steps}
              for k, v in corr.stage_times().items() if v[1]}
    log("Yamamoto 512^3 mixed: %.3f ms/correlation" % (1e3 * dt))
    log(json.dumps(stages))


def serving(lib, corr_mod):
    """Poisson arrivals (20 jobs/s, seed 5) of periodic xi_l / P_l jobs, l <= 4:
    grids U{128, 256, 384, 512}^3, log-uniform 0.1M-10M uniform or Thomas
    points (parents N/100, sigma 4) in a 1000 Mpc/h box, float32 weights
    U[0.5, 1.5], CIC; served by 4 slots in the mixed form with resident
    plans.  Reports latency percentiles under the stream, the burst capacity
    (all jobs queued at once), and the FP64 plain form's capacity on the same
    jobs run one after another."""
    from fftcorr_amd.catalog import thomas_catalog, uniform_catalog
    from fftcorr_amd.serve import CorrelationServer, Job, poisson_stream
    c = SERVE_CFG
    rng = np.random.default_rng(c["seed"])
    jobs = []
    for i in range(c["jobs"]):
        n = int(rng.choice(c["sizes"]))
        npart = int(round(np.exp(rng.uniform(np.log(c["npart"][0]), np.log(c["npart"][1])))))
        if rng.uniform() < 0.5:
            pos = uniform_catalog(npart, 0.0, c["box"], c["seed"] * 1000 + i)
        else:
            pos = thomas_catalog(max(1, npart // 100), 100, 4.0, 0.0, c["box"],
                                 c["seed"] * 1000 + i)
        w = rng.uniform(0.5, 1.5, size=pos.shape[0]).astype(np.float32)
        if rng.uniform() < 0.5:
            kw = dict(rmax=200.0, dr=5.0, kmax=0.03, dk=0.01)
        else:
            kw = dict(rmax=20.0, dr=5.0, kmax=0.4, dk=0.01)
        jobs.append(Job(n=n, box=c["box"], positions=pos, weights=w, tag=i, **kw))
    arrivals = poisson_stream(c["rate"], c["jobs"], c["seed"])
    with CorrelationServer(slots=c["slots"], precision="mixed") as srv:
        # Warm every slot with every geometry (resident plans and subgrids).
        geos = {}
        for j in jobs:
            geos.setdefault((j.n, j.kmax), j)
        warm = [Job(n=j.n, box=j.box, positions=j.positions[:1000], rmax=j.rmax, dr=j.dr,
                    kmax=j.kmax, dk=j.dk) for j in geos.values() for _ in range(c["slots"])]
        for f in [srv.submit(j) for j in warm]:
            f.result()
        t0 = time.perf_counter()
        futs = []
        for j, t in zip(jobs, arrivals):
            dt = t0 + t - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            futs.append(srv.submit(j))
        res = [f.result() for f in futs]
        lat = np.array([r.latency_ms for r in res])
        tb = time.perf_counter()
        for f in [srv.submit(j) for j in jobs]:
            f.result()
        burst = len(jobs) / (time.perf_counter() - tb)
    # The same burst through one slot (one stream): what the concurrent slots add.
    with CorrelationServer(slots=1, precision="mixed") as srv:
        for f in [srv.submit(j) for j in warm[::c["slots"]]]:
            f.result()
        tb = time.perf_counter()
        for f in [srv.submit(j) for j in jobs]:
            f.result()
        burst1 = len(jobs) / (time.perf_counter() - tb)
    # FP64 plain, one job after another (correlators cached per geometry).
    cache = {}
    tp = 0.0
    for j in jobs:
        key = (j.n, j.kmax)
        if key not in cache:
            cc = corr_mod.PeriodicCorrelator((j.n,) * 3, j.box / j.n, j.window, j.rmax, j.dr,
                                             j.kmax, j.dk, j.maxell, precision="fp64")
            cc.algorithm = "plain"
            g = lib.DeviceGrid((j.n,) * 3, (0, 0, 0), j.box / j.n, "fp64")
            cache[key] = (cc, g)
        cc, g = cache[key]
        pos = np.ascontiguousarray(j.positions)
        w = np.ascontiguousarray(j.weights, np.float64)
        t1 = time.perf_counter()
        g.clear()
        st = lib.MaStats()
        lib.check(lib.lib().fc_mass_assign(g.handle, lib.dptr(pos), pos.shape[0], 3, lib.dptr(w),
                                           1.0, j.scheme, 1, 0.0, ctypes.byref(st)))
        g.normalize(st.totw / float(j.n) ** 3)
        lib.check(lib.lib().fc_corr_autocorrelate(cc._handle, g.handle))
        cc._fetch()
        tp += time.perf_counter() - t1
    return {"jobs": len(jobs), "arrival_rate_hz": c["rate"], "slots": c["slots"],
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "throughput_jobs_per_s": len(jobs) / (max(r.finished for r in res) - t0),
            "burst_capacity_jobs_per_s": burst,
            "burst_capacity_1_slot_jobs_per_s": burst1,
            "fp64_plain_serial_jobs_per_s": len(jobs) / tp,
            "burst_over_fp64_plain": burst / (len(jobs) / tp)}


def floor_err(x, ref):
    """|x - ref| / max(|ref|, 1e-3 max|ref_l=0|) per multipole row and bin."""
    floor = 1e-3 * np.max(np.abs(ref[0]))
    return np.abs(x - ref) / np.maximum(np.abs(ref), floor)


def multipole_err(res, ref, key):
    """The error metric on the multipoles themselves (histogram / count, as
    the output tables hold them), bins with counts only; counts are equal."""
    c = key.replace("hist", "counts")
    ok = ref[c][0] > 0
    x = res[key][:, ok] / res[c][:, ok]
    r = ref[key][:, ok] / ref[c][:, ok]
    return floor_err(x, r)


def cpu_baseline(grids, cfg, reps=5):
    """FP64 oracle on interlaced correlations of the same grids: with its FFTs
    on all host threads (scipy.fft), 1 warm-up then `reps` timed runs,
    median; and the reference's own threading (its 3D FFTW plan is single
    threaded, cc/grid/fft_grid.cc:81-84), one timed run on one thread."""
    from oracle import fftcorr_oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    n = cfg["n"]
    da = grids[0].download()
    db = grids[1].download()
    ref = O.PeriodicCorrelator((n, n, n), cfg["box"] / n, cfg["window"], cfg["rmax"], cfg["dr"],
                               cfg["kmax"], cfg["dk"], cfg["maxell"])
    ref.fast_hist = True
    old = O.FFT_WORKERS
    try:
        O.FFT_WORKERS = threads
        out = ref.autocorrelate_interlaced(da, db)  # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = ref.autocorrelate_interlaced(da, db)
            times.append(time.perf_counter() - t0)
        O.FFT_WORKERS = 1
        t0 = time.perf_counter()
        ref.autocorrelate_interlaced(da, db)
        t1 = time.perf_counter() - t0
    finally:
        O.FFT_WORKERS = old
    return float(np.median(times)), threads, out, times, t1


C1_CFG = dict(n=128, box=1000.0, npart=100_000, seed=1, rmax=200.0, dr=5.0, kmax=0.03, dk=0.01,
              reps=5)


def config1(lib, corr_mod, ctx, steps):
    """The reference's CPU-runnable case: 100,000 uniform points (seed 1) in a
    periodic 1000 Mpc/h box, CIC onto 128^3, xi_0 with r bins 0-200 by 5 and
    k bins 0-0.03 by 0.01, FP64 -- the GPU FP64 (fused) form against the
    FP64 oracle on one thread (as the reference runs) and on all threads, 1
    warm-up then 5 timed runs each, medians."""
    from fftcorr_amd.catalog import uniform_catalog
    from oracle import fftcorr_oracle as O
    c = C1_CFG
    n = c["n"]
    cell = c["box"] / n
    pos = np.ascontiguousarray(uniform_catalog(c["npart"], 0.0, c["box"], c["seed"]))
    g = lib.DeviceGrid((n, n, n), (0, 0, 0), cell, "fp64", ctx)
    st = lib.MaStats()
    lib.check(lib.lib().fc_mass_assign(g.handle, lib.dptr(pos), pos.shape[0], 3, None, 1.0, 3, 1,
                                       0.0, st))
    g.normalize(st.totw / float(n) ** 3)
    args = ((n, n, n), cell, 0, c["rmax"], c["dr"], c["kmax"], c["dk"], 0)
    corr = corr_mod.PeriodicCorrelator(*args, precision="fp64")
    L = lib.lib()
    lib.check(L.fc_corr_autocorrelate(corr._handle, g.handle))
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.check(L.fc_corr_autocorrelate(corr._handle, g.handle))
    ctx.synchronize()
    gpu = steps / (time.perf_counter() - t0)
    corr._fetch()
    res = corr.results()
    delta = g.download()
    ref = O.PeriodicCorrelator(*args)
    ref.fast_hist = True
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    rates = {}
    old = O.FFT_WORKERS
    try:
        for nt in (1, threads):
            O.FFT_WORKERS = nt
            want = ref.autocorrelate(delta)
            ts = []
            for _ in range(c["reps"]):
                t0 = time.perf_counter()
                want = ref.autocorrelate(delta)
                ts.append(time.perf_counter() - t0)
            rates[nt] = 1.0 / float(np.median(ts))
    finally:
        O.FFT_WORKERS = old
    err = max(float(floor_err(res["r_hist"], want["r_hist"]).max()),
              float(floor_err(res["k_hist"], want["k_hist"]).max()))
    return {"gpu_fp64": gpu, "unit": "correlations/s",
            "cpu_port_1_thread": rates[1], "cpu_port_threads": rates[threads],
            "cpu_threads": threads,
            "counts_bit_exact": bool(np.array_equal(res["r_counts"], want["r_counts"])),
            "gpu_vs_cpu_max_floor_rel_err": err,
            "config": "periodic 128^3 CIC, 100,000 uniform points (seed 1), 1000 Mpc/h, xi_0, "
                      "r 0-200 by 5, k 0-0.03 by 0.01, FP64"}


BATCH512_CFG = dict(n=512, box=2000.0, parents=80_000, children=100, sigma=3.0, members=8,
                    rmax=200.0, dr=5.0, kmax=0.5, dk=0.005, maxell=4, window=3, scheme=3, reps=3)


def autotune_512(lib, corr_mod, ctx, pos, cfg, steps, default_out):
    """The autotuner (the counterpart of the reference's FFTW_MEASURE
    planning) on the 512^3 plans: its radix plan and tile choices and best
    pass times, and the headline form re-timed with the tuned plan against
    the default plan it was measured with."""
    from fftcorr_amd import fftw
    from fftcorr_amd.autotune import autotune
    n = cfg["n"]
    shape = (n, n, n)
    t0 = time.perf_counter()
    res = autotune(shape, "mixed", reps=5)
    tune_s = time.perf_counter() - t0
    res64 = autotune(shape, "fp64", reps=5)
    out, g, c = run_form(lib, corr_mod, ctx, pos, cfg, "mixed", steps, 3, profile=True)
    del g, c
    fftw.forget_wisdom()
    d_ms = 1e3 * default_out["wall_s"] / default_out["steps"]
    t_ms = 1e3 * out["wall_s"] / steps
    err = max(float(floor_err(out["res"][k], default_out["res"][k]).max())
              for k in ("k_hist", "r_hist"))
    return {
        "candidates": "z pass: radix plans 8x8x4 and 16x16, rows per workgroup 1-64; y and x "
                      "passes: radix plans 8x8x8 and 16x16x2, columns per workgroup 1-32; "
                      "z+y chunk sizes",
        "mixed": {"choices": {k: v for k, v in res.items() if k != "ms"},
                  "best_pass_ms": res["ms"]},
        "fp64": {"choices": {k: v for k, v in res64.items() if k != "ms"},
                 "best_pass_ms": res64["ms"]},
        "tune_s": tune_s,
        "default_ms_per_step": d_ms,
        "tuned_ms_per_step": t_ms,
        "tuned_over_default": d_ms / t_ms,
        "tuned_stage_ms": {k: v[0] / v[1] for k, v in out["stages"].items() if v[1]},
        "max_floor_rel_err_tuned_vs_default": err,
    }


def batched_512(lib, corr_mod, ctx, pos0):
    """The north-star form: a batch of 8 periodic 512^3 correlations (l <= 4,
    kmax 0.5 dk 0.005, rmax 200 dr 5, CIC, sinc^4 window; the headline
    catalogue and seven more seeded Thomas catalogues of the same kind) in
    the mixed form, one HIP-graph launch sequence per batch, against the
    FP64 plain form run member by member on the same grids."""
    from fftcorr_amd.catalog import thomas_catalog
    b = BATCH512_CFG
    n = b["n"]
    cell = b["box"] / n
    args = ((n, n, n), cell, b["window"], b["rmax"], b["dr"], b["kmax"], b["dk"], b["maxell"])
    batch = corr_mod.PeriodicBatch(*args, nbatch=b["members"], precision="mixed", use_graph=True)
    try:
        for i in range(b["members"]):
            pos = pos0 if i == 0 else np.ascontiguousarray(
                thomas_catalog(b["parents"], b["children"], b["sigma"], 0.0, b["box"], 40 + i))
            batch.assign(i, pos, scheme=b["scheme"])
        batch.autocorrelate()  # captures the graph
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(b["reps"]):
            batch.autocorrelate()
        ctx.synchronize()
        rate = b["reps"] * b["members"] / (time.perf_counter() - t0)
        plain = corr_mod.PeriodicCorrelator(*args, precision="fp64")
        plain.algorithm = "plain"
        g = lib.DeviceGrid((n, n, n), (0, 0, 0), cell, "fp64", ctx)
        err = 0.0
        t_plain = 0.0
        nplain = 2
        for i in range(nplain):
            g.upload(batch.grid(i).download())
            lib.check(lib.lib().fc_corr_autocorrelate(plain._handle, g.handle))  # warm-up
            ctx.synchronize()
            t0 = time.perf_counter()
            lib.check(lib.lib().fc_corr_autocorrelate(plain._handle, g.handle))
            ctx.synchronize()
            t_plain += time.perf_counter() - t0
            plain._fetch()
            ref = plain.results()
            got = batch.results(i)
            err = max(err, float(multipole_err(got, ref, "k_hist").max()),
                      float(multipole_err(got, ref, "r_hist").max()))
        plain_rate = nplain / t_plain
        del g, plain
    finally:
        batch.close()
    return {"value": rate, "unit": "correlations/s", "members": b["members"], "grid": n,
            "fp64_plain_one_by_one": plain_rate, "batched_mixed_over_fp64_plain": rate / plain_rate,
            "max_floor_rel_err_vs_fp64_plain": err,
            "config": "batch of 8 periodic 512^3 l<=4 correlations (no interlacing), CIC, "
                      "sinc^4 window, kmax 0.5 dk 0.005, rmax 200 dr 5, mixed, HIP graph"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="time the mixed form only (for profiler runs); prints no JSON line")
    ap.add_argument("--yamamoto-only", action="store_true",
                    help="time the mixed 512^3 Yamamoto case only, with stage times (for "
                         "profiler runs); prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the correlation path here is single-GPU "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from fftcorr_amd import _lib as lib
    from fftcorr_amd import correlate as corr_mod
    from fftcorr_amd.catalog import thomas_catalog

    cfg = CFG
    ctx = lib.context(0)
    log("generating %d-parent Thomas catalogue" % cfg["parents"])
    pos = thomas_catalog(cfg["parents"], cfg["children"], cfg["sigma"], 0.0, cfg["box"],
                         cfg["seed"])
    pos = np.ascontiguousarray(pos, np.float64)
    log("particles: %d" % pos.shape[0])

    if args.yamamoto_only:
        yamamoto_stages(lib, corr_mod, ctx, pos, args.steps)
        return
    side = args.fp64_steps or max(3, args.steps // 4)
    log("mixed form (headline): %d warmup + %d timed steps" % (args.warmup, args.steps))
    fm, gm, cm = run_form(lib, corr_mod, ctx, pos, cfg, "mixed", args.steps, args.warmup,
                          profile=True)
    del cm
    if args.headline_only:
        del gm
    else:
        log("mixed form from host grids (PCIe upload per call): 3 steps")
        host_rate = host_grid_rate(lib, corr_mod, gm, cfg, 3)
        del gm
    if args.headline_only:
        log("mixed form: %.4f ms/correlation" % (1e3 * fm["wall_s"] / args.steps))
        return
    forms = {}
    for name, prec, algo in (("fp32_fused", "fp32", "fused"), ("fp64_fused", "fp64", "fused"),
                             ("fp64_plain", "fp64", "plain")):
        log("%s form: 1 warmup + %d timed steps" % (name, side))
        out, g, c = run_form(lib, corr_mod, ctx, pos, cfg, prec, side, 1, profile=False,
                             algorithm=algo)
        forms[name] = out
        if name != "fp64_fused":
            del g
        else:
            g64 = g
        del c

    def rate(out):
        return out["steps"] / out["wall_s"]

    value = rate(fm)
    ref64 = forms["fp64_fused"]["res"]

    def errs(res, sums=False):
        err = {}
        for key, tag in (("k_hist", "P"), ("r_hist", "xi")):
            e = floor_err(res[key], ref64[key]) if sums else multipole_err(res, ref64, key)
            for q in range(e.shape[0]):
                err["%s_%d" % (tag, 2 * q)] = float(e[q].max())
        return err

    def counts_equal(res):
        return bool(np.array_equal(res["k_counts"], ref64["k_counts"])
                    and np.array_equal(res["r_counts"], ref64["r_counts"]))

    err = errs(fm["res"])
    plain_vs_fused = max(float(floor_err(forms["fp64_plain"]["res"][k], ref64[k]).max())
                         for k in ("k_hist", "r_hist"))

    # Roofline of the dominant stage (HIP events on the correlator's stream).
    st = {k: v for k, v in fm["stages"].items() if v[1] > 0}
    # Two grids per correlation: one forward launch per pass covers both.
    fwd_grids = max(1, round(2 * fm["steps"] / st["fft_z_fwd"][1]))
    sb = stage_bytes(cfg, 4, 4, fm["kshape"], fm["rshape"][0], s_inv=8,
                     k_entries=k_half_entries(cfg, fm["kshape"]), fwd_grids=fwd_grids)
    dom = max(st, key=lambda k: st[k][0])
    avg_ms = st[dom][0] / st[dom][1]
    achieved = sb[dom] / (avg_ms * 1e-3) / 1e9 if dom in sb else None
    stage_report = {k: {"ms_per_launch": v[0] / v[1], "launches_per_step": v[1] / args.steps,
                        "GBps": (sb[k] / (v[0] / v[1] * 1e-3) / 1e9) if k in sb else None}
                    for k, v in st.items()}

    # HBM bytes per launch of the dominant kernel, from the committed
    # FETCH_SIZE / WRITE_SIZE counter runs (profiles/traffic.json), when they
    # cover that stage.
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if dom in tj.get("stages", {}):
            traffic = tj["stages"][dom]["bytes"]

    cpu = None
    c1 = None
    if not args.no_cpu:
        log("CPU baseline: FP64 oracle, 1 warm-up + 5 timed correlations, then one on 1 thread")
        dt, threads, ref, times, t1 = cpu_baseline(g64, cfg)
        e_cpu = max(float(floor_err(ref64["k_hist"], ref["k_hist"]).max()),
                    float(floor_err(ref64["r_hist"], ref["r_hist"]).max()))
        cpu = {"value": 1.0 / dt, "unit": "correlations/s", "cores": threads, "kind": "port",
               "sample": "interlaced 512^3 l<=4 correlations of the same grids, FP64 "
                         "numpy/scipy oracle (scipy.fft workers=%d): 1 warm-up + %d timed, "
                         "median %.2f s" % (threads, len(times), dt),
               "times_s": times,
               "faithful_1_thread": {"value": 1.0 / t1, "unit": "correlations/s", "cores": 1,
                                     "sample": "one correlation with every FFT on one thread, "
                                               "as the reference's 3D FFTW plan runs; %.1f s"
                                               % t1},
               "gpu_fp64_vs_cpu_max_floor_rel_err": e_cpu}
        log("config 1: 128^3 CIC xi_0 FP64, GPU and CPU")
        c1 = config1(lib, corr_mod, ctx, 20)

    log("autotuner on the 512^3 plans; headline form re-timed with the tuned plan")
    tune_report = autotune_512(lib, corr_mod, ctx, pos, cfg, args.steps, fm)
    log("Yamamoto correlator at 512^3 (anisotropic stress case)")
    yama_report = yamamoto_512(lib, corr_mod, ctx, pos)
    log("batched mocks: 64 x 256^3 TSC, mixed, HIP graph")
    batch_report = batched_mocks(lib, corr_mod, ctx, 5)
    log("batched 512^3: 8 members, mixed, HIP graph, against FP64 plain")
    batch512_report = batched_512(lib, corr_mod, ctx, pos)
    log("serving: Poisson stream of mixed-size jobs")
    serve_report = serving(lib, corr_mod)

    lat = fm["lat_ms"]

    def side_report(out):
        r = {"value": rate(out), "ms_per_step": 1e3 * out["wall_s"] / out["steps"],
             "p50_ms": float(np.percentile(out["lat_ms"], 50))}
        return r

    fp32 = side_report(forms["fp32_fused"])
    fp32["max_floor_rel_err_vs_fp64"] = errs(forms["fp32_fused"]["res"])
    fp32["counts_bit_exact"] = counts_equal(forms["fp32_fused"]["res"])
    line = {
        "metric": "correlations/s @512^3 l<=4, FP32-mixed vs FP64 plain at <=1e-5 rel err; p50/p99",
        "value": value,
        "unit": "correlations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * fm["wall_s"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "periodic P_l(k) + xi_l(r), l<=4, 512^3, CIC + interlacing, "
                        "8M Thomas points in a 2000 Mpc/h box, kmax 0.5 dk 0.005, "
                        "rmax 200 dr 5",
            "grid": cfg["n"], "maxell": cfg["maxell"], "particles": int(pos.shape[0]),
            "k_subgrid": list(fm["kshape"]), "r_subgrid": list(fm["rshape"]),
            "precision": "mixed: FP32 grids and spectra in HBM and LDS; FP64 arithmetic in "
                         "the z and y forward FFT passes; FP32 x pass and products; FP64 "
                         "pruned inverse FFT, Y_lm, windows, bins and sums",
        },
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_floor_rel_err_vs_fp64": err,
        "max_floor_rel_err_vs_fp64_on_hist_sums": errs(fm["res"], sums=True),
        "counts_bit_exact": counts_equal(fm["res"]),
        "fp64_plain": side_report(forms["fp64_plain"]),
        "fp64_fused": side_report(forms["fp64_fused"]),
        "fp32_fused": fp32,
        "mixed_host_grids_pcie_inclusive": {"value": host_rate, "unit": "correlations/s"},
        "mixed_over_fp64_plain": value / rate(forms["fp64_plain"]),
        "mixed_over_fp64_fused": value / rate(forms["fp64_fused"]),
        "fp64_plain_vs_fused_max_floor_rel_err": plain_vs_fused,
        "mass_assign_ms_two_grids": fm["ma_ms"],
        "end_to_end_with_mass_assignment": {
            "value": 1e3 / (fm["ma_ms"] + 1e3 * fm["wall_s"] / args.steps),
            "unit": "correlations/s",
            "note": "8M points from host memory assigned to both grids (CIC, interlaced), "
                    "then the correlation; not the headline"},
        "yamamoto_512": yama_report,
        "batched_256": batch_report,
        "batched_512": batch512_report,
        "autotune_512": tune_report,
        "config1_128": c1,
        "serving": serve_report,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": (achieved / PEAK_HBM_GBS) if achieved else None,
                     "traffic": traffic, "algorithmic_bytes": sb.get(dom)},
        "stages": stage_report,
        "cpu_baseline": cpu,
    }
    return line


if __name__ == "__main__":
    # Everything but the one JSON line (the grid classes' own messages
    # included) goes to stderr.
    with contextlib.redirect_stdout(sys.stderr):
        result = main()
    if result is not None:
        print(json.dumps(result), flush=True)
