"""ctypes binding of libfftcorr_hip.so (the C ABI declared in
include/fftcorr_hip.h) and the per-device default context.

The library is required: there is no CPU fallback.  Importing the package
works without it (so that the pure-Python pieces and error paths can be
used), but every compute call raises if the library is missing.
"""

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libfftcorr_hip.so")

FC_OK = 0
FC_ERR_INVALID = 1
FC_ERR_RANGE = 2
FC_ERR_RUNTIME = 3
FC_ERR_NOMEM = 4
FC_ERR_HIP = 5

FC_FP64 = 0
FC_FP32 = 1
FC_MIXED = 2
FC_ALGO_FUSED = 0
FC_ALGO_PLAIN = 1
FC_PERIODIC = 0
FC_YAMAMOTO = 1

# Stage ids of fc_corr_stage_times (include/fftcorr_hip.h).
STAGES = ("fft_z_fwd", "fft_y_fwd", "fft_x_fwd", "combine", "hist_k", "fft_x_inv",
          "fft_y_inv", "fft_z_inv", "hist_r", "subgrid")

_ERRORS = {
    FC_ERR_INVALID: ValueError,
    FC_ERR_RANGE: IndexError,
    FC_ERR_RUNTIME: RuntimeError,
    FC_ERR_NOMEM: MemoryError,
    FC_ERR_HIP: RuntimeError,
}

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64


class MaStats(ctypes.Structure):
    _fields_ = [("num_added", ctypes.c_uint64), ("num_skipped", ctypes.c_uint64),
                ("totw", ctypes.c_double), ("totwsq", ctypes.c_double),
                ("sort_time", ctypes.c_double), ("window_time", ctypes.c_double)]


# name -> (restype, argtypes); every entry point of include/fftcorr_hip.h.
SIGNATURES = {
    "fc_last_error": (ctypes.c_char_p, []),
    "fc_version": (ctypes.c_int, []),
    "fc_device_count": (ctypes.c_int, [_ip]),
    "fc_device_memory": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                        ctypes.POINTER(ctypes.c_size_t)]),
    "fc_device_bytes": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(ctypes.c_size_t)]),
    "fc_device_bytes_reset_peak": (ctypes.c_int, [ctypes.c_int]),
    "fc_mass_assign_footprint": (ctypes.c_int, [_vp, _i64, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                                ctypes.POINTER(ctypes.c_size_t)]),
    "fc_corr_footprint": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                         ctypes.POINTER(ctypes.c_size_t)]),
    "fc_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "fc_ctx_destroy": (ctypes.c_int, [_vp]),
    "fc_ctx_footprint": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "fc_ctx_release_scratch": (ctypes.c_int, [_vp]),
    "fc_ctx_synchronize": (ctypes.c_int, [_vp]),
    "fc_grid_create": (ctypes.c_int, [_vp, _ip, _dp, ctypes.c_double, ctypes.c_int,
                                      ctypes.POINTER(_vp)]),
    "fc_grid_destroy": (ctypes.c_int, [_vp]),
    "fc_grid_clear": (ctypes.c_int, [_vp]),
    "fc_grid_upload": (ctypes.c_int, [_vp, _dp]),
    "fc_grid_download": (ctypes.c_int, [_vp, _dp]),
    "fc_grid_stats": (ctypes.c_int, [_vp, _dp, _dp]),
    "fc_grid_add_scalar": (ctypes.c_int, [_vp, ctypes.c_double]),
    "fc_grid_multiply_by": (ctypes.c_int, [_vp, ctypes.c_double]),
    "fc_grid_normalize": (ctypes.c_int, [_vp, ctypes.c_double]),
    "fc_grid_device_ptr": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "fc_mass_assign": (ctypes.c_int, [_vp, _dp, _i64, ctypes.c_int, _dp, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.POINTER(MaStats)]),
    "fc_mass_assign_device": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_int, _vp,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.POINTER(MaStats)]),
    "fc_corr_create": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.c_double, _dp,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, ctypes.POINTER(_vp)]),
    "fc_corr_destroy": (ctypes.c_int, [_vp]),
    "fc_corr_geometry": (ctypes.c_int, [_vp, _ip, _ip, _ip, _ip, _ip]),
    "fc_corr_set_grid2": (ctypes.c_int, [_vp, _vp]),
    "fc_corr_set_grid2_fft": (ctypes.c_int, [_vp, _dp]),
    "fc_corr_autocorrelate": (ctypes.c_int, [_vp, _vp]),
    "fc_corr_cross_correlate": (ctypes.c_int, [_vp, _vp]),
    "fc_corr_autocorrelate_interlaced": (ctypes.c_int, [_vp, _vp, _vp]),
    "fc_corr_results": (ctypes.c_int, [_vp, _dp, _ip, _dp, _dp, _ip, _dp, _dp]),
    "fc_corr_last_time_ms": (ctypes.c_int, [_vp, _dp]),
    "fc_corr_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fc_corr_set_algorithm": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fc_autotune": (ctypes.c_int, [_vp, _ip, ctypes.c_int, ctypes.c_int, _dp]),
    "fc_set_wavelet_table": (ctypes.c_int, [_dp, ctypes.c_int]),
    "fc_apply_displacement": (ctypes.c_int, [_vp, _dp, _i64, _dp, ctypes.c_int, _dp]),
    "fc_add_random_particles": (ctypes.c_int, [_vp, _i64, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int, _dp, ctypes.c_uint64, _i64,
                                               ctypes.POINTER(MaStats)]),
    "fc_random_uniform_device": (ctypes.c_int, [_vp, _vp, _i64, _dp, _dp, ctypes.c_uint64,
                                                ctypes.c_uint64]),
    "fc_random_raw_device": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.c_uint32, ctypes.c_uint32]),
    "fc_random_uniform": (ctypes.c_int, [_vp, _dp, _i64, _dp, _dp, ctypes.c_uint64,
                                         ctypes.c_uint64]),
    "fc_random_raw": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32), _i64,
                                     ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                     ctypes.c_uint32]),
    "fc_batch_create": (ctypes.c_int, [_vp, _ip, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(_vp)]),
    "fc_batch_destroy": (ctypes.c_int, [_vp]),
    "fc_batch_grid": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "fc_batch_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fc_batch_autocorrelate": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fc_batch_results": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _ip, _dp, _ip, _dp]),
    "fc_batch_correlator": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "fc_batch_last_time_ms": (ctypes.c_int, [_vp, _dp]),
    "fc_plan_tiles": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _ip]),
    "fc_plan_variants": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _ip]),
    "fc_fft_radices": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _ip, _ip]),
    "fc_corr_stage_times": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_int64)]),
    "fc_hist_accumulate": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          _dp, _dp, _i64, _ip, _dp]),
    "fc_hist_accumulate_device_noalloc": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                                         ctypes.c_int, _vp, _vp, _i64, _vp,
                                                         _vp]),
    "fc_fft_r2c": (ctypes.c_int, [_vp, _ip, ctypes.c_int, _dp, _dp]),
    "fc_fft_c2r": (ctypes.c_int, [_vp, _ip, ctypes.c_int, _dp, _dp]),
    "fc_plan_cache_export": (ctypes.c_int, [ctypes.c_char_p]),
    "fc_plan_cache_import": (ctypes.c_int, [ctypes.c_char_p]),
    "fc_plan_cache_forget": (ctypes.c_int, []),
}

_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "libfftcorr_hip.so is not built (expected at %s); run "
                    "`python -m fftcorr_amd.build`" % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def check(rc):
    if rc != FC_OK:
        msg = lib().fc_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, RuntimeError)(msg)


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def precision_code(precision):
    """Correlator precision: "fp64" (FC_FP64), "fp32" (FC_FP32) or "mixed"
    (FC_MIXED: FP32 grids and spectra, FP64 arithmetic in the z and y forward
    FFT passes, FP64 inverse)."""
    if isinstance(precision, str):
        key = precision.lower()
    else:
        key = precision
    if key in (FC_FP64, "fp64", "float64", np.float64):
        return FC_FP64
    if key in (FC_FP32, "fp32", "float32", np.float32):
        return FC_FP32
    if key in (FC_MIXED, "mixed"):
        return FC_MIXED
    raise ValueError("precision must be 'fp64', 'fp32' or 'mixed', got %r" % (precision,))


def grid_precision_code(precision):
    """Storage precision of a density grid: FP32 for "fp32" and "mixed"."""
    code = precision_code(precision)
    return FC_FP64 if code == FC_FP64 else FC_FP32


class Context:
    """One device and one HIP stream (fc_ctx)."""

    def __init__(self, device=0):
        h = _vp()
        check(lib().fc_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = int(device)

    def synchronize(self):
        check(lib().fc_ctx_synchronize(self.handle))

    def close(self):
        if self.handle:
            lib().fc_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}


def default_device():
    return int(os.environ.get("FFTCORR_DEVICE", "0"))


def context(device=None):
    """The shared per-device context (created on first use)."""
    device = default_device() if device is None else int(device)
    with _lock:
        c = _contexts.get(device)
    if c is None:
        c = Context(device)
        with _lock:
            c = _contexts.setdefault(device, c)
    return c


def device_bytes(device=None):
    """(bytes the library holds on the device now, peak since the last reset)."""
    device = default_device() if device is None else int(device)
    a, b = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().fc_device_bytes(device, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def reset_device_peak(device=None):
    device = default_device() if device is None else int(device)
    check(lib().fc_device_bytes_reset_peak(device))


def device_count():
    n = ctypes.c_int()
    check(lib().fc_device_count(ctypes.byref(n)))
    return n.value


class DeviceGrid:
    """Owning wrapper of an fc_grid handle."""

    def __init__(self, shape, posmin=(0.0, 0.0, 0.0), cell_size=1.0, precision=FC_FP64,
                 ctx=None):
        self.ctx = ctx or context()
        self.shape = tuple(int(s) for s in shape)
        self.precision = grid_precision_code(precision)
        s = np.ascontiguousarray(self.shape, np.intc)
        pm = np.ascontiguousarray(posmin, np.float64)
        h = _vp()
        check(lib().fc_grid_create(self.ctx.handle, iptr(s), dptr(pm), float(cell_size),
                                   self.precision, ctypes.byref(h)))
        self.handle = h

    def upload(self, host):
        a = np.ascontiguousarray(host, np.float64)
        if a.shape != self.shape:
            raise ValueError("Incompatible shapes for copy.")
        check(lib().fc_grid_upload(self.handle, dptr(a)))

    def download(self, out=None):
        if out is None:
            out = np.empty(self.shape, np.float64)
        check(lib().fc_grid_download(self.handle, dptr(out)))
        return out

    def clear(self):
        check(lib().fc_grid_clear(self.handle))

    def stats(self):
        s, s2 = ctypes.c_double(), ctypes.c_double()
        check(lib().fc_grid_stats(self.handle, ctypes.byref(s), ctypes.byref(s2)))
        return s.value, s2.value

    def add_scalar(self, v):
        check(lib().fc_grid_add_scalar(self.handle, float(v)))

    def multiply_by(self, v):
        check(lib().fc_grid_multiply_by(self.handle, float(v)))

    def normalize(self, mean):
        check(lib().fc_grid_normalize(self.handle, float(mean)))

    def close(self):
        if getattr(self, "handle", None):
            lib().fc_grid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
