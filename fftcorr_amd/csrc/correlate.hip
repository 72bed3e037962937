// Correlators (cc/correlate/correlator.h:18-505) on the device.
//
// PeriodicCorrelator (line of sight z):
//   z-pass R2C of the grid (copy_into_padded_array fused into the load)
//   -> y, x forward passes -> |F|^2 or F conj(G) -> k-space multipoles
//   (extract_fft_submatrix_c2r + Y_l0(k) W^-1(k) + histogram, every l in one
//   pass) -> x, y inverse passes -> z-pass C2R -> r-space multipoles
//   (extract_submatrix + Y_l0(r) + histogram, every l in one pass).
// Correlator (Yamamoto, observer at the survey origin):
//   per (l, m): z-pass R2C with Y_lm(x) * delta fused into the load -> y, x
//   -> conj-multiply with the stored FFT of grid 2 -> accumulate the k
//   subgrid -> inverse -> accumulate the r subgrid; histograms once per l.
// Work arrays are FP64 (FC_FP64) or FP32 (FC_FP32); Y_lm, windows, bin
// indices and every sum stay in FP64.
#include <algorithm>
#include <array>

#include "fc_objects.h"
#include "fft_plan.h"
#include "hist_device.h"
#include "ylm_device.h"

struct fc_corr {
  fc_ctx* ctx = nullptr;
  int kind = FC_PERIODIC;
  int shape[3] = {0, 0, 0};
  double cell = 1.0;
  double posmin[3] = {0, 0, 0};
  int window = 0;
  double rmax = 0, dr = 0, kmax = 0, dk = 0;
  int maxell = 0, nell = 1;
  int prec = FC_FP64;
  // Geometry (correlator.h:109-205).
  int rmax_cells = 0;
  int rshape = 0;
  int kshape[3] = {0, 0, 0};
  int nrbins = 0, nkbins = 0;
  std::vector<double> rbins, kbins;
  std::vector<double> h_rx, h_kx, h_ky, h_kz;
  fc::DeviceBuffer rx, kx, ky, kz, wx, wy, wz;  // FP64 axes and per-axis window W^2
  // FFT plan + work arrays.
  std::shared_ptr<fc::FftPlan> plan;
  fc::DeviceBuffer work, grid2_fft, work_b;
  bool have_grid2 = false;
  fc::DeviceBuffer rsub, ksub;  // FP64 subgrids (Yamamoto accumulation)
  // Outputs.
  fc::DeviceBuffer r_counts, r_hist, k_counts, k_hist, zerolag;
  fc::HistScratch hs;
  double h_zerolag = 0.0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  // Per-stage profiling: event pairs recorded during the last call(s).
  bool profile = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::array<int, 3>> pending;  // (stage, start event, end event)
  double stage_ms[FC_NSTAGES] = {0};
  int64_t stage_n[FC_NSTAGES] = {0};
};

namespace fc {
namespace {

// ---- element-wise ----------------------------------------------------------
// out *= conj(in) (array_ops::multiply_with_conjugation, array_ops.cc:139-163);
// in == nullptr: out = |out|^2.
template <typename T>
__global__ void k_conj_mult(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ in,
                            int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const cplx<T> a = out[i];
    if (in) {
      const cplx<T> b = in[i];
      out[i] = cplx<T>{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
    } else {
      out[i] = cplx<T>{a.x * a.x + a.y * a.y, T(0)};
    }
  }
}

// Interlacing: F = (F_a + F_b exp(i pi (fx/n0 + fy/n1 + fz/n2))) / 2 with
// signed frequency indices in the numpy.fft.fftfreq convention.
template <typename T>
__global__ void k_interlace(cplx<T>* __restrict__ fa, const cplx<T>* __restrict__ fb, int n0,
                            int n1, int n2, int n2h) {
  const int64_t n = (int64_t)n0 * n1 * n2h;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % n2h);
    const int64_t r = i / n2h;
    const int b = (int)(r % n1), a = (int)(r / n1);
    const int fx = a < (n0 + 1) / 2 ? a : a - n0;
    const int fy = b < (n1 + 1) / 2 ? b : b - n1;
    const int fz = c < (n2 + 1) / 2 ? c : c - n2;
    const double ph = M_PI * ((double)fx / n0 + (double)fy / n1 + (double)fz / n2);
    double sn, cs;
    sincos(ph, &sn, &cs);
    const cplx<T> A = fa[i], B = fb[i];
    const double bx = (double)B.x * cs - (double)B.y * sn;
    const double by = (double)B.x * sn + (double)B.y * cs;
    fa[i] = cplx<T>{T(0.5 * ((double)A.x + bx)), T(0.5 * ((double)A.y + by))};
  }
}

template <typename Tdst>
__global__ void k_from_c128(cplx<Tdst>* __restrict__ d, const double2* __restrict__ s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = cplx<Tdst>{(Tdst)s[i].x, (Tdst)s[i].y};
}

// ---- z-pass loader with Y_lm(x) weighting (correlator.h:435-436) -----------
struct LoadYlm {
  int ell, m, n1;
  double c, x0, y0, z0;  // cell centres relative to the observer: x0 + i etc.
  __device__ __forceinline__ double operator()(double v, int64_t row, int k) const {
    const int i = (int)(row / n1), j = (int)(row % n1);
    return ylm_eval(ell, m, x0 + (double)i * 1.0, y0 + (double)j * 1.0, z0 + (double)k * 1.0,
                    c, v);
  }
};

// ---- subgrid index maps (cc/grid/fft_grid.cc:182-257) ----------------------
struct KMap {
  int n0, n1, n2h, ks0, ks1, ks2;
  // Half-complex element (a, b, c) feeding k-subgrid cell (i, j, k).
  __device__ __forceinline__ int64_t src(int i, int j, int k) const {
    const int ox = ks0 / 2, oy = ks1 / 2, oz = ks2 / 2;
    const int ii = ((n0 - ox + i) % n0 + n0) % n0;
    const int jj = ((n1 - oy + j) % n1 + n1) % n1;
    if (k < oz) {  // negative kz: reflected (-kx, -ky, |kz|) row
      const int iin = (n0 - ii) % n0, jjn = (n1 - jj) % n1;
      return ((int64_t)iin * n1 + jjn) * n2h + (oz - k);
    }
    return ((int64_t)ii * n1 + jj) * n2h + (k - oz);
  }
};

struct RMap {
  int n0, n1, n2, ld, rs;  // ld: real row stride of the padded work grid
  __device__ __forceinline__ int64_t src(int i, int j, int k) const {
    // Periodic image of lag (i - o, j - o, k - o); the positive modulo also
    // covers subgrids wider than the grid (rmax beyond half the box).
    const int o = rs / 2;
    const int ii = ((n0 - o + i) % n0 + n0) % n0;
    const int jj = ((n1 - o + j) % n1 + n1) % n1;
    const int kk = ((n2 - o + k) % n2 + n2) % n2;
    return ((int64_t)ii * n1 + jj) * ld + kk;
  }
};

// Periodic k-space multipoles, every l at once (correlator.h:296-320).
template <typename T, int NV>
struct PeriodicKVals {
  const cplx<T>* P;
  KMap map;
  const double *kx, *ky, *kz, *wx, *wy, *wz;
  double c[NV];
  double dk;
  int nbins;
  __device__ __forceinline__ void operator()(int64_t idx, int* bin, double* v) const {
#pragma clang fp contract(off)
    const int k = (int)(idx % map.ks2);
    const int64_t t = idx / map.ks2;
    const int j = (int)(t % map.ks1), i = (int)(t / map.ks1);
    const double x = kx[i], y = ky[j], z = kz[k];
    *bin = hist_bin(norm3(x, y, z), 0.0, dk, nbins);
    if (*bin < 0) return;
    const double re = (double)P[map.src(i, j, k)].x;
    const double invw = 1.0 / (wx[i] * wy[j] * wz[k]);
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = ylm_eval(2 * q, 0, x, y, z, c[q], invw) * re;
  }
};

// Periodic r-space multipoles, every l at once (correlator.h:336-355).
template <typename T, int NV>
struct PeriodicRVals {
  const T* xi;
  RMap map;
  const double* rx;
  double c[NV];
  double dr;
  int nbins;
  double* zerolag;
  __device__ __forceinline__ void operator()(int64_t idx, int* bin, double* v) const {
#pragma clang fp contract(off)
    const int rs = map.rs;
    const int k = (int)(idx % rs);
    const int64_t t = idx / rs;
    const int j = (int)(t % rs), i = (int)(t / rs);
    const double x = rx[i], y = rx[j], z = rx[k];
    const double val = (double)xi[map.src(i, j, k)];
    const int o = rs / 2;
    if (i == o && j == o && k == o) *zerolag = ylm_eval(0, 0, x, y, z, c[0], 1.0) * val;
    *bin = hist_bin(norm3(x, y, z), 0.0, dr, nbins);
    if (*bin < 0) return;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = ylm_eval(2 * q, 0, x, y, z, c[q], 1.0) * val;
  }
};

// Histogram of an FP64 subgrid (one multipole row).
struct SubgridVals {
  const double* sub;
  const double *ax, *ay, *az;
  int s1, s2;
  double binsize;
  int nbins;
  __device__ __forceinline__ void operator()(int64_t idx, int* bin, double* v) const {
    const int k = (int)(idx % s2);
    const int64_t t = idx / s2;
    const int j = (int)(t % s1), i = (int)(t / s1);
    *bin = hist_bin(norm3(ax[i], ay[j], az[k]), 0.0, binsize, nbins);
    v[0] = sub[idx];
  }
};

// Yamamoto accumulation into the k subgrid (correlator.h:453-466).
template <typename T>
__global__ void k_acc_k(double* __restrict__ ksub, const cplx<T>* __restrict__ P, KMap map,
                        const double* __restrict__ kx, const double* __restrict__ ky,
                        const double* __restrict__ kz, const double* __restrict__ wx,
                        const double* __restrict__ wy, const double* __restrict__ wz, int ell,
                        int m, double c) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)map.ks0 * map.ks1 * map.ks2;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(idx % map.ks2);
    const int64_t t = idx / map.ks2;
    const int j = (int)(t % map.ks1), i = (int)(t / map.ks1);
    const double invw = 1.0 / (wx[i] * wy[j] * wz[k]);
    const double y = ylm_eval(ell, m, kx[i], ky[j], kz[k], c, invw);
    ksub[idx] += y * (double)P[map.src(i, j, k)].x;
  }
}

// Yamamoto accumulation into the r subgrid (correlator.h:477-483).
template <typename T>
__global__ void k_acc_r(double* __restrict__ rsub, const T* __restrict__ xi, RMap map,
                        const double* __restrict__ rx, int ell, int m, double c) {
#pragma clang fp contract(off)
  const int rs = map.rs;
  const int64_t n = (int64_t)rs * rs * rs;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(idx % rs);
    const int64_t t = idx / rs;
    const int j = (int)(t % rs), i = (int)(t / rs);
    const double y = ylm_eval(ell, m, rx[i], rx[j], rx[k], c, 1.0);
    rsub[idx] += y * (double)xi[map.src(i, j, k)];
  }
}

// ---- host orchestration -----------------------------------------------------
// Reads back the events of finished stages into the per-stage totals.
void flush_stage_times(fc_corr* c) {
  if (c->pending.empty()) return;
  FC_HIP(hipEventSynchronize(c->ev_pool[c->pending.back()[2]]));
  for (auto& p : c->pending) {
    float t = 0;
    FC_HIP(hipEventElapsedTime(&t, c->ev_pool[p[1]], c->ev_pool[p[2]]));
    c->stage_ms[p[0]] += t;
    c->stage_n[p[0]] += 1;
  }
  c->pending.clear();
  c->ev_used = 0;
}

// Brackets one stage's launches with events when profiling is on.
struct Stage {
  fc_corr* c;
  int id;
  int e0 = -1;
  Stage(fc_corr* c_, int id_) : c(c_), id(id_) {
    if (!c->profile) return;
    if (c->ev_used + 2 > c->ev_pool.size()) {
      for (int i = 0; i < 64; ++i) {
        hipEvent_t e;
        FC_HIP(hipEventCreate(&e));
        c->ev_pool.push_back(e);
      }
    }
    e0 = (int)c->ev_used++;
    FC_HIP(hipEventRecord(c->ev_pool[e0], c->ctx->stream));
  }
  ~Stage() {
    if (e0 < 0) return;
    const int e1 = (int)c->ev_used++;
    if (hipEventRecord(c->ev_pool[e1], c->ctx->stream) == hipSuccess)
      c->pending.push_back({id, e0, e1});
  }
};

void upload(DeviceBuffer& d, const std::vector<double>& h) {
  d.allocate(sizeof(double) * (h.empty() ? 1 : h.size()));
  if (!h.empty())
    FC_HIP(hipMemcpy(d.as<void>(), h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
}

// cc/correlate/window_correct.h:14-60 (+ CIC codes 3, 4).
double window_sq(int kind, double k, double cell) {
  auto sinc = [](double x) { return x == 0 ? 1.0 : std::sin(x) / x; };
  switch (kind) {
    case 0:
      return 1.0;
    case 1:
      return std::pow(sinc(k * cell / 2.0), 6);
    case 2: {
      double s = std::sin(k * cell / 2.0);
      s *= s;
      return 1 - s + 2.0 / 15.0 * s * s;
    }
    case 3:
      return std::pow(sinc(k * cell / 2.0), 4);
    case 4: {
      double s = std::sin(k * cell / 2.0);
      s *= s;
      return 1 - 2.0 / 3.0 * s;
    }
  }
  fail(FC_ERR_INVALID, "unknown window correction " + std::to_string(kind));
  return 0;
}

void setup_geometry(fc_corr* c) {
  // HistogramList(maxell/2+1, 0, max, binsize), histogram_list.h:13-25.
  auto hist_bins = [](double maxval, double binsize, const char* what) {
    if (!(0.0 < maxval)) fail(FC_ERR_INVALID, "maxval must be greater than minval.");
    if (!(binsize > 0)) fail(FC_ERR_INVALID, "binsize must be positive.");
    const double nb = std::floor((maxval - 0.0) / binsize);
    if (!(nb >= 1)) fail(FC_ERR_INVALID, std::string("Shape must be nonnegative. (") + what +
                                             " has no bins)");
    if (nb > 1e7) fail(FC_ERR_INVALID, std::string("too many ") + what + " bins");
    return (int)nb;
  };
  c->nrbins = hist_bins(c->rmax, c->dr, "separation");
  c->nkbins = hist_bins(c->kmax, c->dk, "wavenumber");
  c->rbins.resize(c->nrbins);
  c->kbins.resize(c->nkbins);
  for (int i = 0; i < c->nrbins; ++i) c->rbins[i] = 0.0 + (i + 0.5) * c->dr;
  for (int i = 0; i < c->nkbins; ++i) c->kbins[i] = 0.0 + (i + 0.5) * c->dk;
  // r subgrid (correlator.h:109-140).
  c->rmax_cells = (int)std::ceil(c->rmax / c->cell);
  c->rshape = 2 * c->rmax_cells + 1;
  const double minsep = -c->cell * c->rmax_cells;
  c->h_rx = sequence(minsep, c->cell, c->rshape);
  // k subgrid (correlator.h:142-205).
  const double k_nyq = M_PI / c->cell;
  for (int i = 0; i < 3; ++i) {
    int ks = (int)(2 * std::ceil(c->kmax / (2.0 * k_nyq / c->shape[i])) + 1);
    if (ks > c->shape[i]) ks = 2 * (c->shape[i] / 2) + 1;
    c->kshape[i] = ks;
  }
  std::vector<double>* axes[3] = {&c->h_kx, &c->h_ky, &c->h_kz};
  for (int i = 0; i < 3; ++i)
    *axes[i] = sequence((-c->kshape[i] / 2) * 2.0 * k_nyq / c->shape[i],
                        2.0 * k_nyq / c->shape[i], c->kshape[i]);
  std::vector<double> w[3];
  for (int i = 0; i < 3; ++i) {
    w[i].resize(c->kshape[i]);
    for (int k = 0; k < c->kshape[i]; ++k) w[i][k] = window_sq(c->window, (*axes[i])[k], c->cell);
  }
  upload(c->rx, c->h_rx);
  upload(c->kx, c->h_kx);
  upload(c->ky, c->h_ky);
  upload(c->kz, c->h_kz);
  upload(c->wx, w[0]);
  upload(c->wy, w[1]);
  upload(c->wz, w[2]);
  c->r_counts.allocate(sizeof(int) * c->nell * c->nrbins);
  c->r_hist.allocate(sizeof(double) * c->nell * c->nrbins);
  c->k_counts.allocate(sizeof(int) * c->nell * c->nkbins);
  c->k_hist.allocate(sizeof(double) * c->nell * c->nkbins);
  c->zerolag.allocate(sizeof(double));
}

void reset_hists(fc_corr* c, hipStream_t s) {
  FC_HIP(hipMemsetAsync(c->r_counts.as<void>(), 0, sizeof(int) * c->nell * c->nrbins, s));
  FC_HIP(hipMemsetAsync(c->r_hist.as<void>(), 0, sizeof(double) * c->nell * c->nrbins, s));
  FC_HIP(hipMemsetAsync(c->k_counts.as<void>(), 0, sizeof(int) * c->nell * c->nkbins, s));
  FC_HIP(hipMemsetAsync(c->k_hist.as<void>(), 0, sizeof(double) * c->nell * c->nkbins, s));
  FC_HIP(hipMemsetAsync(c->zerolag.as<void>(), 0, sizeof(double), s));
}

// Orders the correlator's stream after all work queued on the grid's stream.
void join_streams(fc_corr* c, fc_grid* g) {
  if (g->ctx == c->ctx || g->ctx->stream == c->ctx->stream) return;
  hipEvent_t e;
  FC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  FC_HIP(hipEventRecord(e, g->ctx->stream));
  FC_HIP(hipStreamWaitEvent(c->ctx->stream, e, 0));
  FC_HIP(hipEventDestroy(e));
}

void check_grid_shape(fc_corr* c, fc_grid* g) {
  check_grid(g);
  for (int i = 0; i < 3; ++i)
    if (g->shape[i] != c->shape[i]) fail(FC_ERR_INVALID, "Incompatible shapes for copy.");
  if (g->ctx->device != c->ctx->device)
    fail(FC_ERR_INVALID, "grid and correlator live on different devices");
}

template <typename T>
cplx<T>* work_of(fc_corr* c) {
  return c->work.as<cplx<T>>();
}

// Forward 3D FFT of a device grid into dst (complex half layout).
template <typename T, class Loader>
void forward_fft(fc_corr* c, fc_grid* g, cplx<T>* dst, Loader ld) {
  const FftPlan& p = *c->plan;
  hipStream_t s = c->ctx->stream;
  const int64_t ldi = p.n[2];
  const int64_t n2h = p.n2h;
  {
    Stage st(c, FC_STAGE_FFT_Z_FWD);
    if (g->prec == FC_FP64)
      launch_rows_r2c<T, double>(p, g->data.as<double>(), ldi, dst, n2h, ld, s);
    else
      launch_rows_r2c<T, float>(p, g->data.as<float>(), ldi, dst, n2h, ld, s);
  }
  {
    Stage st(c, FC_STAGE_FFT_Y_FWD);
    launch_cols<T, -1>(p.n[1], dst, (int64_t)p.n[1] * n2h, p.n[0], n2h, (int)n2h, p.tile[1],
                       p.tw1.as<cplx<T>>(), s);
  }
  {
    Stage st(c, FC_STAGE_FFT_X_FWD);
    launch_cols<T, -1>(p.n[0], dst, 0, 1, (int64_t)p.n[1] * n2h, (int)(p.n[1] * n2h),
                       p.tile[0], p.tw0.as<cplx<T>>(), s);
  }
}

// Inverse 3D FFT in place: complex half layout -> padded real rows.
template <typename T>
void inverse_fft(fc_corr* c, cplx<T>* data) {
  const FftPlan& p = *c->plan;
  hipStream_t s = c->ctx->stream;
  const int64_t n2h = p.n2h;
  {
    Stage st(c, FC_STAGE_FFT_X_INV);
    launch_cols<T, +1>(p.n[0], data, 0, 1, (int64_t)p.n[1] * n2h, (int)(p.n[1] * n2h),
                       p.tile[0], p.tw0.as<cplx<T>>(), s);
  }
  {
    Stage st(c, FC_STAGE_FFT_Y_INV);
    launch_cols<T, +1>(p.n[1], data, (int64_t)p.n[1] * n2h, p.n[0], n2h, (int)n2h, p.tile[1],
                       p.tw1.as<cplx<T>>(), s);
  }
  {
    Stage st(c, FC_STAGE_FFT_Z_INV);
    launch_rows_c2r<T, T>(p, data, n2h, reinterpret_cast<T*>(data), 2 * n2h, s);
  }
}

template <typename T, int NV>
void periodic_hists(fc_corr* c, const cplx<T>* P) {
  hipStream_t s = c->ctx->stream;
  const FftPlan& p = *c->plan;
  const uint64_t ncells = (uint64_t)c->shape[0] * c->shape[1] * c->shape[2];
  // correlator.h:294 and :310-313: k_rescale = cell^3 / N.
  const double k_rescale = c->cell * c->cell * c->cell / ncells;
  PeriodicKVals<T, NV> kv;
  kv.P = P;
  kv.map = KMap{p.n[0], p.n[1], p.n2h, c->kshape[0], c->kshape[1], c->kshape[2]};
  kv.kx = c->kx.as<double>();
  kv.ky = c->ky.as<double>();
  kv.kz = c->kz.as<double>();
  kv.wx = c->wx.as<double>();
  kv.wy = c->wy.as<double>();
  kv.wz = c->wz.as<double>();
  for (int q = 0; q < NV; ++q)
    kv.c[q] = ylm_coeff(2 * q, 0, std::sqrt((4.0 * M_PI) * (2 * (2 * q) + 1)) * k_rescale);
  kv.dk = c->dk;
  kv.nbins = c->nkbins;
  const int64_t nk = (int64_t)c->kshape[0] * c->kshape[1] * c->kshape[2];
  {
    Stage st(c, FC_STAGE_HIST_K);
    run_hist<NV>(kv, nk, c->nkbins, c->hs, c->k_hist.as<double>(), c->k_counts.as<int>(), 0,
                 0, NV, s);
  }
  inverse_fft<T>(c, const_cast<cplx<T>*>(P));
  // correlator.h:336: r_rescale = 1 / N / N.
  const double r_rescale = (1.0 / ncells / ncells);
  PeriodicRVals<T, NV> rv;
  rv.xi = reinterpret_cast<const T*>(P);
  rv.map = RMap{p.n[0], p.n[1], p.n[2], 2 * p.n2h, c->rshape};
  rv.rx = c->rx.as<double>();
  for (int q = 0; q < NV; ++q)
    rv.c[q] = ylm_coeff(2 * q, 0, std::sqrt((4.0 * M_PI) * (2 * (2 * q) + 1)) * r_rescale);
  rv.dr = c->dr;
  rv.nbins = c->nrbins;
  rv.zerolag = c->zerolag.as<double>();
  const int64_t nr = (int64_t)c->rshape * c->rshape * c->rshape;
  Stage st(c, FC_STAGE_HIST_R);
  run_hist<NV>(rv, nr, c->nrbins, c->hs, c->r_hist.as<double>(), c->r_counts.as<int>(), 0, 0,
               NV, s);
}

template <typename T>
void periodic_from_power(fc_corr* c, const cplx<T>* P) {
  switch (c->nell) {
    case 1:
      periodic_hists<T, 1>(c, P);
      break;
    case 2:
      periodic_hists<T, 2>(c, P);
      break;
    default:
      periodic_hists<T, 3>(c, P);
  }
}

template <typename T>
void periodic_correlate(fc_corr* c, fc_grid* g1, bool cross, fc_grid* g_interlaced) {
  hipStream_t s = c->ctx->stream;
  const FftPlan& p = *c->plan;
  const int64_t nc = p.nrows() * p.n2h;
  cplx<T>* w = work_of<T>(c);
  forward_fft<T>(c, g1, w, LoadCopy{});
  if (g_interlaced) {
    c->work_b.reserve(p.nrows() * p.n2h * sizeof(cplx<T>));
    forward_fft<T>(c, g_interlaced, c->work_b.as<cplx<T>>(), LoadCopy{});
    Stage st(c, FC_STAGE_COMBINE);
    hipLaunchKernelGGL(k_interlace<T>, dim3(grid1d(nc, 256, 65536)), dim3(256), 0, s, w,
                       c->work_b.as<cplx<T>>(), p.n[0], p.n[1], p.n[2], p.n2h);
    FC_LAUNCH_CHECK();
  }
  {
    Stage st(c, FC_STAGE_COMBINE);
    hipLaunchKernelGGL(k_conj_mult<T>, dim3(grid1d(nc, 256, 65536)), dim3(256), 0, s, w,
                       cross ? c->grid2_fft.as<cplx<T>>() : nullptr, nc);
    FC_LAUNCH_CHECK();
  }
  periodic_from_power<T>(c, w);
}

template <typename T>
void yamamoto_correlate(fc_corr* c, fc_grid* g1) {
  hipStream_t s = c->ctx->stream;
  const FftPlan& p = *c->plan;
  const int64_t nc = p.nrows() * p.n2h;
  cplx<T>* w = work_of<T>(c);
  const uint64_t ncells = (uint64_t)c->shape[0] * c->shape[1] * c->shape[2];
  // Observer at the survey origin; cell centres 0.5 - observer + i
  // (correlator.h:381-402).
  double start[3];
  for (int i = 0; i < 3; ++i) start[i] = 0.5 - (-c->posmin[i] / c->cell);
  const int64_t nk = (int64_t)c->kshape[0] * c->kshape[1] * c->kshape[2];
  const int64_t nr = (int64_t)c->rshape * c->rshape * c->rshape;
  c->rsub.reserve(sizeof(double) * nr);
  c->ksub.reserve(sizeof(double) * nk);
  KMap km{p.n[0], p.n[1], p.n2h, c->kshape[0], c->kshape[1], c->kshape[2]};
  RMap rm{p.n[0], p.n[1], p.n[2], 2 * p.n2h, c->rshape};
  for (int ell = 0; ell <= c->maxell; ell += 2) {
    FC_HIP(hipMemsetAsync(c->rsub.as<void>(), 0, sizeof(double) * nr, s));
    FC_HIP(hipMemsetAsync(c->ksub.as<void>(), 0, sizeof(double) * nk, s));
    for (int m = -ell; m <= ell; ++m) {
      LoadYlm ld{ell, m, p.n[1], ylm_coeff(ell, m, 1.0), start[0], start[1], start[2]};
      forward_fft<T>(c, g1, w, ld);
      {
        Stage st(c, FC_STAGE_COMBINE);
        hipLaunchKernelGGL(k_conj_mult<T>, dim3(grid1d(nc, 256, 65536)), dim3(256), 0, s, w,
                           c->grid2_fft.as<cplx<T>>(), nc);
        FC_LAUNCH_CHECK();
      }
      {
        Stage st(c, FC_STAGE_SUBGRID);
        hipLaunchKernelGGL(k_acc_k<T>, dim3(grid1d(nk, 256, 65536)), dim3(256), 0, s,
                           c->ksub.as<double>(), w, km, c->kx.as<double>(), c->ky.as<double>(),
                           c->kz.as<double>(), c->wx.as<double>(), c->wy.as<double>(),
                           c->wz.as<double>(), ell, m, ylm_coeff(ell, m, 4.0 * M_PI));
        FC_LAUNCH_CHECK();
      }
      inverse_fft<T>(c, w);
      {
        Stage st(c, FC_STAGE_SUBGRID);
        hipLaunchKernelGGL(k_acc_r<T>, dim3(grid1d(nr, 256, 65536)), dim3(256), 0, s,
                           c->rsub.as<double>(), reinterpret_cast<const T*>(w), rm,
                           c->rx.as<double>(), ell, m, ylm_coeff(ell, m, 4.0 * M_PI / ncells));
        FC_LAUNCH_CHECK();
      }
    }
    SubgridVals rv{c->rsub.as<double>(), c->rx.as<double>(), c->rx.as<double>(),
                   c->rx.as<double>(), c->rshape, c->rshape, c->dr, c->nrbins};
    SubgridVals kv{c->ksub.as<double>(), c->kx.as<double>(), c->ky.as<double>(),
                   c->kz.as<double>(), c->kshape[1], c->kshape[2], c->dk, c->nkbins};
    {
      Stage st(c, FC_STAGE_HIST_R);
      run_hist<1>(rv, nr, c->nrbins, c->hs, c->r_hist.as<double>(), c->r_counts.as<int>(),
                  ell / 2, ell / 2, ell / 2 + 1, s);
    }
    {
      Stage st(c, FC_STAGE_HIST_K);
      run_hist<1>(kv, nk, c->nkbins, c->hs, c->k_hist.as<double>(), c->k_counts.as<int>(),
                  ell / 2, ell / 2, ell / 2 + 1, s);
    }
    if (ell == 0) {
      const int o = c->rshape / 2;
      const int64_t center = ((int64_t)o * c->rshape + o) * c->rshape + o;
      FC_HIP(hipMemcpyAsync(c->zerolag.as<double>(), c->rsub.as<double>() + center,
                            sizeof(double), hipMemcpyDeviceToDevice, s));
    }
  }
}

void ensure_work(fc_corr* c) {
  c->work.reserve(c->plan->complex_bytes());
}

void set_grid2_device(fc_corr* c, fc_grid* g2) {
  ensure_work(c);
  c->grid2_fft.reserve(c->plan->complex_bytes());
  if (c->prec == FC_FP64)
    forward_fft<double>(c, g2, c->grid2_fft.as<cplx<double>>(), LoadCopy{});
  else
    forward_fft<float>(c, g2, c->grid2_fft.as<cplx<float>>(), LoadCopy{});
  c->have_grid2 = true;
}

void correlate(fc_corr* c, fc_grid* g1, fc_grid* gb, int mode /*0 auto,1 cross,2 interlaced*/) {
  DeviceGuard dg(c->ctx->device);
  hipStream_t s = c->ctx->stream;
  check_grid_shape(c, g1);
  join_streams(c, g1);
  if (gb) {
    check_grid_shape(c, gb);
    join_streams(c, gb);
  }
  ensure_work(c);
  flush_stage_times(c);
  FC_HIP(hipEventRecord(c->ev0, s));
  reset_hists(c, s);
  if (c->kind == FC_YAMAMOTO) {
    if (mode == 0) set_grid2_device(c, g1);  // Correlator::autocorrelate = cross(grid, grid)
    if (!c->have_grid2) fail(FC_ERR_RUNTIME, "grid2 FFT not set: call set_grid2 first");
    if (mode == 2) fail(FC_ERR_INVALID, "interlacing applies to the periodic correlator");
    if (c->prec == FC_FP64)
      yamamoto_correlate<double>(c, g1);
    else
      yamamoto_correlate<float>(c, g1);
  } else {
    if (mode == 1 && !c->have_grid2) fail(FC_ERR_RUNTIME, "grid2 FFT not set: call set_grid2 first");
    if (c->prec == FC_FP64)
      periodic_correlate<double>(c, g1, mode == 1, mode == 2 ? gb : nullptr);
    else
      periodic_correlate<float>(c, g1, mode == 1, mode == 2 ? gb : nullptr);
  }
  FC_HIP(hipEventRecord(c->ev1, s));
  c->timed = true;
}

}  // namespace
}  // namespace fc

extern "C" {

int fc_corr_create(fc_ctx* ctx, int kind, const int shape[3], double cell_size,
                   const double posmin[3], int window_correct, double rmax, double dr,
                   double kmax, double dk, int maxell, int precision, fc_corr** out) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!shape || !out) fc::fail(FC_ERR_INVALID, "null argument");
    if (kind != FC_PERIODIC && kind != FC_YAMAMOTO) fc::fail(FC_ERR_INVALID, "unknown kind");
    for (int i = 0; i < 3; ++i)
      if (shape[i] <= 0) fc::fail(FC_ERR_INVALID, "Shape must be nonnegative.");
    if (!(cell_size > 0)) fc::fail(FC_ERR_INVALID, "cell_size must be positive");
    if (window_correct < 0 || window_correct > 4)
      fc::fail(FC_ERR_INVALID, "unknown window correction " + std::to_string(window_correct));
    if (maxell < 0 || maxell > 4)
      fc::fail(FC_ERR_INVALID, "Unsupported ell: maxell must be 0, 2 or 4");
    if (precision != FC_FP64 && precision != FC_FP32)
      fc::fail(FC_ERR_INVALID, "unknown precision");
    fc::DeviceGuard dg(ctx->device);
    std::unique_ptr<fc_corr> c(new fc_corr());
    c->ctx = ctx;
    c->kind = kind;
    for (int i = 0; i < 3; ++i) {
      c->shape[i] = shape[i];
      c->posmin[i] = posmin ? posmin[i] : 0.0;
    }
    c->cell = cell_size;
    c->window = window_correct;
    c->rmax = rmax;
    c->dr = dr;
    c->kmax = kmax;
    c->dk = dk;
    c->maxell = maxell;
    c->nell = maxell / 2 + 1;
    c->prec = precision;
    fc::setup_geometry(c.get());
    c->plan = fc::get_fft_plan(ctx->device, shape, precision);
    FC_HIP(hipEventCreate(&c->ev0));
    FC_HIP(hipEventCreate(&c->ev1));
    *out = c.release();
  });
}

int fc_corr_destroy(fc_corr* corr) {
  return fc::guard([&] {
    if (!corr) return;
    fc::DeviceGuard dg(corr->ctx->device);
    (void)hipStreamSynchronize(corr->ctx->stream);
    for (auto e : corr->ev_pool) (void)hipEventDestroy(e);
    if (corr->ev0) (void)hipEventDestroy(corr->ev0);
    if (corr->ev1) (void)hipEventDestroy(corr->ev1);
    delete corr;
  });
}

int fc_corr_geometry(fc_corr* c, int* n_rbins, int* n_kbins, int* n_ell, int rshape[3],
                     int kshape[3]) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    if (n_rbins) *n_rbins = c->nrbins;
    if (n_kbins) *n_kbins = c->nkbins;
    if (n_ell) *n_ell = c->nell;
    for (int i = 0; i < 3; ++i) {
      if (rshape) rshape[i] = c->rshape;
      if (kshape) kshape[i] = c->kshape[i];
    }
  });
}

int fc_corr_set_grid2(fc_corr* c, fc_grid* grid2) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    fc::DeviceGuard dg(c->ctx->device);
    fc::check_grid_shape(c, grid2);
    fc::join_streams(c, grid2);
    fc::set_grid2_device(c, grid2);
  });
}

int fc_corr_set_grid2_fft(fc_corr* c, const double* host_complex) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    if (!host_complex) fc::fail(FC_ERR_INVALID, "null array");
    fc::DeviceGuard dg(c->ctx->device);
    hipStream_t s = c->ctx->stream;
    const int64_t n = c->plan->nrows() * c->plan->n2h;
    c->grid2_fft.reserve(c->plan->complex_bytes());
    fc::DeviceBuffer tmp(sizeof(double2) * n);
    FC_HIP(hipMemcpyAsync(tmp.as<void>(), host_complex, sizeof(double2) * n,
                          hipMemcpyHostToDevice, s));
    if (c->prec == FC_FP64)
      FC_HIP(hipMemcpyAsync(c->grid2_fft.as<void>(), tmp.as<void>(), sizeof(double2) * n,
                            hipMemcpyDeviceToDevice, s));
    else
      hipLaunchKernelGGL(fc::k_from_c128<float>, dim3(fc::grid1d(n, 256, 65536)), dim3(256), 0,
                         s, c->grid2_fft.as<float2>(), tmp.as<double2>(), n);
    FC_HIP(hipStreamSynchronize(s));
    c->have_grid2 = true;
  });
}

int fc_corr_autocorrelate(fc_corr* c, fc_grid* grid) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    fc::correlate(c, grid, nullptr, 0);
  });
}

int fc_corr_cross_correlate(fc_corr* c, fc_grid* grid1) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    fc::correlate(c, grid1, nullptr, 1);
  });
}

int fc_corr_autocorrelate_interlaced(fc_corr* c, fc_grid* a, fc_grid* b) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    if (!b) fc::fail(FC_ERR_INVALID, "null interlaced grid");
    fc::correlate(c, a, b, 2);
  });
}

int fc_corr_results(fc_corr* c, double* r_mid, int* r_counts, double* r_hist, double* k_mid,
                    int* k_counts, double* k_hist, double* zerolag) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    fc::DeviceGuard dg(c->ctx->device);
    hipStream_t s = c->ctx->stream;
    const size_t nr = (size_t)c->nell * c->nrbins, nk = (size_t)c->nell * c->nkbins;
    if (r_counts)
      FC_HIP(hipMemcpyAsync(r_counts, c->r_counts.as<void>(), sizeof(int) * nr,
                            hipMemcpyDeviceToHost, s));
    if (r_hist)
      FC_HIP(hipMemcpyAsync(r_hist, c->r_hist.as<void>(), sizeof(double) * nr,
                            hipMemcpyDeviceToHost, s));
    if (k_counts)
      FC_HIP(hipMemcpyAsync(k_counts, c->k_counts.as<void>(), sizeof(int) * nk,
                            hipMemcpyDeviceToHost, s));
    if (k_hist)
      FC_HIP(hipMemcpyAsync(k_hist, c->k_hist.as<void>(), sizeof(double) * nk,
                            hipMemcpyDeviceToHost, s));
    if (zerolag)
      FC_HIP(hipMemcpyAsync(zerolag, c->zerolag.as<void>(), sizeof(double),
                            hipMemcpyDeviceToHost, s));
    FC_HIP(hipStreamSynchronize(s));
    if (r_mid) std::copy(c->rbins.begin(), c->rbins.end(), r_mid);
    if (k_mid) std::copy(c->kbins.begin(), c->kbins.end(), k_mid);
  });
}

int fc_corr_set_profiling(fc_corr* c, int on) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    fc::DeviceGuard dg(c->ctx->device);
    fc::flush_stage_times(c);
    c->profile = on != 0;
    for (int i = 0; i < FC_NSTAGES; ++i) {
      c->stage_ms[i] = 0;
      c->stage_n[i] = 0;
    }
  });
}

int fc_corr_stage_times(fc_corr* c, double* ms, int64_t* launches) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    fc::DeviceGuard dg(c->ctx->device);
    fc::flush_stage_times(c);
    for (int i = 0; i < FC_NSTAGES; ++i) {
      if (ms) ms[i] = c->stage_ms[i];
      if (launches) launches[i] = c->stage_n[i];
    }
  });
}

int fc_corr_last_time_ms(fc_corr* c, double* ms) {
  return fc::guard([&] {
    if (!c || !ms) fc::fail(FC_ERR_INVALID, "null argument");
    if (!c->timed) fc::fail(FC_ERR_RUNTIME, "no correlation has run yet");
    fc::DeviceGuard dg(c->ctx->device);
    FC_HIP(hipEventSynchronize(c->ev1));
    float t = 0;
    FC_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t;
  });
}

}  // extern "C"
