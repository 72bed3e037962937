// Correlators (cc/correlate/correlator.h:18-505) on the device.
//
// PeriodicCorrelator (line of sight z), per call:
//   z pass R2C of the grid (copy_into_padded_array fused into the load),
//   y pass; [the same for the half-cell-shifted grid when interlacing];
//   one fused x pass: forward x FFT, interlacing combine, |F|^2 (or
//     F conj(G)), the real part of every element the k subgrid reads
//     written to a compact array, inverse x FFT, and only the x lags the r
//     subgrid reads written back (pruned inverse);
//   k-space multipoles from the compact array (extract_fft_submatrix_c2r +
//     Y_l0(k) W^-1(k) + histogram, every l in one pass);
//   pruned inverse y pass, pruned z C2R straight into the r subgrid;
//   r-space multipoles (extract_submatrix + Y_l0(r) + histogram).
// Correlator (Yamamoto, observer at the survey origin), per (l, m):
//   z pass with Y_lm(x) * delta fused into the load -> y -> the same fused x
//   pass with F conj(G2) against the stored FFT of grid 2 -> pruned inverse
//   y, z -> accumulate the k and r subgrids with Y_lm(k), Y_lm(r);
//   histograms once per l.
// Work arrays are FP64 (FC_FP64) or FP32 (FC_FP32); Y_lm, windows, bin
// indices and every sum stay in FP64.
#include <algorithm>
#include <array>
#include <thread>
#include <type_traits>

#include "fc_objects.h"
#include "fft_plan.h"
#include "fused_x.h"
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
  fc::BinnedCells kbin, rbin;                   // subgrid cells sorted by bin
  // k histograms over the compact values (KEntryHalf, KEntryHalfLm): their
  // indices sorted by bin, when the geometry allows it (khalf_ok).
  fc::BinnedCells khalf;
  bool khalf_ok = false;
  fc::DeviceBuffer rw;                          // periodic: r weights per entry
  double kcoef[3] = {0, 0, 0};                  // periodic k-space Y_l0 coefficients
  // FFT plan + work arrays (complex rows of plan->ldc).
  std::shared_ptr<fc::FftPlan> plan;      // forward precision
  std::shared_ptr<fc::FftPlan> plan_inv;  // inverse precision (FP64 plan for FC_MIXED)
  int algo = FC_ALGO_FUSED;
  fc::DeviceBuffer work, grid2_fft, work_b;
  bool have_grid2 = false;
  // Periodic: compact k-subgrid values, pruned inverse stages, r subgrid.
  fc::DeviceBuffer pk, q1, q2, xisub;
  fc::DeviceBuffer rsub, ksub;  // FP64 subgrids (Yamamoto accumulation)
  fc::DeviceBuffer ph[3];       // interlacing phase per axis
  // Outputs.
  fc::DeviceBuffer r_counts, r_hist, k_counts, k_hist, zerolag;
  fc::HistScratch hs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  // Periodic correlations run the k histogram on a second stream (default
  // priority) beside the pruned inverse; created on first use.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  fc::DeviceBuffer part_k;  // the k histogram's partials (the r histogram's are hs.part_seg)
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
// Complex rows of ldc (n2h used): out *= conj(in)
// (array_ops::multiply_with_conjugation, array_ops.cc:139-163);
// in == nullptr: out = |out|^2.
template <typename T>
__global__ void k_conj_mult(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ in,
                            int64_t n) {
#pragma clang fp contract(off)
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
// signed frequency indices in the numpy.fft.fftfreq convention;
// ph0/ph1/ph2[f] = exp(i pi fsigned / n) per axis (host tables, FP64).
template <typename T>
__global__ void k_interlace(cplx<T>* __restrict__ fa, const cplx<T>* __restrict__ fb, int n0,
                            int n1, int64_t ldc, const double2* __restrict__ ph0,
                            const double2* __restrict__ ph1, const double2* __restrict__ ph2) {
  const int64_t n = (int64_t)n0 * n1 * ldc;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % ldc);
    const int64_t r = i / ldc;
    const int b = (int)(r % n1), a = (int)(r / n1);
    const double2 p = cmul(cmul(ph0[a], ph1[b]), ph2[c]);
    const cplx<T> A = fa[i], B = fb[i];
    const double bx = (double)B.x * p.x - (double)B.y * p.y;
    const double by = (double)B.x * p.y + (double)B.y * p.x;
    fa[i] = cplx<T>{T(0.5 * ((double)A.x + bx)), T(0.5 * ((double)A.y + by))};
  }
}

// Host complex128 [rows][n2h] -> device rows of ldc.
template <typename Tdst>
__global__ void k_from_c128(cplx<Tdst>* __restrict__ d, const double2* __restrict__ s,
                            int64_t rows, int n2h, int64_t ldc) {
  const int64_t n = rows * n2h;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / n2h;
    const int c = (int)(i - r * n2h);
    d[r * ldc + c] = cplx<Tdst>{(Tdst)s[i].x, (Tdst)s[i].y};
  }
}

// Fallback (non-fast x length): gather the compact array from the full grid.
template <typename T>
__global__ void k_gather_pk(T* __restrict__ pk, const cplx<T>* __restrict__ P, KSlots ks, int n0,
                            int n1, int64_t ldc) {
  const int oz = ks.ks2 / 2;
  const int64_t n = (int64_t)ks.ks0 * ks.ks1 * (oz + 1);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % (oz + 1));
    const int64_t t = i / (oz + 1);
    const int sy = (int)(t % ks.ks1), sx = (int)(t / ks.ks1);
    const int e = ((sx - ks.ks0 / 2) % n0 + n0) % n0;
    const int y = ((sy - ks.ks1 / 2) % n1 + n1) % n1;
    pk[i] = P[((int64_t)e * n1 + y) * ldc + c].x;
  }
}

// ---- fused periodic x pass --------------------------------------------------
// One workgroup per (y row, tile of tc kz columns); block = tc * (L / E).
// A: grid a after its z and y passes; B: grid b likewise (interlaced) or the
// fully transformed grid 2 (cross).  Writes pk and the pruned inverse-x
// output q[sx][y][c] for x lags sx - rs/2, sx < rs.
template <typename T, int L, int MODE>
__global__ void __launch_bounds__(1024)
    k_periodic_x(const cplx<T>* __restrict__ A, const cplx<T>* __restrict__ B, int n1, int n2h,
                 int64_t ldc, int tilesC, int tc, const cplx<T>* __restrict__ tw,
                 const double2* __restrict__ ph0, const double2* __restrict__ ph1,
                 const double2* __restrict__ ph2, T* __restrict__ pk, KSlots ks,
                 cplx<T>* __restrict__ q, int rs) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  constexpr int E = FftCfg<L>::E;
  constexpr int TCOL = L / E;
  const int y = blockIdx.x / tilesC;
  const int c0 = (blockIdx.x - y * tilesC) * tc;
  const int ncol = min(tc, n2h - c0);
  const int64_t sL = (int64_t)n1 * ldc;
  const int64_t base = (int64_t)y * ldc + c0;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int c = tid % tc, t = tid / tc;
  C rb[E];
  // L * tc elements over tc * (L / E) threads: E loads per thread, issued
  // before the LDS stores.
  auto load_tile = [&](const C* __restrict__ src) {
    C tmp[E];
#pragma unroll
    for (int qq = 0; qq < E; ++qq) {
      const int idx = tid + qq * nthr;
      const int e = idx / tc, cc = idx - e * tc;
      tmp[qq] = cc < ncol ? src[base + e * sL + cc] : C{T(0), T(0)};
    }
#pragma unroll
    for (int qq = 0; qq < E; ++qq) lds[tid + qq * nthr] = tmp[qq];
  };
  if (MODE == kModeInterlaced) {
    load_tile(B);
    __syncthreads();
    fft_lds<T, L, -1>(lds, tw, c, t, LayColFast{tc});
#pragma unroll
    for (int qq = 0; qq < E; ++qq) rb[qq] = lds[(t + qq * TCOL) * tc + c];
    __syncthreads();
  }
  load_tile(A);
  __syncthreads();
  fft_lds<T, L, -1>(lds, tw, c, t, LayColFast{tc});
  {
#pragma clang fp contract(off)
    const int cg = c0 + c;
#pragma unroll
    for (int qq = 0; qq < E; ++qq) {
      const int e = t + qq * TCOL;
      const C fa = lds[e * tc + c];
      C p;
      if (MODE == kModeInterlaced) {
        const double2 ph = cmul(cmul(ph0[e], ph1[y]), ph2[c < ncol ? cg : 0]);
        const C fb = rb[qq];
        const double fx = 0.5 * ((double)fa.x + ((double)fb.x * ph.x - (double)fb.y * ph.y));
        const double fy = 0.5 * ((double)fa.y + ((double)fb.x * ph.y + (double)fb.y * ph.x));
        p = C{T(fx * fx + fy * fy), T(0)};
      } else if (MODE == kModeCross) {
        const C g = c < ncol ? B[base + e * sL + c] : C{T(0), T(0)};
        p = C{fa.x * g.x + fa.y * g.y, fa.y * g.x - fa.x * g.y};
      } else {
        p = C{fa.x * fa.x + fa.y * fa.y, T(0)};
      }
      lds[e * tc + c] = p;
      if (c < ncol) put_pk<T>(pk, ks, L, n1, e, y, cg, p.x);
    }
  }
  __syncthreads();
  fft_lds<T, L, +1>(lds, tw, c, t, LayColFast{tc});
  const int o = rs / 2;
  for (int idx = tid; idx < rs * tc; idx += nthr) {
    const int sx = idx / tc, cc = idx - sx * tc;
    const int e = ((sx - o) % L + L) % L;
    if (cc < ncol) q[((int64_t)sx * n1 + y) * ldc + c0 + cc] = lds[e * tc + cc];
  }
}

// ---- z-pass loader with Y_lm(x) weighting (correlator.h:435-436) -----------
struct LoadYlm {
  int ell, m, n1;
  double c, x0, y0, z0;  // cell centres relative to the observer: x0 + i etc.
  int i0 = 0;            // x slab of row 0
  void shift_x(int dx) { i0 += dx; }
  struct Row {
    int ell, m;
    double c, x, y, z0;
    __device__ __forceinline__ double operator()(double v, int k) const {
      return ylm_eval(ell, m, x, y, z0 + (double)k * 1.0, c, v);
    }
  };
  __device__ __forceinline__ Row row(int64_t r) const {
    const int i = i0 + (int)(r / n1), j = (int)(r % n1);
    return Row{ell, m, c, x0 + (double)i * 1.0, y0 + (double)j * 1.0, z0};
  }
};

// Histogram entries.  A histogram sample is weight * value with the weight
// Y_l0 (x) sqrt(4 pi (2l+1)) times the rescaling (and W^-1(k) in k space),
// formed by the same FP64 operations, in the same order, as the reference's
// make_ylm followed by extraction (correlator.h:296-320, :336-355).
//
// Periodic k space: the value is Re P at the cell's slot of the compact
// array (the reflected slot for kz < 0, fft_grid.cc:238-245) and the weight
// is evaluated on the fly from the subgrid axes.
template <typename T, int NV>
struct KEntry {
  const uint32_t* cells;
  const T* pk;
  KSlots ks;
  FastDiv d2, d1;  // division by ks2 and ks1
  const double *kx, *ky, *kz, *wx, *wy, *wz;
  double coef[NV];
  int64_t bstride = 0;  // pk elements per batch member
  __device__ __forceinline__ void offset(int m) { pk += m * bstride; }
  __device__ __forceinline__ void operator()(int64_t pos, double* v) const {
#pragma clang fp contract(off)
    const uint32_t cell = cells[pos];
    const uint32_t t = d2.div(cell);
    const int k = (int)(cell - t * (uint32_t)ks.ks2);
    const uint32_t iu = d1.div(t);
    const int j = (int)(t - iu * (uint32_t)ks.ks1), i = (int)iu;
    const int oz = ks.oz();
    const int64_t src = k >= oz ? ks.at(i, j, k - oz) : ks.at(ks.ks0 - 1 - i, ks.ks1 - 1 - j, oz - k);
    const double x = (double)pk[src];
    const double invw = 1.0 / (wx[i] * wy[j] * wz[k]);
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = ylm_eval(2 * q, 0, kx[i], ky[j], kz[k], coef[q], invw) * x;
  }
};

// Yamamoto k space, one (l, m) transform: the value is Re P at the cell's
// compact slot (the reflected slot for kz < 0) and the weight
// Y_lm(k) coef W^-1(k), evaluated on the fly with the operations of the
// subgrid accumulation.  The histograms of the (l, m) transforms of one l add
// up to the histogram of the k subgrid summed over m (correlator.h:453-466),
// without the FP64 k subgrid in memory.
template <typename T>
struct KEntryLm {
  const uint32_t* cells;
  const T* pk;
  KSlots ks;
  FastDiv d2, d1;  // division by ks2 and ks1
  const double *kx, *ky, *kz, *wx, *wy, *wz;
  int ell, m;
  double coef;
  __device__ __forceinline__ void offset(int) {}
  __device__ __forceinline__ void operator()(int64_t pos, double* v) const {
#pragma clang fp contract(off)
    const uint32_t cell = cells[pos];
    const uint32_t t = d2.div(cell);
    const int k = (int)(cell - t * (uint32_t)ks.ks2);
    const uint32_t iu = d1.div(t);
    const int j = (int)(t - iu * (uint32_t)ks.ks1), i = (int)iu;
    const int oz = ks.oz();
    const int64_t src = k >= oz ? ks.at(i, j, k - oz) : ks.at(ks.ks0 - 1 - i, ks.ks1 - 1 - j, oz - k);
    const double x = (double)pk[src];
    const double invw = 1.0 / (wx[i] * wy[j] * wz[k]);
    v[0] = ylm_eval(ell, m, kx[i], ky[j], kz[k], coef, invw) * x;
  }
};

// Periodic k space over the half of the k subgrid that the compact values
// cover: entry e is the compact value pk[e] = pk[sx][sy][c] itself (the cell
// (sx, sy, oz + c)).  A value with c > 0 is also read by the mirror cell
// (ks0-1-sx, ks1-1-sy, oz-c); with odd subgrid axes the two have the same |k|
// up to rounding and the even multipoles weight them alike, so the value
// enters once with weight 2 (the host uses these entries only when every
// mirror pair shares its bin).  Half the entries of KEntry, each value
// fetched once.
template <typename T, int NV>
struct KEntryHalf {
  const uint32_t* cells;  // compact indices, sorted by bin
  const T* pk;
  FastDiv dz, d1;  // division by oz + 1 and ks1
  int nz, ks1;
  const double *kx, *ky, *kz, *wx, *wy, *wz;  // kz, wz from the kz = 0 plane on
  double coef[NV];
  int64_t bstride = 0;  // pk elements per batch member
  __device__ __forceinline__ void offset(int m) { pk += m * bstride; }
  __device__ __forceinline__ void operator()(int64_t pos, double* v) const {
#pragma clang fp contract(off)
    const uint32_t e = cells[pos];
    const uint32_t t = dz.div(e);
    const int c = (int)(e - t * (uint32_t)nz);
    const uint32_t iu = d1.div(t);
    const int j = (int)(t - iu * (uint32_t)ks1), i = (int)iu;
    const double x = (double)pk[e];
    const double invw = 1.0 / (wx[i] * wy[j] * wz[c]);
    const double mult = c > 0 ? 2.0 : 1.0;
#pragma unroll
    for (int q = 0; q < NV; ++q)
      v[q] = (ylm_eval(2 * q, 0, kx[i], ky[j], kz[c], coef[q], invw) * x) * mult;
  }
};

// Yamamoto k space over the half-subgrid entries (see KEntryHalf), one (l, m)
// transform: every real Y_lm of even l is an even polynomial in the unit
// vector, so a compact value and its mirror cell share the weight.
template <typename T>
struct KEntryHalfLm {
  const uint32_t* cells;
  const T* pk;
  FastDiv dz, d1;
  int nz, ks1;
  const double *kx, *ky, *kz, *wx, *wy, *wz;  // kz, wz from the kz = 0 plane on
  int ell, m;
  double coef;
  __device__ __forceinline__ void offset(int) {}
  __device__ __forceinline__ void operator()(int64_t pos, double* v) const {
#pragma clang fp contract(off)
    const uint32_t e = cells[pos];
    const uint32_t t = dz.div(e);
    const int c = (int)(e - t * (uint32_t)nz);
    const uint32_t iu = d1.div(t);
    const int j = (int)(t - iu * (uint32_t)ks1), i = (int)iu;
    const double x = (double)pk[e];
    const double invw = 1.0 / (wx[i] * wy[j] * wz[c]);
    v[0] = (ylm_eval(ell, m, kx[i], ky[j], kz[c], coef, invw) * x) * (c > 0 ? 2.0 : 1.0);
  }
};

// Precomputed weights wt[q][pos] (periodic r space).
template <typename V, int NV>
struct WEntry {
  const V* vals;
  const uint32_t* cells;
  const double* wt;
  int64_t n;
  int64_t bstride = 0;  // values per batch member
  __device__ __forceinline__ void offset(int m) { vals += m * bstride; }
  __device__ __forceinline__ void operator()(int64_t pos, double* v) const {
    const double x = (double)vals[cells[pos]];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wt[(int64_t)q * n + pos] * x;
  }
};

// Plain gather (subgrids that already hold weighted values).
template <typename V>
struct GEntry {
  const V* vals;
  const uint32_t* cells;
  int64_t bstride = 0;
  __device__ __forceinline__ void offset(int m) { vals += m * bstride; }
  __device__ __forceinline__ void operator()(int64_t pos, double* v) const {
    v[0] = (double)vals[cells[pos]];
  }
};

// r-space weights, one per sorted entry and multipole (built once).
template <int NV>
__global__ void k_setup_rweights(const uint32_t* __restrict__ cells, int64_t n, int rs,
                                 const double* __restrict__ rx, const double* __restrict__ coef,
                                 double* __restrict__ wt) {
#pragma clang fp contract(off)
  for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < n;
       pos += (int64_t)gridDim.x * blockDim.x) {
    const int64_t cell = cells[pos];
    const int k = (int)(cell % rs);
    const int64_t t = cell / rs;
    const int j = (int)(t % rs), i = (int)(t / rs);
#pragma unroll
    for (int q = 0; q < NV; ++q)
      wt[(int64_t)q * n + pos] = ylm_eval(2 * q, 0, rx[i], rx[j], rx[k], coef[q], 1.0);
  }
}

// Yamamoto accumulation into the k subgrid (correlator.h:453-466) from the
// compact array: ksub += Y_lm(k) 4 pi W^-1(k) Re P.
template <typename T>
__global__ void k_acc_k(double* __restrict__ ksub, const T* __restrict__ pk, KSlots ks,
                        const double* __restrict__ kx, const double* __restrict__ ky,
                        const double* __restrict__ kz, const double* __restrict__ wx,
                        const double* __restrict__ wy, const double* __restrict__ wz, int ell,
                        int m, double c) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)ks.ks0 * ks.ks1 * ks.ks2;
  const int oz = ks.oz();
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(idx % ks.ks2);
    const int64_t t = idx / ks.ks2;
    const int j = (int)(t % ks.ks1), i = (int)(t / ks.ks1);
    const int64_t src = k >= oz ? ks.at(i, j, k - oz)
                                : ks.at(ks.ks0 - 1 - i, ks.ks1 - 1 - j, oz - k);
    const double invw = 1.0 / (wx[i] * wy[j] * wz[k]);
    const double y = ylm_eval(ell, m, kx[i], ky[j], kz[k], c, invw);
    ksub[idx] += y * (double)pk[src];
  }
}

// Yamamoto accumulation into the r subgrid (correlator.h:477-483) from the
// pruned inverse: rsub += Y_lm(r) 4 pi / N xi.
template <typename T>
__global__ void k_acc_r(double* __restrict__ rsub, const T* __restrict__ xi, int rs,
                        const double* __restrict__ rx, int ell, int m, double c) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)rs * rs * rs;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(idx % rs);
    const int64_t t = idx / rs;
    const int j = (int)(t % rs), i = (int)(t / rs);
    const double y = ylm_eval(ell, m, rx[i], rx[j], rx[k], c, 1.0);
    rsub[idx] += y * (double)xi[idx];
  }
}

// r-subgrid extraction from the full real correlation grid (rows of ldr
// reals): rsub += Y_lm(r) c xi[(i - o) mod n0][(j - o) mod n1][(k - o) mod n2]
// with o = rs / 2 (FftGrid::extract_submatrix, fft_grid.cc:182-210).
template <typename T>
__global__ void k_acc_r_full(double* __restrict__ rsub, const T* __restrict__ xi, int n0, int n1,
                             int n2, int64_t ldr, int rs, const double* __restrict__ rx, int ell,
                             int m, double c) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)rs * rs * rs;
  const int o = rs / 2;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(idx % rs);
    const int64_t t = idx / rs;
    const int j = (int)(t % rs), i = (int)(t / rs);
    const int a = ((i - o) % n0 + n0) % n0;
    const int b = ((j - o) % n1 + n1) % n1;
    const int d = ((k - o) % n2 + n2) % n2;
    const double y = ylm_eval(ell, m, rx[i], rx[j], rx[k], c, 1.0);
    rsub[idx] += y * (double)xi[((int64_t)a * n1 + b) * ldr + d];
  }
}

// Fill kernel (used instead of hipMemsetAsync so call sequences stay
// capturable into HIP graphs).
template <typename V>
__global__ void k_fill(V* __restrict__ p, int64_t n, V v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename V>
void fill(V* p, int64_t n, V v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<V>, dim3(grid1d(n, 256, 4096)), dim3(256), 0, s, p, n, v);
  FC_LAUNCH_CHECK();
}

// Zeroes up to 5 arrays of 32-bit words in one launch (the per-correlation
// histogram outputs), block y = array.
struct ZeroSegs {
  uint32_t* p[5];
  int64_t n[5];
};
__global__ void k_zero_segs(ZeroSegs z) {
  uint32_t* p = z.p[blockIdx.y];
  const int64_t n = z.n[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
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

// k histograms over the compact values (KEntryHalf, KEntryHalfLm): allowed when
// the k subgrid axes are odd (every compact value with c > 0 has its mirror
// cell) and every mirror pair shares its bin (cb: the bin of every k cell).
// The entries are then the compact indices of the half subgrid kz >= 0,
// sorted by bin and in index order within a bin; the counts stay those of
// the whole subgrid.
void setup_khalf(fc_corr* c, const std::vector<int16_t>& cb) {
  c->khalf_ok = false;
  const int s0 = c->kshape[0], s1 = c->kshape[1], s2 = c->kshape[2];
  if (s0 % 2 == 0 || s1 % 2 == 0 || s2 % 2 == 0) return;
  const int oz = s2 / 2;
  const int64_t ncols = (int64_t)s0 * s1;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<char> ok(nt, 1);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int64_t col = (int64_t)t; col < ncols; col += nt) {
        const int a = (int)(col / s1), b = (int)(col % s1);
        const size_t row = ((size_t)a * s1 + b) * s2;
        const size_t mrow = ((size_t)(s0 - 1 - a) * s1 + (s1 - 1 - b)) * s2;
        for (int k = 1; k <= oz; ++k)
          if (cb[mrow + oz - k] != cb[row + oz + k]) ok[t] = 0;
      }
    });
  for (auto& x : th) x.join();
  for (char o : ok)
    if (!o) return;
  const std::vector<double> kz_half(c->h_kz.begin() + oz, c->h_kz.end());
  build_binned_cells(c->khalf, c->h_kx, c->h_ky, kz_half, c->dk, c->nkbins);
  c->khalf_ok = true;
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
  // k cells ordered within a bin by the compact k value they read (see
  // KEntry): the two Hermitian partners of a value become neighbours.
  const KSlots ks{c->kshape[0], c->kshape[1], c->kshape[2]};
  const std::function<uint64_t(uint32_t)> kkey = [ks](uint32_t cell) -> uint64_t {
    const int k = (int)(cell % (uint32_t)ks.ks2);
    const uint32_t t = cell / (uint32_t)ks.ks2;
    const int j = (int)(t % (uint32_t)ks.ks1), i = (int)(t / (uint32_t)ks.ks1);
    const int oz = ks.oz();
    return (uint64_t)(k >= oz ? ks.at(i, j, k - oz)
                              : ks.at(ks.ks0 - 1 - i, ks.ks1 - 1 - j, oz - k));
  };
  {
    std::vector<int16_t> cell_bins;
    const bool want = c->nkbins < 32768;
    build_binned_cells(c->kbin, c->h_kx, c->h_ky, c->h_kz, c->dk, c->nkbins, -1, &kkey,
                       want ? &cell_bins : nullptr);
    if (want) setup_khalf(c, cell_bins);
  }
  const int64_t o = c->rmax_cells;
  build_binned_cells(c->rbin, c->h_rx, c->h_rx, c->h_rx, c->dr, c->nrbins,
                     (o * c->rshape + o) * c->rshape + o);
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
  const int64_t nr = (int64_t)c->nell * c->nrbins, nk = (int64_t)c->nell * c->nkbins;
  ZeroSegs z{{c->r_counts.as<uint32_t>(), c->r_hist.as<uint32_t>(), c->k_counts.as<uint32_t>(),
              c->k_hist.as<uint32_t>(), c->zerolag.as<uint32_t>()},
             {nr, 2 * nr, nk, 2 * nk, 2}};
  const int64_t most = std::max<int64_t>(2 * std::max(nr, nk), 2);
  hipLaunchKernelGGL(k_zero_segs, dim3(grid1d(most, 256, 4096), 5), dim3(256), 0, s, z);
  FC_LAUNCH_CHECK();
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

// z and y forward passes of a device grid into dst (complex rows of ldc,
// stored in T; arithmetic in A, which is wider than T for the mixed form and
// then takes its twiddles from the FP64 plan plan_inv).  The grid is
// processed in chunks of x slabs (plan zy_chunk) so that each chunk's y pass
// reads the z pass's output from the last-level cache.
template <typename T, typename A = T, class Loader>
void forward_zy(fc_corr* c, fc_grid* g, cplx<T>* dst, Loader ld) {
  const FftPlan& p = *c->plan;
  const FftPlan& pa = *c->plan_inv;
  hipStream_t s = c->ctx->stream;
  const int X = p.zy_chunk > 0 ? p.zy_chunk : p.n[0];
  for (int x0 = 0; x0 < p.n[0]; x0 += X) {
    const int nx = std::min(X, p.n[0] - x0);
    const int64_t rows = (int64_t)nx * p.n[1];
    cplx<T>* d = dst + (int64_t)x0 * p.n[1] * p.ldc;
    Loader lc = ld;
    lc.shift_x(x0);
    {
      Stage st(c, FC_STAGE_FFT_Z_FWD);
      const int64_t in0 = (int64_t)x0 * p.n[1] * p.n[2];
      if (g->prec == FC_FP64)
        launch_rows_r2c<T, double, Loader, A>(p, g->data.as<double>() + in0, p.n[2], d, p.ldc,
                                              lc, s, rows, &pa);
      else
        launch_rows_r2c<T, float, Loader, A>(p, g->data.as<float>() + in0, p.n[2], d, p.ldc, lc,
                                             s, rows, &pa);
    }
    Stage st(c, FC_STAGE_FFT_Y_FWD);
    ColView<T> v = y_view<T>(p, d);
    v.A = nx;
    launch_cols_wide<T, -1, A>(p.n[1], v, p.tile[1], pa.tw1.as<cplx<A>>(), p.tw1.as<cplx<T>>(),
                               s, p.var[1]);
  }
}

// Interlaced pair: the z and y forward passes of grids a and b as one launch
// each, into dst (grid a's spectrum, then grid b's).  Needs the
// register-chained rows and no z+y chunking; taken for FP32 spectra (512^3:
// 1.630 -> 1.608 ms per mixed correlation), not for FP64 ones, whose four
// separate launches measured 0.5% faster.
bool forward_pair_plan_ok(const fc_corr* c) {
  const FftPlan& p = *c->plan;
  return c->prec != FC_FP64 && p.zy_chunk == 0 && p.fast[2] && rc_len_ok(p.n[2] / 2);
}
bool forward_pair_ok(const fc_corr* c, const fc_grid* ga, const fc_grid* gb) {
  return forward_pair_plan_ok(c) && ga->prec == gb->prec;
}

template <typename T, typename A = T>
void forward_zy_pair(fc_corr* c, fc_grid* ga, fc_grid* gb, cplx<T>* dst) {
  const FftPlan& p = *c->plan;
  const FftPlan& pa = *c->plan_inv;
  hipStream_t s = c->ctx->stream;
  const int64_t rows = p.nrows();
  {
    Stage st(c, FC_STAGE_FFT_Z_FWD);
    if (ga->prec == FC_FP64)
      launch_rows_r2c<T, double, LoadCopy, A>(p, ga->data.as<double>(), p.n[2], dst, p.ldc,
                                              LoadCopy{}, s, 2 * rows, &pa,
                                              gb->data.as<double>(), rows);
    else
      launch_rows_r2c<T, float, LoadCopy, A>(p, ga->data.as<float>(), p.n[2], dst, p.ldc,
                                             LoadCopy{}, s, 2 * rows, &pa, gb->data.as<float>(),
                                             rows);
  }
  Stage st(c, FC_STAGE_FFT_Y_FWD);
  ColView<T> v = y_view<T>(p, dst);
  v.A = 2 * p.n[0];
  launch_cols_wide<T, -1, A>(p.n[1], v, p.tile[1], pa.tw1.as<cplx<A>>(), p.tw1.as<cplx<T>>(), s,
                             p.var[1]);
}

template <typename T, typename A = T>
void forward_x(fc_corr* c, cplx<T>* dst) {
  const FftPlan& p = *c->plan;
  Stage st(c, FC_STAGE_FFT_X_FWD);
  launch_cols_wide<T, -1, A>(p.n[0], x_view<T>(p, dst), p.tile[0],
                             c->plan_inv->tw0.as<cplx<A>>(), p.tw0.as<cplx<T>>(),
                             c->ctx->stream, p.var[0]);
}

// Full forward 3D FFT.
template <typename T, typename A = T, class Loader>
void forward_fft(fc_corr* c, fc_grid* g, cplx<T>* dst, Loader ld) {
  forward_zy<T, A>(c, g, dst, ld);
  forward_x<T, A>(c, dst);
}

// exp(i pi f / n) for the signed frequency f of every index along each axis
// (numpy.fft.fftfreq convention), for interlacing.
void ensure_phase_tables(fc_corr* c) {
  if (c->ph[0].bytes()) return;
  const int len[3] = {c->shape[0], c->shape[1], c->plan->ldc};
  for (int d = 0; d < 3; ++d) {
    const int n = c->shape[d];
    std::vector<double2> h(len[d]);
    for (int a = 0; a < len[d]; ++a) {
      const int f = a < (n + 1) / 2 ? a : a - n;
      const long double ph = 3.141592653589793238462643383279502884L * f / n;
      h[a] = make_double2((double)cosl(ph), (double)sinl(ph));
    }
    c->ph[d].allocate(sizeof(double2) * len[d]);
    FC_HIP(hipMemcpy(c->ph[d].as<void>(), h.data(), sizeof(double2) * len[d],
                     hipMemcpyHostToDevice));
  }
}

KSlots kslots(const fc_corr* c) { return KSlots{c->kshape[0], c->kshape[1], c->kshape[2]}; }

// k histogram of nmem members' compact values (member m at pk + m * bstride)
// over the half-subgrid entries; hist / counts rows of member m at
// m * out_bstride.
template <typename T, int NV>
void run_hist_khalf(fc_corr* c, const T* pk, int64_t bstride, DeviceBuffer& part, double* hist,
                    int* counts, int row0, int crow0, int crow1, hipStream_t s, int nmem = 1,
                    int64_t out_bstride = 0) {
  const int nz = c->kshape[2] / 2 + 1;
  const int oz = c->kshape[2] / 2;
  KEntryHalf<T, NV> ke{c->khalf.cells.as<uint32_t>(), pk, FastDiv((uint32_t)nz),
                       FastDiv((uint32_t)c->kshape[1]), nz, c->kshape[1], c->kx.as<double>(),
                       c->ky.as<double>(), c->kz.as<double>() + oz, c->wx.as<double>(),
                       c->wy.as<double>(), c->wz.as<double>() + oz, {}, bstride};
  for (int q = 0; q < NV; ++q) ke.coef[q] = c->kcoef[q];
  run_hist_seg<NV>(ke, c->khalf, part, hist, counts, row0, crow0, crow1, s, -1, nullptr, nmem,
                   out_bstride, c->kbin.counts.as<int>());
}

template <typename T, typename TI>
void ensure_periodic_buffers(fc_corr* c) {
  const FftPlan& pi = *c->plan_inv;
  const int rs = c->rshape;
  const int oz = c->kshape[2] / 2;
  c->pk.reserve(sizeof(T) * (size_t)c->kshape[0] * c->kshape[1] * (oz + 1));
  c->q1.reserve(sizeof(cplx<TI>) * (size_t)rs * pi.n[1] * pi.ldc);
  c->q2.reserve(sizeof(cplx<TI>) * (size_t)rs * rs * pi.ldc);
  c->xisub.reserve(sizeof(TI) * (size_t)rs * rs * rs);
}

// Outputs of one fused x launch: nb batch members, member i reading its
// spectra at i * a_bstride and writing pk / q1 at i * pk_bstride / q_bstride
// elements.
struct PxOut {
  void* pk;
  void* q1;
  int nb = 1;
  int64_t a_bstride = 0, pk_bstride = 0, q_bstride = 0;
};

// Register-chained fused x pass (power-of-two x lengths up to 512).
template <typename T, typename TI, int MODE>
void launch_periodic_x2(fc_corr* c, const cplx<T>* a, const cplx<T>* b, const PxOut& po) {
  const FftPlan& p = *c->plan;
  const FftPlan& pi = *c->plan_inv;
  constexpr int TC = px_tc<T>();
  const int tilesC = (p.n2h + TC - 1) / TC;
  const int64_t per = (int64_t)p.n[1] * tilesC;
  const int64_t nblocks = per * po.nb;
  if (nblocks > INT32_MAX) fail(FC_ERR_INVALID, "grid too large for the fused x pass");
  const double2* ph[3] = {c->ph[0].as<double2>(), c->ph[1].as<double2>(),
                          c->ph[2].as<double2>()};
  const int rs = c->rshape;
  const int offm = (rs / 2) % p.n[0];
  if constexpr (std::is_same<T, float>::value && MODE == kModeCross) {
    // Cross mode, 512-point x over FP32 spectra, one correlation: the
    // persistent prefetching form, when each tile's 64-byte row segments lie
    // inside the row pitch and its in-tile offsets fit 32 bits.
    const uint64_t pitch = (uint64_t)p.n[1] * p.ldc * sizeof(cplx<T>);
    if (p.n[0] == 512 && po.nb == 1 && (int64_t)tilesC * TC <= p.ldc &&
        pitch * 16 < (1ull << 32)) {
      const int grid = (int)std::min<int64_t>(nblocks, current_device_cus());
      hipLaunchKernelGGL((k_periodic_x2_cross_pipe<TI, 512>), dim3((unsigned)grid),
                         dim3(px_threads<T, 512>()), 0, c->ctx->stream, a, b, p.n[1], p.n2h,
                         (int64_t)p.ldc, tilesC, p.tw0.as<cplx<T>>(), pi.tw0.as<cplx<TI>>(),
                         static_cast<T*>(po.pk), kslots(c), static_cast<cplx<TI>*>(po.q1),
                         (int64_t)pi.ldc, rs, offm, (int)nblocks);
      FC_LAUNCH_CHECK();
      return;
    }
  }
#define FC_PX2_CASE(Lv)                                                                       \
  case Lv:                                                                                    \
    hipLaunchKernelGGL((k_periodic_x2<T, TI, Lv, MODE>), dim3((unsigned)nblocks),             \
                       dim3(px_threads<T, Lv>()), (px_lds_bytes<T, TI, Lv, MODE>()),           \
                       c->ctx->stream, a, b, p.n[1], p.n2h, (int64_t)p.ldc, tilesC,            \
                       p.tw0.as<cplx<T>>(), pi.tw0.as<cplx<TI>>(), ph[0], ph[1], ph[2],        \
                       static_cast<T*>(po.pk), kslots(c), static_cast<cplx<TI>*>(po.q1),     \
                       (int64_t)pi.ldc, rs, offm, (int)per, po.a_bstride, po.pk_bstride,      \
                       po.q_bstride);                                                         \
    break;
  switch (p.n[0]) {
    FC_PX2_CASE(8)
    FC_PX2_CASE(16)
    FC_PX2_CASE(32)
    FC_PX2_CASE(64)
    FC_PX2_CASE(128)
    FC_PX2_CASE(256)
    FC_PX2_CASE(512)
    default:
      fail(FC_ERR_RUNTIME, "internal: no register-chained x kernel for this length");
  }
#undef FC_PX2_CASE
  FC_LAUNCH_CHECK();
}

// LDS-staged fused x pass for the longer fast lengths (forward precision only).
template <typename T, int MODE>
void launch_periodic_x(fc_corr* c, cplx<T>* a, const cplx<T>* b) {
  const FftPlan& p = *c->plan;
  const int tc = p.tile[0];
  const int tilesC = (p.n2h + tc - 1) / tc;
  const size_t cs = sizeof(cplx<T>);
  const double2* ph[3] = {c->ph[0].as<double2>(), c->ph[1].as<double2>(),
                          c->ph[2].as<double2>()};
#define FC_PX_CASE(Lv)                                                                      \
  case Lv: {                                                                                \
    const int thr = tc * (Lv / FftCfg<Lv>::E);                                              \
    hipLaunchKernelGGL((k_periodic_x<T, Lv, MODE>), dim3((unsigned)(p.n[1] * tilesC)),     \
                       dim3(thr), (size_t)Lv * tc * cs, c->ctx->stream, a, b, p.n[1], p.n2h, \
                       (int64_t)p.ldc, tilesC, tc, p.tw0.as<cplx<T>>(), ph[0], ph[1], ph[2],  \
                       c->pk.as<T>(), kslots(c), c->q1.as<cplx<T>>(), c->rshape);          \
    break;                                                                                  \
  }
  switch (p.n[0]) {
    FC_PX_CASE(1024)
    FC_PX_CASE(2048)
    FC_PX_CASE(4096)
    FC_PX_CASE(24)
    FC_PX_CASE(48)
    FC_PX_CASE(96)
    FC_PX_CASE(192)
    FC_PX_CASE(384)
    FC_PX_CASE(768)
    FC_PX_CASE(1536)
    FC_PX_CASE(3072)
    default:
      fail(FC_ERR_RUNTIME, "internal: no fused x kernel for this length");
  }
#undef FC_PX_CASE
  FC_LAUNCH_CHECK();
}

// The register-chained kernel addresses its arrays with 32-bit byte offsets.
bool px_fits(const fc_corr* c) {
  const FftPlan& p = *c->plan;
  const FftPlan& pi = *c->plan_inv;
  const uint64_t lim = 0xffffffffull;
  const uint64_t q1 = (uint64_t)c->rshape * pi.n[1] * pi.ldc * pi.csize();
  const uint64_t pk = (uint64_t)c->kshape[0] * c->kshape[1] * (c->kshape[2] / 2 + 1) * 8;
  return px_len_ok(p.n[0]) && p.complex_bytes() <= lim && q1 <= lim && pk <= lim;
}

template <typename T, typename TI, int MODE>
void fused_x(fc_corr* c, cplx<T>* w, const cplx<T>* b) {
  if (px_fits(c))
    launch_periodic_x2<T, TI, MODE>(c, w, b, PxOut{c->pk.as<void>(), c->q1.as<void>()});
  else if constexpr (std::is_same<T, TI>::value)
    launch_periodic_x<T, MODE>(c, w, b);
  else
    fail(FC_ERR_RUNTIME, "internal: fused x pass unavailable");
}

// Forward x pass(es), interlacing and the conjugate product over the whole
// work grid w (the reference's execute_fft + multiply_with_conjugation).
template <typename T, typename A = T>
void full_power(fc_corr* c, cplx<T>* w, const cplx<T>* b, int mode) {
  hipStream_t s = c->ctx->stream;
  const FftPlan& p = *c->plan;
  const int64_t nc = p.nrows() * p.ldc;
  forward_x<T, A>(c, w);
  if (mode == kModeInterlaced) forward_x<T, A>(c, const_cast<cplx<T>*>(b));
  Stage st(c, FC_STAGE_COMBINE);
  if (mode == kModeInterlaced) {
    hipLaunchKernelGGL(k_interlace<T>, dim3(grid1d(nc, 256, 65536)), dim3(256), 0, s, w, b,
                       p.n[0], p.n[1], (int64_t)p.ldc, c->ph[0].as<double2>(),
                       c->ph[1].as<double2>(), c->ph[2].as<double2>());
    FC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_conj_mult<T>, dim3(grid1d(nc, 256, 65536)), dim3(256), 0, s, w,
                     mode == kModeCross ? b : nullptr, nc);
  FC_LAUNCH_CHECK();
}

template <typename T>
void gather_pk(fc_corr* c, const cplx<T>* P) {
  const FftPlan& p = *c->plan;
  const KSlots ks = kslots(c);
  const int64_t npk = (int64_t)ks.ks0 * ks.ks1 * (ks.ks2 / 2 + 1);
  hipLaunchKernelGGL(k_gather_pk<T>, dim3(grid1d(npk, 256, 65536)), dim3(256), 0,
                     c->ctx->stream, c->pk.as<T>(), P, ks, p.n[0], p.n[1], (int64_t)p.ldc);
  FC_LAUNCH_CHECK();
}

// |F|^2 (or F conj(G)) -> compact k values pk, then the pruned inverse
// (x lags in the fused kernel, then y and z, in the inverse precision TI)
// -> r subgrid xisub.
template <typename T, typename TI>
void power_x(fc_corr* c, cplx<T>* w, const cplx<T>* b, int mode) {
  hipStream_t s = c->ctx->stream;
  const FftPlan& p = *c->plan;
  const FftPlan& pi = *c->plan_inv;
  const int64_t ldq = pi.ldc;
  const int rs = c->rshape, o = rs / 2;
  const bool fused = px_fits(c) || (p.fast[0] && !px_len_ok(p.n[0]) &&
                                    std::is_same<T, TI>::value);
  if (fused) {
    Stage st(c, FC_STAGE_COMBINE);
    if (mode == kModeInterlaced)
      fused_x<T, TI, kModeInterlaced>(c, w, b);
    else if (mode == kModeCross)
      fused_x<T, TI, kModeCross>(c, w, b);
    else
      fused_x<T, TI, kModeAuto>(c, w, b);
  } else {
    // Other x lengths: separate x pass, combine, gather, pruned inverse x
    // (reading the power grid in T, transforming in TI).
    full_power<T, TI>(c, w, b, mode);
    gather_pk<T>(c, w);
    Stage st(c, FC_STAGE_FFT_X_INV);
    ColView<TI, T> v{w, 0, (int64_t)p.n[1] * p.ldc, (int64_t)p.ldc, c->q1.as<cplx<TI>>(), 0,
                     (int64_t)p.n[1] * ldq, ldq, 1, p.n[1], p.n2h, rs, o};
    launch_cols<TI, +1, T>(p.n[0], v, pi.tile[0], pi.tw0.as<cplx<TI>>(), s);
  }
}

// Pruned inverse y and z passes: q1 -> q2 -> the r subgrid xisub.
template <typename T, typename TI>
void pruned_inverse_yz(fc_corr* c) {
  hipStream_t s = c->ctx->stream;
  const FftPlan& pi = *c->plan_inv;
  const int64_t ldq = pi.ldc;
  const int rs = c->rshape, o = rs / 2;
  // Pruned inverse y pass: q1[sx][y][c] -> q2[sx][sy][c], only the rs lags.
  {
    Stage st(c, FC_STAGE_FFT_Y_INV);
    ColView<TI> v{c->q1.as<cplx<TI>>(), (int64_t)pi.n[1] * ldq, ldq, 0, c->q2.as<cplx<TI>>(),
                  (int64_t)rs * ldq, ldq, 0, rs, 1, pi.n2h, rs, o};
    launch_cols<TI, +1>(pi.n[1], v, pi.tile[1], pi.tw1.as<cplx<TI>>(), s);
  }
  // Pruned z C2R: rows of q2 -> r subgrid rows xi[sx][sy][sz].
  Stage st(c, FC_STAGE_FFT_Z_INV);
  launch_rows_c2r<TI, TI>(pi, c->q2.as<cplx<TI>>(), ldq, c->xisub.as<TI>(), rs, s,
                          (int64_t)rs * rs, rs, o);
}

// Power spectrum with its compact k values (pk) and pruned inverse x (q1),
// then the pruned inverse y and z passes (q1 -> xisub).
template <typename T, typename TI>
void power_and_pruned_inverse(fc_corr* c, cplx<T>* w, const cplx<T>* b, int mode) {
  power_x<T, TI>(c, w, b, mode);
  pruned_inverse_yz<T, TI>(c);
}

// Builds the periodic histogram coefficients and r weights once per correlator.
template <int NV>
void ensure_periodic_weights(fc_corr* c) {
  if (c->rw.bytes()) return;
  const uint64_t ncells = (uint64_t)c->shape[0] * c->shape[1] * c->shape[2];
  // correlator.h:294, :310-313 (k_rescale = cell^3 / N) and :336 (1 / N / N).
  const double k_rescale = c->cell * c->cell * c->cell / ncells;
  const double r_rescale = (1.0 / ncells / ncells);
  double rc[NV];
  for (int q = 0; q < NV; ++q) {
    c->kcoef[q] = ylm_coeff(2 * q, 0, std::sqrt((4.0 * M_PI) * (2 * (2 * q) + 1)) * k_rescale);
    rc[q] = ylm_coeff(2 * q, 0, std::sqrt((4.0 * M_PI) * (2 * (2 * q) + 1)) * r_rescale);
  }
  DeviceBuffer coef(sizeof(double) * NV);
  FC_HIP(hipMemcpy(coef.as<void>(), rc, sizeof(double) * NV, hipMemcpyHostToDevice));
  const int64_t nr = c->rbin.n;
  c->rw.allocate(sizeof(double) * NV * std::max<int64_t>(nr, 1));
  hipStream_t s = c->ctx->stream;
  if (nr > 0)
    hipLaunchKernelGGL((k_setup_rweights<NV>), dim3(grid1d(nr, 256, 65536)), dim3(256), 0, s,
                       c->rbin.cells.as<uint32_t>(), nr, c->rshape, c->rx.as<double>(),
                       coef.as<double>(), c->rw.as<double>());
  FC_LAUNCH_CHECK();
  FC_HIP(hipStreamSynchronize(s));  // coef is freed on return
}

// Periodic k histogram of pk on stream s, partials in part.
template <typename T, int NV>
void periodic_k_hist(fc_corr* c, hipStream_t s, DeviceBuffer& part) {
  Stage st(c, FC_STAGE_HIST_K);
  if (c->khalf_ok) {
    run_hist_khalf<T, NV>(c, c->pk.as<T>(), 0, part, c->k_hist.as<double>(),
                          c->k_counts.as<int>(), 0, 0, NV, s);
  } else {
    KEntry<T, NV> ke{c->kbin.cells.as<uint32_t>(), c->pk.as<T>(), kslots(c),
                     FastDiv((uint32_t)c->kshape[2]), FastDiv((uint32_t)c->kshape[1]),
                     c->kx.as<double>(), c->ky.as<double>(), c->kz.as<double>(),
                     c->wx.as<double>(), c->wy.as<double>(), c->wz.as<double>(), {}, 0};
    for (int q = 0; q < NV; ++q) ke.coef[q] = c->kcoef[q];
    run_hist_seg<NV>(ke, c->kbin, part, c->k_hist.as<double>(), c->k_counts.as<int>(), 0, 0,
                     NV, s);
  }
}

template <typename TI, int NV>
void periodic_r_hist(fc_corr* c) {
  Stage st(c, FC_STAGE_HIST_R);
  WEntry<TI, NV> we{c->xisub.as<TI>(), c->rbin.cells.as<uint32_t>(), c->rw.as<double>(),
                    c->rbin.n};
  run_hist_seg<NV>(we, c->rbin, c->hs.part_seg, c->r_hist.as<double>(), c->r_counts.as<int>(), 0,
                   0, NV, c->ctx->stream, c->rbin.tracked_pos, c->zerolag.as<double>());
}

// The side stream and its fork / join events (created on first use).
hipStream_t side_stream(fc_corr* c) {
  if (!c->side) {
    FC_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    FC_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    FC_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  return c->side;
}

// After the forward passes: power spectrum and pruned inverse x, then the k
// histogram (which reads only pk) on the side stream while the pruned
// inverse y and z passes and the r histogram run on the correlator's
// stream; the stream waits for the side stream at the end, so the next
// call's clearing of the histograms follows both.  With stage profiling on,
// everything runs on the correlator's stream, in order.
template <typename T, typename TI, int NV>
void periodic_tail(fc_corr* c, cplx<T>* w, const cplx<T>* b, int mode) {
  hipStream_t s = c->ctx->stream;
  ensure_periodic_weights<NV>(c);
  power_x<T, TI>(c, w, b, mode);
  if (c->profile) {
    periodic_k_hist<T, NV>(c, s, c->part_k);
    pruned_inverse_yz<T, TI>(c);
    periodic_r_hist<TI, NV>(c);
    return;
  }
  hipStream_t hs = side_stream(c);
  FC_HIP(hipEventRecord(c->ev_fork, s));
  FC_HIP(hipStreamWaitEvent(hs, c->ev_fork, 0));
  periodic_k_hist<T, NV>(c, hs, c->part_k);
  FC_HIP(hipEventRecord(c->ev_join, hs));
  pruned_inverse_yz<T, TI>(c);
  periodic_r_hist<TI, NV>(c);
  FC_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
}

template <typename T, typename TI>
void periodic_correlate(fc_corr* c, fc_grid* g1, int mode, fc_grid* gb) {
  const FftPlan& p = *c->plan;
  cplx<T>* w = c->work.as<cplx<T>>();
  ensure_periodic_buffers<T, TI>(c);
  ensure_phase_tables(c);
  const cplx<T>* b = nullptr;
  if (mode == kModeInterlaced && forward_pair_ok(c, g1, gb)) {
    // Both grids' spectra in one buffer, each forward pass one launch.
    c->work.reserve(2 * p.complex_bytes());
    w = c->work.as<cplx<T>>();
    b = w + (int64_t)p.n[0] * p.n[1] * p.ldc;
    forward_zy_pair<T, TI>(c, g1, gb, w);
  } else {
    forward_zy<T, TI>(c, g1, w, LoadCopy{});
    if (mode == kModeInterlaced) {
      c->work_b.reserve(p.complex_bytes());
      forward_zy<T, TI>(c, gb, c->work_b.as<cplx<T>>(), LoadCopy{});
      b = c->work_b.as<cplx<T>>();
    } else if (mode == kModeCross) {
      b = c->grid2_fft.as<cplx<T>>();
    }
  }
  switch (c->nell) {
    case 1:
      periodic_tail<T, TI, 1>(c, w, b, mode);
      break;
    case 2:
      periodic_tail<T, TI, 2>(c, w, b, mode);
      break;
    default:
      periodic_tail<T, TI, 3>(c, w, b, mode);
  }
}

// ---- the reference's stage sequence (FC_ALGO_PLAIN) -------------------------
// Full inverse 3D FFT of the power grid in w, in place (real rows of 2 ldc).
template <typename T>
T* full_inverse(fc_corr* c, cplx<T>* w) {
  const FftPlan& p = *c->plan;
  hipStream_t s = c->ctx->stream;
  {
    Stage st(c, FC_STAGE_FFT_X_INV);
    launch_cols<T, +1>(p.n[0], x_view<T>(p, w), p.tile[0], p.tw0.as<cplx<T>>(), s);
  }
  {
    Stage st(c, FC_STAGE_FFT_Y_INV);
    launch_cols<T, +1>(p.n[1], y_view<T>(p, w), p.tile[1], p.tw1.as<cplx<T>>(), s);
  }
  Stage st(c, FC_STAGE_FFT_Z_INV);
  T* xr = reinterpret_cast<T*>(w);
  launch_rows_c2r<T, T>(p, w, p.ldc, xr, 2 * (int64_t)p.ldc, s);
  return xr;
}

// One multipole of the k subgrid: ksub = Y_lm(k) coef W^-1(k) Re P.
template <typename T>
void plain_k_multipole(fc_corr* c, int ell, int m, double coef, bool accumulate) {
  hipStream_t s = c->ctx->stream;
  const int64_t nk = (int64_t)c->kshape[0] * c->kshape[1] * c->kshape[2];
  Stage st(c, FC_STAGE_SUBGRID);
  if (!accumulate) fill<double>(c->ksub.as<double>(), nk, 0.0, s);
  hipLaunchKernelGGL(k_acc_k<T>, dim3(grid1d(nk, 256, 65536)), dim3(256), 0, s,
                     c->ksub.as<double>(), c->pk.as<T>(), kslots(c), c->kx.as<double>(),
                     c->ky.as<double>(), c->kz.as<double>(), c->wx.as<double>(),
                     c->wy.as<double>(), c->wz.as<double>(), ell, m, ylm_coeff(ell, m, coef));
  FC_LAUNCH_CHECK();
}

// One multipole of the r subgrid from the full real correlation grid.
template <typename T>
void plain_r_multipole(fc_corr* c, const T* xr, int ell, int m, double coef, bool accumulate) {
  hipStream_t s = c->ctx->stream;
  const FftPlan& p = *c->plan;
  const int rs = c->rshape;
  const int64_t nr = (int64_t)rs * rs * rs;
  Stage st(c, FC_STAGE_SUBGRID);
  if (!accumulate) fill<double>(c->rsub.as<double>(), nr, 0.0, s);
  hipLaunchKernelGGL(k_acc_r_full<T>, dim3(grid1d(nr, 256, 65536)), dim3(256), 0, s,
                     c->rsub.as<double>(), xr, p.n[0], p.n[1], p.n[2], 2 * (int64_t)p.ldc, rs,
                     c->rx.as<double>(), ell, m, ylm_coeff(ell, m, coef));
  FC_LAUNCH_CHECK();
}

void hist_subgrids(fc_corr* c, int ell, bool k_side, bool r_side) {
  hipStream_t s = c->ctx->stream;
  if (r_side) {
    Stage st(c, FC_STAGE_HIST_R);
    run_hist_seg<1>(GEntry<double>{c->rsub.as<double>(), c->rbin.cells.as<uint32_t>()}, c->rbin,
                    c->hs.part_seg, c->r_hist.as<double>(), c->r_counts.as<int>(), ell / 2,
                    ell / 2, ell / 2 + 1, s);
    if (ell == 0) {
      const int rs = c->rshape, o = rs / 2;
      const int64_t center = ((int64_t)o * rs + o) * rs + o;
      FC_HIP(hipMemcpyAsync(c->zerolag.as<double>(), c->rsub.as<double>() + center,
                            sizeof(double), hipMemcpyDeviceToDevice, s));
    }
  }
  if (k_side) {
    Stage st(c, FC_STAGE_HIST_K);
    run_hist_seg<1>(GEntry<double>{c->ksub.as<double>(), c->kbin.cells.as<uint32_t>()}, c->kbin,
                    c->hs.part_seg, c->k_hist.as<double>(), c->k_counts.as<int>(), ell / 2,
                    ell / 2, ell / 2 + 1, s);
  }
}

void ensure_subgrids(fc_corr* c, bool ksub = true) {
  const int rs = c->rshape;
  c->rsub.reserve(sizeof(double) * (size_t)rs * rs * rs);
  if (ksub) c->ksub.reserve(sizeof(double) * (size_t)c->kshape[0] * c->kshape[1] * c->kshape[2]);
}

// PeriodicCorrelator::correlate_internal (correlator.h:273-357) stage by stage.
template <typename T>
void periodic_plain(fc_corr* c, fc_grid* g1, int mode, fc_grid* gb) {
  const FftPlan& p = *c->plan;
  cplx<T>* w = c->work.as<cplx<T>>();
  ensure_periodic_buffers<T, T>(c);
  ensure_phase_tables(c);
  ensure_subgrids(c);
  forward_zy<T>(c, g1, w, LoadCopy{});
  const cplx<T>* b = nullptr;
  if (mode == kModeInterlaced) {
    c->work_b.reserve(p.complex_bytes());
    forward_zy<T>(c, gb, c->work_b.as<cplx<T>>(), LoadCopy{});
    b = c->work_b.as<cplx<T>>();
  } else if (mode == kModeCross) {
    b = c->grid2_fft.as<cplx<T>>();
  }
  full_power<T>(c, w, b, mode);
  gather_pk<T>(c, w);
  const uint64_t ncells = (uint64_t)c->shape[0] * c->shape[1] * c->shape[2];
  const double k_rescale = c->cell * c->cell * c->cell / ncells;
  const double r_rescale = (1.0 / ncells / ncells);
  for (int ell = 0; ell <= c->maxell; ell += 2) {
    plain_k_multipole<T>(c, ell, 0, std::sqrt((4.0 * M_PI) * (2 * ell + 1)) * k_rescale, false);
    hist_subgrids(c, ell, true, false);
  }
  const T* xr = full_inverse<T>(c, w);
  for (int ell = 0; ell <= c->maxell; ell += 2) {
    plain_r_multipole<T>(c, xr, ell, 0, std::sqrt((4.0 * M_PI) * (2 * ell + 1)) * r_rescale,
                         false);
    hist_subgrids(c, ell, false, true);
  }
}

// Correlator::correlate_internal (correlator.h:405-491) stage by stage.
template <typename T>
void yamamoto_plain(fc_corr* c, fc_grid* g1) {
  const FftPlan& p = *c->plan;
  cplx<T>* w = c->work.as<cplx<T>>();
  const uint64_t ncells = (uint64_t)c->shape[0] * c->shape[1] * c->shape[2];
  ensure_periodic_buffers<T, T>(c);
  ensure_subgrids(c);
  double start[3];
  for (int i = 0; i < 3; ++i) start[i] = 0.5 - (-c->posmin[i] / c->cell);
  for (int ell = 0; ell <= c->maxell; ell += 2) {
    for (int m = -ell; m <= ell; ++m) {
      LoadYlm ld{ell, m, p.n[1], ylm_coeff(ell, m, 1.0), start[0], start[1], start[2]};
      forward_zy<T>(c, g1, w, ld);
      full_power<T>(c, w, c->grid2_fft.as<cplx<T>>(), kModeCross);
      gather_pk<T>(c, w);
      plain_k_multipole<T>(c, ell, m, 4.0 * M_PI, m != -ell);
      const T* xr = full_inverse<T>(c, w);
      plain_r_multipole<T>(c, xr, ell, m, 4.0 * M_PI / ncells, m != -ell);
    }
    hist_subgrids(c, ell, true, true);
  }
}

template <typename T, typename TI>
void yamamoto_correlate(fc_corr* c, fc_grid* g1) {
  hipStream_t s = c->ctx->stream;
  const FftPlan& p = *c->plan;
  cplx<T>* w = c->work.as<cplx<T>>();
  const uint64_t ncells = (uint64_t)c->shape[0] * c->shape[1] * c->shape[2];
  ensure_periodic_buffers<T, TI>(c);
  ensure_phase_tables(c);
  ensure_subgrids(c, false);
  // Observer at the survey origin; cell centres 0.5 - observer + i
  // (correlator.h:381-402).
  double start[3];
  for (int i = 0; i < 3; ++i) start[i] = 0.5 - (-c->posmin[i] / c->cell);
  const int rs = c->rshape;
  const int64_t nr = (int64_t)rs * rs * rs;
  // Each transform's k histogram runs on the side stream (unless stage
  // profiling is on) beside its pruned inverse and the next transform's
  // forward passes; the next power pass, which rewrites pk, waits for it.
  const bool par = !c->profile;
  hipStream_t ks = par ? side_stream(c) : s;
  DeviceBuffer& kpart = par ? c->part_k : c->hs.part_seg;
  bool pending = false;
  for (int ell = 0; ell <= c->maxell; ell += 2) {
    for (int m = -ell; m <= ell; ++m) {
      LoadYlm ld{ell, m, p.n[1], ylm_coeff(ell, m, 1.0), start[0], start[1], start[2]};
      forward_zy<T, TI>(c, g1, w, ld);
      if (pending) FC_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
      power_x<T, TI>(c, w, c->grid2_fft.as<cplx<T>>(), kModeCross);
      if (par) {
        FC_HIP(hipEventRecord(c->ev_fork, s));
        FC_HIP(hipStreamWaitEvent(ks, c->ev_fork, 0));
      }
      {
        // k-space multipole l: this transform's share, straight from the
        // compact k values (counts once per l).
        Stage st(c, FC_STAGE_HIST_K);
        const int row = ell / 2;
        if (c->khalf_ok) {
          const int nz = c->kshape[2] / 2 + 1, oz = c->kshape[2] / 2;
          KEntryHalfLm<T> ke{c->khalf.cells.as<uint32_t>(), c->pk.as<T>(),
                             FastDiv((uint32_t)nz), FastDiv((uint32_t)c->kshape[1]), nz,
                             c->kshape[1], c->kx.as<double>(), c->ky.as<double>(),
                             c->kz.as<double>() + oz, c->wx.as<double>(), c->wy.as<double>(),
                             c->wz.as<double>() + oz, ell, m, ylm_coeff(ell, m, 4.0 * M_PI)};
          run_hist_seg<1>(ke, c->khalf, kpart, c->k_hist.as<double>(), c->k_counts.as<int>(),
                          row, row, m == -ell ? row + 1 : row, ks, -1, nullptr, 1, 0,
                          c->kbin.counts.as<int>());
        } else {
          KEntryLm<T> ke{c->kbin.cells.as<uint32_t>(), c->pk.as<T>(), kslots(c),
                         FastDiv((uint32_t)c->kshape[2]), FastDiv((uint32_t)c->kshape[1]),
                         c->kx.as<double>(), c->ky.as<double>(), c->kz.as<double>(),
                         c->wx.as<double>(), c->wy.as<double>(), c->wz.as<double>(), ell, m,
                         ylm_coeff(ell, m, 4.0 * M_PI)};
          run_hist_seg<1>(ke, c->kbin, kpart, c->k_hist.as<double>(), c->k_counts.as<int>(),
                          row, row, m == -ell ? row + 1 : row, ks);
        }
      }
      if (par) {
        FC_HIP(hipEventRecord(c->ev_join, ks));
        pending = true;
      }
      pruned_inverse_yz<T, TI>(c);
      Stage st(c, FC_STAGE_SUBGRID);
      if (m == -ell) fill<double>(c->rsub.as<double>(), nr, 0.0, s);
      hipLaunchKernelGGL(k_acc_r<TI>, dim3(grid1d(nr, 256, 65536)), dim3(256), 0, s,
                         c->rsub.as<double>(), c->xisub.as<TI>(), rs, c->rx.as<double>(), ell,
                         m, ylm_coeff(ell, m, 4.0 * M_PI / ncells));
      FC_LAUNCH_CHECK();
    }
    hist_subgrids(c, ell, false, true);
  }
  if (pending) FC_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
}

void ensure_work(fc_corr* c) { c->work.reserve(c->plan->complex_bytes()); }

void set_grid2_device(fc_corr* c, fc_grid* g2) {
  ensure_work(c);
  c->grid2_fft.reserve(c->plan->complex_bytes());
  if (c->prec == FC_FP64)
    forward_fft<double>(c, g2, c->grid2_fft.as<cplx<double>>(), LoadCopy{});
  else if (c->prec == FC_MIXED && c->algo != FC_ALGO_PLAIN)
    forward_fft<float, double>(c, g2, c->grid2_fft.as<cplx<float>>(), LoadCopy{});
  else
    forward_fft<float>(c, g2, c->grid2_fft.as<cplx<float>>(), LoadCopy{});
  c->have_grid2 = true;
}

template <typename T, typename TI>
void run_kind(fc_corr* c, fc_grid* g1, fc_grid* gb, int mode) {
  const bool plain = c->algo == FC_ALGO_PLAIN;
  if (c->kind == FC_YAMAMOTO) {
    if constexpr (std::is_same<T, TI>::value) {
      if (plain) return yamamoto_plain<T>(c, g1);
    }
    yamamoto_correlate<T, TI>(c, g1);
  } else {
    if constexpr (std::is_same<T, TI>::value) {
      if (plain) return periodic_plain<T>(c, g1, mode, gb);
    }
    periodic_correlate<T, TI>(c, g1, mode, gb);
  }
}

void correlate(fc_corr* c, fc_grid* g1, fc_grid* gb, int mode) {
  DeviceGuard dg(c->ctx->device);
  hipStream_t s = c->ctx->stream;
  check_grid_shape(c, g1);
  join_streams(c, g1);
  if (gb) {
    check_grid_shape(c, gb);
    join_streams(c, gb);
  }
  if (c->kind == FC_YAMAMOTO) {
    if (mode == kModeInterlaced)
      fail(FC_ERR_INVALID, "interlacing applies to the periodic correlator");
  } else if (mode == kModeCross && !c->have_grid2) {
    fail(FC_ERR_RUNTIME, "grid2 FFT not set: call set_grid2 first");
  }
  ensure_work(c);
  flush_stage_times(c);
  FC_HIP(hipEventRecord(c->ev0, s));
  reset_hists(c, s);
  if (c->kind == FC_YAMAMOTO) {
    if (mode == kModeAuto) set_grid2_device(c, g1);  // Correlator::autocorrelate = cross(grid, grid)
    if (!c->have_grid2) fail(FC_ERR_RUNTIME, "grid2 FFT not set: call set_grid2 first");
  }
  if (c->prec == FC_FP64)
    run_kind<double, double>(c, g1, gb, mode);
  else if (c->prec == FC_FP32)
    run_kind<float, float>(c, g1, gb, mode);
  else
    run_kind<float, double>(c, g1, gb, mode);
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
    if (precision != FC_FP64 && precision != FC_FP32 && precision != FC_MIXED)
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
    c->plan = fc::get_fft_plan(ctx->device, shape, precision == FC_FP64 ? FC_FP64 : FC_FP32);
    c->plan_inv = precision == FC_MIXED ? fc::get_fft_plan(ctx->device, shape, FC_FP64) : c->plan;
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
    if (corr->side) {
      (void)hipStreamSynchronize(corr->side);
      (void)hipStreamDestroy(corr->side);
    }
    if (corr->ev_fork) (void)hipEventDestroy(corr->ev_fork);
    if (corr->ev_join) (void)hipEventDestroy(corr->ev_join);
    for (auto e : corr->ev_pool) (void)hipEventDestroy(e);
    if (corr->ev0) (void)hipEventDestroy(corr->ev0);
    if (corr->ev1) (void)hipEventDestroy(corr->ev1);
    delete corr;
  });
}

int fc_corr_footprint(fc_corr* c, int interlaced, size_t* held, size_t* full) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    // (bytes held now, bytes the buffer reaches once the correlation paths
    // have run) for every device buffer of the correlator.
    size_t now = 0, need = 0;
    auto add = [&](const fc::DeviceBuffer& b, size_t want) {
      now += b.bytes();
      need += std::max(b.bytes(), want);
    };
    for (const fc::DeviceBuffer* b : {&c->rx, &c->kx, &c->ky, &c->kz, &c->wx, &c->wy, &c->wz,
                                      &c->r_counts, &c->r_hist, &c->k_counts, &c->k_hist,
                                      &c->zerolag, &c->hs.part_hist, &c->hs.part_counts})
      add(*b, 0);
    for (const fc::BinnedCells* bc : {&c->kbin, &c->rbin, &c->khalf}) {
      add(bc->cells, 0);
      add(bc->counts, 0);
      add(bc->chunks, 0);
      add(bc->chunk_first, 0);
      add(bc->order, 0);
    }
    const fc::FftPlan& p = *c->plan;
    const fc::FftPlan& pi = *c->plan_inv;
    const size_t cb = p.complex_bytes();
    const size_t ts = c->prec == FC_FP64 ? 8 : 4;            // forward storage
    const size_t tsi = c->prec == FC_FP32 ? 4 : 8;           // inverse arithmetic
    const size_t rs = (size_t)c->rshape;
    const size_t nv = (size_t)c->nell;
    const size_t K = (size_t)c->kshape[0] * c->kshape[1] * c->kshape[2];
    // Interlaced: the fused path with one-launch forward passes keeps both
    // spectra in work, otherwise the second sits in work_b (grids of two
    // precisions take that path too: then this is an upper bound).
    const bool pair = interlaced && c->algo != FC_ALGO_PLAIN && fc::forward_pair_plan_ok(c);
    add(c->work, pair ? 2 * cb : cb);
    add(c->work_b, interlaced && !pair ? cb : 0);
    add(c->grid2_fft, c->kind == FC_YAMAMOTO || c->have_grid2 ? cb : 0);
    add(c->pk, ts * (size_t)c->kshape[0] * c->kshape[1] * (c->kshape[2] / 2 + 1));
    add(c->q1, 2 * tsi * rs * pi.n[1] * pi.ldc);
    add(c->q2, 2 * tsi * rs * rs * pi.ldc);
    add(c->xisub, tsi * rs * rs * rs);
    add(c->rsub, c->kind == FC_YAMAMOTO || c->algo == FC_ALGO_PLAIN ? 8 * rs * rs * rs : 0);
    add(c->ksub, c->algo == FC_ALGO_PLAIN ? 8 * K : 0);
    add(c->rw, 8 * nv * (size_t)std::max<int64_t>(c->rbin.n, 1));
    add(c->hs.part_seg,
        8 * nv * (size_t)std::max({c->kbin.nchunks, c->rbin.nchunks, c->khalf.nchunks}));
    add(c->part_k, 8 * nv * (size_t)(c->khalf_ok ? c->khalf.nchunks : c->kbin.nchunks));
    for (int d = 0; d < 3; ++d)
      add(c->ph[d], sizeof(double2) * (size_t)(d == 2 ? p.ldc : c->shape[d]));
    if (held) *held = now;
    if (full) *full = need;
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
    const fc::FftPlan& p = *c->plan;
    const int64_t rows = p.nrows();
    const int64_t n = rows * p.n2h;
    c->grid2_fft.reserve(p.complex_bytes());
    fc::DeviceBuffer tmp(sizeof(double2) * n);
    FC_HIP(hipMemcpyAsync(tmp.as<void>(), host_complex, sizeof(double2) * n,
                          hipMemcpyHostToDevice, s));
    if (c->prec == FC_FP64)
      hipLaunchKernelGGL(fc::k_from_c128<double>, dim3(fc::grid1d(n, 256, 65536)), dim3(256), 0,
                         s, c->grid2_fft.as<double2>(), tmp.as<double2>(), rows, p.n2h,
                         (int64_t)p.ldc);
    else
      hipLaunchKernelGGL(fc::k_from_c128<float>, dim3(fc::grid1d(n, 256, 65536)), dim3(256), 0,
                         s, c->grid2_fft.as<float2>(), tmp.as<double2>(), rows, p.n2h,
                         (int64_t)p.ldc);
    FC_LAUNCH_CHECK();
    FC_HIP(hipStreamSynchronize(s));
    c->have_grid2 = true;
  });
}

int fc_corr_autocorrelate(fc_corr* c, fc_grid* grid) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    fc::correlate(c, grid, nullptr, fc::kModeAuto);
  });
}

int fc_corr_cross_correlate(fc_corr* c, fc_grid* grid1) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    fc::correlate(c, grid1, nullptr, fc::kModeCross);
  });
}

int fc_corr_autocorrelate_interlaced(fc_corr* c, fc_grid* a, fc_grid* b) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    if (!b) fc::fail(FC_ERR_INVALID, "null interlaced grid");
    fc::correlate(c, a, b, fc::kModeInterlaced);
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

int fc_corr_set_algorithm(fc_corr* c, int algo) {
  return fc::guard([&] {
    if (!c) fc::fail(FC_ERR_INVALID, "null correlator");
    if (algo != FC_ALGO_FUSED && algo != FC_ALGO_PLAIN)
      fc::fail(FC_ERR_INVALID, "unknown algorithm " + std::to_string(algo));
    if (algo == FC_ALGO_PLAIN && c->prec == FC_MIXED)
      fc::fail(FC_ERR_INVALID, "the plain algorithm runs in fp64 or fp32, not mixed");
    c->algo = algo;
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

// ---- batches of same-geometry periodic correlations -------------------------
// Every member shares the correlator's geometry, plans, bin-sorted cells and
// weights; the grids, work arrays and outputs are stacked [member][...] so
// each pass is one launch over all members.  The launch sequence for a given
// member count can be captured once as a HIP graph and replayed.
struct fc_batch {
  fc_corr* corr = nullptr;
  int nb = 0;
  int grid_prec = FC_FP64;
  size_t grid_bytes = 0;
  fc::DeviceBuffer input;  // [nb][n0][n1][n2] grid storage
  std::vector<fc_grid*> grids;
  fc::DeviceBuffer work, pk, q1, q2, xisub, part;
  fc::DeviceBuffer r_hist, r_counts, k_hist, k_counts, zerolag;
  int64_t npk = 0, nq1 = 0, nq2 = 0, nxs = 0;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_n = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  ~fc_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto g : grids) delete g;
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (corr) fc_corr_destroy(corr);
  }
};

namespace fc {
namespace {

template <typename T, typename TI, int NV>
void batch_sequence(fc_batch* bt, int n) {
  fc_corr* c = bt->corr;
  hipStream_t s = c->ctx->stream;
  const FftPlan& p = *c->plan;
  const FftPlan& pi = *c->plan_inv;
  const int rs = c->rshape, o = rs / 2;
  const int64_t ldq = pi.ldc;
  const int64_t hk = (int64_t)c->nell * c->nkbins, hr = (int64_t)c->nell * c->nrbins;
  fill<int>(bt->r_counts.as<int>(), hr * n, 0, s);
  fill<double>(bt->r_hist.as<double>(), hr * n, 0.0, s);
  fill<int>(bt->k_counts.as<int>(), hk * n, 0, s);
  fill<double>(bt->k_hist.as<double>(), hk * n, 0.0, s);
  fill<double>(bt->zerolag.as<double>(), n, 0.0, s);
  cplx<T>* w = bt->work.as<cplx<T>>();
  const int64_t rows = (int64_t)n * p.n[0] * p.n[1];
  // Forward z and y in the arithmetic of TI (FP64 for the mixed form),
  // stored in T.
  if (bt->grid_prec == FC_FP64)
    launch_rows_r2c<T, double, LoadCopy, TI>(p, bt->input.as<double>(), p.n[2], w, p.ldc,
                                             LoadCopy{}, s, rows, &pi);
  else
    launch_rows_r2c<T, float, LoadCopy, TI>(p, bt->input.as<float>(), p.n[2], w, p.ldc,
                                            LoadCopy{}, s, rows, &pi);
  {
    ColView<T> v = y_view<T>(p, w);
    v.A = n * p.n[0];
    launch_cols_wide<T, -1, TI>(p.n[1], v, p.tile[1], pi.tw1.as<cplx<TI>>(),
                                p.tw1.as<cplx<T>>(), s, p.var[1]);
  }
  if (px_fits(c)) {
    PxOut po{bt->pk.as<void>(), bt->q1.as<void>(), n, (int64_t)p.n[0] * p.n[1] * p.ldc, bt->npk,
             bt->nq1};
    launch_periodic_x2<T, TI, kModeAuto>(c, w, nullptr, po);
  } else {
    // x lengths without the fused kernel (e.g. 384 = 3 x 128): per member,
    // the x pass, |F|^2, the compact k values and the pruned inverse x pass
    // (the single correlator's separate-pass sequence) into its slices.
    const int64_t nc = p.nrows() * p.ldc;
    const KSlots ks = kslots(c);
    const int64_t npk = bt->npk;
    for (int i = 0; i < n; ++i) {
      cplx<T>* wi = w + (int64_t)i * nc;
      launch_cols_wide<T, -1, TI>(p.n[0], x_view<T>(p, wi), p.tile[0], pi.tw0.as<cplx<TI>>(),
                                  p.tw0.as<cplx<T>>(), s, p.var[0]);
      hipLaunchKernelGGL(k_conj_mult<T>, dim3(grid1d(nc, 256, 65536)), dim3(256), 0, s, wi,
                         nullptr, nc);
      FC_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gather_pk<T>, dim3(grid1d(npk, 256, 65536)), dim3(256), 0, s,
                         bt->pk.as<T>() + i * npk, wi, ks, p.n[0], p.n[1], (int64_t)p.ldc);
      FC_LAUNCH_CHECK();
      ColView<TI, T> v{wi, 0, (int64_t)p.n[1] * p.ldc, (int64_t)p.ldc,
                       bt->q1.as<cplx<TI>>() + i * bt->nq1, 0, (int64_t)p.n[1] * ldq, ldq, 1,
                       p.n[1], p.n2h, rs, o};
      launch_cols<TI, +1, T>(p.n[0], v, pi.tile[0], pi.tw0.as<cplx<TI>>(), s);
    }
  }
  {
    ColView<TI> v{bt->q1.as<cplx<TI>>(), (int64_t)pi.n[1] * ldq, ldq, 0, bt->q2.as<cplx<TI>>(),
                  (int64_t)rs * ldq, ldq, 0, n * rs, 1, pi.n2h, rs, o};
    launch_cols<TI, +1>(pi.n[1], v, pi.tile[1], pi.tw1.as<cplx<TI>>(), s);
  }
  launch_rows_c2r<TI, TI>(pi, bt->q2.as<cplx<TI>>(), ldq, bt->xisub.as<TI>(), rs, s,
                          (int64_t)n * rs * rs, rs, o);
  if (c->khalf_ok) {
    run_hist_khalf<T, NV>(c, bt->pk.as<T>(), bt->npk, bt->part, bt->k_hist.as<double>(),
                          bt->k_counts.as<int>(), 0, 0, NV, s, n, hk);
  } else {
    KEntry<T, NV> ke{c->kbin.cells.as<uint32_t>(), bt->pk.as<T>(), kslots(c),
                     FastDiv((uint32_t)c->kshape[2]), FastDiv((uint32_t)c->kshape[1]),
                     c->kx.as<double>(), c->ky.as<double>(), c->kz.as<double>(),
                     c->wx.as<double>(), c->wy.as<double>(), c->wz.as<double>(), {}, bt->npk};
    for (int q = 0; q < NV; ++q) ke.coef[q] = c->kcoef[q];
    run_hist_seg<NV>(ke, c->kbin, bt->part, bt->k_hist.as<double>(), bt->k_counts.as<int>(), 0,
                     0, NV, s, -1, nullptr, n, hk);
  }
  WEntry<TI, NV> we{bt->xisub.as<TI>(), c->rbin.cells.as<uint32_t>(), c->rw.as<double>(),
                    c->rbin.n, bt->nxs};
  run_hist_seg<NV>(we, c->rbin, bt->part, bt->r_hist.as<double>(), bt->r_counts.as<int>(), 0, 0,
                   NV, s, c->rbin.tracked_pos, bt->zerolag.as<double>(), n, hr);
}

template <typename T, typename TI>
void batch_dispatch(fc_batch* bt, int n) {
  switch (bt->corr->nell) {
    case 1:
      return batch_sequence<T, TI, 1>(bt, n);
    case 2:
      return batch_sequence<T, TI, 2>(bt, n);
    default:
      return batch_sequence<T, TI, 3>(bt, n);
  }
}

void batch_enqueue(fc_batch* bt, int n) {
  const int prec = bt->corr->prec;
  if (prec == FC_FP64)
    batch_dispatch<double, double>(bt, n);
  else if (prec == FC_FP32)
    batch_dispatch<float, float>(bt, n);
  else
    batch_dispatch<float, double>(bt, n);
}

template <int NV>
void batch_weights(fc_corr* c) {
  ensure_periodic_weights<NV>(c);
}

}  // namespace
}  // namespace fc

extern "C" {

int fc_batch_create(fc_ctx* ctx, const int shape[3], double cell_size, int window_correct,
                    double rmax, double dr, double kmax, double dk, int maxell, int precision,
                    int nbatch, fc_batch** out) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!shape || !out) fc::fail(FC_ERR_INVALID, "null argument");
    if (nbatch < 1) fc::fail(FC_ERR_INVALID, "nbatch must be positive");
    std::unique_ptr<fc_batch> b(new fc_batch());
    if (const int rc = fc_corr_create(ctx, FC_PERIODIC, shape, cell_size, nullptr,
                                      window_correct, rmax, dr, kmax, dk, maxell, precision,
                                      &b->corr))
      fc::fail(rc, fc_last_error());
    fc_corr* c = b->corr;
    fc::DeviceGuard dg(ctx->device);
    const fc::FftPlan& p = *c->plan;
    const fc::FftPlan& pi = *c->plan_inv;
    const size_t csi = pi.csize();
    const size_t ts = precision == FC_FP64 ? 8 : 4, tsi = precision == FC_FP32 ? 4 : 8;
    b->nb = nbatch;
    b->grid_prec = precision == FC_FP64 ? FC_FP64 : FC_FP32;
    const uint64_t ncell = (uint64_t)shape[0] * shape[1] * shape[2];
    b->grid_bytes = ncell * (b->grid_prec == FC_FP64 ? 8 : 4);
    const int rs = c->rshape;
    b->npk = (int64_t)c->kshape[0] * c->kshape[1] * (c->kshape[2] / 2 + 1);
    b->nq1 = (int64_t)rs * pi.n[1] * pi.ldc;
    b->nq2 = (int64_t)rs * rs * pi.ldc;
    b->nxs = (int64_t)rs * rs * rs;
    b->input.allocate(b->grid_bytes * nbatch);
    FC_HIP(hipMemsetAsync(b->input.as<void>(), 0, b->grid_bytes * nbatch, ctx->stream));
    b->work.allocate(p.complex_bytes() * nbatch);
    b->pk.allocate(ts * b->npk * nbatch);
    b->q1.allocate(csi * b->nq1 * nbatch);
    b->q2.allocate(csi * b->nq2 * nbatch);
    b->xisub.allocate(tsi * b->nxs * nbatch);
    const int64_t hk = (int64_t)c->nell * c->nkbins, hr = (int64_t)c->nell * c->nrbins;
    b->k_hist.allocate(sizeof(double) * hk * nbatch);
    b->k_counts.allocate(sizeof(int) * hk * nbatch);
    b->r_hist.allocate(sizeof(double) * hr * nbatch);
    b->r_counts.allocate(sizeof(int) * hr * nbatch);
    b->zerolag.allocate(sizeof(double) * nbatch);
    b->part.allocate(sizeof(double) * 3 *
                     (size_t)std::max({c->kbin.nchunks, c->rbin.nchunks, c->khalf.nchunks}) *
                     nbatch);
    switch (c->nell) {
      case 1:
        fc::batch_weights<1>(c);
        break;
      case 2:
        fc::batch_weights<2>(c);
        break;
      default:
        fc::batch_weights<3>(c);
    }
    fc::ensure_phase_tables(c);
    for (int i = 0; i < nbatch; ++i) {
      fc_grid* g = new fc_grid();
      g->ctx = ctx;
      for (int d = 0; d < 3; ++d) g->shape[d] = shape[d];
      g->cell = cell_size;
      g->prec = b->grid_prec;
      g->data.wrap(b->input.as<char>() + b->grid_bytes * i, b->grid_bytes);
      g->batch_view = true;
      b->grids.push_back(g);
    }
    FC_HIP(hipEventCreate(&b->ev0));
    FC_HIP(hipEventCreate(&b->ev1));
    FC_HIP(hipStreamSynchronize(ctx->stream));
    *out = b.release();
  });
}

int fc_batch_destroy(fc_batch* b) {
  return fc::guard([&] {
    if (!b) return;
    fc::DeviceGuard dg(b->corr->ctx->device);
    (void)hipStreamSynchronize(b->corr->ctx->stream);
    delete b;
  });
}

int fc_batch_grid(fc_batch* b, int i, fc_grid** out) {
  return fc::guard([&] {
    if (!b || !out) fc::fail(FC_ERR_INVALID, "null argument");
    if (i < 0 || i >= b->nb) fc::fail(FC_ERR_RANGE, "batch member out of range");
    *out = b->grids[i];
  });
}

int fc_batch_set_graph(fc_batch* b, int on) {
  return fc::guard([&] {
    if (!b) fc::fail(FC_ERR_INVALID, "null batch");
    b->use_graph = on != 0;
  });
}

int fc_batch_autocorrelate(fc_batch* b, int n) {
  return fc::guard([&] {
    if (!b) fc::fail(FC_ERR_INVALID, "null batch");
    if (n < 1 || n > b->nb) fc::fail(FC_ERR_RANGE, "member count out of range");
    fc_corr* c = b->corr;
    fc::DeviceGuard dg(c->ctx->device);
    hipStream_t s = c->ctx->stream;
    FC_HIP(hipEventRecord(b->ev0, s));
    if (b->use_graph) {
      if (!b->exec || b->graph_n != n) {
        if (b->exec) FC_HIP(hipGraphExecDestroy(b->exec));
        if (b->graph) FC_HIP(hipGraphDestroy(b->graph));
        b->exec = nullptr;
        b->graph = nullptr;
        FC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          fc::batch_enqueue(b, n);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        FC_HIP(hipStreamEndCapture(s, &b->graph));
        FC_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_n = n;
      }
      FC_HIP(hipGraphLaunch(b->exec, s));
    } else {
      fc::batch_enqueue(b, n);
    }
    FC_HIP(hipEventRecord(b->ev1, s));
    b->timed = true;
  });
}

int fc_batch_results(fc_batch* b, int i, double* r_hist, int* r_counts, double* k_hist,
                     int* k_counts, double* zerolag) {
  return fc::guard([&] {
    if (!b) fc::fail(FC_ERR_INVALID, "null batch");
    if (i < 0 || i >= b->nb) fc::fail(FC_ERR_RANGE, "batch member out of range");
    fc_corr* c = b->corr;
    fc::DeviceGuard dg(c->ctx->device);
    hipStream_t s = c->ctx->stream;
    const size_t hk = (size_t)c->nell * c->nkbins, hr = (size_t)c->nell * c->nrbins;
    if (r_hist)
      FC_HIP(hipMemcpyAsync(r_hist, b->r_hist.as<double>() + hr * i, sizeof(double) * hr,
                            hipMemcpyDeviceToHost, s));
    if (r_counts)
      FC_HIP(hipMemcpyAsync(r_counts, b->r_counts.as<int>() + hr * i, sizeof(int) * hr,
                            hipMemcpyDeviceToHost, s));
    if (k_hist)
      FC_HIP(hipMemcpyAsync(k_hist, b->k_hist.as<double>() + hk * i, sizeof(double) * hk,
                            hipMemcpyDeviceToHost, s));
    if (k_counts)
      FC_HIP(hipMemcpyAsync(k_counts, b->k_counts.as<int>() + hk * i, sizeof(int) * hk,
                            hipMemcpyDeviceToHost, s));
    if (zerolag)
      FC_HIP(hipMemcpyAsync(zerolag, b->zerolag.as<double>() + i, sizeof(double),
                            hipMemcpyDeviceToHost, s));
    FC_HIP(hipStreamSynchronize(s));
  });
}

int fc_batch_correlator(fc_batch* b, fc_corr** out) {
  return fc::guard([&] {
    if (!b || !out) fc::fail(FC_ERR_INVALID, "null argument");
    *out = b->corr;
  });
}

int fc_batch_last_time_ms(fc_batch* b, double* ms) {
  return fc::guard([&] {
    if (!b || !ms) fc::fail(FC_ERR_INVALID, "null argument");
    if (!b->timed) fc::fail(FC_ERR_RUNTIME, "no batch has run yet");
    fc::DeviceGuard dg(b->corr->ctx->device);
    FC_HIP(hipEventSynchronize(b->ev1));
    float t = 0;
    FC_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = t;
  });
}

}  // extern "C"
