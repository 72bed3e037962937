// Host side of the 3D R2C/C2R FFT: plan (twiddle tables + tile shapes per
// axis) and launchers.  A plan replaces FftGrid::plan_fft
// (cc/grid/fft_grid.cc:69-115); the execute calls replace execute_fft /
// execute_ifft (:125-178), split into the z (rows) and y/x (columns) passes so
// callers can fuse their own loads into the z pass.
#pragma once

#include <algorithm>
#include <atomic>
#include <type_traits>

#include "fc_internal.h"
#include "fft_kernels.h"

namespace fc {

constexpr int kMaxFftLen = 4096;
// Default z+y chunk (KiB of complex output per chunk; 0 = no chunking).
constexpr int kDefaultZyChunkKiB = 0;
constexpr size_t kLdsBudget = 65536;  // bytes of LDS per workgroup we allow

// Tile shape chosen for one transform axis (autotuned; see fft.hip).
struct AxisTiles {
  int tc = 0;  // columns (or rows) per workgroup
};

struct FftPlan {
  int device = 0;
  int n[3] = {0, 0, 0};
  int n2h = 0;
  // Complex row stride of the work grid: n2/2+1 rounded up to 128 bytes so
  // every row (and every tile of TC columns) starts on a 128-byte boundary.
  int ldc = 0;
  int prec = FC_FP64;
  bool fast[3] = {false, false, false};  // axis 2: the half-length row FFT
  int tile[3] = {0, 0, 0};
  // Register-chained plan variant per axis (FftCfgV; axis 2: the row FFT of
  // length n2/2) of the forward passes, which the autotuner times; 0 unless
  // it or imported wisdom chose another.  Inverse passes use variant 0.
  int var[3] = {0, 0, 0};
  // x slabs per z+y chunk: the y pass of a chunk reads the z pass's output
  // while it is still in the last-level cache (0: the whole grid at once).
  int zy_chunk = 0;
  int zy_chunk_kib = 0;
  // Twiddle tables exp(-2 pi i m / L) in the plan's precision.
  DeviceBuffer tw0, tw1;      // lengths n0, n1
  DeviceBuffer tw2M, tw2N;    // lengths n2/2 (fast rows) and n2 (post-process / DFT)
  size_t csize() const { return prec == FC_FP64 ? sizeof(double2) : sizeof(float2); }
  int64_t nrows() const { return (int64_t)n[0] * n[1]; }
  size_t complex_bytes() const { return (size_t)n[0] * n[1] * ldc * csize(); }
};

bool fft_len_fast(int L);
int fft_len_e(int L);
// Whether the register-chained kernels have plan variant var for length L.
bool fft_var_ok(int L, int var);

// Plan cache keyed by (device, shape, precision).
std::shared_ptr<FftPlan> get_fft_plan(int device, const int shape[3], int prec);
// An uncached plan built with the current wisdom (the autotuner's scratch plan).
std::shared_ptr<FftPlan> make_private_plan(int device, const int shape[3], int prec);
void set_tile_wisdom(int L, int prec, int kind, int tile);
int default_col_tile(int L, size_t cs);
int default_row_tile(int M, size_t cs);

// ---- launchers -------------------------------------------------------------
// Column view [A][L][G][B] for k_fft_cols / k_dft_cols (see fft_kernels.h).
template <typename T, typename Tin = T>
struct ColView {
  const cplx<Tin>* in;
  int64_t inA, inL, inG;
  cplx<T>* out;
  int64_t outA, outL, outG;
  int A, G, B;
  int nout = 0, off = 0;  // nout == 0: all L outputs
};

template <typename T, int DIR, typename Tin, int TC, typename A, int Lv, int V>
void launch_cols_rc_v(const ColView<T, Tin>& v, int64_t nb, int tilesB, int nout, int offm,
                      const cplx<A>* tw, hipStream_t s) {
  hipLaunchKernelGGL((k_fft_cols_rc<A, Lv, DIR, TC, Tin, T, V>), dim3((unsigned)nb),
                     dim3(TC * (Lv / FftCfgV<Lv, V>::E)), (size_t)Lv * TC * sizeof(cplx<T>), s,
                     v.in, v.inA, v.inL, v.inG, v.out, v.outA, v.outL, v.outG, v.B, tilesB, v.G,
                     nout, offm, tw);
}

// Compute units of the current device (queried once per device).
inline int current_device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  FC_HIP(hipGetDevice(&dev));
  int n = dev >= 0 && dev < 64 ? cache[dev].load(std::memory_order_relaxed) : 0;
  if (n <= 0) {
    FC_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    if (dev >= 0 && dev < 64) cache[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

// Whether k_fft_cols_pipe takes a column view: one group per slab, all L
// outputs in order, each tile's 128-byte row segments inside the row pitch
// (the rows of the work-grid layouts are allocated at full pitch), and its
// 32-bit in-tile offsets in range.
template <typename T, typename Tin>
bool pipe_view_ok(const ColView<T, Tin>& v, int L, int tc, int nout, int offm) {
  const uint64_t cs = sizeof(cplx<T>);
  return std::is_same<T, Tin>::value && v.G == 1 && nout == L && offm == 0 &&
         (int64_t)((v.B + tc - 1) / tc) * tc <= v.inL && (uint64_t)v.inL * cs * 8 < (1ull << 32) &&
         (uint64_t)v.outL * cs * 64 < (1ull << 32);
}

// The persistent prefetching column pass (k_fft_cols_pipe): one workgroup
// per CU (its two tiles and the twiddles take most of the CU's LDS).
template <typename T, int DIR, int TC, typename A, int Lv>
void launch_cols_pipe(const ColView<T, T>& v, int tilesB, int64_t nb, const cplx<A>* tw,
                      hipStream_t s) {
  static_assert(2 * pipe_lds_bytes<Lv, TC, T>() > 80 * 1024, "one workgroup per CU");
  const int grid = (int)std::min<int64_t>(nb, current_device_cus());
  hipLaunchKernelGGL((k_fft_cols_pipe<A, Lv, DIR, TC, T>), dim3((unsigned)grid),
                     dim3(TC * (Lv / 8)), 0, s, v.in, v.inA, v.inL, v.out, v.outA, v.outL, v.B,
                     tilesB, (int)nb, tw);
}

// Register-chained column pass with plan variant var: arithmetic in A
// (twiddles tw in A), loads Tin, LDS and stores in T.  Full 512-point passes
// over FP32 storage (16-column tiles, plan variant 0) run as k_fft_cols_pipe;
// with FP64 storage its 8-column tiles leave one 8-wave workgroup per CU,
// which measured no faster than two workgroups of the plain pass.
template <typename T, int DIR, typename Tin, int TC, typename A = T>
void launch_cols_rc(int L, const ColView<T, Tin>& v, const cplx<A>* tw, hipStream_t s,
                    int var = 0) {
  const int nout = v.nout ? v.nout : L;
  const int offm = ((v.off % L) + L) % L;
  const int tilesB = (v.B + TC - 1) / TC;
  const int64_t nb = (int64_t)v.A * v.G * tilesB;
  if (nb > INT32_MAX) fail(FC_ERR_INVALID, "transform too large for one launch");
  if constexpr (std::is_same<T, Tin>::value && std::is_same<T, float>::value &&
                TC * sizeof(cplx<T>) == 128) {
    if (L == 512 && var == 0 && pipe_view_ok(v, L, TC, nout, offm)) {
      launch_cols_pipe<T, DIR, TC, A, 512>(v, tilesB, nb, tw, s);
      FC_LAUNCH_CHECK();
      return;
    }
  }
#define FC_COLS_RC_CASE(Lv)                                                              \
  case Lv:                                                                               \
    if constexpr (FftCfgV<Lv, 1>::fast) {                                                \
      if (var == 1) {                                                                    \
        launch_cols_rc_v<T, DIR, Tin, TC, A, Lv, 1>(v, nb, tilesB, nout, offm, tw, s);   \
        break;                                                                           \
      }                                                                                  \
    }                                                                                    \
    launch_cols_rc_v<T, DIR, Tin, TC, A, Lv, 0>(v, nb, tilesB, nout, offm, tw, s);       \
    break;
  switch (L) {
    FC_COLS_RC_CASE(8)
    FC_COLS_RC_CASE(16)
    FC_COLS_RC_CASE(32)
    FC_COLS_RC_CASE(64)
    FC_COLS_RC_CASE(128)
    FC_COLS_RC_CASE(256)
    FC_COLS_RC_CASE(512)
    default:
      fail(FC_ERR_RUNTIME, "internal: no register-chained column kernel for this length");
  }
#undef FC_COLS_RC_CASE
  FC_LAUNCH_CHECK();
}

template <typename T, int DIR, typename Tin = T>
void launch_cols(int L, const ColView<T, Tin>& v, int tc, const cplx<T>* tw, hipStream_t s,
                 int var = 0) {
  const size_t cs = sizeof(cplx<T>);
  const int nout = v.nout ? v.nout : L;
  if (rc_len_ok(L) && (size_t)L * tc * cs <= kLdsBudget && tc * (L / 8) <= 1024 &&
      rc_offsets_fit<T, Tin>(L, tc, v.inL, v.outL, nout)) {
    const int vr = fft_var_ok(L, var) ? var : 0;
    if (tc == 16) return launch_cols_rc<T, DIR, Tin, 16>(L, v, tw, s, vr);
    if (tc == 8) return launch_cols_rc<T, DIR, Tin, 8>(L, v, tw, s, vr);
    if (tc == 4) return launch_cols_rc<T, DIR, Tin, 4>(L, v, tw, s, vr);
  }
#define FC_COLS_CASE(Lv)                                                                  \
  case Lv: {                                                                              \
    const int thr = tc * (Lv / FftCfg<Lv>::E);                                            \
    const int tilesB = (v.B + tc - 1) / tc;                                               \
    dim3 g((unsigned)(v.G * tilesB), v.A);                                                \
    hipLaunchKernelGGL((k_fft_cols<T, Lv, DIR, Tin>), g, dim3(thr), (size_t)Lv * tc * cs, s,   \
                       v.in, v.inA, v.inL, v.inG, v.out, v.outA, v.outL, v.outG, v.B,     \
                       tilesB, tc, nout, v.off, tw);                                      \
    break;                                                                                \
  }
  switch (L) {
    FC_COLS_CASE(8)
    FC_COLS_CASE(16)
    FC_COLS_CASE(32)
    FC_COLS_CASE(64)
    FC_COLS_CASE(128)
    FC_COLS_CASE(256)
    FC_COLS_CASE(512)
    FC_COLS_CASE(1024)
    FC_COLS_CASE(2048)
    FC_COLS_CASE(4096)
    FC_COLS_CASE(24)
    FC_COLS_CASE(48)
    FC_COLS_CASE(96)
    FC_COLS_CASE(192)
    FC_COLS_CASE(384)
    FC_COLS_CASE(768)
    FC_COLS_CASE(1536)
    FC_COLS_CASE(3072)
    default: {
      dim3 g((unsigned)(v.G * v.B), v.A);
      hipLaunchKernelGGL((k_dft_cols<T, DIR, Tin>), g, dim3(256), (size_t)L * cs, s, v.in, v.inA,
                         v.inL, v.inG, v.out, v.outA, v.outL, v.outG, v.B, L, nout, v.off, tw);
    }
  }
#undef FC_COLS_CASE
  FC_LAUNCH_CHECK();
}

// Column pass stored in T with arithmetic in the wider A where the
// register-chained kernel takes L (tw_a: twiddles in A), else in T (tw_t).
template <typename T, int DIR, typename A>
void launch_cols_wide(int L, const ColView<T>& v, int tc, const cplx<A>* tw_a,
                      const cplx<T>* tw_t, hipStream_t s, int var = 0) {
  if constexpr (!std::is_same<T, A>::value) {
    int t = tc;
    while (t > 4 && ((size_t)L * t * sizeof(cplx<T>) > kLdsBudget || t * (L / 8) > 1024)) t /= 2;
    if (rc_len_ok(L) && (size_t)L * t * sizeof(cplx<T>) <= kLdsBudget && t * (L / 8) <= 1024 &&
        rc_offsets_fit<T, T>(L, t, v.inL, v.outL, v.nout ? v.nout : L)) {
      const int vr = fft_var_ok(L, var) ? var : 0;
      if (t >= 16) return launch_cols_rc<T, DIR, T, 16, A>(L, v, tw_a, s, vr);
      if (t == 8) return launch_cols_rc<T, DIR, T, 8, A>(L, v, tw_a, s, vr);
      return launch_cols_rc<T, DIR, T, 4, A>(L, v, tw_a, s, vr);
    }
  }
  launch_cols<T, DIR>(L, v, tc, tw_t, s, var);
}

// In-place y pass over the work grid [n0][n1][ldc].
template <typename T>
ColView<T> y_view(const FftPlan& p, cplx<T>* d) {
  const int64_t ldc = p.ldc;
  return ColView<T>{d, (int64_t)p.n[1] * ldc, ldc, 0, d, (int64_t)p.n[1] * ldc, ldc, 0,
                    p.n[0], 1, p.n2h};
}
// In-place x pass over the work grid: L = n0, groups = the n1 rows of each slab.
template <typename T>
ColView<T> x_view(const FftPlan& p, cplx<T>* d) {
  const int64_t ldc = p.ldc;
  return ColView<T>{d, 0, (int64_t)p.n[1] * ldc, ldc, d, 0, (int64_t)p.n[1] * ldc, ldc,
                    1, p.n[1], p.n2h};
}

// z pass (R2C rows) stored in T.  With A wider than T, the register-chained
// kernel computes in A with the twiddles of the A-precision plan *pa (same
// shape), exchanging through LDS in T; lengths it does not take compute in T.
// in2 / nrows1 (optional): rows from nrows1 on are read from in2 (a second
// grid of the same shape), register-chained row lengths only.
template <typename T, typename Tin, class Loader, typename A = T>
void launch_rows_r2c(const FftPlan& p, const Tin* in, int64_t ldi, cplx<T>* out, int64_t ldo,
                     Loader ld, hipStream_t s, int64_t nrows = -1, const FftPlan* pa = nullptr,
                     const Tin* in2 = nullptr, int64_t nrows1 = 0) {
  const size_t cs = sizeof(cplx<T>);
  if (nrows < 0) nrows = p.nrows();
  const int M = p.n[2] / 2;
  const int tr = p.tile[2];
  const cplx<T>* twM = p.tw2M.as<cplx<T>>();
  const cplx<T>* twN = p.tw2N.as<cplx<T>>();
  if (in2 && !(p.fast[2] && rc_len_ok(M)))
    fail(FC_ERR_RUNTIME, "internal: two-grid z pass needs the register-chained rows");
  if (p.fast[2] && rc_len_ok(M)) {
    const FftPlan& pt = std::is_same<T, A>::value ? p : *pa;
    const cplx<A>* twMa = pt.tw2M.as<cplx<A>>();
    const cplx<A>* twNa = pt.tw2N.as<cplx<A>>();
    const int var = fft_var_ok(M, p.var[2]) ? p.var[2] : 0;
    // Wider arithmetic: at most two waves per workgroup (4 rows of M = 256):
    // small workgroups wait at few barriers of other waves (the 512^3 mixed
    // z pass: 0.375 ms with two waves against 0.384 with one; four waves
    // were slower than one in an earlier form of the kernel).
    const int cap = std::max(1, 128 / (M / (var == 1 ? 16 : 8)));
    int trc = std::is_same<T, A>::value ? tr : (tr < cap ? tr : cap);
    while (trc > 1 && (size_t)trc * row_pad_len(M) * cs > kLdsBudget) trc /= 2;
    auto go = [&](auto vtag, auto mtag) {
      constexpr int V = decltype(vtag)::value, Mv = decltype(mtag)::value;
      dim3 g((unsigned)((nrows + trc - 1) / trc));
      hipLaunchKernelGGL((k_fft_rows_r2c_rc<A, Tin, Mv, Loader, T, V>), g,
                         dim3(trc * (Mv / FftCfgV<Mv, V>::E)),
                         (size_t)trc * row_pad_len(Mv) * cs, s, in, ldi, out, ldo, nrows, trc,
                         twMa, twNa, ld, in2, nrows1);
    };
#define FC_R2C_RC_CASE(Mv)                                                       \
  case Mv:                                                                       \
    if constexpr (FftCfgV<Mv, 1>::fast) {                                        \
      if (var == 1) {                                                            \
        go(std::integral_constant<int, 1>{}, std::integral_constant<int, Mv>{}); \
        break;                                                                   \
      }                                                                          \
    }                                                                            \
    go(std::integral_constant<int, 0>{}, std::integral_constant<int, Mv>{});     \
    break;
    switch (M) {
      FC_R2C_RC_CASE(8)
      FC_R2C_RC_CASE(16)
      FC_R2C_RC_CASE(32)
      FC_R2C_RC_CASE(64)
      FC_R2C_RC_CASE(128)
      FC_R2C_RC_CASE(256)
      FC_R2C_RC_CASE(512)
    }
#undef FC_R2C_RC_CASE
  } else if (p.fast[2]) {
#define FC_R2C_CASE(Mv)                                                                     \
  case Mv: {                                                                                \
    const int thr = tr * (Mv / FftCfg<Mv>::E);                                              \
    dim3 g((unsigned)((nrows + tr - 1) / tr));                                              \
    hipLaunchKernelGGL((k_fft_rows_r2c<T, Tin, Mv, Loader>), g, dim3(thr),                  \
                       (size_t)tr * (Mv + 1) * cs, s, in, ldi, out, ldo, nrows, tr, twM,     \
                       twN, ld);                                                            \
    break;                                                                                  \
  }
    switch (M) {
      FC_R2C_CASE(1024)
      FC_R2C_CASE(2048)
      FC_R2C_CASE(24)
      FC_R2C_CASE(48)
      FC_R2C_CASE(96)
      FC_R2C_CASE(192)
      FC_R2C_CASE(384)
      FC_R2C_CASE(768)
      FC_R2C_CASE(1536)
      default:
        fail(FC_ERR_RUNTIME, "internal: no fast row kernel for this length");
    }
#undef FC_R2C_CASE
  } else {
    dim3 g((unsigned)nrows);
    hipLaunchKernelGGL((k_dft_rows_r2c<T, Tin, Loader>), g, dim3(256),
                       (size_t)p.n[2] * sizeof(T), s, in, ldi, out, ldo, p.n[2], twN, ld);
  }
  FC_LAUNCH_CHECK();
}

// nrows rows of ldi complex -> real rows of ldo; nsel > 0 keeps only samples
// (o - off) mod n2, o < nsel (pruned inverse).
template <typename T, typename Tout>
void launch_rows_c2r(const FftPlan& p, const cplx<T>* in, int64_t ldi, Tout* out, int64_t ldo,
                     hipStream_t s, int64_t nrows = -1, int nsel = 0, int off = 0) {
  const size_t cs = sizeof(cplx<T>);
  if (nrows < 0) nrows = p.nrows();
  const int M = p.n[2] / 2;
  const int tr = p.tile[2];
  const cplx<T>* twM = p.tw2M.as<cplx<T>>();
  const cplx<T>* twN = p.tw2N.as<cplx<T>>();
  if (p.fast[2] && rc_len_ok(M)) {
    int trc = tr;
    while (trc > 1 && (size_t)trc * row_pad_len(M) * cs > kLdsBudget) trc /= 2;
    const int offm = nsel ? ((off % (2 * M)) + 2 * M) % (2 * M) : 0;
    // The radix-8 plans (plan variants are tuned on the forward passes).
#define FC_C2R_RC_CASE(Mv)                                                                   \
  case Mv: {                                                                                 \
    dim3 g((unsigned)((nrows + trc - 1) / trc));                                             \
    hipLaunchKernelGGL((k_fft_rows_c2r_rc<T, Tout, Mv>), g, dim3(trc * (Mv / FftCfg<Mv>::E)), \
                       (size_t)trc * row_pad_len(Mv) * cs, s, in, ldi, out, ldo, nrows, trc,  \
                       twM, twN, nsel, offm);                                                \
    break;                                                                                   \
  }
    switch (M) {
      FC_C2R_RC_CASE(8)
      FC_C2R_RC_CASE(16)
      FC_C2R_RC_CASE(32)
      FC_C2R_RC_CASE(64)
      FC_C2R_RC_CASE(128)
      FC_C2R_RC_CASE(256)
      FC_C2R_RC_CASE(512)
    }
#undef FC_C2R_RC_CASE
  } else if (p.fast[2]) {
#define FC_C2R_CASE(Mv)                                                                    \
  case Mv: {                                                                               \
    const int thr = tr * (Mv / FftCfg<Mv>::E);                                             \
    dim3 g((unsigned)((nrows + tr - 1) / tr));                                             \
    hipLaunchKernelGGL((k_fft_rows_c2r<T, Tout, Mv>), g, dim3(thr),                        \
                       (size_t)tr * (Mv + 1) * cs, s, in, ldi, out, ldo, nrows, tr, twM,    \
                       twN, nsel, off);                                                    \
    break;                                                                                 \
  }
    switch (M) {
      FC_C2R_CASE(1024)
      FC_C2R_CASE(2048)
      FC_C2R_CASE(24)
      FC_C2R_CASE(48)
      FC_C2R_CASE(96)
      FC_C2R_CASE(192)
      FC_C2R_CASE(384)
      FC_C2R_CASE(768)
      FC_C2R_CASE(1536)
      default:
        fail(FC_ERR_RUNTIME, "internal: no fast row kernel for this length");
    }
#undef FC_C2R_CASE
  } else {
    dim3 g((unsigned)nrows);
    const int nh = p.n[2] / 2 + 1;
    hipLaunchKernelGGL((k_dft_rows_c2r<T, Tout>), g, dim3(256), (size_t)nh * cs, s, in, ldi,
                       out, ldo, p.n[2], twN, nsel, off);
  }
  FC_LAUNCH_CHECK();
}

// y then x forward passes over the complex work grid [n0][n1][ldc].
template <typename T>
void fft_yx_forward(const FftPlan& p, cplx<T>* data, hipStream_t s) {
  launch_cols<T, -1>(p.n[1], y_view<T>(p, data), p.tile[1], p.tw1.as<cplx<T>>(), s, p.var[1]);
  launch_cols<T, -1>(p.n[0], x_view<T>(p, data), p.tile[0], p.tw0.as<cplx<T>>(), s, p.var[0]);
}

// x then y inverse passes.
template <typename T>
void fft_xy_inverse(const FftPlan& p, cplx<T>* data, hipStream_t s) {
  launch_cols<T, +1>(p.n[0], x_view<T>(p, data), p.tile[0], p.tw0.as<cplx<T>>(), s);
  launch_cols<T, +1>(p.n[1], y_view<T>(p, data), p.tile[1], p.tw1.as<cplx<T>>(), s);
}

}  // namespace fc
