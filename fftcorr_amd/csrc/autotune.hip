// FFT autotuner: times the candidate tile shapes of each pass of a 3D plan on
// the device and records the fastest in the plan cache, the GPU analogue of
// FFTW's measuring planner (FftGrid::plan_fft with FFTW_MEASURE,
// cc/grid/fft_grid.cc:69-115; wisdom export/import, fftcorr/fftw/fftw.pyx).
//
// Tuned per plan: the register-chained plan variant (first radix 8 or 16;
// kinds 4 and 3) with the rows per workgroup of the z pass (kind 1) and the
// columns per workgroup of the y and x passes (kind 0), and the z+y chunk
// size (kind 2).  With FC_MIXED the candidates run the mixed form's kernels
// (FP32 storage, FP64 arithmetic) and the choices go to the FP32 plan, which
// the mixed and FP32 forms share.
// Candidates are timed with HIP events on a private plan and scratch grids,
// so correlators that already hold a plan are not disturbed; plans created
// afterwards pick the tuned shapes up.
#include <algorithm>
#include <vector>

#include "fc_objects.h"
#include "fft_plan.h"

namespace fc {
namespace {

template <typename T>
__global__ void k_fill_pattern(T* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    x[i] = T((double)(h & 0xffff) / 65536.0 - 0.5);
  }
}

struct Timer {
  hipStream_t s;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit Timer(hipStream_t st) : s(st) {
    FC_HIP(hipEventCreate(&e0));
    FC_HIP(hipEventCreate(&e1));
  }
  ~Timer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  template <class F>
  double time(F f, int reps) {
    f();  // warm-up (first launch of a kernel instantiation)
    FC_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) f();
    FC_HIP(hipEventRecord(e1, s));
    FC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    FC_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms / reps;
  }
};

bool col_tile_ok(int L, int tc, size_t cs) {
  const int tcol = fft_len_fast(L) ? L / fft_len_e(L) : 1;
  return tc * tcol <= 1024 && (size_t)L * tc * cs <= kLdsBudget;
}

bool row_tile_ok(int M, int tr, size_t cs) {
  const int tcol = M / fft_len_e(M);
  return tr * tcol <= 1024 && (size_t)tr * (M + 1) * cs <= kLdsBudget;
}

// Arithmetic type A: T, or double for the mixed form (FP32 storage, FP64
// arithmetic in the z and y passes, twiddles from the FP64 plan).
template <typename T, typename A>
void tune(fc_ctx* ctx, const int shape[3], int prec, int reps, double* best_ms) {
  hipStream_t s = ctx->stream;
  constexpr bool wide = !std::is_same<T, A>::value;
  auto p = make_private_plan(ctx->device, shape, prec);
  auto pa = wide ? make_private_plan(ctx->device, shape, FC_FP64) : p;
  const size_t cs = sizeof(cplx<T>);
  const int64_t nreal = (int64_t)shape[0] * shape[1] * shape[2];
  DeviceBuffer grid(sizeof(T) * nreal), work(p->complex_bytes());
  hipLaunchKernelGGL(k_fill_pattern<T>, dim3(grid1d(nreal, 256, 65536)), dim3(256), 0, s,
                     grid.as<T>(), nreal);
  FC_LAUNCH_CHECK();
  cplx<T>* w = work.as<cplx<T>>();
  const T* g = grid.as<T>();
  Timer tm(s);
  double out[4] = {0, 0, 0, 0};

  // z pass: plan variant and rows per workgroup.
  if (p->fast[2]) {
    const int M = shape[2] / 2;
    int best = p->tile[2], bestv = p->var[2];
    double bt = 1e300;
    for (int var = 0; var < kFftVariants; ++var) {
      if (!fft_var_ok(M, var) || (var > 0 && !rc_len_ok(M))) continue;
      for (int tr : {1, 2, 4, 8, 16, 32, 64}) {
        if (!row_tile_ok(M, tr, cs)) continue;
        p->tile[2] = tr;
        p->var[2] = var;
        const double t = tm.time(
            [&] {
              launch_rows_r2c<T, T, LoadCopy, A>(*p, g, shape[2], w, p->ldc, LoadCopy{}, s, -1,
                                                 pa.get());
            },
            reps);
        if (t < bt) bt = t, best = tr, bestv = var;
      }
    }
    p->tile[2] = best;
    p->var[2] = bestv;
    out[0] = bt;
    set_tile_wisdom(M, prec, 1, best);
    set_tile_wisdom(M, prec, 4, bestv);
  }
  // y and x passes: plan variant and columns per workgroup (shared by axes of
  // equal length).
  auto col_pass = [&](int ax, int tc, int var) {
    const int L = shape[ax];
    const ColView<T> v = ax == 1 ? y_view<T>(*p, w) : x_view<T>(*p, w);
    const FftPlan& pt = *pa;
    if constexpr (wide)
      launch_cols_wide<T, -1, A>(L, v, tc, (ax == 1 ? pt.tw1 : pt.tw0).template as<cplx<A>>(),
                                 (ax == 1 ? p->tw1 : p->tw0).template as<cplx<T>>(), s, var);
    else
      launch_cols<T, -1>(L, v, tc, (ax == 1 ? p->tw1 : p->tw0).template as<cplx<T>>(), s, var);
  };
  for (int ax = 1; ax >= 0; --ax) {
    const int L = shape[ax];
    if (ax == 0 && shape[0] == shape[1]) {
      p->tile[0] = p->tile[1];
      p->var[0] = p->var[1];
      out[2] = out[1];
      continue;
    }
    int best = p->tile[ax], bestv = p->var[ax];
    double bt = 1e300;
    for (int var = 0; var < kFftVariants; ++var) {
      if (!fft_var_ok(L, var)) continue;
      for (int tc : {1, 2, 4, 8, 16, 32}) {
        if (!col_tile_ok(L, tc, cs)) continue;
        if (var > 0 && !(tc == 4 || tc == 8 || tc == 16)) continue;  // register-chained only
        p->tile[ax] = tc;
        p->var[ax] = var;
        const double t = tm.time([&] { col_pass(ax, tc, var); }, reps);
        if (t < bt) bt = t, best = tc, bestv = var;
      }
    }
    p->tile[ax] = best;
    p->var[ax] = bestv;
    out[2 - ax] = bt;
    set_tile_wisdom(L, prec, 0, best);
    set_tile_wisdom(L, prec, 3, bestv);
  }
  // z+y chunking.
  {
    const size_t slab = (size_t)shape[1] * p->ldc * cs;
    int best = 0;
    double bt = 1e300;
    for (int kib : {0, 16384, 32768, 65536, 131072, 262144}) {
      int X = kib ? (int)std::max<size_t>(1, ((size_t)kib << 10) / slab) : shape[0];
      if (X > shape[0]) X = shape[0];
      if (kib && X == shape[0]) continue;
      const double t = tm.time(
          [&] {
            for (int x0 = 0; x0 < shape[0]; x0 += X) {
              const int nx = std::min(X, shape[0] - x0);
              cplx<T>* d = w + (int64_t)x0 * shape[1] * p->ldc;
              launch_rows_r2c<T, T, LoadCopy, A>(*p, g + (int64_t)x0 * shape[1] * shape[2],
                                                 shape[2], d, p->ldc, LoadCopy{}, s,
                                                 (int64_t)nx * shape[1], pa.get());
              ColView<T> v = y_view<T>(*p, d);
              v.A = nx;
              if constexpr (wide)
                launch_cols_wide<T, -1, A>(shape[1], v, p->tile[1], pa->tw1.as<cplx<A>>(),
                                           p->tw1.as<cplx<T>>(), s, p->var[1]);
              else
                launch_cols<T, -1>(shape[1], v, p->tile[1], p->tw1.as<cplx<T>>(), s, p->var[1]);
            }
          },
          reps);
      if (t < bt) bt = t, best = kib;
    }
    out[3] = bt;
    set_tile_wisdom(0, prec, 2, best);
  }
  FC_HIP(hipStreamSynchronize(s));
  if (best_ms)
    for (int i = 0; i < 4; ++i) best_ms[i] = out[i];
}

}  // namespace
}  // namespace fc

extern "C" {

int fc_autotune(fc_ctx* ctx, const int shape[3], int precision, int reps, double best_ms[4]) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!shape) fc::fail(FC_ERR_INVALID, "null shape");
    if (reps < 1) fc::fail(FC_ERR_INVALID, "reps must be positive");
    if (precision != FC_FP64 && precision != FC_FP32 && precision != FC_MIXED)
      fc::fail(FC_ERR_INVALID, "unknown precision");
    fc::DeviceGuard dg(ctx->device);
    if (precision == FC_FP64)
      fc::tune<double, double>(ctx, shape, FC_FP64, reps, best_ms);
    else if (precision == FC_FP32)
      fc::tune<float, float>(ctx, shape, FC_FP32, reps, best_ms);
    else  // the mixed form's kernels; its choices are the FP32 plan's
      fc::tune<float, double>(ctx, shape, FC_FP32, reps, best_ms);
  });
}

int fc_plan_tiles(int device, const int shape[3], int precision, int tiles[4]) {
  return fc::guard([&] {
    if (!shape || !tiles) fc::fail(FC_ERR_INVALID, "null argument");
    if (precision != FC_FP64 && precision != FC_FP32 && precision != FC_MIXED)
      fc::fail(FC_ERR_INVALID, "unknown precision");
    fc::DeviceGuard dg(device);
    auto p = fc::make_private_plan(device, shape, precision == FC_FP64 ? FC_FP64 : FC_FP32);
    for (int i = 0; i < 3; ++i) tiles[i] = p->tile[i];
    tiles[3] = p->zy_chunk_kib;
  });
}

int fc_plan_variants(int device, const int shape[3], int precision, int vars[3]) {
  return fc::guard([&] {
    if (!shape || !vars) fc::fail(FC_ERR_INVALID, "null argument");
    if (precision != FC_FP64 && precision != FC_FP32 && precision != FC_MIXED)
      fc::fail(FC_ERR_INVALID, "unknown precision");
    fc::DeviceGuard dg(device);
    auto p = fc::make_private_plan(device, shape, precision == FC_FP64 ? FC_FP64 : FC_FP32);
    for (int i = 0; i < 3; ++i) vars[i] = p->var[i];
  });
}

}  // extern "C"
