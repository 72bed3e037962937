// FFT plans (twiddle tables, tile shapes), the plan cache behind the
// wisdom functions (fftcorr/fftw/fftw.pyx:1-14), and the standalone
// fc_fft_r2c / fc_fft_c2r entry points.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <algorithm>

#include "fft_plan.h"

namespace fc {

bool fft_len_fast(int L) {
  switch (L) {
    case 8: case 16: case 32: case 64: case 128: case 256: case 512: case 1024:
    case 2048: case 4096: case 24: case 48: case 96: case 192: case 384: case 768:
    case 1536: case 3072:
      return true;
    default:
      return false;
  }
}

int fft_len_e(int L) { return (L % 3 == 0) ? 24 : 8; }

namespace {
template <int... Rs>
int copy_radices(Radices<Rs...>, int* out) {
  const int r[] = {Rs...};
  for (int i = 0; i < (int)sizeof...(Rs); ++i) out[i] = r[i];
  return (int)sizeof...(Rs);
}
template <int L>
int radices_of(int var, int* out) {
  if (var == 0) return copy_radices(typename FftCfgV<L, 0>::radices{}, out);
  if constexpr (FftCfgV<L, 1>::fast)
    if (var == 1) return copy_radices(typename FftCfgV<L, 1>::radices{}, out);
  return 0;
}
}  // namespace

int fft_radices(int L, int var, int* out) {
  switch (L) {
    case 8: return radices_of<8>(var, out);
    case 16: return radices_of<16>(var, out);
    case 32: return radices_of<32>(var, out);
    case 64: return radices_of<64>(var, out);
    case 128: return radices_of<128>(var, out);
    case 256: return radices_of<256>(var, out);
    case 512: return radices_of<512>(var, out);
    default: return 0;
  }
}

bool fft_var_ok(int L, int var) {
  if (var == 0) return true;
  if (var != 1 || !rc_len_ok(L)) return false;
  switch (L) {
    case 8: return FftCfgV<8, 1>::fast;
    case 16: return FftCfgV<16, 1>::fast;
    case 32: return FftCfgV<32, 1>::fast;
    case 64: return FftCfgV<64, 1>::fast;
    case 128: return FftCfgV<128, 1>::fast;
    case 256: return FftCfgV<256, 1>::fast;
    case 512: return FftCfgV<512, 1>::fast;
    default: return false;
  }
}

namespace {

struct PlanKey {
  int device, n0, n1, n2, prec;
  bool operator<(const PlanKey& o) const {
    return std::tie(device, n0, n1, n2, prec) < std::tie(o.device, o.n0, o.n1, o.n2, o.prec);
  }
};

std::mutex g_plan_mu;
std::map<PlanKey, std::shared_ptr<FftPlan>> g_plans;
// Tile overrides (imported wisdom or the autotuner): (axis length, precision,
// kind) -> tile.  kind 0 = column pass (columns per workgroup), 1 = row pass
// (rows per workgroup), 2 = z+y chunk size in KiB (length 0), 3 = column
// pass plan variant, 4 = row pass plan variant.
std::map<std::tuple<int, int, int>, int> g_tile_wisdom;

// Stage tables of every plan variant of length L, in variant order (see
// stage_twiddle_base): put(index after the plain part, plain index m).
template <int L, class Put>
int stage_entries(Put put) {
  stage_twiddles_<L>(typename FftCfgV<L, 0>::radices{}, put);
  const int b1 = stage_twiddle_base<L, 1>();
  if constexpr (FftCfgV<L, 1>::fast)
    stage_twiddles_<L>(typename FftCfgV<L, 1>::radices{},
                       [&](int idx, int m) { put(b1 + idx, m); });
  return stage_twiddle_base<L, kFftVariants>();
}

template <class Put>
int stage_entries_for(int L, Put put) {
  switch (L) {
    case 8: return stage_entries<8>(put);
    case 16: return stage_entries<16>(put);
    case 32: return stage_entries<32>(put);
    case 64: return stage_entries<64>(put);
    case 128: return stage_entries<128>(put);
    case 256: return stage_entries<256>(put);
    case 512: return stage_entries<512>(put);
    default: return 0;
  }
}

// Plain table exp(-2 pi i m / L), m < L, followed (register-chained lengths)
// by the per-stage tables of the same values regrouped per thread.
template <typename T>
void upload_twiddles(DeviceBuffer& buf, int L) {
  const int ns = rc_len_ok(L) ? stage_entries_for(L, [](int, int) {}) : 0;
  std::vector<cplx<T>> h(L + ns);
  for (int m = 0; m < L; ++m) {
    long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / L;
    h[m].x = (T)cosl(a);
    h[m].y = (T)(-sinl(a));
  }
  if (ns) stage_entries_for(L, [&](int idx, int m) { h[L + idx] = h[m]; });
  buf.allocate(sizeof(cplx<T>) * h.size());
  FC_HIP(hipMemcpy(buf.as<void>(), h.data(), sizeof(cplx<T>) * h.size(), hipMemcpyHostToDevice));
}

}  // namespace

int default_col_tile(int L, size_t cs) {
  // Runs of 128 bytes per global access: 16 FP32 / 8 FP64 complex columns.
  int tc = cs == sizeof(float2) ? 16 : 8;
  const int tcol = fft_len_fast(L) ? L / fft_len_e(L) : 1;
  while (tc > 1 && (tc * tcol > 1024 || (size_t)L * tc * cs > kLdsBudget)) tc /= 2;
  return tc;
}

int default_row_tile(int M, size_t cs) {
  int tr = 16;
  const int tcol = M / fft_len_e(M);
  while (tr > 1 && (tr * tcol > 1024 || (size_t)tr * (M + 1) * cs > kLdsBudget)) tr /= 2;
  // Keep at least a 64-lane wave per block where rows are short.
  while (tr * tcol < 64 && (size_t)(tr * 2) * (M + 1) * cs <= kLdsBudget && tr < 64) tr *= 2;
  return tr;
}


namespace {
// Builds a plan with the current wisdom (caller holds g_plan_mu).
std::shared_ptr<FftPlan> build_plan(int device, const int shape[3], int prec) {
  auto p = std::make_shared<FftPlan>();
  p->device = device;
  for (int i = 0; i < 3; ++i) p->n[i] = shape[i];
  p->n2h = shape[2] / 2 + 1;
  p->prec = prec;
  const size_t cs = p->csize();
  {
    const int align = (int)(128 / cs);
    p->ldc = (p->n2h + align - 1) / align * align;
  }
  p->fast[0] = fft_len_fast(shape[0]);
  p->fast[1] = fft_len_fast(shape[1]);
  p->fast[2] = (shape[2] % 2 == 0) && fft_len_fast(shape[2] / 2) && shape[2] / 2 <= 2048;
  // The direct-DFT fallbacks stage a whole column / row in LDS.
  for (int i = 0; i < 2; ++i)
    if (!p->fast[i] && (size_t)shape[i] * cs > kLdsBudget)
      fail(FC_ERR_INVALID, "FFT length " + std::to_string(shape[i]) +
                               " needs a 2^a or 3*2^a size at this precision");
  auto tile_for = [&](int L, int kind, int dflt) {
    auto w = g_tile_wisdom.find(std::make_tuple(L, kind == 2 ? 0 : prec, kind));
    return w != g_tile_wisdom.end() ? w->second : dflt;
  };
  p->tile[0] = tile_for(shape[0], 0, default_col_tile(shape[0], cs));
  p->tile[1] = tile_for(shape[1], 0, default_col_tile(shape[1], cs));
  p->tile[2] = p->fast[2] ? tile_for(shape[2] / 2, 1, default_row_tile(shape[2] / 2, cs)) : 1;
  for (int ax = 0; ax < 2; ++ax) {
    const int v = tile_for(shape[ax], 3, 0);
    p->var[ax] = fft_var_ok(shape[ax], v) ? v : 0;
  }
  if (p->fast[2]) {
    const int v = tile_for(shape[2] / 2, 4, 0);
    p->var[2] = fft_var_ok(shape[2] / 2, v) ? v : 0;
  }
  p->zy_chunk_kib = tile_for(0, 2, kDefaultZyChunkKiB);
  {
    const size_t slab = (size_t)shape[1] * p->ldc * cs;
    const size_t want = (size_t)p->zy_chunk_kib << 10;
    p->zy_chunk = want ? (int)std::max<size_t>(1, want / slab) : 0;
    if (p->zy_chunk >= shape[0]) p->zy_chunk = 0;
  }
  DeviceGuard g(device);
  if (prec == FC_FP64) {
    upload_twiddles<double>(p->tw0, shape[0]);
    upload_twiddles<double>(p->tw1, shape[1]);
    upload_twiddles<double>(p->tw2N, shape[2]);
    if (p->fast[2]) upload_twiddles<double>(p->tw2M, shape[2] / 2);
  } else {
    upload_twiddles<float>(p->tw0, shape[0]);
    upload_twiddles<float>(p->tw1, shape[1]);
    upload_twiddles<float>(p->tw2N, shape[2]);
    if (p->fast[2]) upload_twiddles<float>(p->tw2M, shape[2] / 2);
  }
  return p;
}

void check_plan_args(const int shape[3], int prec) {
  for (int i = 0; i < 3; ++i) {
    if (shape[i] <= 0) fail(FC_ERR_INVALID, "FFT shape must be positive");
    if (shape[i] > kMaxFftLen)
      fail(FC_ERR_INVALID, "FFT length " + std::to_string(shape[i]) + " exceeds " +
                               std::to_string(kMaxFftLen));
  }
  if (prec != FC_FP64 && prec != FC_FP32) fail(FC_ERR_INVALID, "unknown precision");
}
}  // namespace

std::shared_ptr<FftPlan> get_fft_plan(int device, const int shape[3], int prec) {
  check_plan_args(shape, prec);
  PlanKey key{device, shape[0], shape[1], shape[2], prec};
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) return it->second;
  auto p = build_plan(device, shape, prec);
  g_plans[key] = p;
  return p;
}

std::shared_ptr<FftPlan> make_private_plan(int device, const int shape[3], int prec) {
  check_plan_args(shape, prec);
  std::lock_guard<std::mutex> lk(g_plan_mu);
  return build_plan(device, shape, prec);
}

void set_tile_wisdom(int L, int prec, int kind, int tile) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  g_tile_wisdom[std::make_tuple(L, kind == 2 ? 0 : prec, kind)] = tile;
  g_plans.clear();
}

namespace {

template <typename T>
void fft_roundtrip_r2c(fc_ctx* ctx, const int shape[3], const double* real_in,
                       double* complex_out) {
  auto p = get_fft_plan(ctx->device, shape, std::is_same<T, double>::value ? FC_FP64 : FC_FP32);
  const int64_t nrows = p->nrows();
  const int n2 = shape[2], n2h = p->n2h, ldc = p->ldc;
  DeviceBuffer din(sizeof(double) * nrows * n2);
  DeviceBuffer work(sizeof(cplx<T>) * nrows * ldc);
  FC_HIP(hipMemcpyAsync(din.as<void>(), real_in, sizeof(double) * nrows * n2,
                        hipMemcpyHostToDevice, ctx->stream));
  launch_rows_r2c<T, double>(*p, din.as<double>(), n2, work.as<cplx<T>>(), ldc, LoadCopy{},
                             ctx->stream);
  fft_yx_forward<T>(*p, work.as<cplx<T>>(), ctx->stream);
  std::vector<cplx<T>> h(nrows * ldc);
  FC_HIP(hipMemcpyAsync(h.data(), work.as<void>(), sizeof(cplx<T>) * h.size(),
                        hipMemcpyDeviceToHost, ctx->stream));
  FC_HIP(hipStreamSynchronize(ctx->stream));
  for (int64_t r = 0; r < nrows; ++r)
    for (int k = 0; k < n2h; ++k) {
      complex_out[2 * (r * n2h + k)] = h[r * ldc + k].x;
      complex_out[2 * (r * n2h + k) + 1] = h[r * ldc + k].y;
    }
}

template <typename T>
void fft_roundtrip_c2r(fc_ctx* ctx, const int shape[3], const double* complex_in,
                       double* real_out) {
  auto p = get_fft_plan(ctx->device, shape, std::is_same<T, double>::value ? FC_FP64 : FC_FP32);
  const int64_t nrows = p->nrows();
  const int n2 = shape[2], n2h = p->n2h, ldc = p->ldc;
  std::vector<cplx<T>> h(nrows * ldc, cplx<T>{T(0), T(0)});
  for (int64_t r = 0; r < nrows; ++r)
    for (int k = 0; k < n2h; ++k) {
      h[r * ldc + k].x = (T)complex_in[2 * (r * n2h + k)];
      h[r * ldc + k].y = (T)complex_in[2 * (r * n2h + k) + 1];
    }
  DeviceBuffer work(sizeof(cplx<T>) * nrows * ldc);
  DeviceBuffer dout(sizeof(double) * nrows * n2);
  FC_HIP(hipMemcpyAsync(work.as<void>(), h.data(), sizeof(cplx<T>) * h.size(),
                        hipMemcpyHostToDevice, ctx->stream));
  fft_xy_inverse<T>(*p, work.as<cplx<T>>(), ctx->stream);
  launch_rows_c2r<T, double>(*p, work.as<cplx<T>>(), ldc, dout.as<double>(), n2, ctx->stream);
  FC_HIP(hipMemcpyAsync(real_out, dout.as<void>(), sizeof(double) * nrows * n2,
                        hipMemcpyDeviceToHost, ctx->stream));
  FC_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace
}  // namespace fc

extern "C" {

int fc_fft_r2c(fc_ctx* ctx, const int shape[3], int precision, const double* real_in,
               double* complex_out) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!shape || !real_in || !complex_out) fc::fail(FC_ERR_INVALID, "null argument");
    fc::DeviceGuard g(ctx->device);
    if (precision == FC_FP64)
      fc::fft_roundtrip_r2c<double>(ctx, shape, real_in, complex_out);
    else if (precision == FC_FP32)
      fc::fft_roundtrip_r2c<float>(ctx, shape, real_in, complex_out);
    else
      fc::fail(FC_ERR_INVALID, "unknown precision");
  });
}

int fc_fft_c2r(fc_ctx* ctx, const int shape[3], int precision, const double* complex_in,
               double* real_out) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!shape || !complex_in || !real_out) fc::fail(FC_ERR_INVALID, "null argument");
    fc::DeviceGuard g(ctx->device);
    if (precision == FC_FP64)
      fc::fft_roundtrip_c2r<double>(ctx, shape, complex_in, real_out);
    else if (precision == FC_FP32)
      fc::fft_roundtrip_c2r<float>(ctx, shape, complex_in, real_out);
    else
      fc::fail(FC_ERR_INVALID, "unknown precision");
  });
}

// Wisdom file: one "fftcorr_hip_tiles v1" header line, then lines
// "<length> <precision> <kind> <tile>".
int fc_plan_cache_export(const char* path) {
  return fc::guard([&] {
    if (!path) fc::fail(FC_ERR_INVALID, "null path");
    std::ofstream f(path);
    if (!f) fc::fail(FC_ERR_RUNTIME, std::string("cannot write ") + path);
    f << "fftcorr_hip_tiles v1\n";
    std::lock_guard<std::mutex> lk(fc::g_plan_mu);
    std::map<std::tuple<int, int, int>, int> all = fc::g_tile_wisdom;
    for (auto& kv : fc::g_plans) {
      const auto& p = *kv.second;
      for (int ax = 0; ax < 2; ++ax) {
        all.emplace(std::make_tuple(p.n[ax], p.prec, 0), p.tile[ax]);
        all.emplace(std::make_tuple(p.n[ax], p.prec, 3), p.var[ax]);
      }
      if (p.fast[2]) {
        all.emplace(std::make_tuple(p.n[2] / 2, p.prec, 1), p.tile[2]);
        all.emplace(std::make_tuple(p.n[2] / 2, p.prec, 4), p.var[2]);
      }
      all.emplace(std::make_tuple(0, 0, 2), p.zy_chunk_kib);
    }
    for (auto& kv : all)
      f << std::get<0>(kv.first) << " " << std::get<1>(kv.first) << " "
        << std::get<2>(kv.first) << " " << kv.second << "\n";
    if (!f) fc::fail(FC_ERR_RUNTIME, std::string("cannot write ") + path);
  });
}

int fc_plan_cache_import(const char* path) {
  return fc::guard([&] {
    if (!path) fc::fail(FC_ERR_INVALID, "null path");
    std::ifstream f(path);
    if (!f) fc::fail(FC_ERR_RUNTIME, std::string("cannot read ") + path);
    std::string header;
    std::getline(f, header);
    if (header != "fftcorr_hip_tiles v1")
      fc::fail(FC_ERR_RUNTIME, std::string("not a plan cache file: ") + path);
    int L, prec, kind, tile;
    std::vector<std::array<int, 4>> rows;
    while (f >> L >> prec >> kind >> tile) {
      const bool prec_ok = prec == FC_FP64 || prec == FC_FP32;
      const bool ok = kind == 2   ? (L == 0 && prec == 0 && tile >= 0)
                      : kind >= 3 ? (kind <= 4 && L > 0 && prec_ok && tile >= 0 &&
                                     tile < fc::kFftVariants)
                                  : ((kind == 0 || kind == 1) && L > 0 && tile > 0 && prec_ok);
      if (!ok) fc::fail(FC_ERR_RUNTIME, std::string("corrupt plan cache file: ") + path);
      rows.push_back({L, prec, kind, tile});
    }
    for (auto& r : rows) fc::set_tile_wisdom(r[0], r[1], r[2], r[3]);
  });
}

int fc_fft_radices(int L, int var, int radices[8], int* count) {
  return fc::guard([&] {
    if (!radices || !count) fc::fail(FC_ERR_INVALID, "null argument");
    *count = fc::fft_radices(L, var, radices);
  });
}

int fc_plan_cache_forget(void) {
  return fc::guard([&] {
    std::lock_guard<std::mutex> lk(fc::g_plan_mu);
    fc::g_tile_wisdom.clear();
    fc::g_plans.clear();
  });
}

}  // extern "C"
