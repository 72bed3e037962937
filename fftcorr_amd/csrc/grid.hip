// ConfigSpaceGrid on the device (cc/grid/config_space_grid.h:10-56) and the
// element-wise grid ops of cc/array/array_ops.cc used on the path.
#include "fc_objects.h"

namespace fc {
namespace {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
    k_grid_stats_partial(const T* __restrict__ g, uint64_t n, double* __restrict__ part) {
  __shared__ double scratch[kRedThreads / kWave];
  // Contiguous chunk per block: the partition (hence rounding) depends only
  // on n and the fixed block count.
  const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * chunk;
  const uint64_t hi = lo + chunk < n ? lo + chunk : n;
  double s = 0.0, s2 = 0.0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const double v = (double)g[i];
    s += v;
    s2 += v * v;
  }
  s = block_sum(s, scratch);
  s2 = block_sum(s2, scratch);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s;
    part[2 * blockIdx.x + 1] = s2;
  }
}

__global__ void k_sum_partials(const double* __restrict__ part, int n, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0, s2 = 0.0;
    for (int i = 0; i < n; ++i) {
      s += part[2 * i];
      s2 += part[2 * i + 1];
    }
    out[0] = s;
    out[1] = s2;
  }
}

// d = (d + a) * b in FP64 (a = -mean, b = 1/mean for the delta normalisation of
// cc/fftcorr.cc:369-375; a = s, b = 1 for add_scalar; a = 0, b = s for
// multiply_by).  mode 0: both steps, 1: add only, 2: multiply only.
template <typename T>
__global__ void k_grid_affine(T* __restrict__ g, uint64_t n, double a, double b, int mode) {
#pragma clang fp contract(off)
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double v = (double)g[i];
    if (mode != 2) v += a;
    if (mode != 1) v *= b;
    g[i] = (T)v;
  }
}

template <typename Tdst, typename Tsrc>
__global__ void k_convert(Tdst* __restrict__ d, const Tsrc* __restrict__ s, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    d[i] = (Tdst)s[i];
}

void grid_affine(fc_grid* g, double a, double b, int mode) {
  DeviceGuard dg(g->ctx->device);
  const uint64_t n = g->size();
  const unsigned nb = grid1d(n, 256, 65536);
  if (g->prec == FC_FP64)
    hipLaunchKernelGGL(k_grid_affine<double>, dim3(nb), dim3(256), 0, g->ctx->stream,
                       g->data.as<double>(), n, a, b, mode);
  else
    hipLaunchKernelGGL(k_grid_affine<float>, dim3(nb), dim3(256), 0, g->ctx->stream,
                       g->data.as<float>(), n, a, b, mode);
  FC_LAUNCH_CHECK();
}

}  // namespace

void grid_stats(fc_grid* g, double* sum, double* sumsq) {
  DeviceGuard dg(g->ctx->device);
  const uint64_t n = g->size();
  DeviceBuffer part(sizeof(double) * 2 * (kRedBlocks + 1));
  double* p = part.as<double>();
  if (g->prec == FC_FP64)
    hipLaunchKernelGGL(k_grid_stats_partial<double>, dim3(kRedBlocks), dim3(kRedThreads), 0,
                       g->ctx->stream, g->data.as<double>(), n, p);
  else
    hipLaunchKernelGGL(k_grid_stats_partial<float>, dim3(kRedBlocks), dim3(kRedThreads), 0,
                       g->ctx->stream, g->data.as<float>(), n, p);
  FC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, g->ctx->stream, p, kRedBlocks,
                     p + 2 * kRedBlocks);
  FC_LAUNCH_CHECK();
  double out[2];
  FC_HIP(hipMemcpyAsync(out, p + 2 * kRedBlocks, sizeof(out), hipMemcpyDeviceToHost,
                        g->ctx->stream));
  FC_HIP(hipStreamSynchronize(g->ctx->stream));
  if (sum) *sum = out[0];
  if (sumsq) *sumsq = out[1];
}

}  // namespace fc

extern "C" {

int fc_grid_create(fc_ctx* ctx, const int shape[3], const double posmin[3], double cell_size,
                   int precision, fc_grid** out) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!shape || !out) fc::fail(FC_ERR_INVALID, "null argument");
    for (int i = 0; i < 3; ++i)
      if (shape[i] <= 0) fc::fail(FC_ERR_INVALID, "Shape must be nonnegative.");
    if (!(cell_size > 0)) fc::fail(FC_ERR_INVALID, "cell_size must be positive");
    if (precision != FC_FP64 && precision != FC_FP32)
      fc::fail(FC_ERR_INVALID, "unknown precision");
    fc::DeviceGuard dg(ctx->device);
    std::unique_ptr<fc_grid> g(new fc_grid());
    g->ctx = ctx;
    for (int i = 0; i < 3; ++i) {
      g->shape[i] = shape[i];
      g->posmin[i] = posmin ? posmin[i] : 0.0;
    }
    g->cell = cell_size;
    g->prec = precision;
    g->data.allocate(g->size() * g->elem());
    FC_HIP(hipMemsetAsync(g->data.as<void>(), 0, g->size() * g->elem(), ctx->stream));
    *out = g.release();
  });
}

int fc_grid_destroy(fc_grid* grid) {
  return fc::guard([&] {
    if (!grid) return;
    if (grid->batch_view)
      fc::fail(FC_ERR_INVALID, "this grid belongs to a batch; destroy the batch instead");
    fc::DeviceGuard dg(grid->ctx->device);
    (void)hipStreamSynchronize(grid->ctx->stream);
    delete grid;
  });
}

int fc_grid_clear(fc_grid* grid) {
  return fc::guard([&] {
    fc::check_grid(grid);
    fc::DeviceGuard dg(grid->ctx->device);
    FC_HIP(hipMemsetAsync(grid->data.as<void>(), 0, grid->size() * grid->elem(),
                          grid->ctx->stream));
  });
}

int fc_grid_upload(fc_grid* grid, const double* host) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (!host) fc::fail(FC_ERR_INVALID, "null host array");
    fc::DeviceGuard dg(grid->ctx->device);
    const uint64_t n = grid->size();
    if (grid->prec == FC_FP64) {
      FC_HIP(hipMemcpyAsync(grid->data.as<void>(), host, n * sizeof(double),
                            hipMemcpyHostToDevice, grid->ctx->stream));
    } else {
      fc::DeviceBuffer tmp(n * sizeof(double));
      FC_HIP(hipMemcpyAsync(tmp.as<void>(), host, n * sizeof(double), hipMemcpyHostToDevice,
                            grid->ctx->stream));
      hipLaunchKernelGGL((fc::k_convert<float, double>), dim3(fc::grid1d(n, 256, 65536)),
                         dim3(256), 0, grid->ctx->stream, grid->data.as<float>(),
                         tmp.as<double>(), n);
      FC_LAUNCH_CHECK();
    }
    FC_HIP(hipStreamSynchronize(grid->ctx->stream));
  });
}

int fc_grid_download(fc_grid* grid, double* host) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (!host) fc::fail(FC_ERR_INVALID, "null host array");
    fc::DeviceGuard dg(grid->ctx->device);
    const uint64_t n = grid->size();
    if (grid->prec == FC_FP64) {
      FC_HIP(hipMemcpyAsync(host, grid->data.as<void>(), n * sizeof(double),
                            hipMemcpyDeviceToHost, grid->ctx->stream));
    } else {
      fc::DeviceBuffer tmp(n * sizeof(double));
      hipLaunchKernelGGL((fc::k_convert<double, float>), dim3(fc::grid1d(n, 256, 65536)),
                         dim3(256), 0, grid->ctx->stream, tmp.as<double>(),
                         grid->data.as<float>(), n);
      FC_LAUNCH_CHECK();
      FC_HIP(hipMemcpyAsync(host, tmp.as<void>(), n * sizeof(double), hipMemcpyDeviceToHost,
                            grid->ctx->stream));
      FC_HIP(hipStreamSynchronize(grid->ctx->stream));
    }
    FC_HIP(hipStreamSynchronize(grid->ctx->stream));
  });
}

int fc_grid_stats(fc_grid* grid, double* sum, double* sumsq) {
  return fc::guard([&] {
    fc::check_grid(grid);
    fc::grid_stats(grid, sum, sumsq);
  });
}

int fc_grid_add_scalar(fc_grid* grid, double s) {
  return fc::guard([&] {
    fc::check_grid(grid);
    fc::grid_affine(grid, s, 1.0, 1);
  });
}

int fc_grid_multiply_by(fc_grid* grid, double s) {
  return fc::guard([&] {
    fc::check_grid(grid);
    fc::grid_affine(grid, 0.0, s, 2);
  });
}

int fc_grid_normalize(fc_grid* grid, double mean) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (!(mean != 0.0)) fc::fail(FC_ERR_INVALID, "mean density must be nonzero");
    fc::grid_affine(grid, -mean, 1.0 / mean, 0);
  });
}

int fc_grid_device_ptr(fc_grid* grid, void** ptr) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (!ptr) fc::fail(FC_ERR_INVALID, "null output");
    *ptr = grid->data.as<void>();
  });
}

}  // extern "C"
