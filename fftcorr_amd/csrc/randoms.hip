// Random catalogues and displacement fields on the device
// (fftcorr/catalog/random_catalog.py:8-87, fftcorr/grid/utils.pyx:19-55).
//
// Randoms come from a counter-based generator, Philox4x32-10 (Salmon,
// Moraes, Dror and Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC 2011): 10 rounds of two 32x32->64-bit multiplies with the key bumped by
// the Weyl constants between rounds.  Particle i of stream `seed` uses the
// counters (i, 0, 0) and (i, 1, 0) -- two 4-word draws, six words -> three
// 53-bit uniforms -- so any range of particles can be generated
// independently and the catalogue is the same however it is batched.
#include "fc_objects.h"

namespace fc {
namespace {

struct Philox {
  uint32_t c[4];
};

__host__ __device__ inline void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  const uint64_t p = (uint64_t)a * b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

__host__ __device__ inline Philox philox4x32_10(Philox x, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, x.c[0], &hi0, &lo0);
    mulhilo(M1, x.c[2], &hi1, &lo1);
    x = Philox{{hi1 ^ x.c[1] ^ k0, lo1, hi0 ^ x.c[3] ^ k1, lo0}};
  }
  return x;
}

// Uniform in [0, 1) from two words: 27 + 26 bits (53 in all).
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__global__ void k_random_raw(uint32_t* __restrict__ out, int64_t n, uint64_t seed,
                             uint64_t first, uint32_t w2, uint32_t w3) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t ctr = first + (uint64_t)i;
    const Philox r = philox4x32_10(Philox{{(uint32_t)ctr, (uint32_t)(ctr >> 32), w2, w3}},
                                   (uint32_t)seed, (uint32_t)(seed >> 32));
    for (int j = 0; j < 4; ++j) out[4 * i + j] = r.c[j];
  }
}

// pos[i] = lo + u * (hi - lo) per axis, u from particle (first + i).
__global__ void k_random_uniform(double* __restrict__ pos, int64_t n, double3 lo, double3 span,
                                 uint64_t seed, uint64_t first) {
#pragma clang fp contract(off)
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t ctr = first + (uint64_t)i;
    const uint32_t c0 = (uint32_t)ctr, c1 = (uint32_t)(ctr >> 32);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const Philox a = philox4x32_10(Philox{{c0, c1, 0u, 0u}}, k0, k1);
    const Philox b = philox4x32_10(Philox{{c0, c1, 1u, 0u}}, k0, k1);
    pos[3 * i + 0] = lo.x + u53(a.c[0], a.c[1]) * span.x;
    pos[3 * i + 1] = lo.y + u53(a.c[2], a.c[3]) * span.y;
    pos[3 * i + 2] = lo.z + u53(b.c[0], b.c[1]) * span.z;
  }
}

// out[i] = pos[i] + disp[cell of pos[i]] (utils.pyx:19-44): the cell is the
// truncated grid coordinate (ConfigSpaceGrid::get_grid_coords, with the
// periodic wrap when asked); particles outside the grid count into *bad.
__global__ void k_apply_disp(const double* __restrict__ pos, int64_t n,
                             const double* __restrict__ disp, int n0, int n1, int n2, double p0,
                             double p1, double p2, double cell, int wrap,
                             double* __restrict__ out, unsigned long long* __restrict__ bad) {
#pragma clang fp contract(off)
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double s[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double pm[3] = {p0, p1, p2};
    const int nn[3] = {n0, n1, n2};
    int ic[3];
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
      double v = (s[d] - pm[d]) / cell;
      const double xmax = (double)nn[d];
      if (v < 0 || v >= xmax) {
        if (!wrap) {
          ok = false;
          break;
        }
        v = fmod(v, xmax);
        if (v < 0) v += xmax;
        if (v >= xmax) v = 0.0;
      }
      ic[d] = (int)v;
    }
    if (!ok) {
      atomicAdd(bad, 1ull);  // an integer count: order does not matter
      for (int d = 0; d < 3; ++d) out[3 * i + d] = s[d];
      continue;
    }
    const double* dv = disp + 3 * (((int64_t)ic[0] * n1 + ic[1]) * n2 + ic[2]);
    for (int d = 0; d < 3; ++d) out[3 * i + d] = s[d] + dv[d];
  }
}

}  // namespace

// Shared with the mass-assignment source.
void mass_assign_device(fc_grid* grid, const double* pos_dev, int64_t n, int ncol,
                        const double* w_dev, double w_scalar, int scheme, int wrap, double shift,
                        fc_ma_stats* stats);

}  // namespace fc

extern "C" {

int fc_random_raw_device(fc_ctx* ctx, uint32_t* out, int64_t n, uint64_t seed, uint64_t first,
                         uint32_t w2, uint32_t w3) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (n < 0 || (n > 0 && !out)) fc::fail(FC_ERR_INVALID, "bad output buffer");
    if (n == 0) return;
    fc::DeviceGuard dg(ctx->device);
    hipLaunchKernelGGL(fc::k_random_raw, dim3(fc::grid1d(n, 256, 65536)), dim3(256), 0,
                       ctx->stream, out, n, seed, first, w2, w3);
    FC_LAUNCH_CHECK();
  });
}

int fc_random_uniform_device(fc_ctx* ctx, double* pos, int64_t n, const double lo[3],
                             const double hi[3], uint64_t seed, uint64_t first) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!lo || !hi || n < 0 || (n > 0 && !pos)) fc::fail(FC_ERR_INVALID, "bad argument");
    if (n == 0) return;
    fc::DeviceGuard dg(ctx->device);
    hipLaunchKernelGGL(fc::k_random_uniform, dim3(fc::grid1d(n, 256, 65536)), dim3(256), 0,
                       ctx->stream, pos, n, make_double3(lo[0], lo[1], lo[2]),
                       make_double3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]), seed, first);
    FC_LAUNCH_CHECK();
  });
}

int fc_random_raw(fc_ctx* ctx, uint32_t* out, int64_t n, uint64_t seed, uint64_t first,
                  uint32_t w2, uint32_t w3) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (n < 0 || (n > 0 && !out)) fc::fail(FC_ERR_INVALID, "bad output buffer");
    if (n == 0) return;
    fc::DeviceGuard dg(ctx->device);
    fc::DeviceBuffer d(sizeof(uint32_t) * 4 * n);
    hipLaunchKernelGGL(fc::k_random_raw, dim3(fc::grid1d(n, 256, 65536)), dim3(256), 0,
                       ctx->stream, d.as<uint32_t>(), n, seed, first, w2, w3);
    FC_LAUNCH_CHECK();
    FC_HIP(hipMemcpyAsync(out, d.as<void>(), sizeof(uint32_t) * 4 * n, hipMemcpyDeviceToHost,
                          ctx->stream));
    FC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int fc_random_uniform(fc_ctx* ctx, double* pos, int64_t n, const double lo[3], const double hi[3],
                      uint64_t seed, uint64_t first) {
  return fc::guard([&] {
    fc::check_ctx(ctx);
    if (!lo || !hi || n < 0 || (n > 0 && !pos)) fc::fail(FC_ERR_INVALID, "bad argument");
    if (n == 0) return;
    fc::DeviceGuard dg(ctx->device);
    fc::DeviceBuffer d(sizeof(double) * 3 * n);
    hipLaunchKernelGGL(fc::k_random_uniform, dim3(fc::grid1d(n, 256, 65536)), dim3(256), 0,
                       ctx->stream, d.as<double>(), n, make_double3(lo[0], lo[1], lo[2]),
                       make_double3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]), seed, first);
    FC_LAUNCH_CHECK();
    FC_HIP(hipMemcpyAsync(pos, d.as<void>(), sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                          ctx->stream));
    FC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int fc_apply_displacement(fc_grid* grid, const double* pos, int64_t n, const double* disp,
                          int periodic_wrap, double* out) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (n < 0 || (n > 0 && (!pos || !disp || !out))) fc::fail(FC_ERR_INVALID, "bad argument");
    if (n == 0) return;
    fc::DeviceGuard dg(grid->ctx->device);
    hipStream_t s = grid->ctx->stream;
    const size_t nd = (size_t)grid->size() * 3;
    fc::DeviceBuffer dpos(sizeof(double) * 3 * n), dout(sizeof(double) * 3 * n),
        ddisp(sizeof(double) * nd), bad(sizeof(unsigned long long));
    FC_HIP(hipMemcpyAsync(dpos.as<void>(), pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          s));
    FC_HIP(hipMemcpyAsync(ddisp.as<void>(), disp, sizeof(double) * nd, hipMemcpyHostToDevice, s));
    FC_HIP(hipMemsetAsync(bad.as<void>(), 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(fc::k_apply_disp, dim3(fc::grid1d(n, 256, 65536)), dim3(256), 0, s,
                       dpos.as<double>(), n, ddisp.as<double>(), grid->shape[0], grid->shape[1],
                       grid->shape[2], grid->posmin[0], grid->posmin[1], grid->posmin[2],
                       grid->cell, periodic_wrap, dout.as<double>(),
                       bad.as<unsigned long long>());
    FC_LAUNCH_CHECK();
    unsigned long long nbad = 0;
    FC_HIP(hipMemcpyAsync(&nbad, bad.as<void>(), sizeof(nbad), hipMemcpyDeviceToHost, s));
    FC_HIP(hipStreamSynchronize(s));
    if (nbad)
      fc::fail(FC_ERR_INVALID,
               "Failed to apply displacement field: coordinates out of bounds? "
               "Try setting periodic_wrap=True.");
    FC_HIP(hipMemcpy(out, dout.as<void>(), sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  });
}

int fc_add_random_particles(fc_grid* grid, int64_t n, double weight, int scheme,
                            int periodic_wrap, const double* disp, uint64_t seed, int64_t batch,
                            fc_ma_stats* stats) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (n < 0 || batch <= 0) fc::fail(FC_ERR_INVALID, "bad particle or batch count");
    if (n == 0) return;
    fc_ctx* ctx = grid->ctx;
    fc::DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    const int64_t nb = std::min<int64_t>(n, batch);
    fc::DeviceBuffer pos(sizeof(double) * 3 * nb), moved, ddisp, bad;
    const double lo[3] = {grid->posmin[0], grid->posmin[1], grid->posmin[2]};
    double3 span = make_double3(grid->cell * grid->shape[0], grid->cell * grid->shape[1],
                                grid->cell * grid->shape[2]);
    if (disp) {
      const size_t nd = (size_t)grid->size() * 3;
      ddisp.allocate(sizeof(double) * nd);
      moved.allocate(sizeof(double) * 3 * nb);
      bad.allocate(sizeof(unsigned long long));
      FC_HIP(hipMemcpyAsync(ddisp.as<void>(), disp, sizeof(double) * nd, hipMemcpyHostToDevice,
                            s));
      FC_HIP(hipMemsetAsync(bad.as<void>(), 0, sizeof(unsigned long long), s));
    }
    for (int64_t first = 0; first < n; first += nb) {
      const int64_t m = std::min(nb, n - first);
      hipLaunchKernelGGL(fc::k_random_uniform, dim3(fc::grid1d(m, 256, 65536)), dim3(256), 0, s,
                         pos.as<double>(), m, make_double3(lo[0], lo[1], lo[2]), span, seed,
                         (uint64_t)first);
      FC_LAUNCH_CHECK();
      const double* p = pos.as<double>();
      if (disp) {
        hipLaunchKernelGGL(fc::k_apply_disp, dim3(fc::grid1d(m, 256, 65536)), dim3(256), 0, s,
                           pos.as<double>(), m, ddisp.as<double>(), grid->shape[0],
                           grid->shape[1], grid->shape[2], grid->posmin[0], grid->posmin[1],
                           grid->posmin[2], grid->cell, periodic_wrap, moved.as<double>(),
                           bad.as<unsigned long long>());
        FC_LAUNCH_CHECK();
        p = moved.as<double>();
      }
      fc::mass_assign_device(grid, p, m, 3, nullptr, weight, scheme, periodic_wrap, 0.0, stats);
    }
    if (disp) {
      unsigned long long nbad = 0;
      FC_HIP(hipMemcpyAsync(&nbad, bad.as<void>(), sizeof(nbad), hipMemcpyDeviceToHost, s));
      FC_HIP(hipStreamSynchronize(s));
      if (nbad)
        fc::fail(FC_ERR_INVALID,
                 "Failed to apply displacement field: coordinates out of bounds? "
                 "Try setting periodic_wrap=True.");
    }
  });
}

}  // extern "C"
