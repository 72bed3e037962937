// Deterministic mass assignment (MassAssignor, cc/particle_mesh/mass_assignor.h:30-234
// with the NGP/TSC stencils of cc/particle_mesh/distribution_functions.h:25-160,
// plus CIC).
//
// The reference buffers particles, merge-sorts them by cell and scatters each
// particle's 27 weights into the grid slab by slab.  Here:
//   1. k_ma_prepare   grid coordinates / skip test (config_space_grid.h:31-46),
//                     cell key, in-cell offsets, FP64 weight sums per block;
//   2. radix sort of (cell key, particle index) — stable, so the order is
//      fixed by the input order;
//   3. k_cell_starts  first sorted particle of every cell (a dense table,
//                     each entry written once, by the particle that opens the
//                     cell or the next non-empty one);
//   4. k_ma_gather    one thread per target cell sums, in FP64 and in a fixed
//                     order (27 stencil offsets, then sorted particles), the
//                     contributions of the particles in its neighbour cells.
// No floating-point atomics: results do not depend on scheduling.
#include <hipcub/hipcub.hpp>

#include <mutex>
#include <vector>

#include "fc_objects.h"

namespace fc {
namespace {

constexpr int kPrepThreads = 256;
constexpr int TX = 2, TY = 4, TZ = 32;  // target tile (256 cells, one per thread)
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;

struct GridGeom {
  int n0, n1, n2;
  double p0, p1, p2;
  double cell;
};

__device__ __forceinline__ bool to_grid(double p, double pmin, double cell, int n, bool wrap,
                                        double* x) {
#pragma clang fp contract(off)
  double v = (p - pmin) / cell;
  const double xmax = (double)n;
  if (v < 0 || v >= xmax) {
    if (!wrap) return false;
    v = fmod(v, xmax);
    if (v < 0) v += xmax;
    // fmod(-eps, n) + n can round up to n: that is the periodic image 0.
    if (v >= xmax) v = 0.0;
  }
  *x = v;
  return true;
}

// Particles: pos [n, ncol] (ncol 3 or 4), weights [n] or null.
__global__ void __launch_bounds__(kPrepThreads)
    k_ma_prepare(const double* __restrict__ pos, int64_t n, int ncol,
                 const double* __restrict__ weights, double w_scalar, GridGeom g, int wrap,
                 double shift, uint64_t sentinel, uint64_t* __restrict__ keys,
                 uint32_t* __restrict__ idx, double4* __restrict__ rec,
                 double* __restrict__ part_w, unsigned long long* __restrict__ part_n) {
#pragma clang fp contract(off)
  __shared__ double sd[kPrepThreads / kWave];
  __shared__ unsigned long long su[kPrepThreads / kWave];
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  double tw = 0.0, tw2 = 0.0;
  unsigned long long added = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const double* p = pos + i * ncol;
    const double w = ncol == 4 ? p[3] : (weights ? weights[i] : w_scalar);
    double x, y, z;
    const bool ok = to_grid(p[0], g.p0, g.cell, g.n0, wrap, &x) &&
                    to_grid(p[1], g.p1, g.cell, g.n1, wrap, &y) &&
                    to_grid(p[2], g.p2, g.cell, g.n2, wrap, &z);
    if (!ok) {
      keys[i] = sentinel;
      idx[i] = (uint32_t)i;
      rec[i] = make_double4(0, 0, 0, 0);
      continue;
    }
    if (shift != 0.0) {  // interlaced grid: displace by `shift` cells, wrap
      x += shift;
      y += shift;
      z += shift;
      if (x >= g.n0) x -= g.n0;
      if (y >= g.n1) y -= g.n1;
      if (z >= g.n2) z -= g.n2;
    }
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const uint64_t ix = (uint64_t)fx, iy = (uint64_t)fy, iz = (uint64_t)fz;
    keys[i] = (ix * g.n1 + iy) * g.n2 + iz;
    idx[i] = (uint32_t)i;
    rec[i] = make_double4(x - fx, y - fy, z - fz, w);
    tw += w;
    tw2 += w * w;
    added += 1;
  }
  tw = block_sum(tw, sd);
  tw2 = block_sum(tw2, sd);
  added = block_sum_u64(added, su);
  if (threadIdx.x == 0) {
    part_w[2 * blockIdx.x] = tw;
    part_w[2 * blockIdx.x + 1] = tw2;
    part_n[blockIdx.x] = added;
  }
}

// One block of kWave threads: lane l sums partials l, l + 64, ... in order,
// then the fixed butterfly of block_sum; the result does not depend on
// scheduling.
__global__ void __launch_bounds__(kWave)
    k_ma_finish_stats(const double* __restrict__ part_w,
                      const unsigned long long* __restrict__ part_n, int nb,
                      double* __restrict__ out_w, unsigned long long* out_n) {
#pragma clang fp contract(off)
  __shared__ double sd[1];
  __shared__ unsigned long long su[1];
  double a = 0.0, b = 0.0;
  unsigned long long c = 0;
  for (int i = threadIdx.x; i < nb; i += kWave) {
    a += part_w[2 * i];
    b += part_w[2 * i + 1];
    c += part_n[i];
  }
  a = block_sum(a, sd);
  b = block_sum(b, sd);
  c = block_sum_u64(c, su);
  if (threadIdx.x == 0) {
    out_w[0] = a;
    out_w[1] = b;
    out_n[0] = c;
  }
}

__global__ void k_permute(const uint32_t* __restrict__ order, const double4* __restrict__ rec,
                          double4* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = rec[order[i]];
}

// start[c] = first sorted particle with key >= c, c in [0, ncells].  Sorted
// particle i (i <= nvalid; particle nvalid stands for the end, key ncells)
// writes the entries (key[i-1], key[i]].
__global__ void k_cell_starts(const uint64_t* __restrict__ keys, const unsigned long long* nvalid_p,
                              uint64_t ncells, int32_t* __restrict__ start) {
  const int64_t nvalid = (int64_t)*nvalid_p;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= nvalid;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t prev = i == 0 ? -1 : (int64_t)keys[i - 1];
    const int64_t cur = i == nvalid ? (int64_t)ncells : (int64_t)keys[i];
    for (int64_t c = prev + 1; c <= cur; ++c) start[c] = (int32_t)i;
  }
}

// v mod n for the halo of a tile: one compare in the common case, the full
// modulo only on grids smaller than the halo.
__device__ __forceinline__ int wrap_index(int v, int n) {
  if ((unsigned)v < (unsigned)n) return v;
  if (v < 0 && v + n >= 0) return v + n;
  if (v >= n && v - n < n) return v - n;
  return ((v % n) + n) % n;
}

// 1D stencil weight for offset o (target = source + o) and in-cell offset r.
template <int SCHEME>
__device__ __forceinline__ double stencil(int o, double r) {
#pragma clang fp contract(off)
  if (SCHEME == 1) {  // TSC, distribution_functions.h:59-67
    if (o < 0) return 0.5 * (1 - r) * (1 - r);
    if (o == 0) return 0.5 + r - r * r;
    return 0.5 * r * r;
  }
  // CIC about cell centres (cell i centred at i + 1/2).
  if (o < 0) return r < 0.5 ? 0.5 - r : 0.0;
  if (o == 0) return r < 0.5 ? r + 0.5 : 1.5 - r;
  return r < 0.5 ? 0.0 : r - 0.5;
}

template <typename Tg, int SCHEME>
__global__ void __launch_bounds__(TX* TY* TZ)
    k_ma_gather(Tg* __restrict__ grid, GridGeom g, const int32_t* __restrict__ start,
                const double4* __restrict__ srec, int tiles_y, int tiles_z) {
#pragma clang fp contract(off)
  __shared__ int32_t cs[HX * HY * HZ];
  __shared__ int32_t ce[HX * HY * HZ];
  int b = blockIdx.x;
  const int bz = b % tiles_z;
  b /= tiles_z;
  const int by = b % tiles_y;
  const int bx = b / tiles_y;
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
  for (int h = threadIdx.x; h < HX * HY * HZ; h += blockDim.x) {
    const int hz = h % HZ, hy = (h / HZ) % HY, hx = h / (HZ * HY);
    const int gx = wrap_index(x0 + hx - 1, g.n0);
    const int gy = wrap_index(y0 + hy - 1, g.n1);
    const int gz = wrap_index(z0 + hz - 1, g.n2);
    const int64_t c = ((int64_t)gx * g.n1 + gy) * g.n2 + gz;
    cs[h] = start[c];
    ce[h] = start[c + 1];
  }
  __syncthreads();
  const int tz = threadIdx.x % TZ, ty = (threadIdx.x / TZ) % TY, tx = threadIdx.x / (TZ * TY);
  const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
  if (x >= g.n0 || y >= g.n1 || z >= g.n2) return;
  double sum = 0.0;
  bool touched = false;
  if (SCHEME == 0) {  // NGP: only the cell itself
    const int h = ((tx + 1) * HY + (ty + 1)) * HZ + (tz + 1);
    touched = cs[h] < ce[h];
    for (int p = cs[h]; p < ce[h]; ++p) sum += srec[p].w;
  } else {
    for (int ox = -1; ox <= 1; ++ox)
      for (int oy = -1; oy <= 1; ++oy)
        for (int oz = -1; oz <= 1; ++oz) {
          const int h = ((tx - ox + 1) * HY + (ty - oy + 1)) * HZ + (tz - oz + 1);
          touched |= cs[h] < ce[h];
          for (int p = cs[h]; p < ce[h]; ++p) {
            const double4 q = srec[p];
            const double wx = stencil<SCHEME>(ox, q.x) * q.w;
            const double wy = stencil<SCHEME>(oy, q.y);
            const double wz = stencil<SCHEME>(oz, q.z);
            sum += wx * wy * wz;
          }
        }
  }
  if (touched) {
    Tg* cell = grid + ((int64_t)x * g.n1 + y) * g.n2 + z;
    *cell = (Tg)((double)*cell + sum);
  }
}

// D12 wavelet (distribution_functions.h:162-217): a particle in cell s with
// sub-cell sample (sx, sy, sz) = floor(64 r) adds
// tab[ox][sx] w * tab[oy][sy] * tab[oz][sz] to cell s - 1 + o, o in [0, 8)^3,
// tab being the caller's 8 x 64 sample table (d12.h's wavesample order).
// Target t therefore gathers from source cells t - 6 .. t + 1 per axis.
constexpr int WC = 8, WS = 64;
constexpr int DX = TX + WC - 1, DY = TY + WC - 1, DZ = TZ + WC - 1;

template <typename Tg>
__global__ void __launch_bounds__(TX* TY* TZ)
    k_ma_gather_wavelet(Tg* __restrict__ grid, GridGeom g, const int32_t* __restrict__ start,
                        const double4* __restrict__ srec, const double* __restrict__ table,
                        int tiles_y, int tiles_z) {
#pragma clang fp contract(off)
  __shared__ int32_t cs[DX * DY * DZ];
  __shared__ int32_t ce[DX * DY * DZ];
  __shared__ double tab[WC * WS];
  int b = blockIdx.x;
  const int bz = b % tiles_z;
  b /= tiles_z;
  const int by = b % tiles_y;
  const int bx = b / tiles_y;
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
  for (int i = threadIdx.x; i < WC * WS; i += blockDim.x) tab[i] = table[i];
  for (int h = threadIdx.x; h < DX * DY * DZ; h += blockDim.x) {
    const int hz = h % DZ, hy = (h / DZ) % DY, hx = h / (DZ * DY);
    const int gx = wrap_index(x0 + hx - (WC - 2), g.n0);
    const int gy = wrap_index(y0 + hy - (WC - 2), g.n1);
    const int gz = wrap_index(z0 + hz - (WC - 2), g.n2);
    const int64_t c = ((int64_t)gx * g.n1 + gy) * g.n2 + gz;
    cs[h] = start[c];
    ce[h] = start[c + 1];
  }
  __syncthreads();
  const int tz = threadIdx.x % TZ, ty = (threadIdx.x / TZ) % TY, tx = threadIdx.x / (TZ * TY);
  const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
  if (x >= g.n0 || y >= g.n1 || z >= g.n2) return;
  double sum = 0.0;
  bool touched = false;
  for (int ox = 0; ox < WC; ++ox)
    for (int oy = 0; oy < WC; ++oy)
      for (int oz = 0; oz < WC; ++oz) {
        const int h = ((tx + WC - 1 - ox) * DY + (ty + WC - 1 - oy)) * DZ + (tz + WC - 1 - oz);
        touched |= cs[h] < ce[h];
        for (int p = cs[h]; p < ce[h]; ++p) {
          const double4 q = srec[p];
          const int sx = (int)floor(q.x * WS), sy = (int)floor(q.y * WS), sz = (int)floor(q.z * WS);
          const double dxw = tab[ox * WS + sx] * q.w;
          const double dxy = dxw * tab[oy * WS + sy];
          sum += tab[oz * WS + sz] * dxy;
        }
      }
  if (touched) {
    Tg* cell = grid + ((int64_t)x * g.n1 + y) * g.n2 + z;
    *cell = (Tg)((double)*cell + sum);
  }
}

// Process-wide D12 table (fc_set_wavelet_table), copied to each context's
// device on first use after a change.
std::mutex g_wavelet_mu;
std::vector<double> g_wavelet;
uint64_t g_wavelet_ver = 0;

const double* wavelet_on_device(fc_ctx* ctx) {
  std::lock_guard<std::mutex> lk(g_wavelet_mu);
  if (g_wavelet.empty())
    fail(FC_ERR_INVALID,
         "distribution scheme 2 (D12 wavelet) needs its sample table: call "
         "fc_set_wavelet_table first");
  if (ctx->wavelet_ver != g_wavelet_ver) {
    ctx->wavelet.reserve(sizeof(double) * g_wavelet.size());
    FC_HIP(hipMemcpyAsync(ctx->wavelet.as<void>(), g_wavelet.data(),
                          sizeof(double) * g_wavelet.size(), hipMemcpyHostToDevice, ctx->stream));
    FC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->wavelet_ver = g_wavelet_ver;
  }
  return ctx->wavelet.as<double>();
}

template <typename Tg>
void launch_gather(int scheme, Tg* grid, const GridGeom& gg, const int32_t* start,
                   const double4* srec, hipStream_t s, const double* table = nullptr) {
  const int tx = (gg.n0 + TX - 1) / TX, ty = (gg.n1 + TY - 1) / TY, tz = (gg.n2 + TZ - 1) / TZ;
  const uint64_t nb = (uint64_t)tx * ty * tz;
  if (nb > 0x7fffffffULL) fail(FC_ERR_INVALID, "grid too large for mass assignment");
  dim3 grd((unsigned)nb), blk(TX * TY * TZ);
  switch (scheme) {
    case 0:
      hipLaunchKernelGGL((k_ma_gather<Tg, 0>), grd, blk, 0, s, grid, gg, start, srec, ty, tz);
      break;
    case 1:
      hipLaunchKernelGGL((k_ma_gather<Tg, 1>), grd, blk, 0, s, grid, gg, start, srec, ty, tz);
      break;
    case 3:
      hipLaunchKernelGGL((k_ma_gather<Tg, 3>), grd, blk, 0, s, grid, gg, start, srec, ty, tz);
      break;
    case 2:
      hipLaunchKernelGGL((k_ma_gather_wavelet<Tg>), grd, blk, 0, s, grid, gg, start, srec,
                         table, ty, tz);
      break;
  }
  FC_LAUNCH_CHECK();
}

int key_bits(uint64_t sentinel) {
  int b = 1;
  while (b < 64 && (sentinel >> b) != 0) ++b;
  return b;
}

void mass_assign(fc_grid* grid, const double* pos_dev, int64_t n, int ncol,
                 const double* w_dev, double w_scalar, int scheme, int wrap, double shift,
                 fc_ma_stats* stats) {
  if (scheme < 0 || scheme > 3)
    fail(FC_ERR_INVALID, "unknown distribution scheme " + std::to_string(scheme));
  if (ncol != 3 && ncol != 4) fail(FC_ERR_INVALID, "Rows must be 3D or 4D.");
  if (n < 0) fail(FC_ERR_INVALID, "negative particle count");
  if (n > 0x7fffffffLL) fail(FC_ERR_INVALID, "more than 2^31-1 particles in one call");
  if (!(shift >= 0.0 && shift < 1.0)) fail(FC_ERR_INVALID, "shift must be in [0, 1)");
  fc_ctx* ctx = grid->ctx;
  hipStream_t s = ctx->stream;
  DeviceGuard dg(ctx->device);
  const uint64_t ncells = grid->size();
  const uint64_t sentinel = ncells;
  GridGeom gg{grid->shape[0], grid->shape[1], grid->shape[2], grid->posmin[0],
              grid->posmin[1], grid->posmin[2], grid->cell};
  const double* wtab = scheme == 2 ? wavelet_on_device(ctx) : nullptr;
  if (scheme == 2)
    for (int d = 0; d < 3; ++d)
      if (grid->shape[d] < WC)
        fail(FC_ERR_INVALID, "the D12 wavelet needs at least 8 cells per axis");
  if (n == 0) return;

  hipEvent_t ev[3];
  for (auto& e : ev) FC_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]);
    }
  } evg{ev};

  if (ncells >= 0x7fffffffULL) fail(FC_ERR_INVALID, "grid too large for mass assignment");
  const int nbp = (int)std::min<int64_t>(1024, (n + kPrepThreads - 1) / kPrepThreads);
  // Work buffers live in the context and are reused by later calls.
  DeviceBuffer* sc = ctx->scratch;
  sc[1].reserve(sizeof(uint64_t) * n);
  sc[2].reserve(sizeof(uint64_t) * n);
  sc[3].reserve(sizeof(uint32_t) * n);
  sc[4].reserve(sizeof(uint32_t) * n);
  sc[5].reserve(sizeof(double4) * n);
  sc[6].reserve(sizeof(double4) * n);
  sc[7].reserve(sizeof(double) * 2 * (nbp + 1));
  sc[8].reserve(sizeof(unsigned long long) * (nbp + 1));
  sc[9].reserve(sizeof(int32_t) * (ncells + 1));
  DeviceBuffer &keys = sc[1], &keys2 = sc[2], &idx = sc[3], &idx2 = sc[4], &rec = sc[5],
               &srec = sc[6], &partw = sc[7], &partn = sc[8], &starts = sc[9];

  hipLaunchKernelGGL(k_ma_prepare, dim3(nbp), dim3(kPrepThreads), 0, s, pos_dev, n, ncol, w_dev,
                     w_scalar, gg, wrap, shift, sentinel, keys.as<uint64_t>(),
                     idx.as<uint32_t>(), rec.as<double4>(), partw.as<double>(),
                     partn.as<unsigned long long>());
  FC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ma_finish_stats, dim3(1), dim3(64), 0, s, partw.as<double>(),
                     partn.as<unsigned long long>(), nbp, partw.as<double>() + 2 * nbp,
                     partn.as<unsigned long long>() + nbp);
  FC_LAUNCH_CHECK();

  FC_HIP(hipEventRecord(ev[0], s));
  const int bits = key_bits(sentinel);
  size_t temp_bytes = 0;
  FC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys.as<uint64_t>(),
                                            keys2.as<uint64_t>(), idx.as<uint32_t>(),
                                            idx2.as<uint32_t>(), (int)n, 0, bits, s));
  ctx->scratch[0].reserve(temp_bytes + 256);
  FC_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->scratch[0].as<void>(), temp_bytes,
                                            keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                            idx.as<uint32_t>(), idx2.as<uint32_t>(), (int)n, 0,
                                            bits, s));
  hipLaunchKernelGGL(k_permute, dim3(grid1d(n, 256, 65536)), dim3(256), 0, s,
                     idx2.as<uint32_t>(), rec.as<double4>(), srec.as<double4>(), n);
  FC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cell_starts, dim3(grid1d(n + 1, 256, 65536)), dim3(256), 0, s,
                     keys2.as<uint64_t>(), partn.as<unsigned long long>() + nbp, ncells,
                     starts.as<int32_t>());
  FC_LAUNCH_CHECK();
  FC_HIP(hipEventRecord(ev[1], s));
  if (grid->prec == FC_FP64)
    launch_gather<double>(scheme, grid->data.as<double>(), gg, starts.as<int32_t>(),
                          srec.as<double4>(), s, wtab);
  else
    launch_gather<float>(scheme, grid->data.as<float>(), gg, starts.as<int32_t>(),
                         srec.as<double4>(), s, wtab);
  FC_HIP(hipEventRecord(ev[2], s));

  double wsum[2];
  unsigned long long nadd = 0;
  FC_HIP(hipMemcpyAsync(wsum, partw.as<double>() + 2 * nbp, sizeof(wsum),
                        hipMemcpyDeviceToHost, s));
  FC_HIP(hipMemcpyAsync(&nadd, partn.as<unsigned long long>() + nbp, sizeof(nadd),
                        hipMemcpyDeviceToHost, s));
  FC_HIP(hipStreamSynchronize(s));
  float t_sort = 0, t_win = 0;
  FC_HIP(hipEventElapsedTime(&t_sort, ev[0], ev[1]));
  FC_HIP(hipEventElapsedTime(&t_win, ev[1], ev[2]));
  if (stats) {
    stats->num_added += nadd;
    stats->num_skipped += (uint64_t)n - nadd;
    stats->totw += wsum[0];
    stats->totwsq += wsum[1];
    stats->sort_time += t_sort * 1e-3;
    stats->window_time += t_win * 1e-3;
  }
}

}  // namespace

void mass_assign_device(fc_grid* grid, const double* pos_dev, int64_t n, int ncol,
                        const double* w_dev, double w_scalar, int scheme, int wrap, double shift,
                        fc_ma_stats* stats) {
  mass_assign(grid, pos_dev, n, ncol, w_dev, w_scalar, scheme, wrap, shift, stats);
}

}  // namespace fc

extern "C" {

int fc_set_wavelet_table(const double* table, int n) {
  return fc::guard([&] {
    if (table && n != fc::WC * fc::WS)
      fc::fail(FC_ERR_INVALID, "the D12 table has 8 x 64 = 512 samples");
    std::lock_guard<std::mutex> lk(fc::g_wavelet_mu);
    if (table)
      fc::g_wavelet.assign(table, table + n);
    else
      fc::g_wavelet.clear();
    ++fc::g_wavelet_ver;
  });
}

int fc_mass_assign_device(fc_grid* grid, const double* pos_dev, int64_t n, int ncol,
                          const double* weights_dev, double w_scalar, int scheme,
                          int periodic_wrap, double shift, fc_ma_stats* stats) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (n > 0 && !pos_dev) fc::fail(FC_ERR_INVALID, "null positions");
    fc::mass_assign(grid, pos_dev, n, ncol, weights_dev, w_scalar, scheme, periodic_wrap,
                    shift, stats);
  });
}

int fc_mass_assign(fc_grid* grid, const double* pos, int64_t n, int ncol,
                   const double* weights, double w_scalar, int scheme, int periodic_wrap,
                   double shift, fc_ma_stats* stats) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (n > 0 && !pos) fc::fail(FC_ERR_INVALID, "null positions");
    if (ncol != 3 && ncol != 4) fc::fail(FC_ERR_INVALID, "Rows must be 3D or 4D.");
    if (n <= 0) return;
    fc::DeviceGuard dg(grid->ctx->device);
    hipStream_t s = grid->ctx->stream;
    fc::DeviceBuffer dpos(sizeof(double) * n * ncol);
    fc::DeviceBuffer dw(weights ? sizeof(double) * n : 0);
    FC_HIP(hipMemcpyAsync(dpos.as<void>(), pos, sizeof(double) * n * ncol,
                          hipMemcpyHostToDevice, s));
    if (weights)
      FC_HIP(hipMemcpyAsync(dw.as<void>(), weights, sizeof(double) * n, hipMemcpyHostToDevice,
                            s));
    fc::mass_assign(grid, dpos.as<double>(), n, ncol, weights ? dw.as<double>() : nullptr,
                    w_scalar, scheme, periodic_wrap, shift, stats);
  });
}

}  // extern "C"
