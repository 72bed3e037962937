// Deterministic mass assignment (MassAssignor, cc/particle_mesh/mass_assignor.h:30-234
// with the NGP/TSC stencils of cc/particle_mesh/distribution_functions.h:25-160,
// plus CIC).
//
// The reference buffers particles, merge-sorts them by cell and scatters each
// particle's 27 weights into the grid slab by slab.  Here:
//   1. k_ma_prepare   grid coordinates / skip test (config_space_grid.h:31-46),
//                     the key cell (the cell holding the particle; for CIC the
//                     lower corner of its 2x2x2 block of target cells), the
//                     offsets from it and FP64 weight sums per block;
//   2. stable radix sort of (key, particle index) (radix_sort.h: passes of
//      at most 8 bits over the bits the largest key needs): within a key
//      cell the particles keep their input order;
//   3. k_permute      records in sorted order;
//   4. k_cell_runs    per occupied key cell, its run {first sorted particle,
//                     count} (one 8-byte entry) and its bit in an occupancy
//                     bitmap (cleared first; the run table is read only where
//                     the bit is set, so it needs no clearing);
//   5. k_ma_tile      one workgroup per 4 x 4 x 32 tile of target cells lists
//                     the particles of every key cell whose stencil reaches
//                     the tile, in key-cell order; each wave takes a contiguous
//                     quarter of the list, 64 particles at a time, forms their
//                     stencil weights and sums them per key cell with a
//                     segmented scan across lanes (a fixed tree).  The last
//                     lane of each key cell's segment adds the sums to the
//                     wave's own FP64 copy of the tile in LDS (distinct key
//                     cells, so no two lanes hit one target cell in a step);
//                     the four copies are added in wave order and each touched
//                     grid cell is read and written once.
// No floating-point atomics and no order that depends on scheduling: repeated
// runs give identical bits.  Each particle record is read by the tiles its
// stencil reaches (1.3-1.5x on average), never once per target cell.

#include <mutex>
#include <vector>

#include "fc_objects.h"
#include "radix_sort.h"

namespace fc {
namespace {

constexpr int kPrepThreads = 256;
#ifndef FC_MA_TILE_Y
#define FC_MA_TILE_Y 4
#endif
constexpr int TX = 4, TY = FC_MA_TILE_Y, TZ = 32;  // target tile of k_ma_tile
constexpr int kTileCells = TX * TY * TZ;
constexpr int kTileThreads = 256;
constexpr int kTileWaves = kTileThreads / kWave;

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

// Key cell and offset along one axis: a = floor(x - anchor), wrapped into
// [0, n), and r = (x - anchor) - floor(x - anchor) in [0, 1).  anchor is 0
// (NGP, TSC, D12: the cell holding the particle) or 1/2 (CIC: the lower of the
// two cells whose centres bracket the particle).
__device__ __forceinline__ int key_axis(double x, double anchor, int n, double* r) {
#pragma clang fp contract(off)
  const double xa = x - anchor;
  const double fa = floor(xa);
  *r = xa - fa;
  int a = (int)fa;
  if (a < 0) a += n;
  return a;
}

// Particles: pos [n, ncol] (ncol 3 or 4), weights [n] or null.  Skipped
// particles get the key `sentinel` (sorted after every cell).
__global__ void __launch_bounds__(kPrepThreads)
    k_ma_prepare(const double* __restrict__ pos, int64_t n, int ncol,
                 const double* __restrict__ weights, double w_scalar, GridGeom g, int wrap,
                 double shift, double anchor, uint32_t sentinel, uint32_t* __restrict__ keys,
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
    idx[i] = (uint32_t)i;
    if (!ok) {
      keys[i] = sentinel;
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
    double rx, ry, rz;
    const int ix = key_axis(x, anchor, g.n0, &rx);
    const int iy = key_axis(y, anchor, g.n1, &ry);
    const int iz = key_axis(z, anchor, g.n2, &rz);
    keys[i] = (uint32_t)(((uint64_t)ix * g.n1 + iy) * g.n2 + iz);
    rec[i] = make_double4(rx, ry, rz, w);
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

__global__ void k_zero_i32(int32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Runs of equal keys in the sorted keys[0, nvalid): the first particle of
// each run writes runs[key] = {its position, the run length}, the length
// found by an exponential then a binary search (a long run costs its first
// thread O(log length), not O(length)), and sets the key's occupancy bit.
__global__ void k_cell_runs(const uint32_t* __restrict__ keys,
                            const unsigned long long* __restrict__ nvalid_p,
                            int2* __restrict__ runs, uint32_t* __restrict__ occ) {
  const int64_t nvalid = (int64_t)*nvalid_p;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvalid;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t k = keys[i];
    if (i > 0 && keys[i - 1] == k) continue;
    int64_t lo = i, hi = i + 1, step = 1;  // keys[lo] == k
    while (hi < nvalid && keys[hi] == k) {
      lo = hi;
      step <<= 1;
      hi = lo + step < nvalid ? lo + step : nvalid;
    }
    while (hi - lo > 1) {  // keys[lo] == k; hi == nvalid or keys[hi] != k
      const int64_t mid = lo + (hi - lo) / 2;
      if (keys[mid] == k)
        lo = mid;
      else
        hi = mid;
    }
    runs[k] = make_int2((int)i, (int)(lo + 1 - i));
    atomicOr(&occ[k >> 5], 1u << (k & 31));
  }
}

// Run {first, count} of key cell c: {0, 0} unless its occupancy bit is set.
__device__ __forceinline__ int2 cell_run(const uint32_t* __restrict__ occ,
                                         const int2* __restrict__ runs, int64_t c) {
  return ((occ[c >> 5] >> (c & 31)) & 1u) ? runs[c] : make_int2(0, 0);
}

// v mod n for the halo of a tile: one compare in the common case, the full
// modulo only on grids smaller than the halo.
__device__ __forceinline__ int wrap_index(int v, int n) {
  if ((unsigned)v < (unsigned)n) return v;
  if (v < 0 && v + n >= 0) return v + n;
  if (v >= n && v - n < n) return v - n;
  return ((v % n) + n) % n;
}

// Stencils: target cell = key cell + o, o in [OMIN, OMAX] per axis, with
// 1D weights wt[o - OMIN] of the offset r from the key cell.
template <int SCHEME>
struct Stencil;
template <>
struct Stencil<0> {  // NGP, distribution_functions.h:25-32
  static constexpr int OMIN = 0, OMAX = 0;
  static constexpr double anchor = 0.0;
  __device__ static void weights(double, double* wt) { wt[0] = 1.0; }
};
template <>
struct Stencil<1> {  // TSC, distribution_functions.h:59-67
  static constexpr int OMIN = -1, OMAX = 1;
  static constexpr double anchor = 0.0;
  __device__ static void weights(double r, double* wt) {
#pragma clang fp contract(off)
    wt[0] = 0.5 * (1 - r) * (1 - r);
    wt[1] = 0.5 + r - r * r;
    wt[2] = 0.5 * r * r;
  }
};
template <>
struct Stencil<3> {  // CIC about cell centres: key cell a <= x - 1/2 < a + 1
  static constexpr int OMIN = 0, OMAX = 1;
  static constexpr double anchor = 0.5;
  __device__ static void weights(double r, double* wt) {
#pragma clang fp contract(off)
    wt[0] = 1.0 - r;
    wt[1] = r;
  }
};

template <typename Tg, int SCHEME>
__global__ void __launch_bounds__(kTileThreads)
    k_ma_tile(Tg* __restrict__ grid, GridGeom g, const int2* __restrict__ runs,
              const uint32_t* __restrict__ occ, const double4* __restrict__ srec, int tiles_y,
              int tiles_z) {
#pragma clang fp contract(off)
  using S = Stencil<SCHEME>;
  constexpr int W = S::OMAX - S::OMIN + 1;
  // Key cells whose stencil reaches the tile: [t0 - OMAX, t0 + T - 1 - OMIN].
  constexpr int RX = TX + W - 1, RY = TY + W - 1, RZ = TZ + W - 1;
  constexpr int RC = RX * RY * RZ;
  constexpr int kPer = (RC + kTileThreads - 1) / kTileThreads;
  __shared__ int32_t pre[RC + 1];  // counts, then their exclusive prefix sums
  __shared__ int32_t bg[RC];
  __shared__ double acc[kTileWaves][kTileCells];
  __shared__ uint32_t touched[kTileCells / 32];
  __shared__ int32_t wtot[kTileWaves];

  int b = blockIdx.x;
  const int bz = b % tiles_z;
  b /= tiles_z;
  const int by = b % tiles_y;
  const int bx = b / tiles_y;
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
  const int tid = threadIdx.x, lane = tid % kWave, wv = tid / kWave;

  // Run tables of the key cells: every occupancy word is requested before the
  // first is used, then the runs of the occupied cells, so a thread waits for
  // two rounds of loads rather than one per cell.
  {
    int64_t cell[kPer];
    uint32_t word[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int rc = tid + q * kTileThreads;
      cell[q] = -1;
      word[q] = 0u;
      if (rc < RC) {
        const int rz = rc % RZ, ry = (rc / RZ) % RY, rx = rc / (RZ * RY);
        const int gx = wrap_index(x0 - S::OMAX + rx, g.n0);
        const int gy = wrap_index(y0 - S::OMAX + ry, g.n1);
        const int gz = wrap_index(z0 - S::OMAX + rz, g.n2);
        cell[q] = ((int64_t)gx * g.n1 + gy) * g.n2 + gz;
        word[q] = occ[cell[q] >> 5];
      }
    }
    int2 run[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const bool on = cell[q] >= 0 && ((word[q] >> (cell[q] & 31)) & 1u);
      run[q] = on ? runs[cell[q]] : make_int2(0, 0);
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int rc = tid + q * kTileThreads;
      if (rc < RC) {
        pre[rc] = run[q].y;
        bg[rc] = run[q].x;
      }
    }
  }
  __syncthreads();
  // Exclusive scan of the counts: thread t owns entries [t kPer, (t+1) kPer).
  int loc[kPer];
  int s = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int i = tid * kPer + q;
    loc[q] = i < RC ? pre[i] : 0;
    s += loc[q];
  }
  int incl = s;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int t = __shfl_up(incl, d, kWave);
    if (lane >= d) incl += t;
  }
  if (lane == kWave - 1) wtot[wv] = incl;
  __syncthreads();
  int run = incl - s, total = 0;
#pragma unroll
  for (int w = 0; w < kTileWaves; ++w) {
    if (w < wv) run += wtot[w];
    total += wtot[w];
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int i = tid * kPer + q;
    if (i < RC) pre[i] = run;
    run += loc[q];
  }
  if (tid == 0) pre[RC] = total;
  if (total == 0) return;  // uniform: no particle reaches this tile
  for (int i = tid; i < kTileWaves * kTileCells; i += kTileThreads) (&acc[0][0])[i] = 0.0;
  for (int i = tid; i < kTileCells / 32; i += kTileThreads) touched[i] = 0u;
  __syncthreads();

  // Wave wv sums list entries [p0, p1) into its copy acc[wv].
  const int p0 = (int)((int64_t)total * wv / kTileWaves);
  const int p1 = (int)((int64_t)total * (wv + 1) / kTileWaves);
  double* mine = acc[wv];
  for (int base = p0; base < p1; base += kWave) {
    const int p = base + lane;
    const bool valid = p < p1;
    int rc = -1;
    double wx[W], wy[W], wz[W];
#pragma unroll
    for (int o = 0; o < W; ++o) wx[o] = wy[o] = wz[o] = 0.0;
    if (valid) {
      int lo = 0, hi = RC;  // pre[lo] <= p < pre[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= p)
          lo = mid;
        else
          hi = mid;
      }
      rc = lo;
      const double4 q = srec[bg[rc] + (p - pre[rc])];
      S::weights(q.x, wx);
      S::weights(q.y, wy);
      S::weights(q.z, wz);
#pragma unroll
      for (int o = 0; o < W; ++o) wx[o] = wx[o] * q.w;  // weight folded into x, as the reference
    }
    // Segments: runs of lanes with the same key cell (rc); sstart = first lane.
    const int prev = __shfl_up(rc, 1, kWave);
    const int next = __shfl_down(rc, 1, kWave);
    const bool head = lane == 0 || prev != rc;
    const bool tail = valid && (lane == kWave - 1 || next != rc);
    int sstart = head ? lane : 0;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int t = __shfl_up(sstart, d, kWave);
      if (lane >= d) sstart = max(sstart, t);
    }
    const bool singles = __ballot(valid && !head) == 0ull;  // wave-uniform
    const int rz = rc % RZ, ry = (rc / RZ) % RY, rx = rc / (RZ * RY);
#pragma unroll
    for (int a = 0; a < W; ++a) {
      const int tx = rx + a - (W - 1);
#pragma unroll
      for (int bb = 0; bb < W; ++bb) {
        const int ty = ry + bb - (W - 1);
        const double vxy = wx[a] * wy[bb];
#pragma unroll
        for (int c = 0; c < W; ++c) {
          const int tz = rz + c - (W - 1);
          double v = vxy * wz[c];
          if (!singles) {
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
              const double t = __shfl_up(v, d, kWave);
              if (lane - d >= sstart) v += t;
            }
          }
          if (tail && tx >= 0 && tx < TX && ty >= 0 && ty < TY && tz >= 0 && tz < TZ &&
              x0 + tx < g.n0 && y0 + ty < g.n1 && z0 + tz < g.n2) {
            const int t = (tx * TY + ty) * TZ + tz;
            mine[t] += v;
            atomicOr(&touched[t >> 5], 1u << (t & 31));
          }
        }
      }
    }
  }
  __syncthreads();
  // Each thread's touched cells: all grid loads first, then the stores.
  constexpr int kOut = kTileCells / kTileThreads;
  Tg old[kOut];
  Tg* dst[kOut];
#pragma unroll
  for (int q = 0; q < kOut; ++q) {
    const int t = tid + q * kTileThreads;
    const int tz = t % TZ, ty = (t / TZ) % TY, tx = t / (TZ * TY);
    const bool on = (touched[t >> 5] >> (t & 31)) & 1u;
    dst[q] = on ? grid + ((int64_t)(x0 + tx) * g.n1 + (y0 + ty)) * g.n2 + (z0 + tz) : nullptr;
    old[q] = on ? *dst[q] : Tg(0);
  }
#pragma unroll
  for (int q = 0; q < kOut; ++q) {
    if (!dst[q]) continue;
    const int t = tid + q * kTileThreads;
    double sum = acc[0][t];
#pragma unroll
    for (int w = 1; w < kTileWaves; ++w) sum += acc[w][t];
    *dst[q] = (Tg)((double)old[q] + sum);
  }
}

// D12 wavelet (distribution_functions.h:162-217): a particle in cell s with
// sub-cell sample (sx, sy, sz) = floor(64 r) adds
// tab[ox][sx] w * tab[oy][sy] * tab[oz][sz] to cell s - 1 + o, o in [0, 8)^3,
// tab being the caller's 8 x 64 sample table (d12.h's wavesample order).
// Target t therefore gathers from source cells t - 6 .. t + 1 per axis.
constexpr int WC = 8, WS = 64;
constexpr int WTX = 2, WTY = 4, WTZ = 32;  // target tile (256 cells, one per thread)
constexpr int DX = WTX + WC - 1, DY = WTY + WC - 1, DZ = WTZ + WC - 1;

template <typename Tg>
__global__ void __launch_bounds__(WTX* WTY* WTZ)
    k_ma_gather_wavelet(Tg* __restrict__ grid, GridGeom g, const int2* __restrict__ runs,
                        const uint32_t* __restrict__ occ,
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
  const int x0 = bx * WTX, y0 = by * WTY, z0 = bz * WTZ;
  for (int i = threadIdx.x; i < WC * WS; i += blockDim.x) tab[i] = table[i];
  for (int h = threadIdx.x; h < DX * DY * DZ; h += blockDim.x) {
    const int hz = h % DZ, hy = (h / DZ) % DY, hx = h / (DZ * DY);
    const int gx = wrap_index(x0 + hx - (WC - 2), g.n0);
    const int gy = wrap_index(y0 + hy - (WC - 2), g.n1);
    const int gz = wrap_index(z0 + hz - (WC - 2), g.n2);
    const int64_t c = ((int64_t)gx * g.n1 + gy) * g.n2 + gz;
    const int2 r = cell_run(occ, runs, c);
    cs[h] = r.x;
    ce[h] = r.x + r.y;
  }
  __syncthreads();
  const int tz = threadIdx.x % WTZ, ty = (threadIdx.x / WTZ) % WTY, tx = threadIdx.x / (WTZ * WTY);
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
void launch_gather(int scheme, Tg* grid, const GridGeom& gg, const int2* runs,
                   const uint32_t* occ, const double4* srec, hipStream_t s,
                   const double* table = nullptr) {
  const bool wav = scheme == 2;
  const int bx = wav ? WTX : TX, by = wav ? WTY : TY, bz = wav ? WTZ : TZ;
  const int tx = (gg.n0 + bx - 1) / bx, ty = (gg.n1 + by - 1) / by, tz = (gg.n2 + bz - 1) / bz;
  const uint64_t nb = (uint64_t)tx * ty * tz;
  if (nb > 0x7fffffffULL) fail(FC_ERR_INVALID, "grid too large for mass assignment");
  dim3 grd((unsigned)nb), blk(wav ? WTX * WTY * WTZ : kTileThreads);
  switch (scheme) {
    case 0:
      hipLaunchKernelGGL((k_ma_tile<Tg, 0>), grd, blk, 0, s, grid, gg, runs, occ, srec, ty, tz);
      break;
    case 1:
      hipLaunchKernelGGL((k_ma_tile<Tg, 1>), grd, blk, 0, s, grid, gg, runs, occ, srec, ty, tz);
      break;
    case 3:
      hipLaunchKernelGGL((k_ma_tile<Tg, 3>), grd, blk, 0, s, grid, gg, runs, occ, srec, ty, tz);
      break;
    case 2:
      hipLaunchKernelGGL((k_ma_gather_wavelet<Tg>), grd, blk, 0, s, grid, gg, runs, occ, srec,
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
  sc[1].reserve(sizeof(uint32_t) * n);
  sc[2].reserve(sizeof(uint32_t) * n);
  sc[3].reserve(sizeof(uint32_t) * n);
  sc[4].reserve(sizeof(uint32_t) * n);
  sc[5].reserve(sizeof(double4) * n);
  sc[6].reserve(sizeof(double4) * n);
  sc[7].reserve(sizeof(double) * 2 * (nbp + 1));
  sc[8].reserve(sizeof(unsigned long long) * (nbp + 1));
  const uint64_t nwords = (ncells + 31) / 32;
  sc[9].reserve(sizeof(int2) * ncells);
  sc[11].reserve(sizeof(uint32_t) * nwords);
  DeviceBuffer &keys = sc[1], &keys2 = sc[2], &idx = sc[3], &idx2 = sc[4], &rec = sc[5],
               &srec = sc[6], &partw = sc[7], &partn = sc[8], &runs = sc[9], &occ = sc[11];
  const double anchor = scheme == 3 ? 0.5 : 0.0;  // CIC keys by its 2x2x2 block

  hipLaunchKernelGGL(k_ma_prepare, dim3(nbp), dim3(kPrepThreads), 0, s, pos_dev, n, ncol, w_dev,
                     w_scalar, gg, wrap, shift, anchor, (uint32_t)sentinel, keys.as<uint32_t>(),
                     idx.as<uint32_t>(), rec.as<double4>(), partw.as<double>(),
                     partn.as<unsigned long long>());
  FC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ma_finish_stats, dim3(1), dim3(64), 0, s, partw.as<double>(),
                     partn.as<unsigned long long>(), nbp, partw.as<double>() + 2 * nbp,
                     partn.as<unsigned long long>() + nbp);
  FC_LAUNCH_CHECK();

  FC_HIP(hipEventRecord(ev[0], s));
  const int bits = key_bits(sentinel);
  ctx->scratch[0].reserve(radix_plan(n, bits).scratch_bytes());
  const bool second = radix_sort_pairs(keys.as<uint32_t>(), keys2.as<uint32_t>(),
                                       idx.as<uint32_t>(), idx2.as<uint32_t>(), n, bits,
                                       ctx->scratch[0].as<void>(), s);
  const uint32_t* skeys = second ? keys2.as<uint32_t>() : keys.as<uint32_t>();
  const uint32_t* sidx = second ? idx2.as<uint32_t>() : idx.as<uint32_t>();
  hipLaunchKernelGGL(k_permute, dim3(grid1d(n, 256, 65536)), dim3(256), 0, s, sidx,
                     rec.as<double4>(), srec.as<double4>(), n);
  FC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_zero_i32, dim3(grid1d(nwords, 256, 65536)), dim3(256), 0, s,
                     occ.as<int32_t>(), (int64_t)nwords);
  FC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cell_runs, dim3(grid1d(n, 256, 65536)), dim3(256), 0, s, skeys,
                     partn.as<unsigned long long>() + nbp, runs.as<int2>(), occ.as<uint32_t>());
  FC_LAUNCH_CHECK();
  FC_HIP(hipEventRecord(ev[1], s));
  if (grid->prec == FC_FP64)
    launch_gather<double>(scheme, grid->data.as<double>(), gg, runs.as<int2>(),
                          occ.as<uint32_t>(), srec.as<double4>(), s, wtab);
  else
    launch_gather<float>(scheme, grid->data.as<float>(), gg, runs.as<int2>(),
                         occ.as<uint32_t>(), srec.as<double4>(), s, wtab);
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

int fc_mass_assign_footprint(fc_grid* grid, int64_t n, int ncol, int has_weights,
                             int host_input, size_t* grow, size_t* transient) {
  return fc::guard([&] {
    fc::check_grid(grid);
    if (n < 0) fc::fail(FC_ERR_INVALID, "negative particle count");
    if (ncol != 3 && ncol != 4) fc::fail(FC_ERR_INVALID, "Rows must be 3D or 4D.");
    fc_ctx* ctx = grid->ctx;
    fc::DeviceGuard dg(ctx->device);
    size_t g = 0;
    if (n > 0) {
      const uint64_t ncells = grid->size();
      const int nbp = (int)std::min<int64_t>(1024, (n + fc::kPrepThreads - 1) / fc::kPrepThreads);
      const size_t temp = fc::radix_plan(n, fc::key_bits(ncells)).scratch_bytes();
      const size_t want[12] = {temp,
                               sizeof(uint32_t) * n,
                               sizeof(uint32_t) * n,
                               sizeof(uint32_t) * n,
                               sizeof(uint32_t) * n,
                               sizeof(double4) * n,
                               sizeof(double4) * n,
                               sizeof(double) * 2 * (nbp + 1),
                               sizeof(unsigned long long) * (nbp + 1),
                               sizeof(int2) * ncells,
                               0,
                               sizeof(uint32_t) * ((ncells + 31) / 32)};
      for (int i = 0; i < 12; ++i)
        if (want[i] > ctx->scratch[i].bytes()) g += want[i] - ctx->scratch[i].bytes();
    }
    if (grow) *grow = g;
    if (transient)
      *transient = host_input && n > 0
                       ? sizeof(double) * (size_t)n * (ncol + (has_weights ? 1 : 0))
                       : 0;
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
