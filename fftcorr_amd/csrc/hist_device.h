// Deterministic wavefront-reduced histogramming (HistogramList::accumulate,
// cc/histogram/histogram_list.h:33-45) for NV values per sample.
//
// Within a 64-lane wave, the lanes that share a bin are summed with a fixed
// xor butterfly and the leader lane adds the sum to that wave's private LDS
// bins; waves are then combined in wave order into one partial per block, and
// blocks (fixed count, contiguous sample ranges) in block order by
// k_hist_final.  Counts are integers (exact in any order).  The summation
// order therefore depends only on the problem shape, never on scheduling.
#pragma once

#include <functional>

#include "fc_objects.h"
#include "xcd.h"

namespace fc {

constexpr int kHistBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Adds (bin, v[NV]) of every lane of the calling wave; bin < 0 skips.
// wbins: this wave's [NV][nbins] LDS bins; lcounts: the block's [nbins] counts.
template <int NV>
__device__ __forceinline__ void wave_hist_add(int bin, const double* v, double* wbins,
                                              int* lcounts, int nbins) {
  const int lane = threadIdx.x % kWave;
  unsigned long long pending = __ballot(bin >= 0);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lb = __shfl(bin, leader, kWave);
    const bool mine = (bin == lb);
    const unsigned long long m = __ballot(mine);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double s = wave_sum(mine ? v[q] : 0.0);
      if (lane == leader) wbins[q * nbins + lb] += s;
    }
    if (lane == leader) atomicAdd(&lcounts[lb], (int)__popcll(m));  // shared by waves
    pending &= ~m;
  }
}

// LDS bytes for a block of `threads` threads.
inline size_t hist_lds_bytes(int nv, int nbins, int threads) {
  return sizeof(double) * (size_t)(threads / kWave) * nv * nbins + sizeof(int) * (size_t)nbins;
}

// Picks 256 threads if the per-wave bins fit, else 64; 0 if even one wave's
// bins do not fit in 64 KB.
inline int hist_threads(int nv, int nbins) {
  if (hist_lds_bytes(nv, nbins, 256) <= 65536) return 256;
  if (hist_lds_bytes(nv, nbins, 64) <= 65536) return 64;
  return 0;
}

// Generic histogram kernel: F(idx, &bin, v) fills bin (or -1) and NV values
// for sample idx < n.  part_hist [gridDim][NV][nbins], part_counts [gridDim][nbins].
template <int NV, class F>
__global__ void __launch_bounds__(256)
    k_hist(F f, int64_t n, int nbins, double* __restrict__ part_hist,
           int* __restrict__ part_counts) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nw = blockDim.x / kWave;
  double* bins = reinterpret_cast<double*>(smem_raw);
  int* lcounts = reinterpret_cast<int*>(bins + (size_t)nw * NV * nbins);
  for (int i = threadIdx.x; i < nw * NV * nbins; i += blockDim.x) bins[i] = 0.0;
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) lcounts[i] = 0;
  __syncthreads();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  double* wb = bins + (size_t)(threadIdx.x / kWave) * NV * nbins;
  const int64_t rounds = hi > lo ? (hi - lo + blockDim.x - 1) / blockDim.x : 0;
  for (int64_t r = 0; r < rounds; ++r) {
    const int64_t idx = lo + r * blockDim.x + threadIdx.x;
    int bin = -1;
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    if (idx < hi) f(idx, &bin, v);
    wave_hist_add<NV>(bin, v, wb, lcounts, nbins);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NV * nbins; i += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += bins[(size_t)w * NV * nbins + i];
    part_hist[(size_t)blockIdx.x * NV * nbins + i] = s;
  }
  for (int i = threadIdx.x; i < nbins; i += blockDim.x)
    part_counts[(size_t)blockIdx.x * nbins + i] = lcounts[i];
}

// Cells of an [s0, s1, s2] subgrid grouped by histogram bin.  The bin of a
// cell depends only on the subgrid axes (correlator.h:129-135, :180-186 and
// histogram_list.h:57), so it is computed once on the host, in FP64 exactly
// as the reference does, and the cells are stored sorted by bin (raster
// order within a bin).  Each bin's entries are then one contiguous range,
// cut into chunks of at most kSegChunk entries: a histogram is a segmented
// sum, one workgroup per chunk, with no per-entry bin and the per-bin counts
// known without counting.
constexpr int kSegThreads = 256;
constexpr int kSegChunk = kSegThreads * 16;

struct BinnedCells {
  DeviceBuffer cells;        // uint32 linear index (i*s1 + j)*s2 + k, sorted by bin
  DeviceBuffer counts;       // int32 [nbins]
  DeviceBuffer chunks;       // int4 {bin, first entry, entries, 0} per chunk
  // int32 [nchunks]: chunks in launch order, by the first cell they hold, so
  // that chunks of neighbouring bins reading the same cache lines run
  // together (and, dealt to the XCDs in contiguous ranges, share an L2).
  DeviceBuffer order;
  DeviceBuffer chunk_first;  // int32 [nbins + 1]: first chunk of each bin
  std::vector<int> h_counts;
  int64_t n = 0;             // entries (cells whose bin is in range)
  int nbins = 0;
  int nchunks = 0;
  int64_t tracked_pos = -1;  // entry position of the tracked cell (or -1)
};

// track_cell >= 0: record the entry position of that cell in tracked_pos.
// order (optional): within each bin, cells are ordered by (order(cell), cell)
// instead of raster order -- e.g. by the address the histogram kernel reads
// for the cell, so that neighbouring entries read neighbouring values.
// cell_bins (optional): receives the bin of every cell in raster order (-1
// outside the histogram); needs nbins < 32768.
void build_binned_cells(BinnedCells& bc, const std::vector<double>& ax,
                        const std::vector<double>& ay, const std::vector<double>& az,
                        double binsize, int nbins, int64_t track_cell = -1,
                        const std::function<uint64_t(uint32_t)>* order = nullptr,
                        std::vector<int16_t>* cell_bins = nullptr);

// Unsigned 32-bit division by a runtime constant d: q = (umulhi(x, m) + x) >> s
// with the round-up magic m = floor(2^32 (2^s - d) / d) + 1, s = ceil(log2 d)
// (Granlund and Montgomery, PLDI 1994), the add done in 64 bits.
struct FastDiv {
  uint32_t m = 0;
  int s = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t d) {
    if (d <= 1) return;
    int l = 0;
    while ((1ull << l) < d) ++l;
    m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    s = l;
  }
  __device__ __forceinline__ uint32_t div(uint32_t x) const {
    const uint64_t t = __umulhi(x, m);
    return (uint32_t)((t + x) >> s);
  }
};

// Segmented histogram over bin-sorted entries: workgroup b sums the NV values
// g(pos, v) of the entries of chunk b (each thread in entry order, then a
// fixed tree over lanes and waves) into part[b][NV].  When pos == zpos the
// first value is also stored to *zval (the zero-lag cell).  blockIdx.y is
// the batch member: g.offset(m) selects its values, and its partials and
// zero-lag value go to part[m][nchunks][NV] and zval[m].
template <int NV, class G>
__global__ void __launch_bounds__(kSegThreads)
    k_hist_seg(G g, const int4* __restrict__ chunks, const int* __restrict__ order,
               double* __restrict__ part, int64_t zpos, double* __restrict__ zval) {
  __shared__ double sw[kSegThreads / kWave][NV];
  g.offset(blockIdx.y);
  part += (size_t)blockIdx.y * gridDim.x * NV;
  if (zval) zval += blockIdx.y;
  const int cid = order[xcd_tile(blockIdx.x, gridDim.x)];
  const int4 ch = chunks[cid];
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int i = threadIdx.x; i < ch.z; i += kSegThreads) {
    const int64_t pos = (int64_t)ch.y + i;
    double v[NV];
    g(pos, v);
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += v[q];
    if (pos == zpos) *zval = v[0];
  }
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double s = wave_sum(acc[q]);
    if (lane == 0) sw[w][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int k = 0; k < kSegThreads / kWave; ++k) s += sw[k][threadIdx.x];
    part[(size_t)cid * NV + threadIdx.x] = s;
  }
}

// One wave per (multipole, bin): hist[row0 + q][bin] += the chunk partials
// of the bin (lane l sums chunks l, l + 64, ... in order, then a fixed xor
// tree); counts rows [crow0, crow1) += counts of the bin.
template <int NV>
__global__ void __launch_bounds__(kWave)
    k_hist_seg_final(const double* __restrict__ part, const int* __restrict__ chunk_first,
                     const int* __restrict__ bin_counts, int nbins, double* __restrict__ hist,
                     int* __restrict__ counts, int row0, int crow0, int crow1, int nchunks,
                     int64_t out_bstride) {
  const int i = blockIdx.x;
  part += (size_t)blockIdx.y * nchunks * NV;
  hist += blockIdx.y * out_bstride;
  counts += blockIdx.y * out_bstride;
  const int lane = threadIdx.x;
  if (i < NV * nbins) {
    const int q = i / nbins, bin = i - q * nbins;
    double s = 0.0;
    for (int c = chunk_first[bin] + lane; c < chunk_first[bin + 1]; c += kWave)
      s += part[(size_t)c * NV + q];
    s = wave_sum(s);
    if (lane == 0) hist[(size_t)(row0 + q) * nbins + bin] += s;
  }
  if (i < nbins && lane == 0)
    for (int r = crow0; r < crow1; ++r) counts[(size_t)r * nbins + i] += bin_counts[i];
}

// One block per (multipole, bin): hist[row0 + q][bin] += the block partials
// summed in a fixed order (strided per thread, then a fixed tree); counts
// rows [crow0, crow1) += static_counts[bin] or the summed block counts.
template <int NV>
__global__ void __launch_bounds__(256)
    k_hist_final(const double* __restrict__ part_hist, const int* __restrict__ part_counts,
                 const int* __restrict__ static_counts, int nblocks, int nbins,
                 double* __restrict__ hist, int* __restrict__ counts, int row0, int crow0,
                 int crow1) {
  __shared__ double sd[256 / kWave];
  __shared__ unsigned long long su[256 / kWave];
  const int i = blockIdx.x;
  if (i < NV * nbins) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
      s += part_hist[(size_t)b * NV * nbins + i];
    s = block_sum(s, sd);
    if (threadIdx.x == 0) {
      const int q = i / nbins, bin = i % nbins;
      hist[(size_t)(row0 + q) * nbins + bin] += s;
    }
  }
  if (i < nbins) {
    unsigned long long c = 0;
    if (static_counts) {
      c = (unsigned long long)static_counts[i];
    } else {
      for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        c += (unsigned long long)part_counts[(size_t)b * nbins + i];
      c = block_sum_u64(c, su);
    }
    if (threadIdx.x == 0)
      for (int r = crow0; r < crow1; ++r) counts[(size_t)r * nbins + i] += (int)c;
  }
}

// Scratch for one histogram pass.
struct HistScratch {
  DeviceBuffer part_hist, part_counts, part_seg;
  void reserve(int nv, int nbins) {
    part_hist.reserve(sizeof(double) * (size_t)kHistBlocks * nv * nbins);
    part_counts.reserve(sizeof(int) * (size_t)kHistBlocks * nbins);
  }
};

// nb batch members: member m's histogram rows start at hist + m * out_bstride
// (counts likewise).
// bin_counts (optional): the counts added per bin, instead of bc's own.
template <int NV, class G>
void run_hist_seg(G g, const BinnedCells& bc, DeviceBuffer& part, double* hist, int* counts,
                  int row0, int crow0, int crow1, hipStream_t s, int64_t zpos = -1,
                  double* zval = nullptr, int nb = 1, int64_t out_bstride = 0,
                  const int* bin_counts = nullptr) {
  const int nbins = bc.nbins;
  if (bc.nchunks > 0) {
    part.reserve(sizeof(double) * (size_t)bc.nchunks * NV * nb);
    hipLaunchKernelGGL((k_hist_seg<NV, G>), dim3(bc.nchunks, nb), dim3(kSegThreads), 0, s, g,
                       bc.chunks.as<int4>(), bc.order.as<int>(), part.as<double>(), zpos, zval);
    FC_LAUNCH_CHECK();
  }
  const int tot = NV * nbins > nbins ? NV * nbins : nbins;
  hipLaunchKernelGGL((k_hist_seg_final<NV>), dim3(tot, nb), dim3(kWave), 0, s,
                     part.as<double>(), bc.chunk_first.as<int>(),
                     bin_counts ? bin_counts : bc.counts.as<int>(), nbins,
                     hist, counts, row0, crow0, crow1, bc.nchunks, out_bstride);
  FC_LAUNCH_CHECK();
}

template <int NV, class F>
void run_hist(F f, int64_t n, int nbins, HistScratch& hs, double* hist, int* counts, int row0,
              int crow0, int crow1, hipStream_t s) {
  const int thr = hist_threads(NV, nbins);
  if (thr == 0) fail(FC_ERR_INVALID, "too many histogram bins (" + std::to_string(nbins) + ")");
  hs.reserve(NV, nbins);
  int nb = (int)std::min<int64_t>(kHistBlocks, (n + thr - 1) / thr);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL((k_hist<NV, F>), dim3(nb), dim3(thr), hist_lds_bytes(NV, nbins, thr), s, f,
                     n, nbins, hs.part_hist.as<double>(), hs.part_counts.as<int>());
  FC_LAUNCH_CHECK();
  const int tot = NV * nbins > nbins ? NV * nbins : nbins;
  hipLaunchKernelGGL((k_hist_final<NV>), dim3(tot), dim3(256), 0, s, hs.part_hist.as<double>(),
                     hs.part_counts.as<int>(), (const int*)nullptr, nb, nbins, hist, counts,
                     row0, crow0, crow1);
  FC_LAUNCH_CHECK();
}

}  // namespace fc
