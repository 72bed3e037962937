// Stable LSD radix sort of (32-bit key, 32-bit value) pairs by the low `bits`
// bits of the key, for the mass assignment's (cell key, particle index) pairs
// (the counterpart of the reference's merge sort by cell, merge_sort_omp.h).
//
// Keys below 2^bits are sorted in ceil(bits / 8) passes of at most 8 bits
// (the 28-bit keys of a 512^3 grid, cell index or the skip sentinel 2^27: 4
// passes of 7 bits; wider digits cost more per tile in counters than the
// pass they save).  Each pass is a stable counting sort over tiles of 4096
// pairs:
//   k_rs_count    digit counts per tile, [digit][tile];
//   k_rs_scan     one workgroup per digit: the exclusive prefix of its row
//                 over tiles (in place) and the digit total;
//   k_rs_scatter  each wave ranks its keys in input order (the lanes of a
//                 64-key round that share a digit found with ballots), the
//                 tile is staged in LDS digit-major, and each digit's run is
//                 written to its place: digit base + earlier tiles' count +
//                 rank in the tile, so equal keys keep their input order and
//                 the stores are contiguous runs.
// Integer counting only: the result depends on the input alone.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "fc_internal.h"

namespace fc {

constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / 64;
constexpr int kRsRounds = 16;                    // 64-key rounds per wave
constexpr int kRsTile = kRsThreads * kRsRounds;  // pairs per tile
constexpr int kRsMaxBits = 8;
constexpr int kRsMaxDigits = 1 << kRsMaxBits;

struct RadixPlan {
  int passes = 0, dbits = 0, nb = 0;
  int64_t tiles = 0;
  // Scratch bytes: counts [nb][tiles] and totals [nb] (int32).
  size_t scratch_bytes() const { return sizeof(int32_t) * ((size_t)tiles * nb + nb) + 256; }
};

inline RadixPlan radix_plan(int64_t n, int bits) {
  RadixPlan p;
  p.passes = std::max(1, (bits + kRsMaxBits - 1) / kRsMaxBits);
  p.dbits = (bits + p.passes - 1) / p.passes;
  if (p.dbits < 1) p.dbits = 1;
  p.nb = 1 << p.dbits;
  p.tiles = (n + kRsTile - 1) / kRsTile;
  return p;
}

// Exclusive scan of a[0, m) in LDS (m <= kRsMaxDigits) by a 256-thread block;
// tmp holds kRsWaves ints.  Ends with a barrier.
__device__ __forceinline__ void rs_block_scan(int32_t* a, int m, int32_t* tmp) {
  constexpr int PER = kRsMaxDigits / kRsThreads;
  const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
  int32_t v[PER], sum = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int i = tid * PER + q;
    v[q] = i < m ? a[i] : 0;
    sum += v[q];
  }
  // Inclusive scan of the thread sums over the wave, then over the waves.
  int32_t inc = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) tmp[w] = inc;
  __syncthreads();
  int32_t wbase = 0;
  for (int q = 0; q < w; ++q) wbase += tmp[q];
  int32_t run = wbase + inc - sum;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int i = tid * PER + q;
    if (i < m) a[i] = run;
    run += v[q];
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kRsThreads)
    k_rs_count(const uint32_t* __restrict__ keys, int64_t n, int shift, int nb,
               int32_t* __restrict__ counts) {
  __shared__ int32_t c[kRsMaxDigits];
  for (int i = threadIdx.x; i < nb; i += kRsThreads) c[i] = 0;
  __syncthreads();
  const int64_t t0 = (int64_t)blockIdx.x * kRsTile;
  const uint32_t mask = (uint32_t)nb - 1u;
  for (int i = threadIdx.x; i < kRsTile; i += kRsThreads) {
    const int64_t k = t0 + i;
    if (k < n) atomicAdd(&c[(keys[k] >> shift) & mask], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += kRsThreads)
    counts[(int64_t)i * gridDim.x + blockIdx.x] = c[i];
}

// Block d: the row of digit d over all tiles, in chunks of 1024 with a
// running carry.
__global__ void __launch_bounds__(kRsThreads)
    k_rs_scan(int32_t* __restrict__ counts, int64_t tiles, int32_t* __restrict__ totals) {
  __shared__ int32_t a[kRsMaxDigits];
  __shared__ int32_t tmp[kRsWaves];
  __shared__ int32_t carry;
  int32_t* row = counts + (int64_t)blockIdx.x * tiles;
  if (threadIdx.x == 0) carry = 0;
  for (int64_t c0 = 0; c0 < tiles; c0 += kRsMaxDigits) {
    const int m = (int)std::min<int64_t>(kRsMaxDigits, tiles - c0);
    for (int i = threadIdx.x; i < m; i += kRsThreads) a[i] = row[c0 + i];
    __syncthreads();
    const int32_t last = a[m - 1];
    rs_block_scan(a, m, tmp);
    const int32_t cin = carry;
    for (int i = threadIdx.x; i < m; i += kRsThreads) row[c0 + i] = cin + a[i];
    __syncthreads();
    if (threadIdx.x == 0) carry = cin + a[m - 1] + last;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(kRsThreads)
    k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                 uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n, int shift,
                 int dbits, const int32_t* __restrict__ prefix, const int32_t* __restrict__ totals) {
  __shared__ int32_t wcnt[kRsWaves][kRsMaxDigits];
  __shared__ int32_t lstart[kRsMaxDigits];
  __shared__ int32_t gstart[kRsMaxDigits];
  __shared__ uint32_t sk[kRsTile], sv[kRsTile];
  __shared__ int32_t tmp[kRsWaves];
  const int nb = 1 << dbits;
  const uint32_t mask = (uint32_t)nb - 1u;
  const int tid = threadIdx.x, w = tid / 64, lane = tid % 64;
  for (int i = tid; i < nb; i += kRsThreads) {
#pragma unroll
    for (int q = 0; q < kRsWaves; ++q) wcnt[q][i] = 0;
    gstart[i] = totals[i];
  }
  __syncthreads();
  rs_block_scan(gstart, nb, tmp);  // digit bases
  for (int i = tid; i < nb; i += kRsThreads)
    gstart[i] += prefix[(int64_t)i * gridDim.x + blockIdx.x];  // + earlier tiles of the digit
  const int64_t t0 = (int64_t)blockIdx.x * kRsTile;
  const int64_t wk0 = t0 + (int64_t)w * (kRsRounds * 64);
  uint32_t key[kRsRounds], val[kRsRounds];
  int32_t pos[kRsRounds];
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    const int64_t k = wk0 + r * 64 + lane;
    key[r] = k < n ? kin[k] : 0u;
    val[r] = k < n ? vin[k] : 0u;
  }
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    const bool in = wk0 + r * 64 + lane < n;
    const uint32_t d = (key[r] >> shift) & mask;
    unsigned long long peers = __ballot(in);
    for (int b = 0; b < dbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __popcll(peers & below);
    pos[r] = in ? wcnt[w][d] + rank : 0;  // every lane reads before the leaders add
    if (in && rank == 0) wcnt[w][d] += __popcll(peers);
  }
  __syncthreads();
  // Per digit: the waves' offsets (in wave order) and the tile's count.
  for (int i = tid; i < nb; i += kRsThreads) {
    int32_t run = 0;
#pragma unroll
    for (int q = 0; q < kRsWaves; ++q) {
      const int32_t c = wcnt[q][i];
      wcnt[q][i] = run;
      run += c;
    }
    lstart[i] = run;
  }
  __syncthreads();
  rs_block_scan(lstart, nb, tmp);  // tile-local digit starts
#pragma unroll
  for (int r = 0; r < kRsRounds; ++r) {
    if (wk0 + r * 64 + lane < n) {
      const uint32_t d = (key[r] >> shift) & mask;
      const int lp = lstart[d] + wcnt[w][d] + pos[r];
      if (lp >= 0 && lp < kRsTile) {
        sk[lp] = key[r];
        sv[lp] = val[r];
      }
    }
  }
  __syncthreads();
  const int tn = (int)std::min<int64_t>(kRsTile, n - t0);
  for (int i = tid; i < tn; i += kRsThreads) {
    const uint32_t kk = sk[i];
    const uint32_t d = (kk >> shift) & mask;
    const int64_t gp = (int64_t)gstart[d] + (i - lstart[d]);
    if (gp >= 0 && gp < n) {
      kout[gp] = kk;
      vout[gp] = sv[i];
    }
  }
}

// Sorts n pairs (k1, v1) by the low `bits` key bits with (k2, v2) as the
// second buffer and scratch (RadixPlan::scratch_bytes) for the counts.
// Returns true when the result is in (k2, v2), false when in (k1, v1).
inline bool radix_sort_pairs(uint32_t* k1, uint32_t* k2, uint32_t* v1, uint32_t* v2, int64_t n,
                             int bits, void* scratch, hipStream_t s) {
  if (n <= 0) return false;
  if (n > INT32_MAX) fail(FC_ERR_INVALID, "too many particles for one sort");
  if (bits < 1 || bits > 32) fail(FC_ERR_RUNTIME, "internal: radix sort bit count");
  const RadixPlan p = radix_plan(n, bits);
  if (p.tiles > INT32_MAX / kRsMaxDigits) fail(FC_ERR_INVALID, "too many particles for one sort");
  int32_t* counts = static_cast<int32_t*>(scratch);
  int32_t* totals = counts + (size_t)p.tiles * p.nb;
  uint32_t *ki = k1, *ko = k2, *vi = v1, *vo = v2;
  for (int q = 0; q < p.passes; ++q) {
    const int shift = q * p.dbits;
    hipLaunchKernelGGL(k_rs_count, dim3((unsigned)p.tiles), dim3(kRsThreads), 0, s, ki, n, shift,
                       p.nb, counts);
    FC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rs_scan, dim3((unsigned)p.nb), dim3(kRsThreads), 0, s, counts,
                       p.tiles, totals);
    FC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)p.tiles), dim3(kRsThreads), 0, s, ki, vi, ko,
                       vo, n, shift, p.dbits, counts, totals);
    FC_LAUNCH_CHECK();
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  return p.passes % 2 == 1;
}

}  // namespace fc
