// Handle types shared by the grid, mass-assignment and correlator sources,
// plus deterministic FP64 block reductions.
#pragma once

#include "fc_internal.h"

// Device-resident density grid [n0][n1][n2] (cc/grid/config_space_grid.h:10-56).
struct fc_grid {
  fc_ctx* ctx = nullptr;
  int shape[3] = {0, 0, 0};
  double posmin[3] = {0, 0, 0};
  double cell = 1.0;
  int prec = FC_FP64;  // storage: FC_FP64 -> double, FC_FP32 -> float
  fc::DeviceBuffer data;
  bool batch_view = false;  // data is a slice of a batch's input buffer
  uint64_t size() const { return (uint64_t)shape[0] * shape[1] * shape[2]; }
  size_t elem() const { return prec == FC_FP64 ? sizeof(double) : sizeof(float); }
};

namespace fc {

constexpr int kWave = 64;

// Sum over the block in a fixed order (wave butterflies, then waves in
// order), so repeated runs give identical bits.  All threads must call it;
// the result is valid in thread 0.  scratch: >= blockDim/64 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + kWave - 1) / kWave;
    for (int w = 0; w < nw; ++w) s += scratch[w];
  }
  __syncthreads();
  return s;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v,
                                                            unsigned long long* scratch) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  unsigned long long s = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + kWave - 1) / kWave;
    for (int w = 0; w < nw; ++w) s += scratch[w];
  }
  __syncthreads();
  return s;
}

inline void check_grid(const fc_grid* g) {
  if (!g) fail(FC_ERR_INVALID, "null grid");
}

}  // namespace fc
