// Shared device helpers: wave64 / block reductions in a fixed order, the
// last-block-done pattern that finishes a grid-wide reduction inside the
// kernel that produced it, and the XCD-aware tile order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace msim {

constexpr int kWave = 64;

// Sum over the 64 lanes of a wavefront by xor butterflies: a fixed pattern,
// so the result does not depend on scheduling.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, kWave);
  return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v = fmax(v, __shfl_xor(v, offset, kWave));
  return v;
}

// Sum over a block of NT threads; the result is valid in thread 0. lds must
// hold NT/64 doubles. Waves are combined in wave order.
template <int NT>
__device__ inline double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double total = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) total += lds[w];
  }
  __syncthreads();
  return total;
}

template <int NT>
__device__ inline double block_max(double v, double* lds) {
  v = wave_max(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double total = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) total = fmax(total, lds[w]);
  }
  __syncthreads();
  return total;
}

// Publishes this block's K partial values (held by thread 0 in `values`) at
// partials[k * blocks + block] and takes a ticket. Returns true in every
// thread of the block that arrives last; that block then sees every other
// block's partials (agent-scope release/acquire through the fences).
template <int K>
__device__ inline bool publish_partials(const double (&values)[K], double* partials,
                                        uint32_t* ticket, uint32_t block, uint32_t blocks,
                                        bool* last_flag) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[static_cast<size_t>(k) * blocks + block] = values[k];
    __threadfence();
    const uint32_t arrived = atomicAdd(ticket, 1u);
    *last_flag = arrived + 1 == blocks;
  }
  __syncthreads();
  const bool last = *last_flag;
  if (last) __threadfence();
  return last;
}

// In the last block: sums partials[k*blocks .. k*blocks+blocks) for k < K in
// a fixed order (thread t takes t, t+NT, ... in order, then the block tree).
// Plain loads suffice: the acquire fence publish_partials executed in this
// block, paired with every other block's release fence through the ticket,
// makes their partials visible (it invalidates this CU's L1 and its XCD's
// L2). Result valid in thread 0.
template <int NT, int K>
__device__ inline void sum_partials(const double* partials, uint32_t blocks, double (&out)[K],
                                    double* lds) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double v = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += NT) {
      v += partials[static_cast<size_t>(k) * blocks + b];
    }
    out[k] = block_sum<NT>(v, lds);
  }
}

template <int NT, int K>
__device__ inline void max_partials(const double* partials, uint32_t blocks, double (&out)[K],
                                    double* lds) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double v = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += NT) {
      v = fmax(v, partials[static_cast<size_t>(k) * blocks + b]);
    }
    out[k] = block_max<NT>(v, lds);
  }
}

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2.
// Reorder so that XCD k works through a contiguous run of tiles: neighbouring
// tiles (which share halo rows) then meet in one L2.
__device__ inline uint32_t xcd_tile(uint32_t block, uint32_t blocks) {
  constexpr uint32_t kXcds = 8;
  if (blocks % kXcds != 0) return block;
  return (block % kXcds) * (blocks / kXcds) + block / kXcds;
}

}  // namespace msim
