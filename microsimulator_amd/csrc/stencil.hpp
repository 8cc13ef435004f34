// Device helpers shared by the pressure-solve kernels: the harmonic mean of
// the assembly, the marching-tile geometry, and the per-block partial sums of
// the Krylov kernels with the fixed-order sum every block re-computes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace msim {

constexpr int kStream = 256;  // threads per block of the streaming kernels

// cpp/core/flow_system.hpp:171-174.
__device__ __host__ inline float harmonic(float a, float b) {
  const float sum = a + b;
  return sum > 0.0f ? 2.0f * a * b / sum : 0.0f;
}

struct Tile {
  uint32_t ty, x0, x1, y;
  bool valid, lo_halo, hi_halo;
  size_t ylo, yhi;
};

template <int TY>
__device__ inline Tile tile_of(uint32_t nx, uint32_t ny, uint32_t rows) {
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t nty = (ny + TY - 1) / TY;
  Tile tile;
  tile.ty = t % nty;
  tile.x0 = (t / nty) * rows;
  tile.x1 = min(nx, tile.x0 + rows);
  tile.y = tile.ty * TY + threadIdx.x;
  tile.valid = tile.y < ny;
  // Thread 0 also carries the site below the tile, thread TY-1 the one above.
  tile.lo_halo = threadIdx.x == 0 && tile.ty > 0;
  tile.hi_halo = threadIdx.x == TY - 1 && tile.ty * TY + TY < ny;
  tile.ylo = tile.ty * TY - 1;
  tile.yhi = static_cast<size_t>(tile.ty) * TY + TY;
  return tile;
}

// Per-block partial sums of the Krylov kernels, two generations of each
// (indexed by iteration parity): p . A p from the operator kernel (p1 blocks),
// r . r and r . M^-1 r from the update kernel and the set-up (p2 blocks).
struct CgPartials {
  double* pq;
  double* rz;
  double* rr;
  uint32_t p1, p2;
};

// Sum of `count` partials in a fixed order, the same in every thread of every
// block that calls it: thread t adds t, t+NT, ... in order, then the block
// combines waves in order.
template <int NT>
__device__ inline double fixed_sum(const double* v, uint32_t count, double* lds, double* bcast) {
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < count; i += NT) s += v[i];
  s = block_sum<NT>(s, lds);
  if (threadIdx.x == 0) *bcast = s;
  __syncthreads();
  const double out = *bcast;
  __syncthreads();
  return out;
}

}  // namespace msim
