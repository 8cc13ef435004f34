// Biomass deposition and colony mobility on the device: colony_volume_fraction
// and colony_mobility (python/src/microsimulator/flow.py:224-382).
//
// k_colony_deposit gives each cell one wave (a persistent loop over cells).
// The wave builds the cell's tent-kernel voxel integrals over its support
// (the separable per-axis integrals in LDS, their product per voxel). With an
// obstacle mask it zeroes solid voxels, finds the strongest voxel (the first
// in the support's x, y, z order) and keeps only the face-connected fluid
// component around it: on one layer as per-row bit masks grown by shifts and
// wave shuffles, otherwise by label propagation in LDS, to a fixpoint. It then
// adds the cell's biomass volume times the normalised weights into the grid
// with FP64 atomics. k_colony_finish divides by the voxel volume and, for the
// mobility, applies the Kozeny-Carman drag of each column's mean volume
// fraction. The atomics make the order of a voxel's additions differ from the
// reference's cell order, so results agree to FP64 rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kDepositThreads = 64;  // one wave per cell

struct DepositParams {
  uint32_t dims[3];
  double origin[3];
  double spacing[3];
  double radius;
  uint32_t cells;
  uint32_t has_mask;
  uint32_t axis_cap;  // LDS slots for the three axes' weights
  uint32_t region_cap;
};

// The tent kernel's cumulative weight at a scaled edge (flow.py:245-254).
__device__ inline double tent_cdf(double e) {
  if (e <= -1.0) return 0.0;
  if (e < 0.0) {
    const double t = e + 1.0;
    return 0.5 * (t * t);
  }
  if (e < 1.0) {
    const double t = 1.0 - e;
    return 1.0 - 0.5 * (t * t);
  }
  return 1.0;
}

__device__ inline void wave_argmax(double& value, uint32_t& index) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    const double v = __shfl_xor(value, offset, kWave);
    const uint32_t i = __shfl_xor(index, offset, kWave);
    if (v > value || (v == value && i < index)) {
      value = v;
      index = i;
    }
  }
}

// One cell per wave, the waves looping over the cells. A cell whose support
// has no connected fluid records itself in `failure` (the least such cell).
// With `momentum` (colony_flow_coupling, flow.py:384-424): each voxel's share
// times the cell's in-plane velocity (vx, vy) is also added into its column's
// momentum (x, y).
__global__ __launch_bounds__(kDepositThreads) void k_colony_deposit(
    DepositParams P, const uint8_t* __restrict__ solid, const float* __restrict__ px,
    const float* __restrict__ py, const float* __restrict__ pz, const float* __restrict__ lengths,
    const float* __restrict__ radii, const float* __restrict__ vx, const float* __restrict__ vy,
    double* volume, double* momentum, unsigned int* failure) {
  extern __shared__ double s_mem[];
  double* s_axis = s_mem;                   // axis weights: x, then y, then z
  double* s_kernel = s_mem + P.axis_cap;    // the support's weights
  uint8_t* s_connected = reinterpret_cast<uint8_t*>(s_kernel + P.region_cap);
  __shared__ int s_changed;
  const uint32_t tid = threadIdx.x;
  for (uint32_t c = blockIdx.x; c < P.cells; c += gridDim.x) {
    const double position[3] = {static_cast<double>(px[c]), static_cast<double>(py[c]),
                                static_cast<double>(pz[c])};
    // Outside the modelled volume: omitted (flow.py:237-241).
    bool outside = false;
    uint32_t lo[3], count[3];
    for (int a = 0; a < 3; ++a) {
      const double p = position[a], o = P.origin[a], h = P.spacing[a];
      const uint32_t n = P.dims[a];
      if (p < o - h / 2.0 || p >= o + (static_cast<double>(n) - 0.5) * h) outside = true;
      const double low = floor((p - P.radius - o) / h + 0.5);
      const double high = ceil((p + P.radius - o) / h + 0.5);
      const double l = fmax(low, 0.0), u = fmin(high, static_cast<double>(n));
      lo[a] = outside ? 0 : static_cast<uint32_t>(l);
      count[a] = outside || u <= l ? 0 : static_cast<uint32_t>(u - l);
    }
    if (outside) continue;
    const uint32_t offset[3] = {0, count[0], count[0] + count[1]};
    const uint32_t axes = count[0] + count[1] + count[2];
    if (axes > P.axis_cap || count[0] * count[1] * count[2] > P.region_cap) {
      // Beyond the LDS the launch sized (the host's bound makes this unreachable).
      if (tid == 0) atomicMin(&failure[1], c);
      continue;
    }
    // Per-axis voxel integrals: differences of the cumulative weight at the
    // voxel edges (o + (k - 1/2) h - p) / R (flow.py:242-256).
    for (uint32_t k = tid; k < axes; k += kDepositThreads) {
      const int a = k < offset[1] ? 0 : (k < offset[2] ? 1 : 2);
      const uint32_t index = lo[a] + (k - offset[a]);
      const double o = P.origin[a], h = P.spacing[a], p = position[a];
      const double e0 = (o + (static_cast<double>(index) - 0.5) * h - p) / P.radius;
      const double e1 = (o + (static_cast<double>(index + 1) - 0.5) * h - p) / P.radius;
      s_axis[k] = tent_cdf(e1) - tent_cdf(e0);
    }
    __syncthreads();
    const uint32_t region = count[0] * count[1] * count[2];
    const uint32_t plane = count[1] * count[2];
    double best = -1.0;
    uint32_t seed = 0xffffffffu;
    for (uint32_t v = tid; v < region; v += kDepositThreads) {
      const uint32_t i = v / plane, j = (v / count[2]) % count[1], l = v % count[2];
      double k = (s_axis[i] * s_axis[offset[1] + j]) * s_axis[offset[2] + l];
      const size_t site = (static_cast<size_t>(lo[0] + i) * P.dims[1] + (lo[1] + j)) *
                              P.dims[2] + (lo[2] + l);
      if (P.has_mask != 0 && solid[site] != 0) k = 0.0;
      s_kernel[v] = k;
      if (k > best) {
        best = k;
        seed = v;
      }
    }
    if (P.has_mask != 0 && count[2] == 1 && count[0] <= kWave && count[1] <= 64) {
      // The face-connected component of the strongest voxel (flow.py:262-276)
      // on one layer: lane i holds row i's fluid and connected voxels as bits
      // and the component grows by shifts within rows and shuffles across
      // them until no lane changes.
      wave_argmax(best, seed);
      __syncthreads();
      const uint32_t rows = count[0], cols = count[1];
      uint64_t fluid = 0, connected = 0;
      if (tid < rows) {
        for (uint32_t j = 0; j < cols; ++j) {
          if (s_kernel[tid * cols + j] > 0.0) fluid |= 1ull << j;
        }
        if (seed / cols == tid) connected = fluid & (1ull << (seed % cols));
      }
      for (;;) {
        const uint64_t below = __shfl_up(connected, 1, kWave);
        const uint64_t above = __shfl_down(connected, 1, kWave);
        uint64_t grown = connected | (connected << 1) | (connected >> 1);
        if (tid >= 1) grown |= below;
        if (tid + 1 < rows) grown |= above;
        grown &= fluid;
        const bool changed = grown != connected;
        connected = grown;
        if (!__any(changed)) break;
      }
      if (tid < rows) {
        for (uint32_t j = 0; j < cols; ++j) {
          if (((connected >> j) & 1ull) == 0) s_kernel[tid * cols + j] = 0.0;
        }
      }
      __syncthreads();
    } else if (P.has_mask != 0) {
      // The face-connected component of the strongest voxel (flow.py:262-276).
      wave_argmax(best, seed);
      for (uint32_t v = tid; v < region; v += kDepositThreads) {
        s_connected[v] = v == seed && s_kernel[v] > 0.0 ? 1 : 0;
      }
      __syncthreads();
      for (;;) {
        if (tid == 0) s_changed = 0;
        __syncthreads();
        bool grew = false;
        for (uint32_t v = tid; v < region; v += kDepositThreads) {
          if (s_connected[v] != 0 || !(s_kernel[v] > 0.0)) continue;
          const uint32_t i = v / plane, j = (v / count[2]) % count[1], l = v % count[2];
          const bool next = (i > 0 && s_connected[v - plane] != 0) ||
                            (i + 1 < count[0] && s_connected[v + plane] != 0) ||
                            (j > 0 && s_connected[v - count[2]] != 0) ||
                            (j + 1 < count[1] && s_connected[v + count[2]] != 0) ||
                            (l > 0 && s_connected[v - 1] != 0) ||
                            (l + 1 < count[2] && s_connected[v + 1] != 0);
          if (next) {
            s_connected[v] = 1;
            grew = true;
          }
        }
        if (grew) s_changed = 1;
        __syncthreads();
        const int more = s_changed;
        __syncthreads();
        if (more == 0) break;
      }
      for (uint32_t v = tid; v < region; v += kDepositThreads) {
        if (s_connected[v] == 0) s_kernel[v] = 0.0;
      }
    }
    double total = 0.0;
    for (uint32_t v = tid; v < region; v += kDepositThreads) total += s_kernel[v];
    total = wave_sum(total);
    if (!(total > 0.0)) {
      if (tid == 0) atomicMin(&failure[0], c);
      __syncthreads();
      continue;
    }
    // biomass_volume (biomass.py:6-14), then its share per voxel.
    const double r = static_cast<double>(radii[c]);
    const double amount = 3.14159265358979323846 * r * r * (static_cast<double>(lengths[c]) + 2.0 * r);
    const double velocity[2] = {momentum != nullptr ? static_cast<double>(vx[c]) : 0.0,
                                momentum != nullptr ? static_cast<double>(vy[c]) : 0.0};
    for (uint32_t v = tid; v < region; v += kDepositThreads) {
      const double k = s_kernel[v];
      if (k == 0.0) continue;
      const uint32_t i = v / plane, j = (v / count[2]) % count[1], l = v % count[2];
      const size_t column = static_cast<size_t>(lo[0] + i) * P.dims[1] + (lo[1] + j);
      const size_t site = column * P.dims[2] + (lo[2] + l);
      const double share = amount * (k / total);
      unsafeAtomicAdd(&volume[site], share);
      if (momentum != nullptr) {
        unsafeAtomicAdd(&momentum[2 * column], share * velocity[0]);
        unsafeAtomicAdd(&momentum[2 * column + 1], share * velocity[1]);
      }
    }
    __syncthreads();
  }
}

// Volume per voxel volume; with `mobility`, per column the mean fraction over
// its fluid voxels capped at `cap`, the drag C f^2 / (1 - f)^3 and
// base / (1 + base * drag) at each fluid voxel (flow.py:359-382); with
// `force`, the column's body force momentum * drag / mass, mass the column's
// deposited volume, zero where it has none (flow.py:421-423).
__global__ __launch_bounds__(256) void k_colony_finish(
    uint32_t nx, uint32_t ny, uint32_t nz, double voxel, const double* __restrict__ volume,
    double* __restrict__ density, const uint8_t* __restrict__ solid, const double* __restrict__ base,
    uint32_t per_site_base, double drag_coefficient, double cap, double* __restrict__ mobility,
    const double* __restrict__ momentum, double* __restrict__ force) {
  const size_t column = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (column >= static_cast<size_t>(nx) * ny) return;
  const size_t first = column * nz;
  double sum = 0.0;
  uint32_t fluid = 0;
  for (uint32_t z = 0; z < nz; ++z) {
    const double d = volume[first + z] / voxel;
    if (density != nullptr) density[first + z] = d;
    sum += d;
    if (solid == nullptr || solid[first + z] == 0) ++fluid;
  }
  if (mobility == nullptr) return;
  const double fraction = fmin(sum / static_cast<double>(fluid > 0 ? fluid : 1), cap);
  const double gap = 1.0 - fraction;
  const double drag = drag_coefficient * fraction * fraction / pow(gap, 3.0);
  for (uint32_t z = 0; z < nz; ++z) {
    const double b = per_site_base != 0 ? base[first + z] : base[0];
    const bool open = solid == nullptr || solid[first + z] == 0;
    mobility[first + z] = open ? b / (1.0 + b * drag) : 0.0;
  }
  if (force == nullptr) return;
  double mass = 0.0;
  for (uint32_t z = 0; z < nz; ++z) mass += volume[first + z];
  const double per_mass = mass > 0.0 ? drag / mass : 0.0;
  force[2 * column] = momentum[2 * column] * per_mass;
  force[2 * column + 1] = momentum[2 * column + 1] * per_mass;
}

__global__ void k_fill_u32_two(unsigned int* out, unsigned int value) {
  if (threadIdx.x < 2 && blockIdx.x == 0) out[threadIdx.x] = value;
}

__global__ __launch_bounds__(256) void k_fill_f64(double* __restrict__ out, size_t count) {
  const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i < count) out[i] = 0.0;
}

}  // namespace

uint32_t colony_region_cap(const msim_grid& grid, double radius) {
  const uint32_t dims[3] = {grid.nx, grid.ny, grid.nz};
  const double spacing[3] = {grid.dx, grid.dy, grid.dz};
  double region = 1.0;
  for (int a = 0; a < 3; ++a) {
    const double along = std::floor(2.0 * radius / spacing[a]) + 4.0;
    region *= std::min(along, static_cast<double>(dims[a]));
  }
  return region > 1e9 ? 0xffffffffu : static_cast<uint32_t>(region);
}

uint32_t colony_deposit(Context& ctx, const msim_grid& grid, const float* origin,
                        const msim_cells& cells, double radius, const double* base,
                        uint32_t base_count, double drag_coefficient, double cap,
                        double* density_host, double* mobility_host, const float* velocities,
                        double* force_host) {
  auto& d = ctx.colony;
  const size_t n = static_cast<size_t>(grid.nx) * grid.ny * grid.nz;
  const size_t count = cells.count;
  const auto put = [&](DeviceBuffer<float>& buffer, const float* host) {
    buffer.reserve(count);
    if (count > 0) {
      check(hipMemcpyAsync(buffer.data(), host, count * sizeof(float), hipMemcpyHostToDevice,
                           ctx.stream),
            "upload cells");
    }
  };
  put(d.position_x, cells.position_x);
  put(d.position_y, cells.position_y);
  put(d.position_z, cells.position_z);
  put(d.lengths, cells.lengths);
  put(d.radii, cells.radii);
  d.solid.reserve(grid.obstacles != nullptr ? n : 1);
  if (grid.obstacles != nullptr) {
    check(hipMemcpyAsync(d.solid.data(), grid.obstacles, n, hipMemcpyHostToDevice, ctx.stream),
          "upload mask");
  }
  const bool coupling = force_host != nullptr;
  const size_t columns = static_cast<size_t>(grid.nx) * grid.ny;
  if (coupling) {
    std::vector<float> split(2 * count);
    for (size_t c = 0; c < count; ++c) {
      split[c] = velocities[2 * c];
      split[count + c] = velocities[2 * c + 1];
    }
    put(d.velocity_x, split.data());
    put(d.velocity_y, split.data() + count);
    d.momentum.reserve(2 * columns);
    d.force.reserve(2 * columns);
    k_fill_f64<<<static_cast<uint32_t>((2 * columns + 255) / 256), 256, 0, ctx.stream>>>(
        d.momentum.data(), 2 * columns);
    check(hipGetLastError(), "launch k_fill_f64");
  }
  d.volume.reserve(n);
  d.out.reserve(n);
  d.failure.reserve(2);
  k_fill_f64<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, ctx.stream>>>(d.volume.data(), n);
  check(hipGetLastError(), "launch k_fill_f64");
  k_fill_u32_two<<<1, 64, 0, ctx.stream>>>(d.failure.data(), 0xffffffffu);
  check(hipGetLastError(), "launch k_fill_u32_two");
  DepositParams P{};
  P.dims[0] = grid.nx;
  P.dims[1] = grid.ny;
  P.dims[2] = grid.nz;
  P.spacing[0] = grid.dx;
  P.spacing[1] = grid.dy;
  P.spacing[2] = grid.dz;
  for (int a = 0; a < 3; ++a) P.origin[a] = static_cast<double>(origin[a]);
  P.radius = radius;
  P.cells = cells.count;
  P.has_mask = grid.obstacles != nullptr ? 1u : 0u;
  P.region_cap = colony_region_cap(grid, radius);
  P.axis_cap = 0;
  const uint32_t dims[3] = {grid.nx, grid.ny, grid.nz};
  const double spacing[3] = {grid.dx, grid.dy, grid.dz};
  for (int a = 0; a < 3; ++a) {
    P.axis_cap += static_cast<uint32_t>(
        std::min(std::floor(2.0 * radius / spacing[a]) + 4.0, static_cast<double>(dims[a])));
  }
  const size_t lds = (P.axis_cap + P.region_cap) * sizeof(double) + P.region_cap;
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  if (count > 0) {
    const uint32_t blocks = static_cast<uint32_t>(std::min<size_t>(count, 256 * 64));
    k_colony_deposit<<<blocks, kDepositThreads, lds, ctx.stream>>>(
        P, d.solid.data(), d.position_x.data(), d.position_y.data(), d.position_z.data(),
        d.lengths.data(), d.radii.data(), coupling ? d.velocity_x.data() : nullptr,
        coupling ? d.velocity_y.data() : nullptr, d.volume.data(),
        coupling ? d.momentum.data() : nullptr, d.failure.data());
    check(hipGetLastError(), "launch k_colony_deposit");
  }
  const double voxel = static_cast<double>(grid.dx * grid.dy * grid.dz);
  const bool want_mobility = mobility_host != nullptr;
  if (want_mobility) {
    d.base.reserve(base_count);
    check(hipMemcpyAsync(d.base.data(), base, base_count * sizeof(double), hipMemcpyHostToDevice,
                         ctx.stream),
          "upload base");
  }
  k_colony_finish<<<static_cast<uint32_t>((columns + 255) / 256), 256, 0, ctx.stream>>>(
      grid.nx, grid.ny, grid.nz, voxel, d.volume.data(), want_mobility ? nullptr : d.out.data(),
      grid.obstacles != nullptr ? d.solid.data() : nullptr, d.base.data(), base_count > 1 ? 1u : 0u,
      drag_coefficient, cap, want_mobility ? d.out.data() : nullptr,
      coupling ? d.momentum.data() : nullptr, coupling ? d.force.data() : nullptr);
  check(hipGetLastError(), "launch k_colony_finish");
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  unsigned int failed[2] = {0, 0};
  check(hipMemcpyAsync(failed, d.failure.data(), sizeof(failed), hipMemcpyDeviceToHost,
                       ctx.stream),
        "read failure");
  check(hipStreamSynchronize(ctx.stream), "colony deposit");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  d.seconds = ms * 1.0e-3;
  if (failed[1] != 0xffffffffu) {
    throw Error(MSIM_RUNTIME_ERROR, "a deposit support exceeded the bound its launch was sized for");
  }
  if (failed[0] == 0xffffffffu) {
    check(hipMemcpyAsync(want_mobility ? mobility_host : density_host, d.out.data(),
                         n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream),
          "download colony field");
    if (coupling) {
      check(hipMemcpyAsync(force_host, d.force.data(), 2 * columns * sizeof(double),
                           hipMemcpyDeviceToHost, ctx.stream),
            "download colony force");
    }
    check(hipStreamSynchronize(ctx.stream), "download colony field");
  }
  return failed[0];
}

}  // namespace msim
