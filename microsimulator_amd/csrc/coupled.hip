// Coupled cell/grid exchange on the device (advance_coupled_cpu,
// cpp/cpu/cpu_coupled.cpp:125-231), FP32 levels, on one layer or several.
//
// The device keeps the cells in spatial order: a stable sort at upload by
// the site of each cell's lower stencil corner, with `order` mapping a slot
// to the cell's index in the caller's order and `offsets` giving each lower
// corner's run of slots. Cells do not move between the steps of a run.
//
// A step is these launches:
//  * k_coupled_cells, one thread per slot: the cell's bilinear stencil (with
//    obstacles, only the corners face-connected to the strongest fluid one,
//    renormalised), the signals it samples from the step's starting levels,
//    the dilution of its species, the rate plan run by an interpreter whose
//    working values live in LDS (a column per thread), the species update,
//    the exchange rate of each signal, and the weights of its 2 x 2 corners
//    (2 x 2 x 2 on a grid of several layers);
//  * forward Euler: the explicit candidate (adr.hip), then k_coupled_commit,
//    which adds dt * source to each level and judges the levels
//    (accept_step); the implicit kinds: k_coupled_sources writes the sources
//    out for the solve's right-hand side.
// A site's source is the reference's float sum in the caller's cell order
// (cpu_coupled.cpp:198-207). The cells that reach site (x, y) are those whose
// lower corner is (x, y), (x, y - 1), (x - 1, y) or (x - 1, y - 1): four runs
// of slots (eight, with z and z - 1, on several layers), each in caller order
// (the stable sort kept it), merged by the caller's index as the sum goes.
// Every kernel returns at once when an earlier failure set the transport
// flags, so a run stops at its first refused step.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kSiteBlock = 256;
// Corners of a stencil: 2 x 2 on one layer, 2 x 2 x 2 on several (KC).
constexpr uint32_t kCorners2 = 4, kCorners3 = 8;
constexpr unsigned long long kNoFailure = ~0ull;
constexpr uint32_t kSpeciesFailure = 0xffffffffu;  // low word of a species-update failure
constexpr uint32_t kStencilFailure = 0xfffffffeu;  // a position outside the grid (checked on the host)

struct CellParams {
  uint32_t nx, ny, nz;
  float dx, dy, dz;
  float ox, oy, oz;
  float volume;  // dx * dy * dz
  float dt;
  uint32_t cells, species, signals, instructions;
  uint32_t has_mask;
  uint32_t step;
};

// interpolation_axis (signals.cpp:98-127) for a position already known to lie
// inside the grid.
struct Axis {
  uint32_t index[2];
  float weight[2];
  uint32_t count;
  bool inside;
};

__device__ inline Axis axis_of(float position, float origin, float spacing, uint32_t dimension) {
  Axis a{{0, 0}, {1.0f, 0.0f}, 1, true};
  if (dimension == 1) return a;
  const double coordinate =
      (static_cast<double>(position) - static_cast<double>(origin)) / static_cast<double>(spacing);
  if (!(coordinate >= 0.0 && coordinate <= static_cast<double>(dimension - 1))) {
    a.inside = false;
    return a;
  }
  const uint32_t lower = static_cast<uint32_t>(floor(coordinate));
  if (lower == dimension - 1) {
    a.index[0] = a.index[1] = lower;
    return a;
  }
  const float fraction = static_cast<float>(coordinate - static_cast<double>(lower));
  a.index[0] = lower;
  a.index[1] = lower + 1;
  a.weight[0] = 1.0f - fraction;
  a.weight[1] = fraction;
  a.count = 2;
  return a;
}

// effective_cell_volume and effective_cell_surface_area (species.cpp:134-142).
constexpr float kPi = static_cast<float>(3.14159265358979323846);
__device__ inline float cell_volume(float length, float radius) {
  return kPi * radius * radius * (length + 2.0f * radius);
}
__device__ inline float cell_surface(float length, float radius) {
  return 2.0f * kPi * radius * (length + 2.0f * radius);
}

// exp, log and pow in FP64, rounded to float. Kept out of line: their FP64
// code would otherwise hold registers across the whole interpreter loop.
__device__ __attribute__((noinline)) float transcendental(uint32_t operation, float a, float b) {
  if (operation == MSIM_RATE_POWER) {
    return static_cast<float>(pow(static_cast<double>(a), static_cast<double>(b)));
  }
  if (operation == MSIM_RATE_EXPONENTIAL) return static_cast<float>(exp(static_cast<double>(a)));
  return static_cast<float>(log(static_cast<double>(a)));
}

__device__ inline void fail_cell(unsigned long long* failure, uint32_t* flags, uint32_t cell,
                                 uint32_t what, uint32_t step) {
  atomicMin(failure, (static_cast<unsigned long long>(cell) << 32) | what);
  atomicOr(&flags[0], kFlagCell);
  atomicMin(&flags[1], step);
}

// One thread per cell (cpu_coupled.cpp:142-197 for one cell). The rate plan's
// values, the samples and the diluted species sit in LDS, column threadIdx.x
// of a [slots][blockDim] array.
// The exponential, logarithm and power are evaluated in FP64 and rounded to
// float.
template <uint32_t KC>
__global__ void k_coupled_cells(CellParams P, const uint8_t* __restrict__ solid,
                                const float* __restrict__ levels, const float* __restrict__ px,
                                const float* __restrict__ py, const float* __restrict__ pz,
                                const float* __restrict__ lengths, const float* __restrict__ radii,
                                const float* __restrict__ growth, const int32_t* __restrict__ types,
                                const float* __restrict__ previous,
                                const uint32_t* __restrict__ order,
                                const msim_rate_instruction* __restrict__ plan,
                                const uint32_t* __restrict__ outputs,
                                const float* __restrict__ species_in, float* species_out,
                                float* __restrict__ exchange, float* __restrict__ weights,
                                unsigned long long* failure, uint32_t* flags) {
  extern __shared__ float s_work[];
  if (flags[0] != 0) return;  // an earlier step failed: the run stops there
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;  // the cell's slot
  if (c >= P.cells) return;
  const uint32_t caller = order[c];  // its index in the caller's order
  const size_t n = static_cast<size_t>(P.nx) * P.ny * P.nz;
  // The stencil (signal_grid_stencil, signals.cpp:357-420): corners x-major,
  // then y, then z.
  const Axis ax = axis_of(px[c], P.ox, P.dx, P.nx);
  const Axis ay = axis_of(py[c], P.oy, P.dy, P.ny);
  const Axis az = axis_of(pz[c], P.oz, P.dz, P.nz);
  if (!ax.inside || !ay.inside || !az.inside) {
    fail_cell(failure, flags, caller, kStencilFailure, P.step);
    return;
  }
  uint32_t site[KC], cx[KC], cy[KC], cz[KC];
  float w[KC];
  const uint32_t count = ax.count * ay.count * az.count;
#pragma unroll
  for (uint32_t e = 0; e < KC; ++e) {
    // Corner e is (a, b, d) for the y and z axes' yc and zc sites.
    const uint32_t a = e < count ? e / (ay.count * az.count) : 0;
    const uint32_t b = e < count ? (e / az.count) % ay.count : 0;
    const uint32_t d = e < count ? e % az.count : 0;
    const uint32_t xi = a != 0 ? ax.index[1] : ax.index[0];
    const uint32_t yi = b != 0 ? ay.index[1] : ay.index[0];
    const uint32_t zi = d != 0 ? az.index[1] : az.index[0];
    cx[e] = xi;
    cy[e] = yi;
    cz[e] = zi;
    site[e] = (xi * P.ny + yi) * P.nz + zi;
    w[e] = e < count ? ((a != 0 ? ax.weight[1] : ax.weight[0]) *
                        (b != 0 ? ay.weight[1] : ay.weight[0])) *
                           (d != 0 ? az.weight[1] : az.weight[0])
                     : 0.0f;
  }
  if (P.has_mask != 0) {
    uint32_t seed = count;
    float best = 0.0f;
#pragma unroll
    for (uint32_t e = 0; e < KC; ++e) {
      if (e < count) {
        if (solid[site[e]] != 0) w[e] = 0.0f;
        if (w[e] > 0.0f && (seed == count || w[e] > best)) {
          seed = e;
          best = w[e];
        }
      }
    }
    if (seed == count) {
      fail_cell(failure, flags, caller, kStencilFailure, P.step);
      return;
    }
    bool connected[KC];
#pragma unroll
    for (uint32_t e = 0; e < KC; ++e) connected[e] = e == seed;
    for (bool grew = true; grew;) {
      grew = false;
#pragma unroll
      for (uint32_t i = 0; i < KC; ++i) {
        if (i >= count || w[i] <= 0.0f || connected[i]) continue;
#pragma unroll
        for (uint32_t j = 0; j < KC; ++j) {
          if (j >= count || !connected[j] || connected[i]) continue;
          const uint32_t dist = (cx[i] > cx[j] ? cx[i] - cx[j] : cx[j] - cx[i]) +
                                (cy[i] > cy[j] ? cy[i] - cy[j] : cy[j] - cy[i]) +
                                (cz[i] > cz[j] ? cz[i] - cz[j] : cz[j] - cz[i]);
          if (dist == 1) {
            connected[i] = true;
            grew = true;
          }
        }
      }
    }
    float total = 0.0f;
#pragma unroll
    for (uint32_t e = 0; e < KC; ++e) {
      if (e < count) {
        if (!connected[e]) w[e] = 0.0f;
        total += w[e];
      }
    }
#pragma unroll
    for (uint32_t e = 0; e < KC; ++e) {
      if (e < count) w[e] /= total;
    }
  }
  // The thread's LDS column: the plan's values, then the sampled signals,
  // then the diluted species.
  float* work = s_work + threadIdx.x;
  const uint32_t T = blockDim.x;
  const uint32_t S = P.signals, K = P.species;
  float* sampled = work + static_cast<size_t>(P.instructions) * T;
  float* held = sampled + static_cast<size_t>(S) * T;
  // The sampled signals (cpu_coupled.cpp:150-160).
  for (uint32_t s = 0; s < S; ++s) {
    float v = 0.0f;
#pragma unroll
    for (uint32_t e = 0; e < KC; ++e) {
      if (e < count) v += w[e] * levels[s * n + site[e]];
    }
    sampled[s * T] = v;
  }
  // Dilution by the growth since the previous step (162-170).
  const float length = lengths[c], radius = radii[c];
  const float previous_volume = cell_volume(previous[c], radius);
  const float current_volume = cell_volume(length, radius);
  const float dilution = previous_volume / current_volume;
  const size_t base = static_cast<size_t>(c) * K;
  for (uint32_t k = 0; k < K; ++k) held[k * T] = species_in[base + k] * dilution;
  const float volume_change_rate =
      P.dt == 0.0f ? 0.0f : (current_volume - previous_volume) / P.dt;
  // The rate plan (evaluate_instruction, cpu_coupled.cpp:20-85).
  for (uint32_t index = 0; index < P.instructions; ++index) {
    const msim_rate_instruction in = plan[index];
    const float a = in.operation >= MSIM_RATE_ADD && in.operation <= MSIM_RATE_SELECT
                        ? work[in.first * T]
                        : 0.0f;
    const float b = in.operation >= MSIM_RATE_ADD && in.operation <= MSIM_RATE_SELECT &&
                            in.operation != MSIM_RATE_NEGATE &&
                            in.operation != MSIM_RATE_EXPONENTIAL &&
                            in.operation != MSIM_RATE_LOGARITHM
                        ? work[in.second * T]
                        : 0.0f;
    float v;
    switch (in.operation) {
      case MSIM_RATE_CONSTANT: v = in.value; break;
      case MSIM_RATE_SPECIES: v = held[in.first * T]; break;
      case MSIM_RATE_SIGNAL: v = sampled[in.first * T]; break;
      case MSIM_RATE_POSITION_X: v = px[c]; break;
      case MSIM_RATE_POSITION_Y: v = py[c]; break;
      case MSIM_RATE_POSITION_Z: v = pz[c]; break;
      case MSIM_RATE_LENGTH: v = length; break;
      case MSIM_RATE_RADIUS: v = radius; break;
      case MSIM_RATE_GROWTH_RATE: v = growth[c]; break;
      case MSIM_RATE_CELL_TYPE: v = static_cast<float>(types[c]); break;
      case MSIM_RATE_VOLUME_CHANGE_RATE: v = volume_change_rate; break;
      case MSIM_RATE_VOLUME: v = current_volume; break;
      case MSIM_RATE_SURFACE_AREA: v = cell_surface(length, radius); break;
      case MSIM_RATE_ADD: v = a + b; break;
      case MSIM_RATE_SUBTRACT: v = a - b; break;
      case MSIM_RATE_MULTIPLY: v = a * b; break;
      case MSIM_RATE_DIVIDE: v = a / b; break;
      case MSIM_RATE_POWER: v = transcendental(in.operation, a, b); break;
      case MSIM_RATE_MINIMUM: v = b < a ? b : a; break;
      case MSIM_RATE_MAXIMUM: v = a < b ? b : a; break;
      case MSIM_RATE_NEGATE: v = -a; break;
      case MSIM_RATE_EXPONENTIAL: v = transcendental(in.operation, a, b); break;
      case MSIM_RATE_LOGARITHM: v = transcendental(in.operation, a, b); break;
      case MSIM_RATE_LESS: v = a < b ? 1.0f : 0.0f; break;
      case MSIM_RATE_LESS_EQUAL: v = a <= b ? 1.0f : 0.0f; break;
      case MSIM_RATE_GREATER: v = a > b ? 1.0f : 0.0f; break;
      case MSIM_RATE_GREATER_EQUAL: v = a >= b ? 1.0f : 0.0f; break;
      case MSIM_RATE_EQUAL: v = a == b ? 1.0f : 0.0f; break;
      default: v = a != 0.0f ? b : work[in.third * T]; break;  // MSIM_RATE_SELECT
    }
    work[index * T] = v;
    if (!isfinite(v)) {
      fail_cell(failure, flags, caller, index, P.step);
      return;
    }
  }
  // The species update (cpu_coupled.cpp:180-187) and the exchange (188-191).
  for (uint32_t k = 0; k < K; ++k) {
    const float v = held[k * T] + P.dt * work[outputs[k] * T];
    if (!isfinite(v)) {
      fail_cell(failure, flags, caller, kSpeciesFailure, P.step);
      return;
    }
    species_out[base + k] = v;
  }
  for (uint32_t s = 0; s < S; ++s) {
    exchange[static_cast<size_t>(c) * S + s] = work[outputs[K + s] * T] / P.volume;
  }
  // The corners' weights at (a, b) = (0, 0), (0, 1), (1, 0), (1, 1) (with d
  // the last bit on several layers); -1 where the stencil has no such corner
  // (an axis of one site, or the upper edge).
#pragma unroll
  for (uint32_t f = 0; f < KC; ++f) {
    const uint32_t fa = KC == kCorners3 ? f >> 2 : f >> 1;
    const uint32_t fb = KC == kCorners3 ? (f >> 1) & 1 : f & 1;
    const uint32_t fd = KC == kCorners3 ? f & 1 : 0;
    const uint32_t e = (fa * ay.count + fb) * az.count + fd;
    float pick = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < KC; ++k) {
      if (k == e) pick = w[k];
    }
    weights[static_cast<size_t>(c) * KC + f] =
        fa < ax.count && fb < ay.count && fd < az.count ? pick : -1.0f;
  }
}

// The sources of site (x, y, z): per signal, from 0, the sum of weight *
// exchange over the cells whose stencil reaches the site, in the caller's
// cell order (cpu_coupled.cpp:198-207). Corner (da, db, dd) of the cells with
// lower corner (x - da, y - db, z - dd) is this site; the runs are merged by
// the caller's index. A corner the stencil lacks (weight -1) adds nothing; a
// present corner of weight 0 adds its product, as the reference does.
template <int S, uint32_t KC>
__device__ inline void site_sources(size_t site, uint32_t ny, uint32_t nz,
                                    const uint32_t* __restrict__ offsets,
                                    const uint32_t* __restrict__ order,
                                    const float* __restrict__ weights,
                                    const float* __restrict__ exchange, float (&src)[S]) {
  const uint32_t z = static_cast<uint32_t>(site % nz);
  const size_t column = site / nz;
  const uint32_t x = static_cast<uint32_t>(column / ny), y = static_cast<uint32_t>(column % ny);
#pragma unroll
  for (int s = 0; s < S; ++s) src[s] = 0.0f;
  uint32_t at[KC], end[KC], head[KC];
#pragma unroll
  for (uint32_t g = 0; g < KC; ++g) {
    const uint32_t da = KC == kCorners3 ? g >> 2 : g >> 1;
    const uint32_t db = KC == kCorners3 ? (g >> 1) & 1 : g & 1;
    const uint32_t dd = KC == kCorners3 ? g & 1 : 0;
    at[g] = end[g] = 0;
    if (x >= da && y >= db && z >= dd) {
      const size_t k = (static_cast<size_t>(x - da) * ny + (y - db)) * nz + (z - dd);
      at[g] = offsets[k];
      end[g] = offsets[k + 1];
    }
    head[g] = at[g] < end[g] ? order[at[g]] : 0xffffffffu;
  }
  for (;;) {
    int pick = -1;
    uint32_t least = 0xffffffffu;
#pragma unroll
    for (uint32_t g = 0; g < KC; ++g) {
      if (at[g] < end[g] && head[g] < least) {
        least = head[g];
        pick = static_cast<int>(g);
      }
    }
    if (pick < 0) break;
    uint32_t slot = 0;
#pragma unroll
    for (uint32_t g = 0; g < KC; ++g) {
      if (static_cast<int>(g) == pick) {
        slot = at[g]++;
        head[g] = at[g] < end[g] ? order[at[g]] : 0xffffffffu;
      }
    }
    // That lower corner reaches this site with its corner g, the cell's entry g.
    const float w = weights[static_cast<size_t>(slot) * KC + pick];
    if (w >= 0.0f) {
      const float* ex = exchange + static_cast<size_t>(slot) * S;
#pragma unroll
      for (int s = 0; s < S; ++s) src[s] += w * ex[s];
    }
  }
}

// Forward Euler's commit: level + dt * source for every level of the site,
// obstacle sites included (cpu_coupled.cpp:216-222), then the levels judged
// as accept_step does: under upwind each must be finite and non-negative;
// under van Leer the block's largest level goes to `largest` for the settle.
// Writes nothing once a failure is flagged.
template <int S, bool kLimited, uint32_t KC>
__global__ __launch_bounds__(kSiteBlock) void k_coupled_commit(
    size_t n, uint32_t ny, uint32_t nz, const float* in, float* out, float dt,
    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ order,
    const float* __restrict__ weights, const float* __restrict__ exchange,
    double* __restrict__ largest, uint32_t* flags, uint32_t step) {
  __shared__ double s_red[kSiteBlock / kWave];
  const bool failed = flags[0] != 0;
  const size_t site = static_cast<size_t>(blockIdx.x) * kSiteBlock + threadIdx.x;
  double top = 0.0;
  bool bad = false;
  if (site < n && !failed) {
    float src[S];
    site_sources<S, KC>(site, ny, nz, offsets, order, weights, exchange, src);
    {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const float v = in[s * n + site] + dt * src[s];
        out[s * n + site] = v;
        if constexpr (kLimited) {
          top = fmax(top, static_cast<double>(v));
        } else {
          bad = bad || !isfinite(v) || v < 0.0f;
        }
      }
    }
  }
  if constexpr (kLimited) {
    top = block_max<kSiteBlock>(top, s_red);
    if (threadIdx.x == 0) largest[blockIdx.x] = top;
  } else if (bad) {
    atomicOr(&flags[0], kFlagLevels);
    atomicMin(&flags[1], step);
  }
}

// The implicit kinds' sources, S*N (a non-finite one refuses the step,
// signals.cpp:1295-1310).
template <int S, uint32_t KC>
__global__ __launch_bounds__(kSiteBlock) void k_coupled_sources(
    size_t n, uint32_t ny, uint32_t nz, float* __restrict__ sources,
    const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ order, const float* __restrict__ weights,
    const float* __restrict__ exchange, uint32_t* flags, uint32_t step) {
  const size_t site = static_cast<size_t>(blockIdx.x) * kSiteBlock + threadIdx.x;
  if (site >= n) return;
  float src[S];
  site_sources<S, KC>(site, ny, nz, offsets, order, weights, exchange, src);
  bool bad = false;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    sources[s * n + site] = src[s];
    bad = bad || !isfinite(src[s]);
  }
  if (bad) {
    atomicOr(&flags[0], kFlagSources);
    atomicMin(&flags[1], step);
  }
}

// The site of each cell's stencil corner below it (clamped into the grid):
// the key of the spatial order, and the cell's index.
__global__ __launch_bounds__(kSiteBlock) void k_cell_keys(uint32_t nx, uint32_t ny, uint32_t nz,
                                                          float dx, float dy, float dz, float ox,
                                                          float oy, float oz,
                                                          const float* __restrict__ px,
                                                          const float* __restrict__ py,
                                                          const float* __restrict__ pz,
                                                          uint32_t cells, uint32_t* __restrict__ keys,
                                                          uint32_t* __restrict__ index) {
  const uint32_t c = blockIdx.x * kSiteBlock + threadIdx.x;
  if (c >= cells) return;
  const auto lower = [](float p, float o, float h, uint32_t dimension) {
    if (dimension == 1) return 0u;
    const double f = floor((static_cast<double>(p) - o) / h);
    return static_cast<uint32_t>(fmin(fmax(f, 0.0), static_cast<double>(dimension - 1)));
  };
  const uint32_t x = lower(px[c], ox, dx, nx), y = lower(py[c], oy, dy, ny);
  const uint32_t z = lower(pz[c], oz, dz, nz);
  keys[c] = (x * ny + y) * nz + z;
  index[c] = c;
}

// out[slot * width + k] = in[order[slot] * width + k] (to the spatial order),
// or the reverse with `back` (to the caller's order).
__global__ __launch_bounds__(kSiteBlock) void k_permute(const uint32_t* __restrict__ order,
                                                        size_t count, uint32_t width, bool back,
                                                        const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kSiteBlock + threadIdx.x;
  if (i >= count * width) return;
  const size_t slot = i / width, k = i % width;
  const size_t caller = static_cast<size_t>(order[slot]) * width + k;
  if (back) {
    out[caller] = in[i];
  } else {
    out[i] = in[caller];
  }
}

// Cells per lower corner (integer counts: their order does not matter).
__global__ __launch_bounds__(kSiteBlock) void k_count_sites(const uint32_t* __restrict__ keys,
                                                            uint32_t cells,
                                                            uint32_t* __restrict__ counts) {
  const uint32_t p = blockIdx.x * kSiteBlock + threadIdx.x;
  if (p < cells) atomicAdd(&counts[keys[p]], 1u);
}

__global__ __launch_bounds__(kSiteBlock) void k_fill_u32(uint32_t* __restrict__ out, size_t count,
                                                         uint32_t value) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kSiteBlock + threadIdx.x;
  if (i < count) out[i] = value;
}

__global__ void k_fill_u64(unsigned long long* out, unsigned long long value) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = value;
}

uint32_t blocks_of(size_t count) {
  return static_cast<uint32_t>((count + kSiteBlock - 1) / kSiteBlock);
}

// Threads per block of the cell kernel: as many as fit each thread's LDS
// column (the plan's values, the samples and the species) in 64 KiB.
uint32_t cell_threads(uint32_t slots) {
  if (slots <= 64) return 256;
  if (slots <= 128) return 128;
  return 64;
}

template <int S, uint32_t KC>
void launch_commit_kc(Context& ctx, bool limited, const float* in, float* out, float dt,
                      uint32_t step) {
  auto& c = ctx.coupled;
  const size_t n = ctx.adr.sites();
  const uint32_t ny = ctx.adr.params.ny, nz = ctx.adr.params.nz;
  if (limited) {
    k_coupled_commit<S, true, KC><<<blocks_of(n), kSiteBlock, 0, ctx.stream>>>(
        n, ny, nz, in, out, dt, c.offsets.data(), c.order.data(), c.weights.data(),
        c.exchange.data(), c.largest.data(), ctx.adr.flags.data(), step);
  } else {
    k_coupled_commit<S, false, KC><<<blocks_of(n), kSiteBlock, 0, ctx.stream>>>(
        n, ny, nz, in, out, dt, c.offsets.data(), c.order.data(), c.weights.data(),
        c.exchange.data(), c.largest.data(), ctx.adr.flags.data(), step);
  }
}

template <int S>
void launch_commit(Context& ctx, bool limited, const float* in, float* out, float dt,
                   uint32_t step) {
  if (ctx.adr.params.nz > 1) {
    launch_commit_kc<S, kCorners3>(ctx, limited, in, out, dt, step);
  } else {
    launch_commit_kc<S, kCorners2>(ctx, limited, in, out, dt, step);
  }
}

template <int S>
void launch_sources(Context& ctx, uint32_t step) {
  auto& c = ctx.coupled;
  const size_t n = ctx.adr.sites();
  const uint32_t ny = ctx.adr.params.ny, nz = ctx.adr.params.nz;
  if (nz > 1) {
    k_coupled_sources<S, kCorners3><<<blocks_of(n), kSiteBlock, 0, ctx.stream>>>(
        n, ny, nz, c.sources.data(), c.offsets.data(), c.order.data(), c.weights.data(),
        c.exchange.data(), ctx.adr.flags.data(), step);
  } else {
    k_coupled_sources<S, kCorners2><<<blocks_of(n), kSiteBlock, 0, ctx.stream>>>(
        n, ny, nz, c.sources.data(), c.offsets.data(), c.order.data(), c.weights.data(),
        c.exchange.data(), ctx.adr.flags.data(), step);
  }
}

void commit(Context& ctx, bool limited, const float* in, float* out, float dt, uint32_t step) {
  switch (ctx.adr.params.signals) {
    case 1: launch_commit<1>(ctx, limited, in, out, dt, step); break;
    case 2: launch_commit<2>(ctx, limited, in, out, dt, step); break;
    case 3: launch_commit<3>(ctx, limited, in, out, dt, step); break;
    case 4: launch_commit<4>(ctx, limited, in, out, dt, step); break;
    case 5: launch_commit<5>(ctx, limited, in, out, dt, step); break;
    case 6: launch_commit<6>(ctx, limited, in, out, dt, step); break;
    case 7: launch_commit<7>(ctx, limited, in, out, dt, step); break;
    default: launch_commit<8>(ctx, limited, in, out, dt, step); break;
  }
  check(hipGetLastError(), "launch k_coupled_commit");
}

void sources(Context& ctx, uint32_t step) {
  switch (ctx.adr.params.signals) {
    case 1: launch_sources<1>(ctx, step); break;
    case 2: launch_sources<2>(ctx, step); break;
    case 3: launch_sources<3>(ctx, step); break;
    case 4: launch_sources<4>(ctx, step); break;
    case 5: launch_sources<5>(ctx, step); break;
    case 6: launch_sources<6>(ctx, step); break;
    case 7: launch_sources<7>(ctx, step); break;
    default: launch_sources<8>(ctx, step); break;
  }
  check(hipGetLastError(), "launch k_coupled_sources");
}

// The cell kernel of one step.
void cell_phase(Context& ctx, const float* levels, const float* previous, float dt,
                uint32_t step) {
  auto& c = ctx.coupled;
  auto& a = ctx.adr;
  const size_t n = a.sites();
  CellParams P{};
  P.nx = a.params.nx;
  P.ny = a.params.ny;
  P.nz = a.params.nz;
  P.dx = a.params.dx;
  P.dy = a.params.dy;
  P.dz = a.params.dz;
  P.ox = c.origin[0];
  P.oy = c.origin[1];
  P.oz = c.origin[2];
  P.volume = a.params.dx * a.params.dy * a.params.dz;
  P.dt = dt;
  P.cells = c.cells;
  P.species = c.species;
  P.signals = c.signals;
  P.instructions = c.instructions;
  P.has_mask = c.has_mask ? 1u : 0u;
  P.step = step;
  const float* species_in = c.species_current == 0 ? c.species_a.data() : c.species_b.data();
  float* species_out = c.species_current == 0 ? c.species_b.data() : c.species_a.data();
  if (c.cells > 0) {
    const uint32_t slots = c.instructions + c.signals + c.species;
    const uint32_t threads = cell_threads(slots);
    const size_t lds = static_cast<size_t>(slots) * threads * sizeof(float);
    const auto kernel = P.nz > 1 ? k_coupled_cells<kCorners3> : k_coupled_cells<kCorners2>;
    kernel<<<(c.cells + threads - 1) / threads, threads, lds, ctx.stream>>>(
        P, a.solid.data(), levels, c.position_x.data(), c.position_y.data(),
        c.position_z.data(), c.lengths.data(), c.radii.data(), c.growth.data(), c.types.data(),
        previous, c.order.data(), c.plan.data(), c.outputs.data(), species_in, species_out,
        c.exchange.data(), c.weights.data(), c.failure.data(), a.flags.data());
    check(hipGetLastError(), "launch k_coupled_cells");
  }
}

// The failure the flags and the failure word name, as the reference reports it.
[[noreturn]] void throw_failure(uint32_t flag_bits, unsigned long long failure) {
  if ((flag_bits & kFlagCell) != 0) {
    const uint32_t what = static_cast<uint32_t>(failure & 0xffffffffull);
    if (what == kSpeciesFailure) {
      throw Error(MSIM_DOMAIN_ERROR, "coupled species update produced a non-finite level");
    }
    if (what == kStencilFailure) {
      throw Error(MSIM_INVALID_ARGUMENT, "signal sample position is outside the grid");
    }
    throw Error(MSIM_DOMAIN_ERROR,
                "coupled rate instruction " + std::to_string(what) + " produced a non-finite value");
  }
  if ((flag_bits & kFlagSources) != 0) {
    throw Error(MSIM_INVALID_ARGUMENT, "signal source rates must be finite");
  }
  if ((flag_bits & kFlagNegative) != 0) {
    throw Error(MSIM_DOMAIN_ERROR,
                "signal grid update produced a non-finite or negative concentration");
  }
  throw Error(MSIM_INVALID_ARGUMENT, "signal grid levels must be finite and non-negative");
}

}  // namespace

void coupled_upload(Context& ctx, const float* origin, const msim_cells& cells,
                    const float* species, const msim_rate_plan& plan, bool has_mask) {
  auto& c = ctx.coupled;
  const auto& a = ctx.adr;
  c.uploaded = false;
  c.cells = cells.count;
  c.species = plan.species_count;
  c.signals = plan.signal_count;
  c.instructions = plan.instruction_count;
  c.has_mask = has_mask;
  std::memcpy(c.origin, origin, sizeof(c.origin));
  const size_t count = cells.count, n = a.sites();
  const uint32_t K = plan.species_count;
  // keys holds the sort's keys, then the counts per site (N + 1).
  c.keys.reserve(std::max<size_t>(count, n + 1));
  for (auto* b : {&c.entries, &c.sorted_keys}) b->reserve(std::max<size_t>(count, 1));
  c.order.reserve(count);
  c.offsets.reserve(n + 1);
  c.staging.reserve(std::max<size_t>(3 * count, count * K));
  size_t bytes = 0;
  check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, c.keys.data(), c.sorted_keys.data(),
                                           c.entries.data(), c.order.data(),
                                           static_cast<int>(std::max<size_t>(count, 1)), 0, 32,
                                           ctx.stream),
        "size coupled sort");
  size_t scan = 0;
  check(hipcub::DeviceScan::ExclusiveSum(nullptr, scan, c.keys.data(), c.offsets.data(),
                                         static_cast<int>(n + 1), ctx.stream),
        "size offsets scan");
  c.sort_scratch.reserve(std::max(bytes, scan));
  auto* staging = reinterpret_cast<uint32_t*>(c.staging.data());
  const auto stage = [&](const void* host, size_t items, size_t at) {
    check(hipMemcpyAsync(staging + at, host, items * sizeof(uint32_t), hipMemcpyHostToDevice,
                         ctx.stream),
          "upload cells");
  };
  // The spatial order: a stable sort of the cells by their lower corner's site.
  if (count > 0) {
    stage(cells.position_x, count, 0);
    stage(cells.position_y, count, count);
    stage(cells.position_z, count, 2 * count);
    k_cell_keys<<<blocks_of(count), kSiteBlock, 0, ctx.stream>>>(
        a.params.nx, a.params.ny, a.params.nz, a.params.dx, a.params.dy, a.params.dz, origin[0],
        origin[1], origin[2], c.staging.data(), c.staging.data() + count,
        c.staging.data() + 2 * count, static_cast<uint32_t>(count), c.keys.data(),
        c.entries.data());
    check(hipGetLastError(), "launch k_cell_keys");
    const int end_bit = std::max(1, static_cast<int>(std::bit_width(static_cast<uint64_t>(n))));
    size_t need = 0;
    check(hipcub::DeviceRadixSort::SortPairs(nullptr, need, c.keys.data(), c.sorted_keys.data(),
                                             c.entries.data(), c.order.data(),
                                             static_cast<int>(count), 0, end_bit, ctx.stream),
          "size cell order");
    if (need > c.sort_scratch.size()) c.sort_scratch.reserve(need);
    need = c.sort_scratch.size();
    check(hipcub::DeviceRadixSort::SortPairs(c.sort_scratch.data(), need, c.keys.data(),
                                             c.sorted_keys.data(), c.entries.data(),
                                             c.order.data(), static_cast<int>(count), 0, end_bit,
                                             ctx.stream),
          "sort cell order");
  }
  // Each lower corner's run of slots: counts per site, then their exclusive sum.
  k_fill_u32<<<blocks_of(n + 1), kSiteBlock, 0, ctx.stream>>>(c.keys.data(), n + 1, 0u);
  check(hipGetLastError(), "launch k_fill_u32");
  if (count > 0) {
    k_count_sites<<<blocks_of(count), kSiteBlock, 0, ctx.stream>>>(
        c.sorted_keys.data(), static_cast<uint32_t>(count), c.keys.data());
    check(hipGetLastError(), "launch k_count_sites");
  }
  size_t scan_bytes = c.sort_scratch.size();
  check(hipcub::DeviceScan::ExclusiveSum(c.sort_scratch.data(), scan_bytes, c.keys.data(),
                                         c.offsets.data(), static_cast<int>(n + 1), ctx.stream),
        "scan offsets");
  // Every per-cell array to the spatial order.
  const auto place = [&](auto& buffer, const void* host, uint32_t width) {
    buffer.reserve(count * width);
    if (count == 0 || width == 0) return;
    const size_t items = count * width;
    const bool positions =
        host == cells.position_x || host == cells.position_y || host == cells.position_z;
    const size_t at = host == cells.position_y ? count : (host == cells.position_z ? 2 * count : 0);
    if (!positions) stage(host, items, 0);
    k_permute<<<blocks_of(items), kSiteBlock, 0, ctx.stream>>>(
        c.order.data(), count, width, false, staging + at,
        reinterpret_cast<uint32_t*>(buffer.data()));
    check(hipGetLastError(), "launch k_permute");
  };
  place(c.position_x, cells.position_x, 1);
  place(c.position_y, cells.position_y, 1);
  place(c.position_z, cells.position_z, 1);
  place(c.lengths, cells.lengths, 1);
  place(c.radii, cells.radii, 1);
  place(c.growth, cells.growth_rates, 1);
  place(c.types, cells.cell_types, 1);
  place(c.species_a, species, K);
  c.previous.reserve(count);
  c.species_b.reserve(count * K);
  c.species_current = 0;
  c.plan.reserve(plan.instruction_count);
  check(hipMemcpyAsync(c.plan.data(), plan.instructions,
                       plan.instruction_count * sizeof(msim_rate_instruction),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload plan");
  std::vector<uint32_t> outputs(plan.species_outputs, plan.species_outputs + plan.species_count);
  outputs.insert(outputs.end(), plan.signal_outputs, plan.signal_outputs + plan.signal_count);
  c.outputs.reserve(outputs.size());
  check(hipMemcpyAsync(c.outputs.data(), outputs.data(), outputs.size() * sizeof(uint32_t),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload plan");
  c.weights.reserve(count * (a.params.nz > 1 ? kCorners3 : kCorners2));
  c.exchange.reserve(count * plan.signal_count);
  c.sources.reserve(static_cast<size_t>(plan.signal_count) * n);
  c.largest.reserve(blocks_of(n));
  c.failure.reserve(1);
  check(hipStreamSynchronize(ctx.stream), "upload coupled state");
  c.uploaded = true;
}

void coupled_run(Context& ctx, const float* previous_lengths, float dt, uint32_t steps) {
  auto& c = ctx.coupled;
  auto& a = ctx.adr;
  const size_t n = a.sites();
  const size_t count = static_cast<size_t>(a.params.signals) * n;
  const bool limited = a.params.limited != 0;
  const bool implicit = a.integration != MSIM_FORWARD_EULER;
  const float* xf = a.flow_field ? ctx.flow.faces.data() : a.xf.data();
  const float* yf = a.flow_field ? ctx.flow.faces.data() + ctx.flow.x_faces() : a.yf.data();
  // The first step's previous lengths; later steps' are the cells' own.
  if (c.cells > 0) {
    if (previous_lengths != nullptr) {
      check(hipMemcpyAsync(c.staging.data(), previous_lengths, c.cells * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload previous lengths");
      k_permute<<<blocks_of(c.cells), kSiteBlock, 0, ctx.stream>>>(
          c.order.data(), c.cells, 1, false, reinterpret_cast<const uint32_t*>(c.staging.data()),
          reinterpret_cast<uint32_t*>(c.previous.data()));
      check(hipGetLastError(), "launch k_permute");
    } else {
      check(hipMemcpyAsync(c.previous.data(), c.lengths.data(), c.cells * sizeof(float),
                           hipMemcpyDeviceToDevice, ctx.stream),
            "copy lengths");
    }
  }
  const uint32_t init_flags[2] = {0u, 0xffffffffu};
  check(hipMemcpyAsync(a.flags.data(), init_flags, sizeof(init_flags), hipMemcpyHostToDevice,
                       ctx.stream),
        "clear flags");
  k_fill_u64<<<1, 64, 0, ctx.stream>>>(c.failure.data(), kNoFailure);
  check(hipGetLastError(), "launch k_fill_u64");
  a.last_report = msim_signal_report{1, 0, 0.0f};
  const int species_start = c.species_current, levels_start = a.current;
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  uint32_t flag_bits[2] = {0, 0};
  unsigned long long failure = kNoFailure;
  const auto read_flags = [&]() {
    check(hipMemcpyAsync(flag_bits, a.flags.data(), sizeof(flag_bits), hipMemcpyDeviceToHost,
                         ctx.stream),
          "read flags");
    check(hipMemcpyAsync(&failure, c.failure.data(), sizeof(failure), hipMemcpyDeviceToHost,
                         ctx.stream),
          "read failure");
    check(hipStreamSynchronize(ctx.stream), "coupled step");
  };
  for (uint32_t step = 0; step < steps; ++step) {
    float* level_bufs[2] = {reinterpret_cast<float*>(a.levels_a.data()),
                            reinterpret_cast<float*>(a.levels_b.data())};
    float* old = level_bufs[a.current];
    float* next = level_bufs[a.current ^ 1];
    if (step == 1 && c.cells > 0) {
      check(hipMemcpyAsync(c.previous.data(), c.lengths.data(), c.cells * sizeof(float),
                           hipMemcpyDeviceToDevice, ctx.stream),
            "copy lengths");
    }
    cell_phase(ctx, old, c.previous.data(), dt, step);
    if (!implicit) {
      if (dt > 0.0f) {
        if (a.params.nz > 1) {
          adr3_candidate(ctx, dt, old, next, step, xf, yf, a.zf.data());
        } else {
          fe_candidate_f32(ctx, dt, old, next, step, xf, yf);
        }
        commit(ctx, limited, next, next, dt, step);
      } else {
        commit(ctx, limited, old, next, dt, step);
      }
      if (limited) launch_settle_f32(ctx, next, count, c.largest.data(), blocks_of(n), step);
      a.current ^= 1;
      c.species_current ^= 1;
      continue;
    }
    // The implicit kinds: the sources, judged on the host before the solve.
    sources(ctx, step);
    read_flags();
    if (flag_bits[0] != 0) throw_failure(flag_bits[0], failure);
    if (dt > 0.0f) {
      implicit_run(ctx, dt, 1, xf, yf, c.sources.data(), "Implicit coupled signal solve");
    }
    c.species_current ^= 1;
  }
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  read_flags();
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  c.seconds = ms * 1.0e-3;
  if (flag_bits[0] != 0) {
    // Every step from the failed one on did nothing: the state is the one
    // that step started from.
    const uint32_t failed = flag_bits[1];
    a.current = levels_start ^ static_cast<int>(failed & 1u);
    c.species_current = species_start ^ static_cast<int>(failed & 1u);
    throw_failure(flag_bits[0], failure);
  }
}

void coupled_download(Context& ctx, float* levels, float* species) {
  auto& c = ctx.coupled;
  if (levels != nullptr) adr_download(ctx, levels);
  if (species != nullptr && c.cells > 0 && c.species > 0) {
    const float* src = c.species_current == 0 ? c.species_a.data() : c.species_b.data();
    k_permute<<<blocks_of(static_cast<size_t>(c.cells) * c.species), kSiteBlock, 0, ctx.stream>>>(
        c.order.data(), c.cells, c.species, true, reinterpret_cast<const uint32_t*>(src),
        reinterpret_cast<uint32_t*>(c.staging.data()));
    check(hipGetLastError(), "launch k_permute");
    check(hipMemcpyAsync(species, c.staging.data(),
                         static_cast<size_t>(c.cells) * c.species * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download species");
    check(hipStreamSynchronize(ctx.stream), "download species");
  }
}

}  // namespace msim
