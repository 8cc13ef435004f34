// Transport on a grid of several layers (nz > 1), FP32 levels: the explicit
// forward-Euler step, the survey of an upload (greatest outflow and mass),
// and the implicit step's set-up and Jacobi sweep, one thread per site with
// every signal of it, in the reference's three-axis float order
// (adr_site3.hpp; cpp/core/signals.cpp:1109-1134 forward Euler, 1145-1291
// transport, 305-320 affine reaction, 1293-1375 Crank-Nicolson and backward
// Euler). One-layer grids take the marching kernels of adr.hip and
// implicit.hip; these kernels exist for the reference's layered grids
// (its conformance grids are 9 x 7 x 5), where each layer is a plane of
// ny * nz sites with z contiguous, so a wave's neighbour loads stay within a
// few cache lines.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "adr_site3.hpp"
#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kBlock3 = 256;

// The site's geometry once (closure and, with a field, the face
// velocities); constant advection is the signal's own.
__device__ inline Geometry3 signal_geometry(const AdrParams& P, const Geometry3& base, int s) {
  Geometry3 g = base;
  if (P.has_field == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) g.ulo[a] = g.uup[a] = P.advection[s][a];
  }
  return g;
}

// One forward-Euler step (signal_grid_forward_euler_candidate): the rate
// with the affine reaction and Monod uptake (as the affine loss
// vmax*phi/(K+c) it equals under forward Euler), c + dt * rate. With
// `judge`, a non-finite, or under upwind negative, candidate raises the
// flag, and under van Leer each block leaves the largest level it wrote for
// the settle; without, the candidate is only written (a coupled step adds
// its sources and judges the sum).
template <int S, bool kLimited, bool kPatched>
__global__ __launch_bounds__(kBlock3) void k_adr3_step(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ zf, const float* __restrict__ phi,
    const float* __restrict__ source, const float* __restrict__ loss, const float* __restrict__ in,
    float* __restrict__ out, float dt, uint32_t* flags, uint32_t step, double* largest,
    int judge) {
  __shared__ double s_red[kBlock3 / kWave];
  if (flags[0] != 0) return;  // an earlier step failed: the run stops there
  const size_t n = static_cast<size_t>(P.nx) * P.ny * P.nz;
  const size_t i = static_cast<size_t>(blockIdx.x) * kBlock3 + threadIdx.x;
  bool bad = false;
  double top = 0.0;
  if (i < n) {
    if (solid[i] != 0) {
#pragma unroll
      for (int s = 0; s < S; ++s) out[s * n + i] = in[s * n + i];
    } else {
      const Site3 site = site3(P, i);
      const Geometry3 base = geometry3(P, solid, xf, yf, zf, site, 0);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const float* lv = in + s * n;
        const Geometry3 g = signal_geometry(P, base, s);
        const float c = lv[i];
        float rate = transport_rate3_f32<kLimited, kPatched>(P, solid, g, lv, site, s);
        if (P.has_affine != 0) rate += source[s * n + i] - (loss[s * n + i] * c);
        if (P.has_monod != 0) {
          const float uptake = P.vmax[s] * phi[i] / (P.monod_k[s] + c);
          rate += 0.0f - (uptake * c);
        }
        const float next = c + (dt * rate);
        if (!isfinite(next) || (!kLimited && next < 0.0f)) bad = true;
        out[s * n + i] = next;
        top = fmax(top, static_cast<double>(next));
      }
    }
  }
  if (judge == 0) return;
  if (bad) {
    atomicOr(&flags[0], kFlagNegative);
    atomicMin(&flags[1], step);
  }
  if constexpr (kLimited) {
    const double m = block_max<kBlock3>(top, s_red);
    if (threadIdx.x == 0) largest[blockIdx.x] = m;
  }
}

// Sum of levels per signal (FP64) and the greatest outflow over fluid sites
// (SignalGrid::transport_cache, signals.cpp:870-896) over three axes.
__global__ __launch_bounds__(kBlock3) void k_adr3_survey(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ zf, const float* __restrict__ levels,
    double* partials, CgState* st, double* mass) {
  __shared__ double s_red[kBlock3 / kWave];
  __shared__ bool s_last;
  const size_t n = static_cast<size_t>(P.nx) * P.ny * P.nz;
  double acc[kMaxSignals + 1];
#pragma unroll
  for (int k = 0; k <= kMaxSignals; ++k) acc[k] = 0.0;
  const uint32_t dims[3] = {P.nx, P.ny, P.nz};
  const float spacing[3] = {P.dx, P.dy, P.dz};
  const size_t stride = static_cast<size_t>(gridDim.x) * kBlock3;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock3 + threadIdx.x; i < n; i += stride) {
    for (uint32_t s = 0; s < P.signals && s < kMaxSignals; ++s)
      acc[s] += static_cast<double>(levels[s * n + i]);
    if (P.has_field != 0 && solid[i] == 0) {
      const Geometry3 g = geometry3(P, solid, xf, yf, zf, site3(P, i), 0);
      double outflow = 0.0;
      for (int a = 0; a < 3; ++a) {
        if (dims[a] == 1) continue;
        const double inverse_spacing = 1.0 / static_cast<double>(spacing[a]);
        outflow += (fmax(static_cast<double>(g.uup[a]), 0.0) -
                    fmin(static_cast<double>(g.ulo[a]), 0.0)) *
                   inverse_spacing;
      }
      acc[kMaxSignals] = fmax(acc[kMaxSignals], outflow);
    }
  }
  double part[kMaxSignals + 1];
#pragma unroll
  for (int k = 0; k < kMaxSignals; ++k) part[k] = block_sum<kBlock3>(acc[k], s_red);
  part[kMaxSignals] = block_max<kBlock3>(acc[kMaxSignals], s_red);
  if (publish_partials<kMaxSignals + 1>(part, partials, &st->ticket[kTicketMisc], blockIdx.x,
                                        gridDim.x, &s_last)) {
    double total[kMaxSignals];
    sum_partials<kBlock3, kMaxSignals>(partials, gridDim.x, total, s_red);
    double outflow[1];
    max_partials<kBlock3, 1>(partials + static_cast<size_t>(kMaxSignals) * gridDim.x, gridDim.x,
                             outflow, s_red);
    if (threadIdx.x == 0) {
      st->ticket[kTicketMisc] = 0;
      const double volume =
          static_cast<double>(P.dx) * static_cast<double>(P.dy) * static_cast<double>(P.dz);
      for (int s = 0; s < kMaxSignals; ++s) mass[s] = total[s] * volume;
      st->peak = outflow[0];
    }
  }
}

// The implicit step's fixed parts from the levels it starts from
// (signals.cpp:1320-1343): with Monod the lagged loss into `lagged`, the
// right-hand side old + explicit_dt * rate(old) + dt * source, and per block
// the sum of squares of the right-hand side and the least operator diagonal.
template <int S, bool kLimited, bool kPatched>
__global__ __launch_bounds__(kBlock3) void k_implicit3_setup(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ zf, const float* __restrict__ phi,
    const float* __restrict__ source, const float* __restrict__ loss,
    const float* __restrict__ coupled, const float* __restrict__ old, float* __restrict__ rhs,
    float* __restrict__ lagged, float explicit_dt, float dt, double* __restrict__ sums,
    float* __restrict__ least) {
  __shared__ double s_red[kBlock3 / kWave];
  __shared__ float s_min[kBlock3 / kWave];
  const size_t n = static_cast<size_t>(P.nx) * P.ny * P.nz;
  const size_t i = static_cast<size_t>(blockIdx.x) * kBlock3 + threadIdx.x;
  const bool affine = P.has_affine != 0, monod = P.has_monod != 0;
  double acc = 0.0;
  float low = 0.0f;
  if (i < n && solid[i] == 0) {
    const Site3 site = site3(P, i);
    const Geometry3 base = geometry3(P, solid, xf, yf, zf, site, 0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float* lv = old + s * n;
      const Geometry3 g = signal_geometry(P, base, s);
      const float c = lv[i];
      float rate = transport_rate3_f32<kLimited, kPatched>(P, solid, g, lv, site, s);
      float L = affine ? loss[s * n + i] : 0.0f;
      if (monod) {
        const float uptake = P.vmax[s] * phi[i] / (P.monod_k[s] + c);
        L = affine ? L + uptake : uptake;
        lagged[s * n + i] = L;
      }
      const bool reaction = affine || monod;
      if (reaction) rate += (affine ? source[s * n + i] : 0.0f) - (L * c);
      const float r =
          (c + (explicit_dt * rate)) + (dt * (coupled != nullptr ? coupled[s * n + i] : 0.0f));
      rhs[s * n + i] = r;
      acc += static_cast<double>(r) * static_cast<double>(r);
      low = fminf(low, operator_diagonal3_f32(P, g, s, reaction, L));
    }
  }
  const double total = block_sum<kBlock3>(acc, s_red);
  float m = low;
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) m = fminf(m, __shfl_xor(m, offset, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) s_min[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kBlock3 / kWave; ++w) m = fminf(m, s_min[w]);
    sums[blockIdx.x] = total;
    least[blockIdx.x] = m;
  }
}

// One Jacobi iteration (signals.cpp:1348-1373) from iterate `cur`: per fluid
// level the operator rate, the residual rhs - c + half_dt * rate (its square
// summed per block) and the damped update into `nxt`.
template <int S, bool kLimited, bool kPatched>
__global__ __launch_bounds__(kBlock3) void k_implicit3_sweep(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ zf, const float* __restrict__ source,
    const float* __restrict__ loss, const float* __restrict__ rhs, const float* __restrict__ cur,
    float* __restrict__ nxt, float half_dt, float relaxation, ImplicitState* state,
    double* __restrict__ sums) {
  __shared__ double s_red[kBlock3 / kWave];
  if (state->halted != 0) return;
  const size_t n = static_cast<size_t>(P.nx) * P.ny * P.nz;
  const size_t i = static_cast<size_t>(blockIdx.x) * kBlock3 + threadIdx.x;
  const bool affine = P.has_affine != 0, reaction = affine || P.has_monod != 0;
  double acc = 0.0;
  bool bad = false;
  if (i < n && solid[i] == 0) {
    const Site3 site = site3(P, i);
    const Geometry3 base = geometry3(P, solid, xf, yf, zf, site, 0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float* lv = cur + s * n;
      const Geometry3 g = signal_geometry(P, base, s);
      const float v = lv[i];
      if (!isfinite(v)) bad = true;
      const float L = reaction ? loss[s * n + i] : 0.0f;
      float rate = transport_rate3_f32<kLimited, kPatched>(P, solid, g, lv, site, s);
      if (reaction) rate += (affine ? source[s * n + i] : 0.0f) - (L * v);
      const float b = rhs[s * n + i];
      const float r = (b - v) + (half_dt * rate);
      acc += static_cast<double>(r) * static_cast<double>(r);
      const float diagonal = operator_diagonal3_f32(P, g, s, reaction, L);
      const float remainder = rate - (diagonal * v);
      const float update = (b + (half_dt * remainder)) / (1.0f - (half_dt * diagonal));
      nxt[s * n + i] = v + (relaxation * (update - v));
    }
  }
  const double total = block_sum<kBlock3>(acc, s_red);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
  if (bad) atomicOr(&state->nonfinite, 1u);
}

uint32_t blocks3(size_t n) { return static_cast<uint32_t>((n + kBlock3 - 1) / kBlock3); }

// Calls body.template operator()<S, kLimited, kPatched>() for the uploaded
// spec's signal count, scheme and patches.
template <typename Body>
void dispatch3(const AdrParams& P, Body&& body) {
  const bool limited = P.limited != 0, patched = P.edge != nullptr;
  auto by_kind = [&](auto s_tag) {
    constexpr int S = decltype(s_tag)::value;
    if (limited) {
      if (patched) body.template operator()<S, true, true>();
      else body.template operator()<S, true, false>();
    } else {
      if (patched) body.template operator()<S, false, true>();
      else body.template operator()<S, false, false>();
    }
  };
  switch (P.signals) {
    case 1: by_kind(std::integral_constant<int, 1>{}); break;
    case 2: by_kind(std::integral_constant<int, 2>{}); break;
    case 3: by_kind(std::integral_constant<int, 3>{}); break;
    case 4: by_kind(std::integral_constant<int, 4>{}); break;
    case 5: by_kind(std::integral_constant<int, 5>{}); break;
    case 6: by_kind(std::integral_constant<int, 6>{}); break;
    case 7: by_kind(std::integral_constant<int, 7>{}); break;
    default: by_kind(std::integral_constant<int, 8>{}); break;
  }
}

}  // namespace

void adr3_survey(Context& ctx, const float* xf, const float* yf, const float* zf,
                 const float* levels, CgState* state, double* partials, double* mass) {
  const auto& P = ctx.adr.params;
  const size_t n = static_cast<size_t>(P.nx) * P.ny * P.nz;
  const uint32_t blocks = std::min<uint32_t>(blocks3(n), 2048);
  k_adr3_survey<<<blocks, kBlock3, 0, ctx.stream>>>(P, ctx.adr.solid.data(), xf, yf, zf, levels,
                                                    partials, state, mass);
  check(hipGetLastError(), "launch k_adr3_survey");
}

void adr3_run(Context& ctx, float dt, uint32_t steps, const float* xf, const float* yf,
              const float* zf) {
  auto& a = ctx.adr;
  const auto& P = a.params;
  const size_t n = a.sites();
  const size_t count = static_cast<size_t>(P.signals) * n;
  const uint32_t blocks = blocks3(n);
  a.partials.reserve(std::max<size_t>(a.partials.size(), blocks));
  double* largest = a.partials.data();
  for (uint32_t k = 0; k < steps; ++k) {
    const float* src = reinterpret_cast<const float*>(a.current == 0 ? a.levels_a.data()
                                                                     : a.levels_b.data());
    float* dst = reinterpret_cast<float*>(a.current == 0 ? a.levels_b.data() : a.levels_a.data());
    dispatch3(P, [&]<int S, bool kLimited, bool kPatched>() {
      k_adr3_step<S, kLimited, kPatched><<<blocks, kBlock3, 0, ctx.stream>>>(
          P, a.solid.data(), xf, yf, zf, a.phi.data(), a.source.data(), a.loss.data(), src, dst,
          dt, a.flags.data(), k, largest, 1);
    });
    check(hipGetLastError(), "launch k_adr3_step");
    if (P.limited != 0) launch_settle_f32(ctx, dst, count, largest, blocks, k);
    a.current ^= 1;
  }
}

void adr3_candidate(Context& ctx, float dt, const float* in, float* out, uint32_t step,
                    const float* xf, const float* yf, const float* zf) {
  auto& a = ctx.adr;
  const auto& P = a.params;
  const uint32_t blocks = blocks3(a.sites());
  dispatch3(P, [&]<int S, bool kLimited, bool kPatched>() {
    k_adr3_step<S, kLimited, kPatched><<<blocks, kBlock3, 0, ctx.stream>>>(
        P, a.solid.data(), xf, yf, zf, a.phi.data(), a.source.data(), a.loss.data(), in, out, dt,
        a.flags.data(), step, nullptr, 0);
  });
  check(hipGetLastError(), "launch k_adr3_step");
}

void implicit3_setup(Context& ctx, uint32_t blocks, const float* xf, const float* yf,
                     const float* zf, const float* sources, const float* old, float explicit_dt,
                     float dt) {
  auto& a = ctx.adr;
  dispatch3(a.params, [&]<int S, bool kLimited, bool kPatched>() {
    k_implicit3_setup<S, kLimited, kPatched><<<blocks, kBlock3, 0, ctx.stream>>>(
        a.params, a.solid.data(), xf, yf, zf, a.phi.data(), a.source.data(), a.loss.data(),
        sources, old, a.rhs.data(), a.lagged.data(), explicit_dt, dt, a.implicit_sums.data(),
        a.implicit_least.data());
  });
  check(hipGetLastError(), "launch k_implicit3_setup");
}

void implicit3_sweep(Context& ctx, uint32_t blocks, const float* xf, const float* yf,
                     const float* zf, const float* cur, float* nxt, float half_dt,
                     float relaxation) {
  auto& a = ctx.adr;
  const float* loss = a.params.has_monod != 0 ? a.lagged.data() : a.loss.data();
  dispatch3(a.params, [&]<int S, bool kLimited, bool kPatched>() {
    k_implicit3_sweep<S, kLimited, kPatched><<<blocks, kBlock3, 0, ctx.stream>>>(
        a.params, a.solid.data(), xf, yf, zf, a.source.data(), loss, a.rhs.data(), cur, nxt,
        half_dt, relaxation, a.implicit_state.data(), a.implicit_sums.data());
  });
  check(hipGetLastError(), "launch k_implicit3_sweep");
}

}  // namespace msim
