// Implicit (Crank-Nicolson or backward Euler) transport steps on the device,
// FP32 levels (grids of several layers take adr3.hip's set-up and sweep): the reference's damped Jacobi iteration
// (signal_grid_crank_nicolson_candidate, cpp/core/signals.cpp:1293-1375)
// under its stopping rule (solve_signal_jacobi,
// cpp/core/signal_jacobi.hpp:53-84), then the step's settle and validation
// (SignalGrid::accept_step, signals.cpp:1003-1011).
//
// One kernel per iteration does what the reference does in two passes: from
// iterate k it computes every level's operator rate (transport plus affine
// reaction), the residual of iterate k for the stopping rule, and the Jacobi
// update that is iterate k + 1, which needs only the site's own rate. A
// single-block kernel then turns the residual's partial sums into its RMS
// and decides. The host enqueues a batch of iterations at a time; once the
// rule is met (or the residual is not finite) every later kernel of the
// batch returns at once, so the iterate the rule accepted stays in its
// buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "adr_site.hpp"
#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kImplicitBlock = 256;
constexpr uint32_t kImplicitBatch = 16;  // iterations between host checks

// The iteration's site work shared by the setup and the sweep: the site's
// geometry (closure once, face velocities per signal) and the neighbours of
// signal s in `lv`, with exterior values resolved.
struct SiteView {
  SiteGeometry g;
  float c, l0, u0, l1, u1;
};

template <bool kPatched>
__device__ inline void site_view(const AdrParams& P, const uint8_t* __restrict__ solid,
                                 const float* __restrict__ xf, const float* __restrict__ yf,
                                 const float* __restrict__ lv, uint32_t x, uint32_t y, int s,
                                 const SiteGeometry& base, SiteView& v) {
  const uint32_t nx = P.nx, ny = P.ny;
  const size_t i = static_cast<size_t>(x) * ny + y;
  v.g = base;
  if (P.has_field == 0) {
    v.g.ulo[0] = v.g.uup[0] = P.advection[s][0];
    v.g.ulo[1] = v.g.uup[1] = P.advection[s][1];
  }
  const float c = lv[i];
  v.c = c;
  v.l0 = x == 0 ? exterior<kPatched>(P, 0, s, c, lv[static_cast<size_t>(nx - 1) * ny + y], y)
                : lv[i - ny];
  v.u0 = x + 1 == nx ? exterior<kPatched>(P, 1, s, c, lv[y], y) : lv[i + ny];
  v.l1 = y == 0 ? exterior<kPatched>(P, 2, s, c, lv[static_cast<size_t>(x) * ny + ny - 1], x)
                : lv[i - 1];
  v.u1 = y + 1 == ny ? exterior<kPatched>(P, 3, s, c, lv[static_cast<size_t>(x) * ny], x)
                     : lv[i + 1];
}

template <bool kLimited, bool kPatched>
__device__ inline float site_rate(const AdrParams& P, const uint8_t* __restrict__ solid,
                                  const float* __restrict__ lv, uint32_t x, uint32_t y, int s,
                                  const SiteView& v) {
  if constexpr (kLimited) {
    Limiter<float> lim;
    site_limiter(P, solid, v.g, lv, x, y, lim);
    return transport_rate_f32<true>(P, v.g, v.c, v.l0, v.u0, v.l1, v.u1, s, &lim);
  } else {
    return transport_rate_f32<false>(P, v.g, v.c, v.l0, v.u0, v.l1, v.u1, s);
  }
}

// The step's fixed parts from the levels it starts from (signals.cpp:
// 1320-1343): with Monod uptake the lagged loss vmax*phi/(K+c) (added to any
// affine loss) into `lagged`; the right-hand side
// old + explicit_dt * rate(old) + dt * source, with the coupled step's
// sources (zero without); and per block the sum of squares of
// the right-hand side (for its RMS) and the least operator diagonal (for the
// operator scale).
template <int S, bool kLimited, bool kPatched>
__global__ __launch_bounds__(kImplicitBlock) void k_implicit_setup(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ phi, const float* __restrict__ source,
    const float* __restrict__ loss, const float* __restrict__ coupled,
    const float* __restrict__ old, float* __restrict__ rhs, float* __restrict__ lagged,
    float explicit_dt, float dt, double* __restrict__ sums, float* __restrict__ least) {
  __shared__ double s_red[kImplicitBlock / kWave];
  __shared__ float s_min[kImplicitBlock / kWave];
  const uint32_t nx = P.nx, ny = P.ny;
  const size_t n = static_cast<size_t>(nx) * ny;
  const size_t i = static_cast<size_t>(blockIdx.x) * kImplicitBlock + threadIdx.x;
  const bool affine = P.has_affine != 0, monod = P.has_monod != 0;
  double acc = 0.0;
  float low = 0.0f;
  if (!kLimited && i < n && solid[i] == 0) {
    // Upwind: every value this site reads is loaded before any is used (as
    // in the sweep); the arithmetic is the general path's below.
    const uint32_t x = static_cast<uint32_t>(i / ny), y = static_cast<uint32_t>(i % ny);
    const size_t row = static_cast<size_t>(x) * ny;
    const size_t at[4] = {x == 0 ? static_cast<size_t>(nx - 1) * ny + y : i - ny,
                          x + 1 == nx ? static_cast<size_t>(y) : i + ny,
                          y == 0 ? row + ny - 1 : i - 1, y + 1 == ny ? row : i + 1};
    float c[S], nb[S][4], L0[S], q[S];
    const float occupancy = monod ? phi[i] : 0.0f;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float* lv = old + s * n;
      c[s] = lv[i];
#pragma unroll
      for (int k = 0; k < 4; ++k) nb[s][k] = lv[at[k]];
      L0[s] = affine ? loss[s * n + i] : 0.0f;
      q[s] = affine ? source[s * n + i] : 0.0f;
    }
    const SiteGeometry base = geometry(P, solid, xf, yf, x, y, 0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      SiteGeometry g = base;
      if (P.has_field == 0) {
        g.ulo[0] = g.uup[0] = P.advection[s][0];
        g.ulo[1] = g.uup[1] = P.advection[s][1];
      }
      const float v = c[s];
      const float l0 = x == 0 ? exterior<kPatched>(P, 0, s, v, nb[s][0], y) : nb[s][0];
      const float u0 = x + 1 == nx ? exterior<kPatched>(P, 1, s, v, nb[s][1], y) : nb[s][1];
      const float l1 = y == 0 ? exterior<kPatched>(P, 2, s, v, nb[s][2], x) : nb[s][2];
      const float u1 = y + 1 == ny ? exterior<kPatched>(P, 3, s, v, nb[s][3], x) : nb[s][3];
      float rate = transport_rate_f32<false>(P, g, v, l0, u0, l1, u1, s);
      float L = L0[s];
      if (monod) {
        const float uptake = P.vmax[s] * occupancy / (P.monod_k[s] + v);
        L = affine ? L + uptake : uptake;
        lagged[s * n + i] = L;
      }
      const bool reaction = affine || monod;
      if (reaction) rate += q[s] - (L * v);
      const float r = (v + (explicit_dt * rate)) +
                      (dt * (coupled != nullptr ? coupled[s * n + i] : 0.0f));
      rhs[s * n + i] = r;
      acc += static_cast<double>(r) * static_cast<double>(r);
      low = fminf(low, operator_diagonal_f32(P, g, s, reaction, L));
    }
  } else if (i < n && solid[i] == 0) {
    const uint32_t x = static_cast<uint32_t>(i / ny), y = static_cast<uint32_t>(i % ny);
    const SiteGeometry base = geometry(P, solid, xf, yf, x, y, 0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float* lv = old + s * n;
      SiteView v;
      site_view<kPatched>(P, solid, xf, yf, lv, x, y, s, base, v);
      float rate = site_rate<kLimited, kPatched>(P, solid, lv, x, y, s, v);
      float L = affine ? loss[s * n + i] : 0.0f;
      if (monod) {
        const float uptake = P.vmax[s] * phi[i] / (P.monod_k[s] + v.c);
        L = affine ? L + uptake : uptake;
        lagged[s * n + i] = L;
      }
      const bool reaction = affine || monod;
      if (reaction) rate += (affine ? source[s * n + i] : 0.0f) - (L * v.c);
      const float r = (v.c + (explicit_dt * rate)) + (dt * (coupled != nullptr ? coupled[s * n + i]
                                                                              : 0.0f));
      rhs[s * n + i] = r;
      acc += static_cast<double>(r) * static_cast<double>(r);
      low = fminf(low, operator_diagonal_f32(P, v.g, s, reaction, L));
    }
  }
  const double total = block_sum<kImplicitBlock>(acc, s_red);
  float m = low;
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) m = fminf(m, __shfl_xor(m, offset, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) s_min[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kImplicitBlock / kWave; ++w) m = fminf(m, s_min[w]);
    sums[blockIdx.x] = total;
    least[blockIdx.x] = m;
  }
}

// One Jacobi iteration (signals.cpp:1348-1373) from iterate `cur`: per fluid
// level the operator rate, the residual rhs - c + half_dt * rate (its square
// summed per block) and the damped update into `nxt`. Levels at obstacle
// sites are never written. A non-finite level of the iterate is recorded: the
// reference refuses to take the rates of such levels.
template <int S, bool kLimited, bool kPatched>
__global__ __launch_bounds__(kImplicitBlock) void k_implicit_sweep(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ source, const float* __restrict__ loss,
    const float* __restrict__ rhs, const float* __restrict__ cur, float* __restrict__ nxt,
    float half_dt, float relaxation, ImplicitState* state, double* __restrict__ sums) {
  __shared__ double s_red[kImplicitBlock / kWave];
  if (state->halted != 0) return;
  const uint32_t nx = P.nx, ny = P.ny;
  const size_t n = static_cast<size_t>(nx) * ny;
  const size_t i = static_cast<size_t>(blockIdx.x) * kImplicitBlock + threadIdx.x;
  const bool affine = P.has_affine != 0, reaction = affine || P.has_monod != 0;
  double acc = 0.0;
  bool bad = false;
  if constexpr (!kLimited) {
    // Upwind: every level this site reads is loaded before any is used, so a
    // wave has all of its loads in flight at once; then the same arithmetic
    // as the general path below.
    if (i < n && solid[i] == 0) {
      const uint32_t x = static_cast<uint32_t>(i / ny), y = static_cast<uint32_t>(i % ny);
      const size_t row = static_cast<size_t>(x) * ny;
      const size_t at[4] = {x == 0 ? static_cast<size_t>(nx - 1) * ny + y : i - ny,
                            x + 1 == nx ? static_cast<size_t>(y) : i + ny,
                            y == 0 ? row + ny - 1 : i - 1, y + 1 == ny ? row : i + 1};
      float c[S], nb[S][4], b[S], L[S], q[S];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const float* lv = cur + s * n;
        c[s] = lv[i];
#pragma unroll
        for (int k = 0; k < 4; ++k) nb[s][k] = lv[at[k]];
        b[s] = rhs[s * n + i];
        L[s] = reaction ? loss[s * n + i] : 0.0f;
        q[s] = affine ? source[s * n + i] : 0.0f;
      }
      const SiteGeometry base = geometry(P, solid, xf, yf, x, y, 0);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        SiteGeometry g = base;
        if (P.has_field == 0) {
          g.ulo[0] = g.uup[0] = P.advection[s][0];
          g.ulo[1] = g.uup[1] = P.advection[s][1];
        }
        const float v = c[s];
        if (!isfinite(v)) bad = true;
        const float l0 = x == 0 ? exterior<kPatched>(P, 0, s, v, nb[s][0], y) : nb[s][0];
        const float u0 = x + 1 == nx ? exterior<kPatched>(P, 1, s, v, nb[s][1], y) : nb[s][1];
        const float l1 = y == 0 ? exterior<kPatched>(P, 2, s, v, nb[s][2], x) : nb[s][2];
        const float u1 = y + 1 == ny ? exterior<kPatched>(P, 3, s, v, nb[s][3], x) : nb[s][3];
        float rate = transport_rate_f32<false>(P, g, v, l0, u0, l1, u1, s);
        if (reaction) rate += q[s] - (L[s] * v);
        const float r = (b[s] - v) + (half_dt * rate);
        acc += static_cast<double>(r) * static_cast<double>(r);
        const float diagonal = operator_diagonal_f32(P, g, s, reaction, L[s]);
        const float remainder = rate - (diagonal * v);
        const float update = (b[s] + (half_dt * remainder)) / (1.0f - (half_dt * diagonal));
        nxt[s * n + i] = v + (relaxation * (update - v));
      }
    }
  } else if (i < n && solid[i] == 0) {
    // van Leer: the same, with the four sites two away (site_limiter's, in
    // adr_site.hpp) loaded up front as well.
    const uint32_t x = static_cast<uint32_t>(i / ny), y = static_cast<uint32_t>(i % ny);
    const size_t row = static_cast<size_t>(x) * ny;
    const size_t at[4] = {x == 0 ? static_cast<size_t>(nx - 1) * ny + y : i - ny,
                          x + 1 == nx ? static_cast<size_t>(y) : i + ny,
                          y == 0 ? row + ny - 1 : i - 1, y + 1 == ny ? row : i + 1};
    const bool xp3 = P.kind[0] == MSIM_PERIODIC && nx >= 3;
    const bool yp3 = P.kind[2] == MSIM_PERIODIC && ny >= 3;
    // Far sites: x - 2, x + 2, y - 2, y + 2 (wrapped on a periodic axis of at
    // least three sites, absent past a boundary otherwise).
    size_t far_at[4];
    bool far_in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int axis = k >> 1;
      const int64_t count = axis == 0 ? nx : ny;
      const int64_t along = static_cast<int64_t>(axis == 0 ? x : y) + ((k & 1) ? 2 : -2);
      far_in[k] = (axis == 0 ? xp3 : yp3) || (along >= 0 && along < count);
      const uint32_t wrapped = static_cast<uint32_t>(((along % count) + count) % count);
      far_at[k] = axis == 0 ? static_cast<size_t>(wrapped) * ny + y : row + wrapped;
    }
    bool far_open[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) far_open[k] = far_in[k] && solid[far_at[k]] == 0;
    float c[S], nb[S][4], fv[S][4], b[S], L[S], q[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float* lv = cur + s * n;
      c[s] = lv[i];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        nb[s][k] = lv[at[k]];
        fv[s][k] = far_in[k] ? lv[far_at[k]] : 0.0f;
      }
      b[s] = rhs[s * n + i];
      L[s] = reaction ? loss[s * n + i] : 0.0f;
      q[s] = affine ? source[s * n + i] : 0.0f;
    }
    const SiteGeometry base = geometry(P, solid, xf, yf, x, y, 0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      SiteGeometry g = base;
      if (P.has_field == 0) {
        g.ulo[0] = g.uup[0] = P.advection[s][0];
        g.ulo[1] = g.uup[1] = P.advection[s][1];
      }
      const float v = c[s];
      if (!isfinite(v)) bad = true;
      const float l0 = x == 0 ? exterior<kPatched>(P, 0, s, v, nb[s][0], y) : nb[s][0];
      const float u0 = x + 1 == nx ? exterior<kPatched>(P, 1, s, v, nb[s][1], y) : nb[s][1];
      const float l1 = y == 0 ? exterior<kPatched>(P, 2, s, v, nb[s][2], x) : nb[s][2];
      const float u1 = y + 1 == ny ? exterior<kPatched>(P, 3, s, v, nb[s][3], x) : nb[s][3];
      Limiter<float> lim;
      lim.lower_site[0] = !g.clo[0] && (x > 0 || xp3);
      lim.upper_site[0] = !g.cup[0] && (x + 1 < nx || xp3);
      lim.lower_site[1] = !g.clo[1] && (y > 0 || yp3);
      lim.upper_site[1] = !g.cup[1] && (y + 1 < ny || yp3);
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        lim.far_lo_ok[a] = far_open[2 * a];
        lim.far_hi_ok[a] = far_open[2 * a + 1];
        lim.far_lo[a] = far_open[2 * a] ? fv[s][2 * a] : 0.0f;
        lim.far_hi[a] = far_open[2 * a + 1] ? fv[s][2 * a + 1] : 0.0f;
      }
      float rate = transport_rate_f32<true>(P, g, v, l0, u0, l1, u1, s, &lim);
      if (reaction) rate += q[s] - (L[s] * v);
      const float r = (b[s] - v) + (half_dt * rate);
      acc += static_cast<double>(r) * static_cast<double>(r);
      const float diagonal = operator_diagonal_f32(P, g, s, reaction, L[s]);
      const float remainder = rate - (diagonal * v);
      const float update = (b[s] + (half_dt * remainder)) / (1.0f - (half_dt * diagonal));
      nxt[s * n + i] = v + (relaxation * (update - v));
    }
  }
  const double total = block_sum<kImplicitBlock>(acc, s_red);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
  if (bad) atomicOr(&state->nonfinite, 1u);
}

// The RMS of the right-hand side and the operator scale
// max(1, 1 - half_dt * least diagonal) (SignalGridLayout::
// implicit_operator_scale, signals.cpp:823-825) from the setup's partials;
// resets the solve's state.
__global__ __launch_bounds__(kImplicitBlock) void k_implicit_begin(
    const double* __restrict__ sums, const float* __restrict__ least, uint32_t blocks,
    double levels, float half_dt, ImplicitState* state) {
  __shared__ double s_red[kImplicitBlock / kWave];
  __shared__ float s_min[kImplicitBlock / kWave];
  double v = 0.0;
  float m = 0.0f;
  for (uint32_t b = threadIdx.x; b < blocks; b += kImplicitBlock) {
    v += sums[b];
    m = fminf(m, least[b]);
  }
  v = block_sum<kImplicitBlock>(v, s_red);
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) m = fminf(m, __shfl_xor(m, offset, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) s_min[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kImplicitBlock / kWave; ++w) m = fminf(m, s_min[w]);
    state->rhs_rms = static_cast<float>(sqrt(v / levels));
    const float scaled = 1.0f - (half_dt * m);
    state->scale = scaled > 1.0f ? scaled : 1.0f;
    state->threshold = 0.0f;
    state->rms = 0.0f;
    state->halted = 0;
    state->converged = 0;
    state->iteration = 0;
    state->nonfinite = 0;
  }
}

// solve_signal_jacobi's decision after iterate k (signal_jacobi.hpp:53-84):
// the residual RMS from the sweep's partials; at k = 0 the threshold
// max(absolute, eps * rhs RMS * scale) + relative * RMS; stop converged when
// the RMS meets it, unconverged when it (or the threshold) is not finite.
__global__ __launch_bounds__(kImplicitBlock) void k_implicit_decide(
    const double* __restrict__ sums, uint32_t blocks, double levels, uint32_t k,
    float absolute_tolerance, float relative_tolerance, ImplicitState* state) {
  __shared__ double s_red[kImplicitBlock / kWave];
  if (state->halted != 0) return;
  double v = 0.0;
  for (uint32_t b = threadIdx.x; b < blocks; b += kImplicitBlock) v += sums[b];
  v = block_sum<kImplicitBlock>(v, s_red);
  if (threadIdx.x != 0) return;
  const float rms = static_cast<float>(sqrt(v / levels));
  state->rms = rms;
  state->iteration = k;
  if (state->nonfinite != 0) {
    state->halted = 1;
    return;
  }
  if (k == 0) {
    const float floor = FLT_EPSILON * state->rhs_rms * state->scale;
    const float threshold =
        (absolute_tolerance < floor ? floor : absolute_tolerance) + (relative_tolerance * rms);
    state->threshold = threshold;
    if (isfinite(rms) && rms <= threshold) {
      state->converged = 1;
      state->halted = 1;
    } else if (!isfinite(rms) || !isfinite(threshold)) {
      state->halted = 1;
    }
    return;
  }
  if (!isfinite(rms)) {
    state->halted = 1;
  } else if (rms <= state->threshold) {
    state->converged = 1;
    state->halted = 1;
  }
}

// The largest level of the accepted iterate, per block (for the settle).
__global__ __launch_bounds__(kImplicitBlock) void k_level_max(const float* __restrict__ levels,
                                                              size_t count, double* largest) {
  __shared__ double s_red[kImplicitBlock / kWave];
  double m = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kImplicitBlock + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * kImplicitBlock) {
    m = fmax(m, static_cast<double>(levels[i]));
  }
  m = block_max<kImplicitBlock>(m, s_red);
  if (threadIdx.x == 0) largest[blockIdx.x] = m;
}

template <int S, bool kLimited, bool kPatched>
void launch_setup(Context& ctx, uint32_t blocks, const float* xf, const float* yf,
                  const float* sources, const float* old, float explicit_dt, float dt) {
  auto& a = ctx.adr;
  if (a.params.nz > 1) {
    implicit3_setup(ctx, blocks, xf, yf, a.zf.data(), sources, old, explicit_dt, dt);
    return;
  }
  k_implicit_setup<S, kLimited, kPatched><<<blocks, kImplicitBlock, 0, ctx.stream>>>(
      a.params, a.solid.data(), xf, yf, a.phi.data(), a.source.data(), a.loss.data(), sources,
      old, a.rhs.data(), a.lagged.data(), explicit_dt, dt, a.implicit_sums.data(),
      a.implicit_least.data());
}

template <int S, bool kLimited, bool kPatched>
void launch_sweep(Context& ctx, uint32_t blocks, const float* xf, const float* yf,
                  const float* cur, float* nxt, float half_dt, float relaxation) {
  auto& a = ctx.adr;
  if (a.params.nz > 1) {
    implicit3_sweep(ctx, blocks, xf, yf, a.zf.data(), cur, nxt, half_dt, relaxation);
    return;
  }
  const float* loss = a.params.has_monod != 0 ? a.lagged.data() : a.loss.data();
  k_implicit_sweep<S, kLimited, kPatched><<<blocks, kImplicitBlock, 0, ctx.stream>>>(
      a.params, a.solid.data(), xf, yf, a.source.data(), loss, a.rhs.data(), cur, nxt, half_dt,
      relaxation, a.implicit_state.data(), a.implicit_sums.data());
}

template <int S, bool kLimited, bool kPatched>
void implicit_steps(Context& ctx, float dt, uint32_t steps, const float* xf, const float* yf,
                    const float* sources, const char* label) {
  auto& a = ctx.adr;
  const size_t n = a.sites();
  const size_t count = static_cast<size_t>(S) * n;
  const double levels = static_cast<double>(count);
  const uint32_t blocks = static_cast<uint32_t>((n + kImplicitBlock - 1) / kImplicitBlock);
  const bool backward = a.integration == MSIM_BACKWARD_EULER;
  const float half_dt = (backward ? 1.0f : 0.5f) * dt;
  const float explicit_dt = backward ? 0.0f : half_dt;
  const float relaxation = kLimited ? 0.5f : 1.0f;
  float* base[2] = {reinterpret_cast<float*>(a.levels_a.data()),
                    reinterpret_cast<float*>(a.levels_b.data())};
  float* third = reinterpret_cast<float*>(a.levels_c.data());
  for (uint32_t step = 0; step < steps; ++step) {
    // The step's levels stay untouched in `old` until it is accepted: the
    // iterates alternate between the two other buffers, iterate k (k >= 1)
    // in work[k & 1]. Both start as copies, so they hold the levels of the
    // obstacle sites, which no kernel writes.
    float* old = base[a.current];
    float* work[2] = {third, base[a.current ^ 1]};
    for (float* w : work) {
      check(hipMemcpyAsync(w, old, count * sizeof(float), hipMemcpyDeviceToDevice, ctx.stream),
            "copy levels");
    }
    auto iterate = [&](uint32_t k) { return k == 0 ? old : work[k & 1u]; };
    launch_setup<S, kLimited, kPatched>(ctx, blocks, xf, yf, sources, old, explicit_dt, dt);
    check(hipGetLastError(), "launch k_implicit_setup");
    k_implicit_begin<<<1, kImplicitBlock, 0, ctx.stream>>>(
        a.implicit_sums.data(), a.implicit_least.data(), blocks, levels, half_dt,
        a.implicit_state.data());
    check(hipGetLastError(), "launch k_implicit_begin");
    // Iterates 0..max_iterations are judged, a batch at a time.
    ImplicitState host{};
    uint32_t k = 0;
    // The first batch reaches the iterate at which the previous step stopped
    // (steps of one run stop at similar iterates), later ones kImplicitBatch
    // further: a launch after the rule is met costs a few microseconds.
    uint32_t batch = std::clamp<uint32_t>(a.predicted_iterations + 1, 2, kImplicitBatch);
    for (;;) {
      const uint32_t last = static_cast<uint32_t>(
          std::min<uint64_t>(static_cast<uint64_t>(k) + batch - 1, a.max_iterations));
      batch = kImplicitBatch;
      for (; k <= last; ++k) {
        launch_sweep<S, kLimited, kPatched>(ctx, blocks, xf, yf, iterate(k), work[(k + 1) & 1u],
                                            half_dt, relaxation);
        check(hipGetLastError(), "launch k_implicit_sweep");
        k_implicit_decide<<<1, kImplicitBlock, 0, ctx.stream>>>(
            a.implicit_sums.data(), blocks, levels, k, a.absolute_tolerance,
            a.relative_tolerance, a.implicit_state.data());
        check(hipGetLastError(), "launch k_implicit_decide");
      }
      check(hipMemcpyAsync(&host, a.implicit_state.data(), sizeof(host), hipMemcpyDeviceToHost,
                           ctx.stream),
            "read implicit state");
      check(hipStreamSynchronize(ctx.stream), "implicit iterations");
      if (host.halted != 0 || k > a.max_iterations) break;
    }
    a.last_report.converged = host.converged;
    a.last_report.iterations = host.halted != 0 ? host.iteration : a.max_iterations;
    a.last_report.residual_rms = host.rms;
    a.predicted_iterations = a.last_report.iterations;
    if (host.nonfinite != 0) {
      throw Error(MSIM_INVALID_ARGUMENT, "signal transport levels must be finite");
    }
    if (host.converged == 0) {
      std::string message = std::string(label) + " did not converge after " +
                            std::to_string(a.last_report.iterations) + " iterations";
      if (kLimited) {
        message +=
            "; under the van Leer scheme, shorten the step until flow crosses only a few sites "
            "in one";
      }
      throw Error(MSIM_NOT_CONVERGED, message);
    }
    if (host.iteration == 0) continue;  // the levels already met the rule
    // accept_step (signals.cpp:1003-1011): the settle under van Leer, then
    // the levels' validation, in place on the accepted iterate.
    float* accepted = work[host.iteration & 1u];
    const uint32_t parts = static_cast<uint32_t>(std::min<size_t>((count + 255) / 256, 1024));
    if (kLimited) {
      k_level_max<<<parts, kImplicitBlock, 0, ctx.stream>>>(accepted, count, a.partials.data());
      check(hipGetLastError(), "launch k_level_max");
    }
    launch_settle_f32(ctx, accepted, count, a.partials.data(), kLimited ? parts : 0, step);
    uint32_t flags[2] = {0, 0};
    check(hipMemcpyAsync(flags, a.flags.data(), sizeof(flags), hipMemcpyDeviceToHost, ctx.stream),
          "read flags");
    check(hipStreamSynchronize(ctx.stream), "settle");
    if (flags[0] != 0) {
      throw Error(MSIM_INVALID_ARGUMENT, "signal grid levels must be finite and non-negative");
    }
    if (accepted == third) {
      check(hipMemcpyAsync(base[a.current ^ 1], third, count * sizeof(float),
                           hipMemcpyDeviceToDevice, ctx.stream),
            "copy accepted levels");
    }
    a.current ^= 1;
  }
}

template <int S>
void implicit_typed(Context& ctx, float dt, uint32_t steps, const float* xf, const float* yf,
                    const float* sources, const char* label) {
  const auto& P = ctx.adr.params;
  const bool limited = P.limited != 0, patched = P.edge != nullptr;
  if (limited) {
    if (patched) implicit_steps<S, true, true>(ctx, dt, steps, xf, yf, sources, label);
    else implicit_steps<S, true, false>(ctx, dt, steps, xf, yf, sources, label);
  } else {
    if (patched) implicit_steps<S, false, true>(ctx, dt, steps, xf, yf, sources, label);
    else implicit_steps<S, false, false>(ctx, dt, steps, xf, yf, sources, label);
  }
}

}  // namespace

void implicit_run(Context& ctx, float dt, uint32_t steps, const float* xf, const float* yf,
                  const float* sources, const char* label) {
  auto& a = ctx.adr;
  const size_t n = a.sites();
  const uint32_t blocks = static_cast<uint32_t>((n + kImplicitBlock - 1) / kImplicitBlock);
  a.rhs.reserve(static_cast<size_t>(a.params.signals) * n);
  a.levels_c.reserve(static_cast<size_t>(a.params.signals) * n * sizeof(float));
  if (a.params.has_monod != 0) a.lagged.reserve(static_cast<size_t>(a.params.signals) * n);
  a.implicit_sums.reserve(blocks);
  a.implicit_least.reserve(blocks);
  a.implicit_state.reserve(1);
  switch (a.params.signals) {
    case 1: implicit_typed<1>(ctx, dt, steps, xf, yf, sources, label); break;
    case 2: implicit_typed<2>(ctx, dt, steps, xf, yf, sources, label); break;
    case 3: implicit_typed<3>(ctx, dt, steps, xf, yf, sources, label); break;
    case 4: implicit_typed<4>(ctx, dt, steps, xf, yf, sources, label); break;
    case 5: implicit_typed<5>(ctx, dt, steps, xf, yf, sources, label); break;
    case 6: implicit_typed<6>(ctx, dt, steps, xf, yf, sources, label); break;
    case 7: implicit_typed<7>(ctx, dt, steps, xf, yf, sources, label); break;
    default: implicit_typed<8>(ctx, dt, steps, xf, yf, sources, label); break;
  }
}

}  // namespace msim
