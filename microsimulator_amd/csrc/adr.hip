// Explicit advection-diffusion-reaction step on the device, one layer.
//
// Two kernels share one site-per-thread structure (each thread advances every
// signal of its site, so the face velocities, mask and occupancy are read once
// for all signals):
//  * k_adr_f32: FP32 levels, the reference's per-site rate in its own float
//    order (cpp/core/signals.cpp:1145-1291 transport, 305-320 affine
//    reaction, 1109-1134 forward Euler), Monod uptake as the affine loss
//    vmax*phi/(K+c) it equals under forward Euler.
//  * k_adr_f64: FP64 levels in face-flux form: every face flux is one
//    expression of the two sites it joins, so what leaves one site enters the
//    other bit for bit; the step's mass, boundary supply and uptake are
//    tallied in FP64 for the conservation ledger.
// A site whose update is negative or non-finite raises a flag (the reference
// throws, signals.cpp:1124-1130); the step is then refused as a whole.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adr_site.hpp"
#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kAdrBlock = 256;
constexpr uint32_t kFlagNegative = 1;

// One FP32 forward-Euler step for S signals; one site per thread.
template <int S>
__global__ __launch_bounds__(kAdrBlock) void k_adr_f32(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ phi, const float* __restrict__ source,
    const float* __restrict__ loss, const float* __restrict__ in, float* __restrict__ out, float dt,
    uint32_t* flags, uint32_t step) {
  if (flags[0] != 0) return;  // an earlier step failed: the run stops there
  const size_t n = static_cast<size_t>(P.nx) * P.ny;
  const size_t i = static_cast<size_t>(blockIdx.x) * kAdrBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t x = static_cast<uint32_t>(i / P.ny), y = static_cast<uint32_t>(i % P.ny);
  if (solid[i] != 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) out[s * n + i] = in[s * n + i];
    return;
  }
  bool bad = false;
  SiteGeometry g = geometry(P, solid, xf, yf, x, y, 0);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (P.has_field == 0 && s > 0) g = geometry(P, solid, xf, yf, x, y, s);
    const size_t index = s * n + i;
    const float candidate = fe_candidate_f32(P, g, in + s * n, x, y, s, dt,
                                             P.has_affine != 0 ? source[index] : 0.0f,
                                             P.has_affine != 0 ? loss[index] : 0.0f, phi);
    if (!isfinite(candidate) || candidate < 0.0f) bad = true;
    out[index] = candidate;
  }
  if (bad) {
    atomicOr(&flags[0], kFlagNegative);
    atomicMin(&flags[1], step);
  }
}

// The ledger of a step lives as per-block partial sums (mass after, boundary
// supply rate, uptake rate, for each signal) in two generations by step
// parity. Blocks j < 3S of the next step add partial j of the previous step
// (in a fixed order) into its ledger row, so no grid-wide fence is needed; a
// last small launch adds the final step's.
struct LedgerParts {
  double* parts;   // [2][3 * kMaxSignals][blocks]
  double* ledger;  // [steps][S][4]
  uint32_t blocks;
};

template <int S, int NT>
__device__ inline void fold_ledger(const AdrParams& P, LedgerParts L, uint32_t step, float dt,
                                   uint32_t j, double* lds, double* bcast) {
  // Partial j of `step` into ledger row `step`.
  const double* v =
      L.parts + (static_cast<size_t>(step & 1) * (3 * kMaxSignals) + j) * L.blocks;
  double total = 0.0;
  for (uint32_t b = threadIdx.x; b < L.blocks; b += NT) total += v[b];
  total = block_sum<NT>(total, lds);
  if (threadIdx.x == 0) {
    const double volume =
        static_cast<double>(P.dx) * static_cast<double>(P.dy) * static_cast<double>(P.dz);
    const uint32_t s = j / 3, what = j % 3;
    double* row = L.ledger + (static_cast<size_t>(step) * S + s) * 4;
    row[1 + what] = what == 0 ? total * volume : static_cast<double>(dt) * total * volume;
  }
  (void)bcast;
}

// One FP64 step in face-flux form for S signals, grid-stride over sites.
template <int S>
__global__ __launch_bounds__(kAdrBlock) void k_adr_f64(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ phi, const float* __restrict__ source,
    const float* __restrict__ loss, const double* __restrict__ in, double* __restrict__ out,
    float dt, uint32_t* flags, uint32_t step, LedgerParts L) {
  __shared__ double s_red[kAdrBlock / kWave];
  __shared__ double s_b;
  if (flags[0] != 0) return;
  if (step > 0 && blockIdx.x < 3 * S) {
    fold_ledger<S, kAdrBlock>(P, L, step - 1, dt, blockIdx.x, s_red, &s_b);
  }
  const size_t n = static_cast<size_t>(P.nx) * P.ny;
  double acc[3 * S];
#pragma unroll
  for (int k = 0; k < 3 * S; ++k) acc[k] = 0.0;
  bool bad = false;
  const double ih[2] = {1.0 / static_cast<double>(P.dx), 1.0 / static_cast<double>(P.dy)};
  const size_t grid_stride = static_cast<size_t>(gridDim.x) * kAdrBlock;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kAdrBlock + threadIdx.x; i < n;
       i += grid_stride) {
    const uint32_t x = static_cast<uint32_t>(i / P.ny), y = static_cast<uint32_t>(i % P.ny);
    if (solid[i] != 0) {
#pragma unroll
      for (int s = 0; s < S; ++s) out[s * n + i] = in[s * n + i];
      continue;
    }
    const uint32_t coord[2] = {x, y};
    const uint32_t dims[2] = {P.nx, P.ny};
    const size_t stride[2] = {P.ny, 1};
    SiteGeometry g = geometry(P, solid, xf, yf, x, y, 0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (P.has_field == 0 && s > 0) g = geometry(P, solid, xf, yf, x, y, s);
      const double* lv = in + s * n;
      const double c = lv[i];
      const double D = static_cast<double>(P.diffusion[s]);
      double rate = 0.0, supply = 0.0;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        if (dims[a] == 1) continue;
        double fl = 0.0, fu = 0.0;
        const double ul = static_cast<double>(g.ulo[a]), uu = static_cast<double>(g.uup[a]);
        // lower face
        if (coord[a] == 0) {
          if (P.kind[2 * a] == MSIM_FIXED) {
            const double cb = static_cast<double>(P.boundary[s][2 * a]);
            fl = (ul >= 0.0 ? ul * cb : ul * c) - D * (c - cb) * ih[a];
            supply += fl * ih[a];
          }
        } else if (!g.clo[a]) {
          const double cl = lv[i - stride[a]];
          fl = (ul >= 0.0 ? ul * cl : ul * c) - D * (c - cl) * ih[a];
        }
        // upper face
        if (coord[a] + 1 == dims[a]) {
          if (P.kind[2 * a + 1] == MSIM_FIXED) {
            const double cb = static_cast<double>(P.boundary[s][2 * a + 1]);
            fu = (uu >= 0.0 ? uu * c : uu * cb) - D * (cb - c) * ih[a];
            supply -= fu * ih[a];
          }
        } else if (!g.cup[a]) {
          const double cu = lv[i + stride[a]];
          fu = (uu >= 0.0 ? uu * c : uu * cu) - D * (cu - c) * ih[a];
        }
        rate -= (fu - fl) * ih[a];
      }
      double lossr = 0.0;
      const size_t index = s * n + i;
      if (P.has_affine != 0) {
        const double src = static_cast<double>(source[index]);
        rate += src;
        supply += src;
        lossr += static_cast<double>(loss[index]);
      }
      if (P.has_monod != 0) {
        lossr += static_cast<double>(P.vmax[s]) * static_cast<double>(phi[i]) /
                 (static_cast<double>(P.monod_k[s]) + c);
      }
      rate -= lossr * c;
      const double next = c + static_cast<double>(dt) * rate;
      if (!isfinite(next) || next < 0.0) bad = true;
      out[index] = next;
      acc[3 * s + 0] += next;
      acc[3 * s + 1] += supply;
      acc[3 * s + 2] += lossr * c;
    }
  }
  if (bad) {
    atomicOr(&flags[0], kFlagNegative);
    atomicMin(&flags[1], step);
  }
  // Solid sites hold their (zero) level; the mass counts only what moved.
#pragma unroll
  for (int k = 0; k < 3 * S; ++k) {
    const double part = block_sum<kAdrBlock>(acc[k], s_red);
    if (threadIdx.x == 0) {
      L.parts[(static_cast<size_t>(step & 1) * (3 * kMaxSignals) + k) * L.blocks + blockIdx.x] =
          part;
    }
  }
}

// Adds the final step's ledger partials.
template <int S>
__global__ __launch_bounds__(kAdrBlock) void k_adr_ledger_tail(AdrParams P, LedgerParts L,
                                                               uint32_t step, float dt,
                                                               const uint32_t* flags) {
  __shared__ double s_red[kAdrBlock / kWave];
  __shared__ double s_b;
  if (flags[0] != 0) return;
  fold_ledger<S, kAdrBlock>(P, L, step, dt, blockIdx.x, s_red, &s_b);
}

// Sum of levels per signal (FP64), and the greatest outflow rate over fluid
// sites (SignalGrid::transport_cache, signals.cpp:870-896) for the explicit
// bound.
template <typename T>
__global__ __launch_bounds__(kAdrBlock) void k_adr_survey(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const T* __restrict__ levels, double* partials, CgState* st,
    double* mass) {
  __shared__ double s_red[kAdrBlock / kWave];
  __shared__ bool s_last;
  const size_t n = static_cast<size_t>(P.nx) * P.ny;
  double acc[kMaxSignals + 1];
#pragma unroll
  for (int k = 0; k <= kMaxSignals; ++k) acc[k] = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kAdrBlock;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kAdrBlock + threadIdx.x; i < n; i += stride) {
    for (uint32_t s = 0; s < P.signals && s < kMaxSignals; ++s)
      acc[s] += static_cast<double>(levels[s * n + i]);
    if (P.has_field != 0 && solid[i] == 0) {
      const uint32_t x = static_cast<uint32_t>(i / P.ny), y = static_cast<uint32_t>(i % P.ny);
      double outflow = 0.0;
      if (P.nx > 1) {
        outflow += (fmax(static_cast<double>(xf[i + P.ny]), 0.0) -
                    fmin(static_cast<double>(xf[i]), 0.0)) *
                   (1.0 / static_cast<double>(P.dx));
      }
      if (P.ny > 1) {
        const size_t yfi = static_cast<size_t>(x) * (P.ny + 1) + y;
        outflow += (fmax(static_cast<double>(yf[yfi + 1]), 0.0) -
                    fmin(static_cast<double>(yf[yfi]), 0.0)) *
                   (1.0 / static_cast<double>(P.dy));
      }
      acc[kMaxSignals] = fmax(acc[kMaxSignals], outflow);
    }
  }
  double part[kMaxSignals + 1];
#pragma unroll
  for (int k = 0; k < kMaxSignals; ++k) part[k] = block_sum<kAdrBlock>(acc[k], s_red);
  part[kMaxSignals] = block_max<kAdrBlock>(acc[kMaxSignals], s_red);
  if (publish_partials<kMaxSignals + 1>(part, partials, &st->ticket[kTicketMisc], blockIdx.x,
                                        gridDim.x, &s_last)) {
    double total[kMaxSignals];
    sum_partials<kAdrBlock, kMaxSignals>(partials, gridDim.x, total, s_red);
    double outflow[1];
    max_partials<kAdrBlock, 1>(partials + static_cast<size_t>(kMaxSignals) * gridDim.x, gridDim.x,
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

uint32_t site_blocks(size_t n) { return static_cast<uint32_t>((n + kAdrBlock - 1) / kAdrBlock); }

}  // namespace

void adr_upload(Context& ctx, const msim_signal_grid& grid, const void* levels,
                uint32_t precision, bool use_flow_field) {
  auto& a = ctx.adr;
  auto& P = a.params;
  const size_t n = static_cast<size_t>(grid.grid.nx) * grid.grid.ny;
  const uint32_t S = grid.signal_count;
  const bool same_spec = a.uploaded && grid.revision != 0 && grid.revision == a.revision &&
                         P.nx == grid.grid.nx && P.ny == grid.grid.ny && P.signals == S &&
                         !use_flow_field && !a.flow_field;
  std::memset(&P, 0, sizeof(P));
  P.nx = grid.grid.nx;
  P.ny = grid.grid.ny;
  P.signals = S;
  P.dx = grid.grid.dx;
  P.dy = grid.grid.dy;
  P.dz = grid.grid.dz;
  for (int f = 0; f < 6; ++f) P.kind[f] = grid.grid.boundary[f];
  for (uint32_t s = 0; s < S; ++s) {
    P.diffusion[s] = grid.diffusion[s];
    for (int f = 0; f < 6; ++f)
      P.boundary[s][f] = grid.boundary_values != nullptr ? grid.boundary_values[6 * s + f] : 0.0f;
    if (grid.monod_vmax != nullptr) {
      P.vmax[s] = grid.monod_vmax[s];
      P.monod_k[s] = grid.monod_k[s];
    }
  }
  P.has_field = (use_flow_field || grid.x_faces != nullptr) ? 1 : 0;
  P.has_affine = grid.loss_rates != nullptr ? 1 : 0;
  P.has_monod = grid.phi != nullptr ? 1 : 0;
  a.precision = precision;
  const size_t bytes = S * n * (precision == 64 ? sizeof(double) : sizeof(float));
  a.levels_a.reserve(bytes);
  a.levels_b.reserve(bytes);
  a.current = 0;
  a.flags.reserve(2);
  a.state.reserve(1);
  a.mass.reserve(kMaxSignals);
  a.partials.reserve(static_cast<size_t>(3 * kMaxSignals) * 2 * 1024 +
                     static_cast<size_t>(kMaxSignals + 1) * 2048);
  if (!same_spec) {
    a.solid.reserve(n);
    if (grid.grid.obstacles != nullptr) {
      check(hipMemcpyAsync(a.solid.data(), grid.grid.obstacles, n, hipMemcpyHostToDevice,
                           ctx.stream),
            "upload mask");
    } else {
      check(hipMemsetAsync(a.solid.data(), 0, n, ctx.stream), "clear mask");
    }
    if (use_flow_field) {
      a.xf.release();
      a.yf.release();
    } else if (grid.x_faces != nullptr) {
      a.xf.reserve(static_cast<size_t>(P.nx + 1) * P.ny);
      a.yf.reserve(static_cast<size_t>(P.nx) * (P.ny + 1));
      check(hipMemcpyAsync(a.xf.data(), grid.x_faces, a.xf.size() * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload x faces");
      check(hipMemcpyAsync(a.yf.data(), grid.y_faces, a.yf.size() * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload y faces");
    }
    if (grid.phi != nullptr) {
      a.phi.reserve(n);
      check(hipMemcpyAsync(a.phi.data(), grid.phi, n * sizeof(float), hipMemcpyHostToDevice,
                           ctx.stream),
            "upload phi");
    }
    if (grid.loss_rates != nullptr) {
      a.source.reserve(S * n);
      a.loss.reserve(S * n);
      check(hipMemcpyAsync(a.loss.data(), grid.loss_rates, S * n * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload loss");
      if (grid.source_rates != nullptr) {
        check(hipMemcpyAsync(a.source.data(), grid.source_rates, S * n * sizeof(float),
                             hipMemcpyHostToDevice, ctx.stream),
              "upload source");
      } else {
        check(hipMemsetAsync(a.source.data(), 0, S * n * sizeof(float), ctx.stream),
              "clear source");
      }
    }
  }
  a.flow_field = use_flow_field;
  a.revision = grid.revision;
  check(hipMemcpyAsync(a.levels_a.data(), levels, bytes, hipMemcpyHostToDevice, ctx.stream),
        "upload levels");
  check(hipMemsetAsync(a.state.data(), 0, sizeof(CgState), ctx.stream), "clear state");
  const float* xf = use_flow_field ? ctx.flow.faces.data() : a.xf.data();
  const float* yf = use_flow_field ? ctx.flow.faces.data() + ctx.flow.x_faces() : a.yf.data();
  const uint32_t blocks = std::min<uint32_t>(site_blocks(n), 2048);
  if (precision == 64) {
    k_adr_survey<double><<<blocks, kAdrBlock, 0, ctx.stream>>>(
        P, a.solid.data(), xf, yf, reinterpret_cast<const double*>(a.levels_a.data()),
        a.partials.data(), a.state.data(), a.mass.data());
  } else {
    k_adr_survey<float><<<blocks, kAdrBlock, 0, ctx.stream>>>(
        P, a.solid.data(), xf, yf, reinterpret_cast<const float*>(a.levels_a.data()),
        a.partials.data(), a.state.data(), a.mass.data());
  }
  check(hipGetLastError(), "launch k_adr_survey");
  CgState host{};
  check(hipMemcpyAsync(&host, a.state.data(), sizeof(CgState), hipMemcpyDeviceToHost, ctx.stream),
        "read survey");
  check(hipStreamSynchronize(ctx.stream), "survey");
  a.greatest_outflow = P.has_field != 0 ? host.peak : 0.0;
  check(hipMemcpyAsync(a.host_mass, a.mass.data(), sizeof(a.host_mass), hipMemcpyDeviceToHost,
                       ctx.stream),
        "read mass");
  check(hipStreamSynchronize(ctx.stream), "read mass");
  a.uploaded = true;
}

template <int S>
void launch_steps(Context& ctx, float dt, uint32_t steps, const float* xf, const float* yf) {
  auto& a = ctx.adr;
  const auto& P = a.params;
  const size_t n = a.sites();
  const uint32_t blocks = site_blocks(n);
  const uint32_t ledger_blocks = std::min<uint32_t>(blocks, 1024);
  LedgerParts L{a.partials.data(), a.ledger.data(), ledger_blocks};
  for (uint32_t k = 0; k < steps; ++k) {
    uint8_t* src = a.current == 0 ? a.levels_a.data() : a.levels_b.data();
    uint8_t* dst = a.current == 0 ? a.levels_b.data() : a.levels_a.data();
    if (a.precision == 64) {
      k_adr_f64<S><<<ledger_blocks, kAdrBlock, 0, ctx.stream>>>(
          P, a.solid.data(), xf, yf, a.phi.data(), a.source.data(), a.loss.data(),
          reinterpret_cast<const double*>(src), reinterpret_cast<double*>(dst), dt,
          a.flags.data(), k, L);
    } else {
      k_adr_f32<S><<<blocks, kAdrBlock, 0, ctx.stream>>>(
          P, a.solid.data(), xf, yf, a.phi.data(), a.source.data(), a.loss.data(),
          reinterpret_cast<const float*>(src), reinterpret_cast<float*>(dst), dt, a.flags.data(),
          k);
    }
    check(hipGetLastError(), "launch transport step");
    a.current ^= 1;
  }
  if (a.precision == 64 && steps > 0) {
    k_adr_ledger_tail<S><<<3 * S, kAdrBlock, 0, ctx.stream>>>(P, L, steps - 1, dt,
                                                             a.flags.data());
    check(hipGetLastError(), "launch ledger tail");
  }
}

void adr_run(Context& ctx, float dt, uint32_t steps, double* ledger) {
  auto& a = ctx.adr;
  const auto& P = a.params;
  const float* xf = a.flow_field ? ctx.flow.faces.data() : a.xf.data();
  const float* yf = a.flow_field ? ctx.flow.faces.data() + ctx.flow.x_faces() : a.yf.data();
  const uint32_t init_flags[2] = {0u, 0xffffffffu};
  check(hipMemcpyAsync(a.flags.data(), init_flags, sizeof(init_flags), hipMemcpyHostToDevice,
                       ctx.stream),
        "clear flags");
  if (a.precision == 64) a.ledger.reserve(static_cast<size_t>(steps) * P.signals * 4 + 1);
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  switch (P.signals) {
    case 1: launch_steps<1>(ctx, dt, steps, xf, yf); break;
    case 2: launch_steps<2>(ctx, dt, steps, xf, yf); break;
    case 3: launch_steps<3>(ctx, dt, steps, xf, yf); break;
    case 4: launch_steps<4>(ctx, dt, steps, xf, yf); break;
    case 5: launch_steps<5>(ctx, dt, steps, xf, yf); break;
    case 6: launch_steps<6>(ctx, dt, steps, xf, yf); break;
    case 7: launch_steps<7>(ctx, dt, steps, xf, yf); break;
    default: launch_steps<8>(ctx, dt, steps, xf, yf); break;
  }
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  uint32_t flags[2] = {0, 0};
  check(hipMemcpyAsync(flags, a.flags.data(), sizeof(flags), hipMemcpyDeviceToHost, ctx.stream),
        "read flags");
  std::vector<double> rows;
  if (a.precision == 64 && steps > 0) {
    rows.resize(static_cast<size_t>(steps) * P.signals * 4);
    check(hipMemcpyAsync(rows.data(), a.ledger.data(), rows.size() * sizeof(double),
                         hipMemcpyDeviceToHost, ctx.stream),
          "read ledger");
  }
  check(hipStreamSynchronize(ctx.stream), "transport steps");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  a.seconds = ms * 1.0e-3;
  if (flags[0] != 0) {
    // Levels stay those before the failed step: step k read buffer (k even ?
    // the starting one : the other).
    const uint32_t failed = flags[1];
    a.current = static_cast<int>((a.current ^ (steps & 1u)) ^ (failed & 1u));
    throw Error(MSIM_DOMAIN_ERROR,
                "signal grid update produced a non-finite or negative concentration (step " +
                    std::to_string(failed) + ")");
  }
  if (!rows.empty()) {
    // Mass before each step: the mass after the previous one.
    for (uint32_t k = 0; k < steps; ++k) {
      for (uint32_t s = 0; s < P.signals; ++s) {
        double* row = rows.data() + (static_cast<size_t>(k) * P.signals + s) * 4;
        row[0] = a.host_mass[s];
        a.host_mass[s] = row[1];
      }
    }
    if (ledger != nullptr) std::memcpy(ledger, rows.data(), rows.size() * sizeof(double));
  }
}

void adr_download(Context& ctx, void* levels) {
  auto& a = ctx.adr;
  const size_t bytes =
      a.params.signals * a.sites() * (a.precision == 64 ? sizeof(double) : sizeof(float));
  const uint8_t* src = a.current == 0 ? a.levels_a.data() : a.levels_b.data();
  check(hipMemcpyAsync(levels, src, bytes, hipMemcpyDeviceToHost, ctx.stream), "download levels");
  check(hipStreamSynchronize(ctx.stream), "download levels");
}

}  // namespace msim
