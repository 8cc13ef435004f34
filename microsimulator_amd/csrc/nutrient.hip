// The nutrient loop on the device (msim_hip_nutrient_run): the steps of the
// reference's nutrient study (scripts/run_nutrient_benchmarks.py:96-120) with
// every array kept in HBM. Per step: the mobility of the occupancy by the
// Brinkman closure, the operator and its multilevel hierarchy rebuilt, the
// pressure solve and velocity, the explicit bound of the new field, one FP64
// transport step with Monod uptake and its ledger row, and the occupancy grown
// by the uptake. The host only reads a few scalars per step (the solve's
// stopping tests, the bound's greatest outflow and largest occupancy).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kNutrientBlock = 256;

// _darcy_coefficients (python/src/microsimulator/flow.py:361-384): the
// Kozeny-Carman drag C f^2 / (1 - f)^3 of the capped occupancy and
// base / (1 + base * drag), in FP64, zero at solid sites.
__global__ __launch_bounds__(kNutrientBlock) void k_nutrient_mobility(
    const float* __restrict__ base, const float* __restrict__ phi,
    const uint8_t* __restrict__ solid, float* __restrict__ mobility, size_t n, double drag_c,
    double cap) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kNutrientBlock + threadIdx.x;
  if (i >= n) return;
  if (solid[i] != 0) {
    mobility[i] = 0.0f;
    return;
  }
  const double f = fmin(static_cast<double>(phi[i]), cap);
  const double drag = drag_c * f * f / ((1.0 - f) * (1.0 - f) * (1.0 - f));
  const double b = static_cast<double>(base[i]);
  mobility[i] = static_cast<float>(b / (1.0 + b * drag));
}

// phi += yield * dt * sum_s vmax_s phi c_s / (K_s + c_s), with c the levels the
// step started from; per block the largest occupancy after.
template <int S>
__global__ __launch_bounds__(kNutrientBlock) void k_nutrient_grow(
    AdrParams P, const double* __restrict__ levels, const uint8_t* __restrict__ solid,
    float* __restrict__ phi, size_t n, double yield_dt, double* __restrict__ maxima) {
  __shared__ double s_red[kNutrientBlock / kWave];
  const size_t i = static_cast<size_t>(blockIdx.x) * kNutrientBlock + threadIdx.x;
  double top = 0.0;
  if (i < n && solid[i] == 0) {
    const double f = static_cast<double>(phi[i]);
    double uptake = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double c = levels[s * n + i];
      uptake += static_cast<double>(P.vmax[s]) * f * c / (static_cast<double>(P.monod_k[s]) + c);
    }
    const float next = static_cast<float>(f + yield_dt * uptake);
    phi[i] = next;
    top = static_cast<double>(next);
  }
  top = block_max<kNutrientBlock>(top, s_red);
  if (threadIdx.x == 0) maxima[blockIdx.x] = top;
}

template <int S>
void launch_grow(Context& ctx, const double* levels, double yield_dt, double* maxima) {
  auto& a = ctx.adr;
  const size_t n = a.sites();
  const uint32_t blocks = static_cast<uint32_t>((n + kNutrientBlock - 1) / kNutrientBlock);
  k_nutrient_grow<S><<<blocks, kNutrientBlock, 0, ctx.stream>>>(
      a.params, levels, a.solid.data(), a.phi.data(), n, yield_dt, maxima);
}

}  // namespace

void nutrient_run(Context& ctx, const msim_nutrient_params& params, uint32_t steps,
                  double* ledger, msim_nutrient_step* per_step) {
  auto& f = ctx.flow;
  auto& a = ctx.adr;
  const auto& P = a.params;
  const size_t n = f.sites();
  const uint32_t S = P.signals;
  const uint32_t blocks = static_cast<uint32_t>((n + kNutrientBlock - 1) / kNutrientBlock);
  f.base_mobility.reserve(n);
  check(hipMemcpyAsync(f.base_mobility.data(), params.base_mobility, n * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload base mobility");
  DeviceBuffer<double> block_maxima;
  block_maxima.reserve(blocks);
  std::vector<double> maxima(blocks);
  std::vector<double> row(static_cast<size_t>(S) * 4);
  // The solve and the transport step time themselves with the context's
  // events; the step's own three are these.
  struct Events {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x != nullptr) static_cast<void>(hipEventDestroy(x));
    }
  } ev;
  for (hipEvent_t& e : ev.e) check(hipEventCreate(&e), "hipEventCreate");
  hipEvent_t begin = ev.e[0], solved = ev.e[1], end = ev.e[2];
  for (uint32_t k = 0; k < steps; ++k) {
    check(hipEventRecord(begin, ctx.stream), "event");
    // 1-2. The mobility of the current occupancy, the operator, the solve.
    k_nutrient_mobility<<<blocks, kNutrientBlock, 0, ctx.stream>>>(
        f.base_mobility.data(), a.phi.data(), f.solid.data(), f.mobility.data(), n,
        static_cast<double>(params.drag_coefficient),
        static_cast<double>(params.max_volume_fraction));
    check(hipGetLastError(), "launch k_nutrient_mobility");
    launch_assembly(ctx.stream, f.solid.data(), f.mobility.data(), f.cond.data(), f.diag.data(),
                    f.rhs.data(), f.bc.data(), f.cxl.data(), f.cyl.data(), f.nx, f.ny, f.dx, f.dy,
                    f.dz, f.edges());
    f.mg.ready = false;
    msim_flow_report report{};
    flow_solve(ctx, params.flow, report);
    check(hipEventRecord(solved, ctx.stream), "event");
    // The transport reads the field this solve left, in place.
    a.flow_generation = f.generation;
    // 3. The explicit bound of the new field (signals.cpp:1066-1102), Monod's
    // loss term from the largest occupancy, then the step.
    const double outflow = adr_refresh_outflow(ctx);
    double max_loss[kMaxSignals] = {};
    for (uint32_t s = 0; s < S; ++s) {
      max_loss[s] = static_cast<double>(P.vmax[s]) * a.phi_max / static_cast<double>(P.monod_k[s]);
    }
    validate_explicit_bound(P, MSIM_FORWARD_EULER, outflow, max_loss, params.dt);
    double* ledger_row = ledger != nullptr ? ledger + static_cast<size_t>(k) * S * 4 : row.data();
    adr_run(ctx, params.dt, 1, ledger_row);
    // 4. The occupancy grows by the uptake at the levels the step started from.
    const double* started = reinterpret_cast<const double*>(
        a.current == 0 ? a.levels_b.data() : a.levels_a.data());
    const double yield_dt = static_cast<double>(params.yield) * static_cast<double>(params.dt);
    switch (S) {
      case 1: launch_grow<1>(ctx, started, yield_dt, block_maxima.data()); break;
      case 2: launch_grow<2>(ctx, started, yield_dt, block_maxima.data()); break;
      case 3: launch_grow<3>(ctx, started, yield_dt, block_maxima.data()); break;
      case 4: launch_grow<4>(ctx, started, yield_dt, block_maxima.data()); break;
      case 5: launch_grow<5>(ctx, started, yield_dt, block_maxima.data()); break;
      case 6: launch_grow<6>(ctx, started, yield_dt, block_maxima.data()); break;
      case 7: launch_grow<7>(ctx, started, yield_dt, block_maxima.data()); break;
      default: launch_grow<8>(ctx, started, yield_dt, block_maxima.data()); break;
    }
    check(hipGetLastError(), "launch k_nutrient_grow");
    check(hipEventRecord(end, ctx.stream), "event");
    check(hipMemcpyAsync(maxima.data(), block_maxima.data(), blocks * sizeof(double),
                         hipMemcpyDeviceToHost, ctx.stream),
          "read occupancy maxima");
    check(hipStreamSynchronize(ctx.stream), "nutrient step");
    a.phi_max = *std::max_element(maxima.begin(), maxima.end());
    if (per_step != nullptr) {
      msim_nutrient_step& out = per_step[k];
      out.iterations = report.iterations;
      out.fell_back = report.fell_back;
      out.true_relative_residual = report.true_relative_residual;
      out.max_speed = report.max_speed;
      out.max_phi = static_cast<float>(a.phi_max);
      double worst = 0.0;
      for (uint32_t s = 0; s < S; ++s) {
        const double* r = ledger_row + static_cast<size_t>(s) * 4;
        const double scale = std::max(std::abs(r[1]), 1e-300);
        worst = std::max(worst, std::abs(r[1] - r[0] - r[2] + r[3]) / scale);
      }
      out.mass_balance = worst;
      float ms = 0.0f, solve_ms = 0.0f;
      check(hipEventElapsedTime(&ms, begin, end), "event time");
      check(hipEventElapsedTime(&solve_ms, begin, solved), "event time");
      out.seconds = ms * 1.0e-3;
      out.solve_seconds = solve_ms * 1.0e-3;
    }
  }
}

}  // namespace msim
