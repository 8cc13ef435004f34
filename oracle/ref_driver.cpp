// C entry points over the reference's own CPU path, compiled against the
// reference sources where they lie (see oracle/Makefile, target `ref`) into
// oracle/_ref/libmsim_ref.so. Test infrastructure only: tests/ use it to pin
// the C oracle, tests/golden/make_golden.py writes fixtures from it, and
// bench.py may time it as the CPU baseline. Nothing in the product loads it.
//
// Reference functions called:
//   solve_depth_averaged_flow_cpu        cpp/cpu/cpu_flow.cpp:338-351
//   detail::reduce_shallow_flow          cpp/core/flow_system.hpp:1255-1266
//   detail::axis_flow_boundary           cpp/core/flow_system.hpp:186-201
//   detail::DepthAveragedFlowSystem      cpp/core/flow_system.hpp:520-748
//   detail::solve_depth_averaged_pressure cpp/cpu/cpu_flow.cpp:172-175
//   SignalGrid + advance_signal_grid_cpu cpp/core/signals.cpp:1377-1389
//   solve_depth_averaged_port_flow_cpu   cpp/cpu/cpu_flow.cpp:353-375
//   WorldState + CoupledRatePlan + advance_coupled_cpu cpp/cpu/cpu_coupled.cpp:125-231
//   set_cpu_thread_count                 cpp/include/microsimulator/threads.hpp:19
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <vector>

#include "core/flow_system.hpp"
#include "microsimulator/coupled_rates.hpp"
#include "microsimulator/flow.hpp"
#include "microsimulator/signals.hpp"
#include "microsimulator/threads.hpp"
#include "microsimulator/world_state.hpp"

namespace ms = microsimulator;

namespace {

// Status numbering shared with oracle/msim_oracle.h.
constexpr int kOk = 0, kInvalid = 1, kRuntime = 2, kOverflow = 4, kDomain = 5;

int report_error(const std::exception& error, int code, char* err, std::size_t errlen) {
  if (err != nullptr && errlen > 0) std::snprintf(err, errlen, "%s", error.what());
  return code;
}

template <typename Body>
int guarded(char* err, std::size_t errlen, Body&& body) {
  try {
    body();
    return kOk;
  } catch (const std::invalid_argument& e) {
    return report_error(e, kInvalid, err, errlen);
  } catch (const std::overflow_error& e) {
    return report_error(e, kOverflow, err, errlen);
  } catch (const std::domain_error& e) {
    return report_error(e, kDomain, err, errlen);
  } catch (const std::logic_error& e) {
    // out_of_range (a sample outside the grid) and the other argument errors.
    return report_error(e, kInvalid, err, errlen);
  } catch (const std::exception& e) {
    return report_error(e, kRuntime, err, errlen);
  }
}

ms::GridBoundary boundary_of(std::uint8_t kind, std::uint32_t signals, const float* values) {
  ms::GridBoundary boundary;
  boundary.kind = static_cast<ms::GridBoundaryKind>(kind);
  if (boundary.kind == ms::GridBoundaryKind::fixed) {
    boundary.values.assign(signals, 0.0F);
    if (values != nullptr) {
      for (std::uint32_t s = 0; s < signals; ++s) boundary.values[s] = values[s];
    }
  }
  return boundary;
}

// boundary_values: 6*S, face-major as in mo_signal_grid (face f, signal s at
// [6*s + f]); may be null for zero reservoirs.
ms::SignalGridSpec make_spec(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx,
                             float dy, float dz, const std::uint8_t* obstacles,
                             const std::uint8_t* kinds, std::uint32_t signals,
                             const float* boundary_values) {
  ms::SignalGridSpec spec;
  spec.signal_count = signals;
  spec.shape = {.x = nx, .y = ny, .z = nz};
  spec.spacing = {dx, dy, dz};
  spec.diffusion.assign(signals, 0.0F);
  spec.advection.assign(signals, ms::Vec3{});
  if (obstacles != nullptr) {
    spec.obstacles.assign(obstacles, obstacles + static_cast<std::size_t>(nx) * ny * nz);
  }
  ms::GridBoundary* faces[6] = {&spec.x_lower, &spec.x_upper, &spec.y_lower,
                                &spec.y_upper, &spec.z_lower, &spec.z_upper};
  for (int f = 0; f < 6; ++f) {
    std::vector<float> values(signals, 0.0F);
    if (boundary_values != nullptr) {
      for (std::uint32_t s = 0; s < signals; ++s) values[s] = boundary_values[6 * s + f];
    }
    *faces[f] = boundary_of(kinds[f], signals, values.data());
  }
  return spec;
}

// A transport spec from the flat arguments the transport entry points take:
// patches as per patch (face, begin[0], begin[1], end[0], end[1]) with
// `signals` values each; advection (x, y, z) per signal or null.
ms::SignalGridSpec signal_spec(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx,
                               float dy, float dz, const std::uint8_t* obstacles,
                               const std::uint8_t* kinds, std::uint32_t signals,
                               const float* diffusion, const float* advection,
                               const float* boundary_values, const float* x_faces,
                               const float* y_faces, const float* z_faces,
                               const float* source_rates, const float* loss_rates,
                               std::uint32_t scheme, std::uint32_t patch_count,
                               const std::uint32_t* patch_geometry, const float* patch_values,
                               std::uint32_t integration, std::uint32_t max_iterations,
                               float absolute_tolerance, float relative_tolerance) {
  auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, signals, boundary_values);
  // Reservoir patches: per patch (face, begin[0], begin[1], end[0], end[1])
  // and `signals` values.
  ms::GridBoundary* faces[6] = {&spec.x_lower, &spec.x_upper, &spec.y_lower,
                                &spec.y_upper, &spec.z_lower, &spec.z_upper};
  for (std::uint32_t k = 0; k < patch_count; ++k) {
    const std::uint32_t* g = patch_geometry + 5 * k;
    ms::GridBoundaryPatch patch;
    patch.region.begin = {g[1], g[2]};
    patch.region.end = {g[3], g[4]};
    patch.values.assign(patch_values + k * signals, patch_values + (k + 1) * signals);
    faces[g[0]]->patches.push_back(std::move(patch));
  }
  const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
  spec.diffusion.assign(diffusion, diffusion + signals);
  if (advection != nullptr) {
    for (std::uint32_t k = 0; k < signals; ++k) {
      spec.advection[k] = ms::Vec3{advection[3 * k], advection[3 * k + 1], advection[3 * k + 2]};
    }
  }
  spec.advection_scheme =
      scheme == 1 ? ms::SignalAdvectionKind::van_leer : ms::SignalAdvectionKind::upwind;
  spec.integration = static_cast<ms::SignalIntegrationKind>(integration);
  spec.solver.max_iterations = max_iterations;
  spec.solver.absolute_tolerance = absolute_tolerance;
  spec.solver.relative_tolerance = relative_tolerance;
  if (x_faces != nullptr) {
    ms::SignalGridVelocityField field;
    field.x_faces.assign(x_faces, x_faces + (nx + 1) * static_cast<std::size_t>(ny) * nz);
    field.y_faces.assign(y_faces, y_faces + static_cast<std::size_t>(nx) * (ny + 1) * nz);
    field.z_faces.assign(z_faces, z_faces + static_cast<std::size_t>(nx) * ny * (nz + 1));
    spec.velocity_field = std::move(field);
  }
  if (loss_rates != nullptr) {
    ms::SignalGridAffineReaction reaction;
    reaction.source_rates.assign(source_rates, source_rates + signals * n);
    reaction.loss_rates.assign(loss_rates, loss_rates + signals * n);
    spec.reaction = std::move(reaction);
  }
  return spec;
}

struct FlowHandle {
  ms::SignalGridSpec spec;
  std::vector<float> mobility;
  std::unique_ptr<ms::detail::ShallowFlowReduction> reduction;
  std::unique_ptr<ms::detail::FlowBoundary> boundary;
  std::unique_ptr<ms::detail::DepthAveragedFlowSystem> system;
};

}  // namespace

extern "C" {

void ref_set_threads(std::uint32_t count) {
  ms::set_cpu_thread_count(count);
}

std::uint32_t ref_threads(void) {
  return static_cast<std::uint32_t>(ms::cpu_thread_count());
}

// solve_depth_averaged_flow_cpu on a one-signal grid with the given boundary
// kinds. faces: x|y|z of the grid. The reference refuses tolerances below
// float epsilon.
int ref_solve_depth_flow(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx, float dy,
                         float dz, const std::uint8_t* obstacles, const std::uint8_t* kinds,
                         const float* mobility, std::uint32_t axis, float mean_inlet_speed,
                         float tolerance, std::uint32_t max_iterations, float* faces,
                         std::uint32_t* iterations, float* relative_residual, float* max_speed,
                         char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::span<const float> mob;
    if (mobility != nullptr) mob = std::span<const float>(mobility, n);
    ms::DepthAveragedFlowParameters parameters;
    parameters.mean_inlet_speed = mean_inlet_speed;
    parameters.axis = static_cast<ms::FlowAxis>(axis);
    parameters.relative_tolerance = tolerance;
    parameters.max_iterations = max_iterations;
    const auto result = ms::solve_depth_averaged_flow_cpu(spec, mob, parameters);
    std::memcpy(faces, result.field.x_faces.data(), result.field.x_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size(), result.field.y_faces.data(),
                result.field.y_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size() + result.field.y_faces.size(),
                result.field.z_faces.data(), result.field.z_faces.size() * sizeof(float));
    *iterations = result.report.iterations;
    *relative_residual = result.report.relative_residual;
    *max_speed = result.report.max_speed;
  });
}

// solve_resolved_flow_cpu (cpp/cpu/cpu_flow.cpp:463-467): the resolved
// Stokes-Brinkman field (x|y|z faces) and its report.
int ref_solve_resolved_flow(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx,
                            float dy, float dz, const std::uint8_t* obstacles,
                            const std::uint8_t* kinds, const float* drag, std::uint32_t axis,
                            float mean_inlet_speed, float tolerance, std::uint32_t max_outer,
                            float inner_tolerance, std::uint32_t max_inner, float* faces,
                            std::uint32_t* outer_iterations, std::uint64_t* inner_iterations,
                            float* report_values, std::uint32_t* min_gap, char* err,
                            std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::span<const float> d;
    if (drag != nullptr) d = std::span<const float>(drag, n);
    ms::ResolvedFlowParameters parameters;
    parameters.mean_inlet_speed = mean_inlet_speed;
    parameters.axis = static_cast<ms::FlowAxis>(axis);
    parameters.relative_tolerance = tolerance;
    parameters.max_outer_iterations = max_outer;
    parameters.inner_relative_tolerance = inner_tolerance;
    parameters.max_inner_iterations = max_inner;
    const auto result = ms::solve_resolved_flow_cpu(spec, d, parameters);
    std::memcpy(faces, result.field.x_faces.data(), result.field.x_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size(), result.field.y_faces.data(),
                result.field.y_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size() + result.field.y_faces.size(),
                result.field.z_faces.data(), result.field.z_faces.size() * sizeof(float));
    *outer_iterations = result.report.outer_iterations;
    *inner_iterations = result.report.inner_iterations;
    report_values[0] = result.report.relative_residual;
    report_values[1] = result.report.momentum_relative_residual;
    report_values[2] = result.report.divergence_rms;
    report_values[3] = result.report.mean_inlet_speed;
    report_values[4] = result.report.max_speed;
    *min_gap = result.report.min_gap_voxels;
  });
}

// The same with DepthAveragedFlowParameters::body_force: force_columns pairs
// (fx, fy), x-major, y-minor.
int ref_solve_depth_flow_forced(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx,
                                float dy, float dz, const std::uint8_t* obstacles,
                                const std::uint8_t* kinds, const float* mobility,
                                std::uint32_t axis, float mean_inlet_speed, float tolerance,
                                std::uint32_t max_iterations, const float* force,
                                std::size_t force_columns, float* faces,
                                std::uint32_t* iterations, float* relative_residual,
                                float* max_speed, char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::span<const float> mob;
    if (mobility != nullptr) mob = std::span<const float>(mobility, n);
    ms::DepthAveragedFlowParameters parameters;
    parameters.mean_inlet_speed = mean_inlet_speed;
    parameters.axis = static_cast<ms::FlowAxis>(axis);
    parameters.relative_tolerance = tolerance;
    parameters.max_iterations = max_iterations;
    for (std::size_t c = 0; c < force_columns; ++c) {
      parameters.body_force.push_back({force[2 * c], force[2 * c + 1]});
    }
    const auto result = ms::solve_depth_averaged_flow_cpu(spec, mob, parameters);
    std::memcpy(faces, result.field.x_faces.data(), result.field.x_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size(), result.field.y_faces.data(),
                result.field.y_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size() + result.field.y_faces.size(),
                result.field.z_faces.data(), result.field.z_faces.size() * sizeof(float));
    *iterations = result.report.iterations;
    *relative_residual = result.report.relative_residual;
    *max_speed = result.report.max_speed;
  });
}

// Builds the reduced, assembled depth-averaged system once (as the reference
// does before each solve) and keeps it for repeated solves.
void* ref_flow_create(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx, float dy,
                      float dz, const std::uint8_t* obstacles, const std::uint8_t* kinds,
                      const float* mobility, std::uint32_t axis, char* err, std::size_t errlen) {
  auto handle = std::make_unique<FlowHandle>();
  const int status = guarded(err, errlen, [&] {
    handle->spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    if (mobility != nullptr) handle->mobility.assign(mobility, mobility + n);
    handle->reduction = std::make_unique<ms::detail::ShallowFlowReduction>(
        ms::detail::reduce_shallow_flow(handle->spec, handle->mobility,
                                        static_cast<ms::FlowAxis>(axis)));
    handle->boundary = std::make_unique<ms::detail::FlowBoundary>(
        ms::detail::axis_flow_boundary(handle->reduction->reduced_layout(), axis));
    handle->system = std::make_unique<ms::detail::DepthAveragedFlowSystem>(
        handle->reduction->grid(), handle->reduction->conductance(), *handle->boundary);
  });
  return status == kOk ? handle.release() : nullptr;
}

void ref_flow_destroy(void* handle) {
  delete static_cast<FlowHandle*>(handle);
}

// The assembled per-site arrays of the reduced grid (nz = 1 reduced grid).
int ref_flow_arrays(void* opaque, float* conductance, float* diagonal, float* rhs,
                    float* boundary_conductance, float* face_conductance, float* face_pressure) {
  const auto& h = *static_cast<FlowHandle*>(opaque);
  const auto& s = *h.system;
  const auto& c = h.reduction->conductance();
  std::memcpy(conductance, c.data(), c.size() * sizeof(float));
  std::memcpy(diagonal, s.diagonal().data(), s.diagonal().size() * sizeof(float));
  std::memcpy(rhs, s.right_hand_side().data(), s.right_hand_side().size() * sizeof(float));
  std::memcpy(boundary_conductance, s.boundary_conductance().data(),
              s.boundary_conductance().size() * sizeof(float));
  std::memcpy(face_conductance, s.face_conductance().data(),
              s.face_conductance().size() * sizeof(float));
  std::memcpy(face_pressure, s.face_pressure().data(), s.face_pressure().size() * sizeof(float));
  return kOk;
}

// Jacobi-PCG on the kept system through detail::solve_depth_averaged_pressure
// (no tolerance floor). pressure: one double per reduced site. seconds: wall
// time of the solve alone. A solve that stops at max_iterations returns
// kRuntime with the reference's "did not converge" message.
int ref_flow_pressure(void* opaque, float tolerance, std::uint32_t max_iterations,
                      double* pressure, double* seconds, char* err, std::size_t errlen) {
  auto& h = *static_cast<FlowHandle*>(opaque);
  const auto start = std::chrono::steady_clock::now();
  const int status = guarded(err, errlen, [&] {
    const auto result =
        ms::detail::solve_depth_averaged_pressure(*h.system, tolerance, max_iterations);
    if (pressure != nullptr) std::memcpy(pressure, result.data(), result.size() * sizeof(double));
  });
  const auto stop = std::chrono::steady_clock::now();
  if (seconds != nullptr) *seconds = std::chrono::duration<double>(stop - start).count();
  return status;
}

// DepthAveragedFlowSystem::apply over the active sites, scattered back to the
// full reduced grid (zero at inactive sites).
int ref_flow_apply(void* opaque, const double* input, double* output) {
  const auto& h = *static_cast<FlowHandle*>(opaque);
  const auto& active = h.system->active();
  std::vector<double> in(active.size()), out;
  for (std::size_t i = 0; i < active.size(); ++i) in[i] = input[active.sites[i]];
  h.system->apply(in, out);
  const auto full = h.system->scatter(out);
  std::memcpy(output, full.data(), full.size() * sizeof(double));
  return kOk;
}

// `steps` steps of advance_signal_grid_cpu. levels: S*N in place. report
// (may be null): the last step's converged, iterations and residual RMS.
// velocity faces may be null (no field); source/loss may be null (no
// reaction). boundary_values: 6*S as in make_spec. seconds: time of the steps.
int ref_advance_signal_grid(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx,
                            float dy, float dz, const std::uint8_t* obstacles,
                            const std::uint8_t* kinds, std::uint32_t signals,
                            const float* diffusion, const float* advection,
                            const float* boundary_values,
                            const float* x_faces, const float* y_faces, const float* z_faces,
                            const float* source_rates, const float* loss_rates,
                            std::uint32_t scheme, std::uint32_t patch_count,
                            const std::uint32_t* patch_geometry, const float* patch_values,
                            std::uint32_t integration, std::uint32_t max_iterations,
                            float absolute_tolerance, float relative_tolerance, float dt,
                            std::uint32_t steps, float* levels, float* report, double* seconds,
                            char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = signal_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, signals, diffusion,
                                  advection, boundary_values, x_faces, y_faces, z_faces,
                                  source_rates, loss_rates, scheme, patch_count, patch_geometry,
                                  patch_values, integration, max_iterations, absolute_tolerance,
                                  relative_tolerance);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    ms::SignalGrid grid(spec, std::vector<float>(levels, levels + signals * n));
    const auto start = std::chrono::steady_clock::now();
    ms::SignalSolveReport last;
    for (std::uint32_t step = 0; step < steps; ++step) {
      last = ms::advance_signal_grid_cpu(grid, dt);
    }
    if (report != nullptr) {
      report[0] = last.converged ? 1.0F : 0.0F;
      report[1] = static_cast<float>(last.iterations);
      report[2] = last.residual_rms;
    }
    const auto stop = std::chrono::steady_clock::now();
    if (seconds != nullptr) *seconds = std::chrono::duration<double>(stop - start).count();
    const auto out = grid.levels();
    std::memcpy(levels, out.data(), out.size() * sizeof(float));
  });
}

// solve_depth_averaged_port_flow_cpu on a one-signal grid. Ports: per port
// (face, begin[0], begin[1], end[0], end[1], kind) in port_geometry and its
// value in port_values. faces: x|y|z of the grid; states: 3 floats per port
// (pressure, flux, mean speed); report: solves, iterations.
int ref_port_flow(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx, float dy,
                  float dz, const std::uint8_t* obstacles, const std::uint8_t* kinds,
                  const float* mobility, std::uint32_t port_count,
                  const std::uint32_t* port_geometry, const float* port_values, float tolerance,
                  std::uint32_t max_iterations, float* faces, float* states,
                  std::uint32_t* counts, float* relative_residual, float* max_speed,
                  double* seconds, char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::span<const float> mob;
    if (mobility != nullptr) mob = std::span<const float>(mobility, n);
    std::vector<ms::FlowPort> ports(port_count);
    for (std::uint32_t k = 0; k < port_count; ++k) {
      const std::uint32_t* g = port_geometry + 6 * k;
      ports[k].face = static_cast<ms::GridFace>(g[0]);
      ports[k].patch.begin = {g[1], g[2]};
      ports[k].patch.end = {g[3], g[4]};
      ports[k].kind = static_cast<ms::FlowPortKind>(g[5]);
      ports[k].value = port_values[k];
    }
    ms::DepthAveragedPortFlowParameters parameters;
    parameters.relative_tolerance = tolerance;
    parameters.max_iterations = max_iterations;
    const auto start = std::chrono::steady_clock::now();
    const auto result = ms::solve_depth_averaged_port_flow_cpu(spec, ports, mob, parameters);
    const auto stop = std::chrono::steady_clock::now();
    if (seconds != nullptr) *seconds = std::chrono::duration<double>(stop - start).count();
    std::memcpy(faces, result.field.x_faces.data(), result.field.x_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size(), result.field.y_faces.data(),
                result.field.y_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size() + result.field.y_faces.size(),
                result.field.z_faces.data(), result.field.z_faces.size() * sizeof(float));
    for (std::uint32_t k = 0; k < port_count; ++k) {
      states[3 * k] = result.ports[k].pressure;
      states[3 * k + 1] = result.ports[k].flux;
      states[3 * k + 2] = result.ports[k].mean_speed;
    }
    counts[0] = result.report.solves;
    counts[1] = result.report.iterations;
    *relative_residual = result.report.relative_residual;
    *max_speed = result.report.max_speed;
  });
}

// solve_resolved_port_flow_cpu (cpp/cpu/cpu_flow.cpp:474-479) on a one-signal
// grid, ports as for ref_port_flow. faces: x|y|z; states: 3 floats per port;
// counts: solves, outer iterations, inner iterations, minimum gap; report:
// relative residual, momentum relative residual, divergence RMS, max speed.
int ref_resolved_port_flow(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx,
                           float dy, float dz, const std::uint8_t* obstacles,
                           const std::uint8_t* kinds, const float* drag,
                           std::uint32_t port_count, const std::uint32_t* port_geometry,
                           const float* port_values, float tolerance, std::uint32_t max_outer,
                           float inner_tolerance, std::uint32_t max_inner, float* faces,
                           float* states, std::uint64_t* counts, float* report_values, char* err,
                           std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::span<const float> d;
    if (drag != nullptr) d = std::span<const float>(drag, n);
    std::vector<ms::FlowPort> ports(port_count);
    for (std::uint32_t k = 0; k < port_count; ++k) {
      const std::uint32_t* g = port_geometry + 6 * k;
      ports[k].face = static_cast<ms::GridFace>(g[0]);
      ports[k].patch.begin = {g[1], g[2]};
      ports[k].patch.end = {g[3], g[4]};
      ports[k].kind = static_cast<ms::FlowPortKind>(g[5]);
      ports[k].value = port_values[k];
    }
    ms::ResolvedPortFlowParameters parameters;
    parameters.relative_tolerance = tolerance;
    parameters.max_outer_iterations = max_outer;
    parameters.inner_relative_tolerance = inner_tolerance;
    parameters.max_inner_iterations = max_inner;
    const auto result = ms::solve_resolved_port_flow_cpu(spec, ports, d, parameters);
    std::memcpy(faces, result.field.x_faces.data(), result.field.x_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size(), result.field.y_faces.data(),
                result.field.y_faces.size() * sizeof(float));
    std::memcpy(faces + result.field.x_faces.size() + result.field.y_faces.size(),
                result.field.z_faces.data(), result.field.z_faces.size() * sizeof(float));
    for (std::uint32_t k = 0; k < port_count; ++k) {
      states[3 * k] = result.ports[k].pressure;
      states[3 * k + 1] = result.ports[k].flux;
      states[3 * k + 2] = result.ports[k].mean_speed;
    }
    counts[0] = result.report.solves;
    counts[1] = result.report.outer_iterations;
    counts[2] = result.report.inner_iterations;
    counts[3] = result.report.min_gap_voxels;
    report_values[0] = result.report.relative_residual;
    report_values[1] = result.report.momentum_relative_residual;
    report_values[2] = result.report.divergence_rms;
    report_values[3] = result.report.max_speed;
  });
}

// advance_coupled_cpu on a one-layer-or-more grid: the spec as
// ref_advance_signal_grid builds it (with `origin`), a WorldState of `cells`
// cells (positions, lengths, radii, growth rates, types, species levels
// cells*species) and a CoupledRatePlan (instructions as 4 uint32 (operation,
// first, second, third) plus a float value each) of plan_signals signals.
// levels and species in place;
// report: converged, iterations, residual RMS, seconds of advance_coupled_cpu.
int ref_advance_coupled(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx, float dy,
                        float dz, const float* origin, const std::uint8_t* obstacles,
                        const std::uint8_t* kinds, std::uint32_t signals, const float* diffusion,
                        const float* boundary_values, const float* x_faces, const float* y_faces,
                        const float* z_faces, const float* source_rates, const float* loss_rates,
                        std::uint32_t scheme, std::uint32_t integration,
                        std::uint32_t max_iterations, float absolute_tolerance,
                        float relative_tolerance, std::uint32_t cells, const float* position_x,
                        const float* position_y, const float* position_z, const float* lengths,
                        const float* radii, const float* growth_rates,
                        const std::int32_t* cell_types, std::uint32_t species_count,
                        float* species, std::uint32_t instruction_count,
                        const std::uint32_t* instruction_codes, const float* instruction_values,
                        const std::uint32_t* species_outputs, std::uint32_t plan_signals,
                        const std::uint32_t* signal_outputs, const float* previous_lengths,
                        float dt, float* levels, float* report,
                        char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, signals, boundary_values);
    spec.origin = {origin[0], origin[1], origin[2]};
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    spec.diffusion.assign(diffusion, diffusion + signals);
    spec.advection_scheme =
        scheme == 1 ? ms::SignalAdvectionKind::van_leer : ms::SignalAdvectionKind::upwind;
    spec.integration = static_cast<ms::SignalIntegrationKind>(integration);
    spec.solver.max_iterations = max_iterations;
    spec.solver.absolute_tolerance = absolute_tolerance;
    spec.solver.relative_tolerance = relative_tolerance;
    if (x_faces != nullptr) {
      ms::SignalGridVelocityField field;
      field.x_faces.assign(x_faces, x_faces + (nx + 1) * static_cast<std::size_t>(ny) * nz);
      field.y_faces.assign(y_faces, y_faces + static_cast<std::size_t>(nx) * (ny + 1) * nz);
      field.z_faces.assign(z_faces, z_faces + static_cast<std::size_t>(nx) * ny * (nz + 1));
      spec.velocity_field = std::move(field);
    }
    if (loss_rates != nullptr) {
      ms::SignalGridAffineReaction reaction;
      reaction.source_rates.assign(source_rates, source_rates + signals * n);
      reaction.loss_rates.assign(loss_rates, loss_rates + signals * n);
      spec.reaction = std::move(reaction);
    }
    ms::SignalGrid grid(spec, std::vector<float>(levels, levels + signals * n));
    ms::WorldState state(cells, species_count);
    for (std::uint32_t c = 0; c < cells; ++c) {
      ms::CellInit cell;
      cell.position = {position_x[c], position_y[c], position_z[c]};
      cell.length = lengths[c];
      cell.radius = radii[c];
      cell.growth_rate = growth_rates[c];
      cell.cell_type = cell_types[c];
      cell.species.assign(species + static_cast<std::size_t>(c) * species_count,
                          species + static_cast<std::size_t>(c + 1) * species_count);
      static_cast<void>(state.add_cell(cell));
    }
    std::vector<ms::RateInstruction> instructions(instruction_count);
    for (std::uint32_t k = 0; k < instruction_count; ++k) {
      instructions[k].operation = static_cast<ms::RateOp>(instruction_codes[4 * k]);
      instructions[k].first = instruction_codes[4 * k + 1];
      instructions[k].second = instruction_codes[4 * k + 2];
      instructions[k].third = instruction_codes[4 * k + 3];
      instructions[k].value = instruction_values[k];
    }
    const ms::CoupledRatePlan plan(
        species_count, plan_signals, std::move(instructions),
        std::vector<std::uint32_t>(species_outputs, species_outputs + species_count),
        std::vector<std::uint32_t>(signal_outputs, signal_outputs + plan_signals));
    const auto start = std::chrono::steady_clock::now();
    const auto result = ms::advance_coupled_cpu(
        state, grid, plan, std::span<const float>(previous_lengths, cells), dt);
    const auto stop = std::chrono::steady_clock::now();
    const auto out = grid.levels();
    std::memcpy(levels, out.data(), out.size() * sizeof(float));
    const auto held = state.species_state();
    std::memcpy(species, held.levels.data(), held.levels.size() * sizeof(float));
    report[0] = result.converged ? 1.0F : 0.0F;
    report[1] = static_cast<float>(result.iterations);
    report[2] = result.residual_rms;
    report[3] = std::chrono::duration<float>(stop - start).count();
  });
}

}  // extern "C"

#ifdef MSIM_WITH_HIP_BINDING
#include "microsimulator/backend.hpp"

// The reference-side `hip` backend (microsimulator_amd/binding/hip_backend.cpp)
// driven through the reference's own ComputeBackend interface, beside the
// reference's CPU functions on the same spec: what the reference's
// conformance tests would compare with `--backend hip`.
namespace microsimulator {
std::unique_ptr<ComputeBackend> make_hip_backend(std::uint32_t device_index);
std::size_t hip_backend_device_count() noexcept;
}  // namespace microsimulator

extern "C" {

std::size_t bind_device_count(void) { return ms::hip_backend_device_count(); }

// Makes the backend on device 0 and reads its info and features: name
// ("hip"), whether it is native, and supports() of signals, coupled_rates,
// depth_averaged_flow and resolved_flow as bits 0-3.
int bind_create(char* name, std::size_t name_length, std::uint32_t* features, char* err,
                std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto backend = ms::make_hip_backend(0);
    const auto info = backend->info();
    std::snprintf(name, name_length, "%s", info.name.c_str());
    *features = (backend->supports(ms::BackendFeature::signals) ? 1u : 0u) |
                (backend->supports(ms::BackendFeature::coupled_rates) ? 2u : 0u) |
                (backend->supports(ms::BackendFeature::depth_averaged_flow) ? 4u : 0u) |
                (backend->supports(ms::BackendFeature::resolved_flow) ? 8u : 0u) |
                (info.native ? 16u : 0u);
  });
}

// solve_depth_averaged_flow through the hip backend (faces_hip) and through
// solve_depth_averaged_flow_cpu (faces_cpu), on a one-signal grid with the
// given boundary kinds; force: (fx, fy) per column or null. report:
// iterations, relative residual, max speed of each (hip first).
int bind_flow(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx, float dy, float dz,
              const std::uint8_t* obstacles, const std::uint8_t* kinds, const float* mobility,
              std::uint32_t axis, float mean_inlet_speed, float tolerance,
              std::uint32_t max_iterations, const float* force, std::size_t force_columns,
              float* faces_hip, float* faces_cpu, float* report, char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::span<const float> mob;
    if (mobility != nullptr) mob = std::span<const float>(mobility, n);
    ms::DepthAveragedFlowParameters parameters;
    parameters.mean_inlet_speed = mean_inlet_speed;
    parameters.axis = static_cast<ms::FlowAxis>(axis);
    parameters.relative_tolerance = tolerance;
    parameters.max_iterations = max_iterations;
    for (std::size_t c = 0; c < force_columns; ++c) {
      parameters.body_force.push_back({force[2 * c], force[2 * c + 1]});
    }
    const auto backend = ms::make_hip_backend(0);
    const auto hip = backend->solve_depth_averaged_flow(spec, mob, parameters);
    const auto cpu = ms::solve_depth_averaged_flow_cpu(spec, mob, parameters);
    float* out[2] = {faces_hip, faces_cpu};
    const ms::DepthAveragedFlowResult* results[2] = {&hip, &cpu};
    for (int k = 0; k < 2; ++k) {
      const auto& field = results[k]->field;
      std::memcpy(out[k], field.x_faces.data(), field.x_faces.size() * sizeof(float));
      std::memcpy(out[k] + field.x_faces.size(), field.y_faces.data(),
                  field.y_faces.size() * sizeof(float));
      std::memcpy(out[k] + field.x_faces.size() + field.y_faces.size(), field.z_faces.data(),
                  field.z_faces.size() * sizeof(float));
      report[3 * k] = static_cast<float>(results[k]->report.iterations);
      report[3 * k + 1] = results[k]->report.relative_residual;
      report[3 * k + 2] = results[k]->report.max_speed;
    }
  });
}

// solve_resolved_flow through the hip backend (faces_hip) and through
// solve_resolved_flow_cpu (faces_cpu). report: outer and inner iterations,
// relative residual, divergence RMS, max speed and minimum gap of each (hip
// first).
int bind_resolved_flow(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx, float dy,
                       float dz, const std::uint8_t* obstacles, const std::uint8_t* kinds,
                       const float* drag, std::uint32_t axis, float mean_inlet_speed,
                       float tolerance, float* faces_hip, float* faces_cpu, double* report,
                       char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::span<const float> d;
    if (drag != nullptr) d = std::span<const float>(drag, n);
    ms::ResolvedFlowParameters parameters;
    parameters.mean_inlet_speed = mean_inlet_speed;
    parameters.axis = static_cast<ms::FlowAxis>(axis);
    parameters.relative_tolerance = tolerance;
    parameters.inner_relative_tolerance = tolerance;
    const auto backend = ms::make_hip_backend(0);
    const auto hip = backend->solve_resolved_flow(spec, d, parameters);
    const auto cpu = ms::solve_resolved_flow_cpu(spec, d, parameters);
    float* out[2] = {faces_hip, faces_cpu};
    const ms::ResolvedFlowResult* results[2] = {&hip, &cpu};
    for (int k = 0; k < 2; ++k) {
      const auto& field = results[k]->field;
      std::memcpy(out[k], field.x_faces.data(), field.x_faces.size() * sizeof(float));
      std::memcpy(out[k] + field.x_faces.size(), field.y_faces.data(),
                  field.y_faces.size() * sizeof(float));
      std::memcpy(out[k] + field.x_faces.size() + field.y_faces.size(), field.z_faces.data(),
                  field.z_faces.size() * sizeof(float));
      const auto& r = results[k]->report;
      report[6 * k] = r.outer_iterations;
      report[6 * k + 1] = static_cast<double>(r.inner_iterations);
      report[6 * k + 2] = r.relative_residual;
      report[6 * k + 3] = r.divergence_rms;
      report[6 * k + 4] = r.max_speed;
      report[6 * k + 5] = r.min_gap_voxels;
    }
  });
}

// solve_resolved_port_flow through the hip backend (faces_hip, states_hip)
// and through solve_resolved_port_flow_cpu (faces_cpu, states_cpu), ports as
// for ref_port_flow. report per backend (hip first): solves, outer and inner
// iterations, relative residual, divergence RMS, max speed, minimum gap.
int bind_resolved_port_flow(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx,
                            float dy, float dz, const std::uint8_t* obstacles,
                            const std::uint8_t* kinds, const float* drag,
                            std::uint32_t port_count, const std::uint32_t* port_geometry,
                            const float* port_values, float tolerance, float* faces_hip,
                            float* faces_cpu, float* states_hip, float* states_cpu,
                            double* report, char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = make_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, 1, nullptr);
    const std::size_t n = static_cast<std::size_t>(nx) * ny * nz;
    std::span<const float> d;
    if (drag != nullptr) d = std::span<const float>(drag, n);
    std::vector<ms::FlowPort> ports(port_count);
    for (std::uint32_t k = 0; k < port_count; ++k) {
      const std::uint32_t* g = port_geometry + 6 * k;
      ports[k].face = static_cast<ms::GridFace>(g[0]);
      ports[k].patch.begin = {g[1], g[2]};
      ports[k].patch.end = {g[3], g[4]};
      ports[k].kind = static_cast<ms::FlowPortKind>(g[5]);
      ports[k].value = port_values[k];
    }
    ms::ResolvedPortFlowParameters parameters;
    parameters.relative_tolerance = tolerance;
    parameters.inner_relative_tolerance = tolerance;
    const auto backend = ms::make_hip_backend(0);
    const auto hip = backend->solve_resolved_port_flow(spec, ports, d, parameters);
    const auto cpu = ms::solve_resolved_port_flow_cpu(spec, ports, d, parameters);
    float* out[2] = {faces_hip, faces_cpu};
    float* st[2] = {states_hip, states_cpu};
    const ms::ResolvedPortFlowResult* results[2] = {&hip, &cpu};
    for (int k = 0; k < 2; ++k) {
      const auto& field = results[k]->field;
      std::memcpy(out[k], field.x_faces.data(), field.x_faces.size() * sizeof(float));
      std::memcpy(out[k] + field.x_faces.size(), field.y_faces.data(),
                  field.y_faces.size() * sizeof(float));
      std::memcpy(out[k] + field.x_faces.size() + field.y_faces.size(), field.z_faces.data(),
                  field.z_faces.size() * sizeof(float));
      for (std::uint32_t p = 0; p < port_count; ++p) {
        st[k][3 * p] = results[k]->ports[p].pressure;
        st[k][3 * p + 1] = results[k]->ports[p].flux;
        st[k][3 * p + 2] = results[k]->ports[p].mean_speed;
      }
      const auto& r = results[k]->report;
      report[7 * k] = r.solves;
      report[7 * k + 1] = r.outer_iterations;
      report[7 * k + 2] = static_cast<double>(r.inner_iterations);
      report[7 * k + 3] = r.relative_residual;
      report[7 * k + 4] = r.divergence_rms;
      report[7 * k + 5] = r.max_speed;
      report[7 * k + 6] = r.min_gap_voxels;
    }
  });
}

// `steps` transport steps through the hip backend's advance_signal_grid
// (levels in place) and through advance_signal_grid_cpu (levels_cpu), the
// spec as ref_advance_signal_grid takes it. report: the last step's
// (converged, iterations, residual RMS) of each (hip first).
int bind_signal_steps(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, float dx, float dy,
                      float dz, const std::uint8_t* obstacles, const std::uint8_t* kinds,
                      std::uint32_t signals, const float* diffusion, const float* advection,
                      const float* boundary_values, const float* x_faces, const float* y_faces,
                      const float* z_faces, const float* source_rates, const float* loss_rates,
                      std::uint32_t scheme, std::uint32_t patch_count,
                      const std::uint32_t* patch_geometry, const float* patch_values,
                      std::uint32_t integration, std::uint32_t max_iterations,
                      float absolute_tolerance, float relative_tolerance, float dt,
                      std::uint32_t steps, float* levels, float* levels_cpu, float* report,
                      char* err, std::size_t errlen) {
  return guarded(err, errlen, [&] {
    const auto spec = signal_spec(nx, ny, nz, dx, dy, dz, obstacles, kinds, signals, diffusion,
                                  advection, boundary_values, x_faces, y_faces, z_faces,
                                  source_rates, loss_rates, scheme, patch_count, patch_geometry,
                                  patch_values, integration, max_iterations, absolute_tolerance,
                                  relative_tolerance);
    const std::size_t count = static_cast<std::size_t>(nx) * ny * nz * signals;
    ms::SignalGrid on_hip(spec, std::vector<float>(levels, levels + count));
    ms::SignalGrid on_cpu(spec, std::vector<float>(levels, levels + count));
    const auto backend = ms::make_hip_backend(0);
    ms::SignalSolveReport a, b;
    for (std::uint32_t step = 0; step < steps; ++step) {
      a = backend->advance_signal_grid(on_hip, dt);
      b = ms::advance_signal_grid_cpu(on_cpu, dt);
    }
    std::memcpy(levels, on_hip.levels().data(), count * sizeof(float));
    std::memcpy(levels_cpu, on_cpu.levels().data(), count * sizeof(float));
    const ms::SignalSolveReport* r[2] = {&a, &b};
    for (int k = 0; k < 2; ++k) {
      report[3 * k] = r[k]->converged ? 1.0F : 0.0F;
      report[3 * k + 1] = static_cast<float>(r[k]->iterations);
      report[3 * k + 2] = r[k]->residual_rms;
    }
  });
}

}  // extern "C"
#endif  // MSIM_WITH_HIP_BINDING
