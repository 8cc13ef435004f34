// The `hip` ComputeBackend a maintainer adds to the reference as
// cpp/hip/hip_backend.cpp: the reference's backend interface
// (cpp/include/microsimulator/backend.hpp:21-56) over this repository's C ABI
// (include/msim_hip.h). It compiles against the reference's own headers
// (oracle/Makefile, target `binding`, builds and links it with the
// reference's CPU sources for the tests).
//
// The reference gains `hip` in BackendKind (types.hpp:59-63) with this file;
// against headers that lack it, MICROSIMULATOR_HAS_HIP_KIND is left undefined
// and info() reports the next enumerator value.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "microsimulator/backend.hpp"
#include "msim_hip.h"

namespace microsimulator {
namespace {

#ifdef MICROSIMULATOR_HAS_HIP_KIND
constexpr BackendKind kHipKind = BackendKind::hip;
#else
constexpr BackendKind kHipKind = static_cast<BackendKind>(3);
#endif

// The C ABI's status and message as the exception the reference throws
// (conjugate_gradient.hpp:91-98 failures and non-convergence are
// runtime_error, as the CPU backend's are).
[[noreturn]] void rethrow(int status, msim_ctx* ctx) {
  const std::string message = msim_hip_last_error(ctx);
  switch (status) {
    case MSIM_INVALID_ARGUMENT: throw std::invalid_argument(message);
    case MSIM_OVERFLOW: throw std::overflow_error(message);
    case MSIM_DOMAIN_ERROR: throw std::domain_error(message);
    default: throw std::runtime_error(message);
  }
}

msim_grid grid_of(const SignalGridSpec& spec) {
  msim_grid g{};
  g.nx = spec.shape.x;
  g.ny = spec.shape.y;
  g.nz = spec.shape.z;
  g.dx = spec.spacing.x;
  g.dy = spec.spacing.y;
  g.dz = spec.spacing.z;
  g.obstacles = spec.obstacles.empty() ? nullptr : spec.obstacles.data();
  const GridBoundary* faces[6] = {&spec.x_lower, &spec.x_upper, &spec.y_lower,
                                  &spec.y_upper, &spec.z_lower, &spec.z_upper};
  for (int f = 0; f < 6; ++f) g.boundary[f] = static_cast<std::uint8_t>(faces[f]->kind);
  return g;
}

// The transport fields of a spec, pointing into `held` for what the C ABI
// lays out differently from the reference (boundary values per face, patches,
// advection vectors).
struct HeldSignalGrid {
  std::vector<float> values, advection;
  std::vector<msim_boundary_patch> patches;
  msim_signal_grid grid{};
};

void fill_signal_grid(const SignalGridSpec& spec, std::uint64_t revision, HeldSignalGrid& held) {
  msim_signal_grid& s = held.grid;
  s = msim_signal_grid{};
  s.grid = grid_of(spec);
  s.signal_count = static_cast<std::uint32_t>(spec.signal_count);
  s.diffusion = spec.diffusion.data();
  held.values.assign(6 * spec.signal_count, 0.0F);  // face f of signal k at [6k + f]
  const GridBoundary* faces[6] = {&spec.x_lower, &spec.x_upper, &spec.y_lower,
                                  &spec.y_upper, &spec.z_lower, &spec.z_upper};
  for (int f = 0; f < 6; ++f) {
    for (std::size_t k = 0; k < spec.signal_count && k < faces[f]->values.size(); ++k) {
      held.values[6 * k + f] = faces[f]->values[k];
    }
  }
  s.boundary_values = held.values.data();
  held.patches.clear();
  for (int f = 0; f < 6; ++f) {
    for (const auto& patch : faces[f]->patches) {
      held.patches.push_back({static_cast<std::uint32_t>(f),
                              {patch.region.begin[0], patch.region.begin[1]},
                              {patch.region.end[0], patch.region.end[1]}, patch.values.data()});
    }
  }
  s.patches = held.patches.empty() ? nullptr : held.patches.data();
  s.patch_count = static_cast<std::uint32_t>(held.patches.size());
  held.advection.clear();
  for (const auto& v : spec.advection) {
    held.advection.insert(held.advection.end(), {v.x, v.y, v.z});
  }
  s.advection = held.advection.empty() ? nullptr : held.advection.data();
  s.advection_scheme =
      spec.advection_scheme == SignalAdvectionKind::van_leer ? MSIM_VAN_LEER : MSIM_UPWIND;
  s.integration = static_cast<std::uint32_t>(spec.integration);  // the same numbering
  s.solver_max_iterations = spec.solver.max_iterations;
  s.solver_absolute_tolerance = spec.solver.absolute_tolerance;
  s.solver_relative_tolerance = spec.solver.relative_tolerance;
  if (spec.velocity_field) {
    s.x_faces = spec.velocity_field->x_faces.data();
    s.y_faces = spec.velocity_field->y_faces.data();
    s.z_faces = spec.velocity_field->z_faces.data();
  }
  if (spec.reaction) {
    s.source_rates = spec.reaction->source_rates.data();
    s.loss_rates = spec.reaction->loss_rates.data();
  }
  s.revision = revision;
}

// The pressure solver form: the reference's FP64 Jacobi-PCG by default (its
// iteration counts, as its conformance suite checks them), or the multilevel
// preconditioner (MICROSIMULATOR_HIP_SOLVER=multigrid) or the learned one
// (=learned, after set_learned_weights).
std::uint32_t solver_from_environment() {
  const char* name = std::getenv("MICROSIMULATOR_HIP_SOLVER");
  if (name == nullptr || std::strcmp(name, "jacobi") == 0) return MSIM_SOLVER_JACOBI_FP64;
  if (std::strcmp(name, "multigrid") == 0) return MSIM_SOLVER_MULTIGRID;
  if (std::strcmp(name, "mixed") == 0) return MSIM_SOLVER_MIXED;
  if (std::strcmp(name, "learned") == 0) return MSIM_SOLVER_LEARNED;
  throw std::invalid_argument("unknown MICROSIMULATOR_HIP_SOLVER");
}

class HipBackend final : public ComputeBackend {
 public:
  HipBackend(std::uint32_t device, std::uint32_t solver) : device_(device), solver_(solver) {
    if (const int s = msim_hip_create(static_cast<int>(device), &ctx_); s != MSIM_OK) {
      throw std::runtime_error("HIP backend device is unavailable");
    }
  }
  ~HipBackend() override { msim_hip_destroy(ctx_); }
  HipBackend(const HipBackend&) = delete;
  HipBackend& operator=(const HipBackend&) = delete;

  [[nodiscard]] BackendInfo info() const override {
    char name[256] = {};
    msim_hip_device_name(ctx_, name, sizeof name);
    return {.kind = kHipKind, .name = "hip", .device = name, .device_index = device_,
            .native = true};
  }
  [[nodiscard]] bool supports(BackendFeature f) const noexcept override {
    return f == BackendFeature::signals || f == BackendFeature::coupled_rates ||
           f == BackendFeature::depth_averaged_flow || f == BackendFeature::resolved_flow;
  }

  [[nodiscard]] DepthAveragedFlowResult solve_depth_averaged_flow(
      const SignalGridSpec& spec, std::span<const float> mobility,
      const DepthAveragedFlowParameters& p) override {
    p.validate();
    const msim_grid g = grid_of(spec);
    msim_flow_params params{};
    params.mean_inlet_speed = p.mean_inlet_speed;
    params.axis = static_cast<std::uint32_t>(p.axis);
    params.relative_tolerance = p.relative_tolerance;
    params.max_iterations = p.max_iterations;
    params.solver = solver_;
    params.scale = 1;
    params.fallback_iterations = 200;
    if (!p.body_force.empty()) {
      // std::array<float, 2> per column: (fx, fy) pairs, contiguous.
      params.body_force = p.body_force.front().data();
      params.body_force_columns = p.body_force.size();
    }
    DepthAveragedFlowResult result;
    result.field.x_faces.resize(spec.x_face_count());
    result.field.y_faces.resize(spec.y_face_count());
    result.field.z_faces.resize(spec.z_face_count());
    msim_flow_report report{};
    if (const int s = msim_hip_solve_depth_averaged_flow(
            ctx_, &g, mobility.empty() ? nullptr : mobility.data(), &params,
            result.field.x_faces.data(), result.field.y_faces.data(),
            result.field.z_faces.data(), &report);
        s != MSIM_OK) {
      rethrow(s, ctx_);
    }
    result.report = {.iterations = report.iterations,
                     .relative_residual = static_cast<float>(report.relative_residual),
                     .mean_inlet_speed = report.mean_inlet_speed,
                     .max_speed = report.max_speed};
    return result;
  }

  [[nodiscard]] DepthAveragedPortFlowResult solve_depth_averaged_port_flow(
      const SignalGridSpec& spec, std::span<const FlowPort> ports,
      std::span<const float> mobility, const DepthAveragedPortFlowParameters& p) override {
    p.validate();
    const msim_grid g = grid_of(spec);
    std::vector<msim_flow_port> held;
    for (const auto& port : ports) {
      held.push_back({static_cast<std::uint32_t>(port.face),
                      {port.patch.begin[0], port.patch.begin[1]},
                      {port.patch.end[0], port.patch.end[1]},
                      port.kind == FlowPortKind::speed ? static_cast<std::uint32_t>(MSIM_PORT_SPEED)
                                                       : static_cast<std::uint32_t>(MSIM_PORT_PRESSURE),
                      port.value});
    }
    msim_port_flow_params params{};
    params.relative_tolerance = p.relative_tolerance;
    params.max_iterations = p.max_iterations;
    // The learned smoother serves axis solves; ports take the multilevel form.
    params.solver = solver_ == MSIM_SOLVER_LEARNED
                        ? static_cast<std::uint32_t>(MSIM_SOLVER_MULTIGRID)
                        : solver_;
    DepthAveragedPortFlowResult result;
    result.field.x_faces.resize(spec.x_face_count());
    result.field.y_faces.resize(spec.y_face_count());
    result.field.z_faces.resize(spec.z_face_count());
    std::vector<msim_port_state> states(ports.size());
    msim_port_flow_report report{};
    if (const int s = msim_hip_solve_depth_averaged_port_flow(
            ctx_, &g, mobility.empty() ? nullptr : mobility.data(), held.data(),
            static_cast<std::uint32_t>(held.size()), &params, result.field.x_faces.data(),
            result.field.y_faces.data(), result.field.z_faces.data(), &report, states.data());
        s != MSIM_OK) {
      rethrow(s, ctx_);
    }
    result.report = {.solves = report.solves, .iterations = report.iterations,
                     .relative_residual = static_cast<float>(report.relative_residual),
                     .max_speed = report.max_speed};
    for (const auto& st : states) {
      result.ports.push_back({.pressure = st.pressure, .flux = st.flux,
                              .mean_speed = st.mean_speed});
    }
    return result;
  }

  SignalSolveReport advance_signal_grid(SignalGrid& grid, float dt) override {
    HeldSignalGrid held;
    fill_signal_grid(grid.spec(), grid.revision(), held);
    std::vector<float> levels(grid.levels().begin(), grid.levels().end());
    msim_signal_report report{};
    if (const int st = msim_hip_advance_signal_grid(ctx_, &held.grid, dt, levels.data(), &report);
        st != MSIM_OK) {
      rethrow(st, ctx_);
    }
    grid.accept_step(std::move(levels));
    return {.converged = report.converged != 0, .iterations = report.iterations,
            .residual_rms = report.residual_rms};
  }

  SignalSolveReport advance_coupled(WorldState& state, SignalGrid& grid,
                                    const CoupledRatePlan& plan,
                                    std::span<const float> previous_lengths, float dt) override {
    HeldSignalGrid held;
    const auto& spec = grid.spec();
    fill_signal_grid(spec, grid.revision(), held);
    const float origin[3] = {spec.origin.x, spec.origin.y, spec.origin.z};
    const auto geometry = state.geometry_state();
    const auto attributes = state.cell_attributes();
    if (previous_lengths.size() != state.size()) {
      throw std::invalid_argument("previous cell lengths and world state cell counts disagree");
    }
    const msim_cells cells{static_cast<std::uint32_t>(state.size()),
                           geometry.position_x.data(),
                           geometry.position_y.data(),
                           geometry.position_z.data(),
                           geometry.lengths.data(),
                           geometry.radii.data(),
                           attributes.growth_rates.data(),
                           attributes.cell_types.data()};
    std::vector<msim_rate_instruction> program;
    for (const auto& in : plan.instructions()) {
      program.push_back({static_cast<std::uint32_t>(in.operation), in.first, in.second,
                         in.third, in.value});
    }
    const msim_rate_plan rates{static_cast<std::uint32_t>(plan.species_count()),
                               static_cast<std::uint32_t>(plan.signal_count()),
                               static_cast<std::uint32_t>(program.size()),
                               program.data(),
                               plan.species_outputs().data(),
                               plan.signal_outputs().data()};
    std::vector<float> levels(grid.levels().begin(), grid.levels().end());
    auto species = state.species_state().levels;  // written only when the step succeeds
    std::vector<float> next(species.begin(), species.end());
    msim_signal_report report{};
    if (const int st = msim_hip_advance_coupled(ctx_, &held.grid, origin, levels.data(), &cells,
                                                next.data(), &rates, previous_lengths.data(),
                                                dt, &report);
        st != MSIM_OK) {
      rethrow(st, ctx_);
    }
    grid.accept_step(std::move(levels));
    std::copy(next.begin(), next.end(), species.begin());
    return {.converged = report.converged != 0, .iterations = report.iterations,
            .residual_rms = report.residual_rms};
  }

  // The rest of ComputeBackend is not on this backend: no silent fallback
  // (docs/architecture/0001-native-backends.md:33).
  void advance_growth(WorldState&, float) override { unsupported(); }
  void advance_species(WorldState&, const SpeciesRatePlan&, std::span<const float>,
                       float) override {
    unsupported();
  }
  ContactGraph find_cell_contacts(const WorldState&, const ContactParameters&) override {
    unsupported();
  }
  ExternalContactGraph find_external_contacts(const WorldState&, const ConstraintSet&,
                                              const ConstraintContactParameters&) override {
    unsupported();
  }
  ContactSolveResult solve_contact_forces(const ContactSystem&, std::span<const float>,
                                          const ContactSolveParameters&) override {
    unsupported();
  }
  ResolvedFlowResult solve_resolved_flow(const SignalGridSpec& spec, std::span<const float> drag,
                                         const ResolvedFlowParameters& p) override {
    p.validate();
    const msim_grid g = grid_of(spec);
    if (!drag.empty() && drag.size() != spec.site_count()) {
      throw std::invalid_argument("resolved-flow drag must hold one value per grid site");
    }
    const msim_resolved_params params{p.mean_inlet_speed, static_cast<std::uint32_t>(p.axis),
                                      p.relative_tolerance, p.max_outer_iterations,
                                      p.inner_relative_tolerance, p.max_inner_iterations};
    ResolvedFlowResult result;
    result.field.x_faces.resize(spec.x_face_count());
    result.field.y_faces.resize(spec.y_face_count());
    result.field.z_faces.resize(spec.z_face_count());
    msim_resolved_report report{};
    if (const int s = msim_hip_solve_resolved_flow(
            ctx_, &g, drag.empty() ? nullptr : drag.data(), &params, result.field.x_faces.data(),
            result.field.y_faces.data(), result.field.z_faces.data(), &report);
        s != MSIM_OK) {
      rethrow(s, ctx_);
    }
    result.report = {.outer_iterations = report.outer_iterations,
                     .inner_iterations = report.inner_iterations,
                     .relative_residual = report.relative_residual,
                     .momentum_relative_residual = report.momentum_relative_residual,
                     .divergence_rms = report.divergence_rms,
                     .mean_inlet_speed = report.mean_inlet_speed,
                     .max_speed = report.max_speed,
                     .min_gap_voxels = report.min_gap_voxels};
    return result;
  }
  ResolvedPortFlowResult solve_resolved_port_flow(const SignalGridSpec& spec,
                                                  std::span<const FlowPort> ports,
                                                  std::span<const float> drag,
                                                  const ResolvedPortFlowParameters& p) override {
    p.validate();
    const msim_grid g = grid_of(spec);
    if (!drag.empty() && drag.size() != spec.site_count()) {
      throw std::invalid_argument("resolved-flow drag must hold one value per grid site");
    }
    std::vector<msim_flow_port> held;
    for (const auto& port : ports) {
      held.push_back({static_cast<std::uint32_t>(port.face),
                      {port.patch.begin[0], port.patch.begin[1]},
                      {port.patch.end[0], port.patch.end[1]},
                      port.kind == FlowPortKind::speed ? static_cast<std::uint32_t>(MSIM_PORT_SPEED)
                                                       : static_cast<std::uint32_t>(MSIM_PORT_PRESSURE),
                      port.value});
    }
    const msim_resolved_port_params params{p.relative_tolerance, p.max_outer_iterations,
                                           p.inner_relative_tolerance, p.max_inner_iterations};
    ResolvedPortFlowResult result;
    result.field.x_faces.resize(spec.x_face_count());
    result.field.y_faces.resize(spec.y_face_count());
    result.field.z_faces.resize(spec.z_face_count());
    std::vector<msim_port_state> states(ports.size());
    msim_resolved_port_report report{};
    if (const int s = msim_hip_solve_resolved_port_flow(
            ctx_, &g, drag.empty() ? nullptr : drag.data(), held.data(),
            static_cast<std::uint32_t>(held.size()), &params, result.field.x_faces.data(),
            result.field.y_faces.data(), result.field.z_faces.data(), &report, states.data());
        s != MSIM_OK) {
      rethrow(s, ctx_);
    }
    result.report = {.solves = report.solves,
                     .outer_iterations = report.outer_iterations,
                     .inner_iterations = report.inner_iterations,
                     .relative_residual = report.relative_residual,
                     .momentum_relative_residual = report.momentum_relative_residual,
                     .divergence_rms = report.divergence_rms,
                     .max_speed = report.max_speed,
                     .min_gap_voxels = report.min_gap_voxels};
    for (const auto& st : states) {
      result.ports.push_back({.pressure = st.pressure, .flux = st.flux,
                              .mean_speed = st.mean_speed});
    }
    return result;
  }

 private:
  [[noreturn]] static void unsupported() {
    throw std::runtime_error("not implemented in the HIP backend");
  }
  std::uint32_t device_;
  std::uint32_t solver_;
  msim_ctx* ctx_ = nullptr;
};

}  // namespace

std::unique_ptr<ComputeBackend> make_hip_backend(std::uint32_t device_index) {
  return std::make_unique<HipBackend>(device_index, solver_from_environment());
}

std::size_t hip_backend_device_count() noexcept {
  return static_cast<std::size_t>(msim_hip_device_count());
}

}  // namespace microsimulator
