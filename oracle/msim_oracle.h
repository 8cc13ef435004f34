/*
 * CPU oracle for the grid flow-and-transport path: a plain-C restatement of
 * the reference's CPU algorithm, used ONLY as a checker by tests/, by the
 * smoke check in __graft_entry__.py and by the cpu_baseline leg of bench.py.
 * Nothing in the product (microsimulator_amd/) links, loads or calls it.
 *
 * Every function names the reference file:line it follows. Each float
 * operation is written in the reference's own order and precision so that,
 * built with -ffp-contract=off, the results are bit-identical to the
 * reference's CPU backend (checked against oracle/_ref and tests/golden).
 *
 * Parity status per function (what pins it):
 *   mo_flow_assemble        bit-exact vs reference assembly fixtures + closed forms
 *   mo_flow_solve           bit-exact vs reference solve_depth_averaged_flow_cpu
 *   mo_flow_solve_forced    bit-exact vs the same with a body force (live _ref)
 *                           fixtures; plug flow 5.0, mobility-split ratio 3,
 *                           dense-Cholesky brute force on tiny grids
 *   mo_flow_apply           symmetry / positivity; bit-exact vs reference apply
 *                           through the solve fixtures
 *   mo_adr_step_f32         bit-exact vs reference advance_signal_grid_cpu
 *                           fixtures (affine and Monod-as-affine); stencil
 *                           closed form 0.25/0.5/0.25, reaction closed form 1.75
 *   mo_adr_step_f64         (no reference counterpart: FP64 state) reduces to
 *                           mo_adr_step_f32 within FP32 rounding; exact mass
 *                           ledger identity to 1e-12
 *   mo_adr_validate_step    the reference's explicit bound, known answers from
 *                           the reference tests (dt 0.25 ok / 0.51 rejected)
 *   mo_adr_step_implicit_f32 bit-exact vs reference advance_signal_grid_cpu under
 *                           Crank-Nicolson and backward Euler (oracle/_ref): levels,
 *                           iterations and residual RMS; the reference tests'
 *                           known answers (2.2, 0.4/0.2/0.4) and non-convergence
 *   mo_advance_coupled_f32  bit-exact vs reference advance_coupled_cpu (oracle/_ref):
 *                           grid levels, species levels and report, forward Euler
 *                           and Crank-Nicolson, every rate operation; the reference
 *                           tests' closed forms and refusals
 *   mo_port_flow_solve      bit-exact vs reference solve_depth_averaged_port_flow_cpu
 *                           (oracle/_ref); the reference tests' known answers
 *                           (port fluxes 1.5/0.5/-2, pressure-driven duct 0.8,
 *                           whole-face ports equal to the single-axis solve);
 *                           the refusals of the reference tests
 */
#ifndef MSIM_ORACLE_H
#define MSIM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the same numbering as the product C-ABI (include/msim_hip.h). */
enum {
  MO_OK = 0,
  MO_INVALID_ARGUMENT = 1,
  MO_RUNTIME_ERROR = 2,
  MO_NOT_CONVERGED = 3,
  MO_OVERFLOW = 4,
  MO_DOMAIN_ERROR = 5,
  MO_CG_CURVATURE = 11,
  MO_CG_STEP = 12,
  MO_CG_RESIDUAL = 13,
  MO_CG_BREAKDOWN = 14,
  MO_CG_DIRECTION = 15
};

enum { MO_NO_FLUX = 0, MO_PERIODIC = 1, MO_FIXED = 2 };

typedef struct {
  uint32_t nx, ny, nz;
  float dx, dy, dz;
  const uint8_t* obstacles; /* nx*ny*nz, or NULL for all fluid */
  /* boundary kinds in the order x_lower, x_upper, y_lower, y_upper, z_lower, z_upper */
  uint8_t boundary[6];
} mo_grid;

typedef struct {
  uint32_t iterations;
  double relative_residual;      /* recursive residual at exit (reference report) */
  double true_relative_residual; /* ||b - A x|| / ||b|| recomputed in FP64 */
  float mean_inlet_speed;
  float max_speed;
  float solved_mean; /* mean inlet speed of the unscaled field */
  float factor;      /* scaling factor applied */
} mo_flow_report;

/* The structured-grid form of the depth-averaged system (nz = 1):
 * per site: column conductance, diagonal, rhs, boundary conductance;
 * per face (x|y|z concatenated, reference face order): face conductance and
 * boundary pressure. Arrays are caller-allocated: N = nx*ny, F = total faces. */
int mo_flow_assemble(const mo_grid* grid, const float* mobility, uint32_t axis,
                     float* conductance, float* diagonal, float* rhs, float* boundary_conductance,
                     float* face_conductance, float* face_pressure, char* err, size_t errlen);

/* y = A x over the whole structured grid (inactive sites give 0), in FP64,
 * following DepthAveragedFlowSystem::apply's per-site arithmetic. */
int mo_flow_apply(const mo_grid* grid, const float* mobility, uint32_t axis, const double* x,
                  double* y, char* err, size_t errlen);

/* Full depth-averaged solve: reduction, assembly, Jacobi-PCG, velocity,
 * lift and scaling. faces: x|y|z face velocities of the original grid.
 * pressure: N doubles (zero at inactive sites), may be NULL.
 * scale = 0 returns the unscaled (Dirichlet 1/0) field. */
int mo_flow_solve(const mo_grid* grid, const float* mobility, uint32_t axis,
                  float mean_inlet_speed, float tolerance, uint32_t max_iterations, int scale,
                  float* faces, double* pressure, mo_flow_report* report, char* err,
                  size_t errlen);

/* mo_flow_solve with DepthAveragedFlowParameters::body_force (flow.hpp:24-25):
 * force holds (fx, fy) per column of the grid (x-major, y-minor), force_columns
 * of them (nx*ny, or the reference's refusal), or NULL for none. */
int mo_flow_solve_forced(const mo_grid* grid, const float* mobility, uint32_t axis,
                         float mean_inlet_speed, float tolerance, uint32_t max_iterations,
                         int scale, const float* force, size_t force_columns, float* faces,
                         double* pressure, mo_flow_report* report, char* err, size_t errlen);

/* ResolvedFlowReport (cpp/include/microsimulator/flow.hpp:53-62), with the
 * scaling's solved mean and factor. */
typedef struct {
  uint32_t outer_iterations;
  uint64_t inner_iterations;
  float relative_residual;
  float momentum_relative_residual;
  float divergence_rms;
  float mean_inlet_speed;
  float max_speed;
  uint32_t min_gap_voxels;
  float solved_mean;
  float factor;
} mo_resolved_report;

/* solve_resolved_flow_cpu (cpp/cpu/cpu_flow.cpp:463-467): the resolved
 * Stokes-Brinkman flow along `axis` (0, 1 or 2), drag one float per site or
 * NULL (zero), FGMRES to `tolerance` with the momentum block by Jacobi-PCG to
 * `inner_tolerance`; faces: x|y|z velocities of every face, scaled to
 * mean_inlet_speed (scale = 0: the unscaled solve). */
int mo_resolved_flow_solve(const mo_grid* grid, const float* drag, uint32_t axis,
                           float mean_inlet_speed, float tolerance, uint32_t max_outer,
                           float inner_tolerance, uint32_t max_inner, int scale, float* faces,
                           mo_resolved_report* report, char* err, size_t errlen);

/* The resolved system's blocks applied to given vectors (faces u, sites p):
 * au = A u, gp = G p, du = D u (ResolvedFlowSystem::apply_momentum, gradient,
 * divergence, flow_system.hpp:881-956). */
int mo_resolved_operator(const mo_grid* grid, const float* drag, uint32_t axis, const double* u,
                         const double* p, double* au, double* gp, double* du, char* err,
                         size_t errlen);

/* FlowPortKind and FlowPort (cpp/include/microsimulator/flow.hpp:69-91): an
 * opening over a rectangle of one fixed face, driven at a mean speed (into
 * the grid positive) or held at a pressure. In-plane indices as for a
 * boundary patch: on an x face (y, z), on a y face (x, z), on a z face (x, y). */
enum { MO_PORT_SPEED = 0, MO_PORT_PRESSURE = 1 };
typedef struct {
  uint32_t face; /* 0..5: x_lower, x_upper, y_lower, y_upper, z_lower, z_upper */
  uint32_t begin[2], end[2];
  uint32_t kind; /* MO_PORT_SPEED or MO_PORT_PRESSURE */
  float value;
} mo_flow_port;

/* FlowPortState (flow.hpp:96-100). */
typedef struct {
  float pressure, flux, mean_speed;
} mo_port_state;

/* DepthAveragedPortFlowReport (flow.hpp:109-114). */
typedef struct {
  uint32_t solves, iterations;
  float relative_residual, max_speed;
} mo_port_report;

/* solve_depth_averaged_port_flow_cpu (cpp/cpu/cpu_flow.cpp:353-375): shallow
 * flow driven through ports, as the superposition of one Jacobi-PCG pressure
 * solve per speed port (and one for the held pressures). faces: x|y|z face
 * velocities of the original grid; states: one per port. */
int mo_port_flow_solve(const mo_grid* grid, const float* mobility, const mo_flow_port* ports,
                       uint32_t count, float tolerance, uint32_t max_iterations, float* faces,
                       mo_port_state* states, mo_port_report* report, char* err, size_t errlen);

/* ResolvedPortFlowReport (cpp/include/microsimulator/flow.hpp). */
typedef struct {
  uint32_t solves;
  uint32_t outer_iterations;
  uint64_t inner_iterations;
  float relative_residual;
  float momentum_relative_residual;
  float divergence_rms;
  float max_speed;
  uint32_t min_gap_voxels;
  /* Solves whose system was shifted by the depth-averaged pressure
   * (shallow_port_pressure found a shallow description). */
  uint32_t shifted_solves;
} mo_resolved_port_report;

/* solve_resolved_port_flow_cpu (cpp/cpu/cpu_flow.cpp:474-479, over
 * detail::solve_resolved_port_flow, flow_system.hpp:1482-1548): the resolved
 * Stokes-Brinkman flow driven through ports, the superposition of one block
 * solve per speed port (and one for the held pressures), each for the
 * correction to the depth-averaged pressure of the same ports where shallow
 * flow describes the geometry. faces: x|y|z velocities; states: one per
 * port. */
int mo_resolved_port_flow_solve(const mo_grid* grid, const float* drag, const mo_flow_port* ports,
                                uint32_t count, float tolerance, uint32_t max_outer,
                                float inner_tolerance, uint32_t max_inner, float* faces,
                                mo_port_state* states, mo_resolved_port_report* report,
                                char* err, size_t errlen);

/* GridBoundaryPatch (signals.hpp:72-84): a rectangle of a fixed face with its
 * own reservoir value per signal. In-plane indices: on an x face (y, z), on a
 * y face (x, z), on a z face (x, y); begin inclusive, end exclusive. */
typedef struct {
  uint32_t face; /* 0..5: x_lower, x_upper, y_lower, y_upper, z_lower, z_upper */
  uint32_t begin[2], end[2];
  const float* values; /* signal_count values */
} mo_boundary_patch;

/* Signal-grid transport description for one explicit step. */
typedef struct {
  mo_grid grid;
  uint32_t signal_count;
  const float* diffusion;       /* S */
  const float* advection;       /* 3*S constant velocity, used when faces are NULL */
  const float* boundary_values; /* 6*S reservoir values of fixed faces (others ignored) */
  const float* x_faces;         /* velocity field (nx+1)*ny*nz, or NULL */
  const float* y_faces;         /* nx*(ny+1)*nz */
  const float* z_faces;         /* nx*ny*(nz+1) */
  const float* source_rates;    /* affine reaction S*N, or NULL */
  const float* loss_rates;      /* S*N, or NULL */
  const float* phi;             /* Monod occupancy N, or NULL (no Monod) */
  const float* monod_vmax;      /* S */
  const float* monod_k;         /* S */
  uint32_t advection_scheme;    /* MO_UPWIND or MO_VAN_LEER (SignalAdvectionKind) */
  const mo_boundary_patch* patches; /* reservoir patches of fixed faces, or NULL */
  uint32_t patch_count;
  uint32_t integration;         /* MO_FORWARD_EULER, MO_CRANK_NICOLSON or MO_BACKWARD_EULER */
  uint32_t solver_max_iterations; /* SignalSolveParameters (signals.hpp:49-55) */
  float solver_absolute_tolerance;
  float solver_relative_tolerance;
} mo_signal_grid;

enum { MO_UPWIND = 0, MO_VAN_LEER = 1 };
enum { MO_FORWARD_EULER = 0, MO_CRANK_NICOLSON = 1, MO_BACKWARD_EULER = 2 };

/* SignalSolveReport (signals.hpp:57-61). */
typedef struct {
  uint32_t converged, iterations;
  float residual_rms;
} mo_signal_report;

/* One implicit step (Crank-Nicolson or backward Euler) in FP32, levels S*N
 * in place: signal_grid_crank_nicolson_candidate (signals.cpp:1293-1375)
 * under solve_signal_jacobi (signal_jacobi.hpp:53-84), then accept_step's
 * settle and validation (signals.cpp:979-1011). Monod uptake enters as the
 * loss vmax*phi/(K+c) of the starting levels, added to any affine loss.
 * report may be NULL. A solve that does not converge returns
 * MO_NOT_CONVERGED with the reference's message. */
int mo_adr_step_implicit_f32(const mo_signal_grid* sg, float dt, float* levels,
                             mo_signal_report* report, char* err, size_t errlen);

/* The explicit stability bound of SignalGrid::validate_step. For Monod the
 * loss bound is vmax * max(phi) / K. Returns MO_OK or MO_INVALID_ARGUMENT. */
int mo_adr_validate_step(const mo_signal_grid* sg, float dt, char* err, size_t errlen);

/* One forward-Euler step in FP32, levels S*N in place (reference order).
 * Under van Leer the step's levels are settled and validated as
 * SignalGrid::accept_step does (signals.cpp:979-1011). */
int mo_adr_step_f32(const mo_signal_grid* sg, float dt, float* levels, char* err, size_t errlen);

/* One forward-Euler step with FP64 state in face-flux form. ledger (may be
 * NULL) receives per signal: [mass_before, mass_after, boundary_inflow*dt,
 * uptake*dt] with mass = sum(c) * voxel volume. */
int mo_adr_step_f64(const mo_signal_grid* sg, float dt, double* levels, double* ledger, char* err,
                    size_t errlen);

/* RateInstruction (species.hpp:18-24) with RateOp's codes
 * (cpp/include/microsimulator/rate_ops.def). */
typedef struct {
  uint32_t operation, first, second, third;
  float value;
} mo_rate_instruction;

/* CoupledRatePlan (coupled_rates.hpp:16-41). */
typedef struct {
  uint32_t species_count, signal_count, instruction_count;
  const mo_rate_instruction* instructions;
  const uint32_t* species_outputs; /* species_count instruction indices */
  const uint32_t* signal_outputs;  /* signal_count instruction indices */
} mo_rate_plan;

/* The cells of a WorldState the coupled step reads (world_state.hpp:58-84). */
typedef struct {
  uint32_t count;
  const float *position_x, *position_y, *position_z, *lengths, *radii, *growth_rates;
  const int32_t* cell_types;
} mo_cells;

/* advance_coupled_cpu (cpp/cpu/cpu_coupled.cpp:125-231) with its checks
 * (validate_coupled_step, 95-123): every cell samples the grid through its
 * stencil (signal_grid_stencil, signals.cpp:357-420), dilutes and advances
 * its species by the plan, and exchanges with the grid; the sources are
 * summed into the sites cell by cell and the grid takes a step of its
 * integration kind with them. levels S*N and species count*species_count in
 * place; origin: the grid's origin (x, y, z). report (may be NULL): the
 * implicit solve's. A refused step changes neither. */
int mo_advance_coupled_f32(const mo_signal_grid* sg, const float* origin, float* levels,
                           const mo_cells* cells, float* species, const mo_rate_plan* plan,
                           const float* previous_lengths, float dt, mo_signal_report* report,
                           char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif
