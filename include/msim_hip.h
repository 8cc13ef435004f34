/*
 * msim_hip.h -- C ABI of the MI355X (gfx950) backend for MicroSimulator's
 * grid flow-and-transport path: the depth-averaged (Hele-Shaw/Darcy) pressure
 * solve (axis- or port-driven), face-velocity reconstruction, the resolved
 * Stokes-Brinkman flow (axis- or port-driven), the conservative
 * advection-diffusion-reaction step (explicit or implicit), and the coupled
 * cell/grid exchange around it.
 *
 * The entry points are what a `hip` ComputeBackend in the reference would
 * bind (cpp/include/microsimulator/backend.hpp:21-56). Plain pointers and
 * sizes only; nothing throws across this boundary. Every call returns an
 * msim_status; msim_hip_last_error() gives the message of the last failure on
 * a context. Arrays are caller-owned host memory and are never retained after
 * a call returns; device memory belongs to the context.
 *
 * Storage follows the reference: site = (x*ny + y)*nz + z; face arrays
 * x: (fx*ny + y)*nz + z, y: (x*(ny+1) + fy)*nz + z, z: (x*ny + y)*(nz+1) + fz
 * (cpp/core/flow_system.hpp:64-96). Levels are signal-major, s*N + site.
 *
 * A context is single-threaded: calls on one context must not overlap.
 * Different contexts may be used from different threads.
 */
#ifndef MSIM_HIP_H
#define MSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. They map onto the exception types the reference's backends
 * throw (cpp/cpu/cpu_flow.cpp, cpp/core/signals.cpp, cpp/cuda/cuda_backend.cu):
 * invalid_argument, runtime_error, overflow_error, domain_error, and the
 * conjugate-gradient failure kinds 1-5 (cpp/core/conjugate_gradient.hpp:91-98)
 * offset by 10. */
typedef enum {
  MSIM_OK = 0,
  MSIM_INVALID_ARGUMENT = 1, /* std::invalid_argument */
  MSIM_RUNTIME_ERROR = 2,    /* std::runtime_error (e.g. outlet unreachable) */
  MSIM_NOT_CONVERGED = 3,    /* runtime_error "conjugate gradient did not converge" */
  MSIM_OVERFLOW = 4,         /* std::overflow_error (uint32 index space) */
  MSIM_DOMAIN_ERROR = 5,     /* negative or non-finite concentration */
  MSIM_DEVICE_ERROR = 6,     /* a HIP runtime call failed */
  MSIM_NO_DEVICE = 7,        /* no such device */
  MSIM_UNSUPPORTED = 8,      /* outside what this backend implements */
  MSIM_CG_CURVATURE = 11,    /* ConjugateGradientFailure::curvature */
  MSIM_CG_STEP = 12,         /* ::step */
  MSIM_CG_RESIDUAL = 13,     /* ::residual */
  MSIM_CG_BREAKDOWN = 14,    /* ::breakdown */
  MSIM_CG_DIRECTION = 15     /* ::direction */
} msim_status;

/* GridBoundaryKind, cpp/include/microsimulator/signals.hpp:17-21. */
enum { MSIM_NO_FLUX = 0, MSIM_PERIODIC = 1, MSIM_FIXED = 2 };

/* Pressure solver forms. */
enum {
  MSIM_SOLVER_JACOBI_FP64 = 0, /* the reference's FP64 Jacobi-PCG, on the device */
  MSIM_SOLVER_MIXED = 1,       /* FP32 Jacobi-PCG inside FP64 iterative refinement */
  MSIM_SOLVER_LEARNED = 2,     /* learned multilevel preconditioner in flexible CG */
  MSIM_SOLVER_MULTIGRID = 3    /* the same multilevel cycle without the learned smoother */
};

typedef struct msim_ctx msim_ctx;

/* The lattice of a SignalGridSpec (signals.hpp:122-156): shape, spacing,
 * obstacle mask (uint8 0/1 per site, or NULL) and the six boundary kinds in
 * the order x_lower, x_upper, y_lower, y_upper, z_lower, z_upper. */
typedef struct {
  uint32_t nx, ny, nz;
  float dx, dy, dz;
  const uint8_t* obstacles;
  uint8_t boundary[6];
} msim_grid;

/* DepthAveragedFlowParameters (cpp/include/microsimulator/flow.hpp:18-28)
 * plus the controls of the device solver forms. */
typedef struct {
  float mean_inlet_speed;      /* target mean speed over open inlet faces (may be 0 with a
                                  body force) */
  uint32_t axis;               /* FlowAxis: 0 = x, 1 = y */
  double relative_tolerance;   /* on ||b - Ax|| / ||b|| */
  uint32_t max_iterations;     /* Krylov iterations, summed over inner solves */
  uint32_t solver;             /* MSIM_SOLVER_* */
  uint32_t scale;              /* 1: scale to mean_inlet_speed as the reference does;
                                  0: return the raw Dirichlet (1 -> 0) field */
  uint32_t check_interval;     /* iterations between host convergence checks; 0 = the
                                  form's own: 64 for Jacobi-PCG (FP64 and MIXED), 8 for
                                  flexible CG (MULTIGRID, LEARNED) */
  double inner_tolerance;      /* MIXED/LEARNED: inner FP32 relative tolerance; 0 = auto */
  uint32_t fallback_iterations;/* LEARNED: iteration cap before the Jacobi-PCG fallback */
  uint32_t profile;            /* 1: time every Krylov kernel launch with events */
  /* DepthAveragedFlowParameters::body_force (flow.hpp:24-25): the in-plane body
   * force per column in pressure-gradient units, (fx, fy) pairs, x-major,
   * y-minor; NULL for none. A forced flow solves the system a second time with
   * every boundary pressure zero and the body flux's divergence as its
   * right-hand side (solve_shallow_axis_flow, flow_system.hpp:1420-1447); the
   * driven field is added after scaling, which then meets the mean inlet speed
   * with both (scale_velocity, 1368-1415). Iterations add over both solves and
   * the residuals are the larger. Needs scale = 1. */
  const float* body_force;
  size_t body_force_columns;   /* pairs in body_force: the grid's nx*ny */
} msim_flow_params;

/* DepthAveragedFlowReport (flow.hpp:30-35) plus what the device forms add. */
typedef struct {
  uint32_t iterations;           /* Krylov iterations, all inner solves */
  uint32_t refinements;          /* FP64 refinement steps (MIXED/LEARNED) */
  double relative_residual;      /* recursive residual where the solve stopped */
  double true_relative_residual; /* ||b - Ax|| / ||b|| recomputed in FP64 */
  float mean_inlet_speed;
  float max_speed;
  float solved_mean;             /* mean inlet speed of the unscaled field */
  float factor;                  /* scale applied to the unscaled field */
  uint32_t fell_back;            /* LEARNED: 1 when Jacobi-PCG finished the solve */
  double seconds;                /* device time of the solve and reconstruction */
  /* With params.profile: summed event time and launch count of the two
   * Krylov kernels (operator+direction, and update+dots). */
  double apply_seconds;
  uint32_t apply_launches;
  double update_seconds;
  uint32_t update_launches;
  /* MULTIGRID/LEARNED with params.profile: summed event time of the finest
   * level's down pass, its up pass, and the rest of the cycle, per cycle. */
  double cycle_down_seconds;
  double cycle_up_seconds;
  double cycle_rest_seconds;
  uint32_t cycles;
} msim_flow_report;

/* Device count through hipGetDeviceCount; 0 on error, as
 * cuda_backend_device_count does (cpp/cuda/cuda_backend.cu:1533-1541).
 * Replaces backend_device_count(BackendKind::cuda), backend.hpp:62. */
int msim_hip_device_count(void);

/* Creates a context on one device (one stream, default priority). Replaces
 * make_cuda_backend(device_index), cpp/include/microsimulator/backend.hpp:60. */
int msim_hip_create(int device, msim_ctx** out);

/* Frees the context and all device memory it holds. */
void msim_hip_destroy(msim_ctx* ctx);

/* Message of the last failure on ctx ("" when none). */
const char* msim_hip_last_error(const msim_ctx* ctx);

/* BackendInfo::device (types.hpp:77-83): the device's name. */
int msim_hip_device_name(msim_ctx* ctx, char* name, size_t length);

/* ComputeBackend::solve_depth_averaged_flow (backend.hpp:44-46): host arrays
 * in, the velocity field out. mobility: one float per site or NULL (the
 * reference's Hele-Shaw default). x/y/z_faces: face velocities. A grid of
 * several layers is solved on its grid of columns and lifted onto the voxel
 * faces (ShallowFlowReduction, flow_system.hpp:257-470). Equivalent to
 * prepare + solve + download. */
int msim_hip_solve_depth_averaged_flow(msim_ctx* ctx, const msim_grid* grid,
                                       const float* mobility, const msim_flow_params* params,
                                       float* x_faces, float* y_faces, float* z_faces,
                                       msim_flow_report* report);

/* The same solve split so that the system stays on the device between
 * solves (as cuda_backend.cu's sync_grid keeps a grid's arrays):
 *  prepare  -- validates (validate_flow_grid, cpu_flow.cpp:288-336), uploads
 *              the mask and mobility, assembles the operator on the device
 *              (flow_system.hpp:522-636);
 *  solve    -- pressure solve, FP64 true residual, velocity reconstruction,
 *              lift and scaling, all on the device;
 *  download -- pressure (N doubles, zero at inactive sites) and/or faces
 *              (any pointer may be NULL). */
int msim_hip_flow_prepare(msim_ctx* ctx, const msim_grid* grid, const float* mobility,
                          uint32_t axis);
int msim_hip_flow_solve(msim_ctx* ctx, const msim_flow_params* params, msim_flow_report* report);
int msim_hip_flow_download(msim_ctx* ctx, double* pressure, float* x_faces, float* y_faces,
                           float* z_faces);

/* DepthAveragedFlowSystem::apply (flow_system.hpp:670-695) on the prepared
 * system: y = A x for host vectors of N doubles (zero rows at inactive
 * sites). For checking the device operator. */
int msim_hip_flow_apply(msim_ctx* ctx, const double* x, double* y);

/* One application of the multilevel preconditioner of the MULTIGRID solver
 * (no counterpart in the reference) to a host residual of N floats: z = M(r).
 * For checking the device cycle against a CPU restatement. */
int msim_hip_flow_precondition(msim_ctx* ctx, const float* r, float* z);

/* The assembled per-site arrays of the prepared system (N floats each, any
 * pointer may be NULL): column conductance, Jacobi diagonal, right-hand side
 * and boundary conductance (flow_system.hpp:549-599). */
int msim_hip_flow_arrays(msim_ctx* ctx, float* conductance, float* diagonal, float* rhs,
                         float* boundary_conductance);

/* FlowPortKind (cpp/include/microsimulator/flow.hpp:69-74). */
enum { MSIM_PORT_SPEED = 0, MSIM_PORT_PRESSURE = 1 };

/* FlowPort (flow.hpp:76-91): an opening over a rectangle of one fixed face.
 * face: 0..5 in the order of msim_grid.boundary. begin/end: the rectangle's
 * in-plane indices (y, z on an x face; x, z on a y face; x, y on a z face),
 * end exclusive. value: for a speed port the mean speed normal to the face
 * over its open area, positive into the grid; for a pressure port the
 * pressure held just outside the face. */
typedef struct {
  uint32_t face;
  uint32_t begin[2];
  uint32_t end[2];
  uint32_t kind; /* MSIM_PORT_SPEED or MSIM_PORT_PRESSURE */
  float value;
} msim_flow_port;

/* FlowPortState (flow.hpp:93-100): the pressure held outside the port, the
 * volume entering the grid through it per unit time (negative when leaving)
 * and that flux over its open area. */
typedef struct {
  float pressure;
  float flux;
  float mean_speed;
} msim_port_state;

/* DepthAveragedPortFlowParameters (flow.hpp:102-107) plus the device form of
 * each pressure solve. */
typedef struct {
  float relative_tolerance;     /* per pressure solve; at least float epsilon */
  uint32_t max_iterations;      /* per pressure solve; positive */
  uint32_t solver;              /* MSIM_SOLVER_*: JACOBI_FP64 is the reference's PCG */
  uint32_t check_interval;      /* as msim_flow_params */
  double inner_tolerance;       /* MIXED/LEARNED inner tolerance; 0 = auto */
  uint32_t fallback_iterations; /* LEARNED: cap before the Jacobi-PCG fallback */
} msim_port_flow_params;

/* DepthAveragedPortFlowReport (flow.hpp:109-114) plus device figures. */
typedef struct {
  uint32_t solves;               /* pressure solves combined into the field */
  uint32_t iterations;           /* Krylov iterations summed over the solves */
  float relative_residual;       /* largest over the solves */
  float max_speed;               /* largest |velocity| of the combined field */
  double true_relative_residual; /* largest FP64 ||b - Ax|| / ||b|| over the solves */
  double seconds;                /* device time of the solves, lifts and combination */
} msim_port_flow_report;

/* ComputeBackend::solve_depth_averaged_port_flow (backend.hpp:47-49), as
 * solve_depth_averaged_port_flow_cpu (cpp/cpu/cpu_flow.cpp:353-375) over
 * ShallowPortFlow (cpp/core/flow_system.hpp:1268-1311): the ports' checks
 * (validate_flow_ports, cpu_flow.cpp:193-286, on the grid and on its grid of
 * columns), one pressure solve on the device per speed port held alone at
 * pressure one plus one for the pressure ports' own pressures when any is
 * nonzero, each lifted onto the voxel faces on the device, and the
 * combination that carries each speed port's stated mean speed
 * (superpose_port_flows, flow_system.hpp:1139-1252): the small dense system
 * for the speed ports' pressures is solved on the host, the field combined
 * on the device. mobility: one float per site or NULL. x/y/z_faces: the
 * combined face velocities (any may be NULL). states: port_count entries
 * (may be NULL). The prepared flow system is replaced. */
int msim_hip_solve_depth_averaged_port_flow(msim_ctx* ctx, const msim_grid* grid,
                                            const float* mobility, const msim_flow_port* ports,
                                            uint32_t port_count,
                                            const msim_port_flow_params* params, float* x_faces,
                                            float* y_faces, float* z_faces,
                                            msim_port_flow_report* report,
                                            msim_port_state* states);

/* ResolvedFlowParameters (cpp/include/microsimulator/flow.hpp:42-51). */
typedef struct {
  float mean_inlet_speed;          /* nonzero */
  uint32_t axis;                   /* 0 x, 1 y, 2 z */
  float relative_tolerance;        /* FGMRES, >= float epsilon */
  uint32_t max_outer_iterations;
  float inner_relative_tolerance;  /* the momentum block's CG, >= float epsilon */
  uint32_t max_inner_iterations;
} msim_resolved_params;

/* ResolvedFlowReport (cpp/include/microsimulator/flow.hpp:53-62), and the
 * device time of the solve. */
typedef struct {
  uint32_t outer_iterations;
  uint64_t inner_iterations;
  float relative_residual;
  float momentum_relative_residual;
  float divergence_rms;
  float mean_inlet_speed;
  float max_speed;
  uint32_t min_gap_voxels;
  double seconds;
} msim_resolved_report;

/* ComputeBackend::solve_resolved_flow (backend.hpp:50-52; the reference's
 * solve_resolved_flow_cpu, cpp/cpu/cpu_flow.cpp:379-467): the resolved
 * Stokes-Brinkman flow of the voxel grid along an axis, velocity on every
 * face and pressure at every site, by flexible GMRES (restarts of 40,
 * cpp/core/flexible_gmres.hpp:33-108) preconditioned by Jacobi-PCG on the
 * momentum block and the scaled pressure diagonal, all vectors FP64 on the
 * device; the field scaled to the mean inlet speed. drag: one float per site
 * (the Brinkman drag) or NULL (zero). x/y/z_faces: the field (any may be
 * NULL). Errors as the reference's: invalid arguments, a grid that carries no
 * flow, and the inner or outer solve failing to converge (MSIM_NOT_CONVERGED
 * or MSIM_RUNTIME_ERROR with its message). */
int msim_hip_solve_resolved_flow(msim_ctx* ctx, const msim_grid* grid, const float* drag,
                                 const msim_resolved_params* params, float* x_faces,
                                 float* y_faces, float* z_faces, msim_resolved_report* report);

/* ResolvedPortFlowParameters (cpp/include/microsimulator/flow.hpp:125-132). */
typedef struct {
  float relative_tolerance;        /* FGMRES, >= float epsilon */
  uint32_t max_outer_iterations;   /* per block solve (the reference's default 5000) */
  float inner_relative_tolerance;  /* the momentum block's CG, >= float epsilon */
  uint32_t max_inner_iterations;
} msim_resolved_port_params;

/* ResolvedPortFlowReport (cpp/include/microsimulator/flow.hpp:134-147):
 * iteration counts summed and residuals the largest over the block solves;
 * the divergence and gap of the combined field; and the device time of the
 * block solves. */
typedef struct {
  uint32_t solves;
  uint32_t outer_iterations;
  uint64_t inner_iterations;
  float relative_residual;
  float momentum_relative_residual;
  float divergence_rms;
  float max_speed;
  uint32_t min_gap_voxels;
  double seconds;
} msim_resolved_port_report;

/* ComputeBackend::solve_resolved_port_flow (backend.hpp:53-55; the
 * reference's solve_resolved_port_flow_cpu, cpp/cpu/cpu_flow.cpp:474-479,
 * over detail::solve_resolved_port_flow, cpp/core/flow_system.hpp:1482-1548):
 * the resolved Stokes-Brinkman flow driven through ports, the combination of
 * one block solve on the device per speed port held alone at pressure one
 * (plus one for the pressure ports' own pressures when any is nonzero) that
 * carries each speed port's stated mean speed (superpose_port_flows,
 * 1146-1252). Where shallow flow describes the geometry (shallow_port_pressure,
 * 1318-1338), each block solve is for the correction to the depth-averaged
 * pressure of the same ports, itself an FP64 Jacobi-PCG solve on the device.
 * drag: one float per site or NULL. x/y/z_faces: the combined field (any may
 * be NULL). states: port_count entries (may be NULL). The context's prepared
 * depth-averaged flow system is replaced. Errors as the reference's. */
int msim_hip_solve_resolved_port_flow(msim_ctx* ctx, const msim_grid* grid, const float* drag,
                                      const msim_flow_port* ports, uint32_t port_count,
                                      const msim_resolved_port_params* params, float* x_faces,
                                      float* y_faces, float* z_faces,
                                      msim_resolved_port_report* report,
                                      msim_port_state* states);

/* A signal grid for the explicit transport step: SignalGridSpec's transport
 * fields (signals.hpp:122-156) flattened. boundary_values: 6*S reservoir
 * values, face f of signal s at [6*s + f] (uniform per face). Velocity field
 * faces may be NULL (no advection). Affine reaction (SignalGridAffineReaction,
 * signals.hpp:109-114) may be NULL. Monod uptake -vmax*phi*c/(K+c), with phi
 * per site and vmax, K per signal, may be NULL. revision: SignalGrid::revision
 * (signals.hpp:217-221); the device copies of the spec arrays are rewritten
 * only when it changes. */
/* GridBoundaryPatch (signals.hpp:72-84): a rectangle of a fixed face that
 * holds its own reservoir value per signal, replacing the face's values there;
 * where patches overlap the later one holds (GridBoundary::values_at,
 * signals.cpp:439-447). In-plane indices: on an x face (y, z), on a y face
 * (x, z), on a z face (x, y); begin inclusive, end exclusive. */
typedef struct {
  uint32_t face; /* 0..5: x_lower, x_upper, y_lower, y_upper, z_lower, z_upper */
  uint32_t begin[2];
  uint32_t end[2];
  const float* values; /* signal_count values */
} msim_boundary_patch;

typedef struct {
  msim_grid grid;
  uint32_t signal_count;
  const float* diffusion;
  const float* boundary_values;
  const float* x_faces;
  const float* y_faces;
  const float* z_faces;
  const float* source_rates;
  const float* loss_rates;
  const float* phi;
  const float* monod_vmax;
  const float* monod_k;
  uint64_t revision;
  uint32_t advection_scheme; /* MSIM_UPWIND or MSIM_VAN_LEER (SignalAdvectionKind, signals.hpp:41-47) */
  const msim_boundary_patch* patches; /* reservoir patches of fixed faces, or NULL */
  uint32_t patch_count;
  /* SignalIntegrationKind (signals.hpp:33-37) and, for the implicit kinds,
   * SignalSolveParameters (signals.hpp:49-55; the reference's defaults are
   * 10000, 1e-7, 1e-5). The implicit kinds hold FP32 levels. */
  uint32_t integration; /* MSIM_FORWARD_EULER, MSIM_CRANK_NICOLSON or MSIM_BACKWARD_EULER */
  uint32_t solver_max_iterations;
  float solver_absolute_tolerance;
  float solver_relative_tolerance;
  /* SignalGridSpec::advection (signals.hpp:128): a constant velocity (x, y, z)
   * per signal, 3*signal_count floats, or NULL for none. Zero with a velocity
   * field. */
  const float* advection;
} msim_signal_grid;

/* SignalIntegrationKind (signals.hpp:33-37). The implicit kinds solve each
 * step by the reference's damped Jacobi iteration
 * (signal_grid_crank_nicolson_candidate, signals.cpp:1293-1375) under its
 * stopping rule (solve_signal_jacobi, signal_jacobi.hpp:53-84); Monod uptake
 * enters them as the loss vmax*phi/(K+c) of the step's starting levels, added
 * to any affine loss. */
enum { MSIM_FORWARD_EULER = 0, MSIM_CRANK_NICOLSON = 1, MSIM_BACKWARD_EULER = 2 };

/* SignalAdvectionKind (signals.hpp:41-47): first-order upwind, or upwind plus
 * the van Leer limited correction (signals.cpp:1227-1286), whose explicit
 * bound halves the advective step (signals.cpp:1089-1094) and whose levels
 * are settled after each step (settle_step_levels, signals.cpp:979-1001). */
enum { MSIM_UPWIND = 0, MSIM_VAN_LEER = 1 };

/* SignalSolveReport (signals.hpp:57-61). */
typedef struct {
  uint32_t converged;
  uint32_t iterations;
  float residual_rms;
} msim_signal_report;

/* ComputeBackend::advance_signal_grid (backend.hpp:30, advance_signal_grid_cpu,
 * signals.cpp:1377-1389): levels S*N in place. A forward-Euler step is checked
 * against the explicit stability bound first (SignalGrid::validate_step,
 * signals.cpp:1066-1102); an implicit step that does not converge fails with
 * MSIM_NOT_CONVERGED and the reference's message. report: the implicit
 * solve's (converged, iterations, residual RMS); {1, 0, 0} for forward Euler.
 * A refused or failed step leaves levels unchanged. */
int msim_hip_advance_signal_grid(msim_ctx* ctx, const msim_signal_grid* grid, float dt,
                                 float* levels, msim_signal_report* report);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): msim_hip_advance_signal_grid's forward-Euler step, with the same
 * arguments, checks, messages and results, that allocates nothing. The spec's
 * arrays, both level buffers and every partial sum live in one static device
 * array of msim_hip_static_scratch_bytes() bytes (64 MiB) that each device
 * holds until the process ends; a grid that needs more is refused with
 * MSIM_INVALID_ARGUMENT. Forward Euler only, no boundary patches (both
 * MSIM_UNSUPPORTED). Every call on one device shares that array and rewrites
 * what it reads: calls from one process are serialised per device, but calls
 * that overlap on one device from outside that rule (another copy of the
 * library in the same process) race. */
int msim_hip_advance_signal_grid_static(msim_ctx* ctx, const msim_signal_grid* grid, float dt,
                                        float* levels, msim_signal_report* report);
size_t msim_hip_static_scratch_bytes(void);

/* The nutrient loop of the reference's nutrient study
 * (scripts/run_nutrient_benchmarks.py:96-120) kept on the device, on a grid of
 * one layer: per step
 *  1. the mobility of the occupancy phi by the reference's Brinkman closure,
 *     base / (1 + base * C * phi^2 / (1 - phi)^3) with phi capped
 *     (_darcy_coefficients, python/src/microsimulator/flow.py:361-384);
 *  2. the operator and its multilevel hierarchy rebuilt from it, the pressure
 *     solve of `flow` (its form, tolerance and scaling) and the velocity;
 *  3. the explicit bound of the new field, then one FP64 forward-Euler step of
 *     the uploaded signals over it, with Monod uptake on phi and the step's FP64
 *     ledger row (as msim_hip_adr_run's);
 *  4. phi grown by yield x dt x the step's uptake, sum over signals of
 *     vmax phi c / (K + c) at the levels the step started from.
 * The reference's script grows cells from the nutrient they took up and
 * recomputes the colony's mobility before each solve; here the biomass is the
 * grid's occupancy. Needs msim_hip_flow_prepare (the grid and axis) and
 * msim_hip_adr_upload (FP64 levels, phi with Monod, no affine reaction,
 * use_flow_field = 1) first. ledger: steps*S*4 doubles or NULL; per_step:
 * `steps` records or NULL. A refused step stops the run, keeping the state the
 * step started from. */
typedef struct {
  msim_flow_params flow;       /* the solve of every step */
  const float* base_mobility;  /* per site: the mobility without biomass (zero at solids) */
  float drag_coefficient;      /* C (the reference's default 100) */
  float max_volume_fraction;   /* phi cap inside the drag (0.9) */
  float yield;                 /* occupancy gained per unit of solute taken up */
  float dt;                    /* transport step */
} msim_nutrient_params;

typedef struct {
  uint32_t iterations;            /* Krylov iterations of the step's pressure solve */
  uint32_t fell_back;             /* the learned solver's Jacobi-PCG fallback ran */
  double true_relative_residual;  /* FP64 true residual of the solve */
  float max_speed;
  float max_phi;                  /* largest occupancy after the step */
  double mass_balance;            /* largest |after - before - supply + uptake| / after */
  double seconds;                 /* device time of the step */
  double solve_seconds;           /* of which the pressure solve and velocity */
} msim_nutrient_step;

int msim_hip_nutrient_run(msim_ctx* ctx, const msim_nutrient_params* params, uint32_t steps,
                          double* ledger, msim_nutrient_step* per_step);

/* No counterpart in the reference: the device's achievable HBM bandwidth, as
 * GB/s read plus written by `repeats` STREAM-style copies of `bytes` (16-byte
 * loads and stores, four in flight per thread), timed with HIP events, or by
 * the runtime's device-to-device copy where that is faster; the ceiling the
 * stencil kernels are measured against beside the published 8 TB/s. */
int msim_hip_copy_bandwidth(msim_ctx* ctx, size_t bytes, uint32_t repeats, double* gb_per_s);

/* The Monod occupancy (N floats) the uploaded transport reads: the one uploaded,
 * or the one a nutrient run grew. */
int msim_hip_adr_occupancy(msim_ctx* ctx, float* phi);

/* Device-resident transport for many coupled steps.
 *  adr_upload   -- spec arrays and initial levels (float when precision is
 *                  32, double when 64). With use_flow_field != 0 the velocity
 *                  field is the one the last flow solve left on the device.
 *  adr_run      -- `steps` steps of dt of the spec's integration kind (the
 *                  implicit kinds need precision 32). ledger (steps*S*4 doubles,
 *                  or NULL) receives per step and signal: mass before, mass
 *                  after, boundary supply and uptake, each in FP64 (mass =
 *                  sum of levels x voxel volume). A run refused at step k
 *                  keeps the levels after steps 0..k-1 and fills their
 *                  ledger rows; the next run's ledger carries on from them.
 *                  A field read in place is refused once any later flow solve
 *                  or preparation has run on the context.
 *  adr_download -- levels in the uploaded precision. */
int msim_hip_adr_upload(msim_ctx* ctx, const msim_signal_grid* grid, const void* levels,
                        uint32_t precision, uint32_t use_flow_field);
int msim_hip_adr_run(msim_ctx* ctx, float dt, uint32_t steps, double* ledger);
int msim_hip_adr_download(msim_ctx* ctx, void* levels);

/* Device time (seconds) of the kernels of the last adr_run. */
double msim_hip_adr_seconds(msim_ctx* ctx);

/* The SignalSolveReport (signals.hpp:57-61) of the last implicit step that
 * adr_run took ({1, 0, 0} after forward-Euler steps). No counterpart in the
 * reference, whose advance returns it per step. */
int msim_hip_adr_report(msim_ctx* ctx, msim_signal_report* report);

/* ---------------------------------------------------------------------------
 * Coupled cell/grid exchange (ComputeBackend::advance_coupled, backend.hpp:31-34;
 * advance_coupled_cpu, cpp/cpu/cpu_coupled.cpp:125-231), on one layer or
 * several, FP32 levels, at most 8 signals, and at most 256 rate instructions,
 * signals and species together.
 *
 * Each cell samples the grid's starting levels through its trilinear stencil
 * (signal_grid_stencil, signals.cpp:357-420: with obstacles only the fluid
 * corners face-connected to the strongest one, renormalised), dilutes its
 * species by previous/current effective volume (pi r^2 (L + 2r),
 * species.cpp:134-137), runs the rate plan, adds dt times its species
 * outputs, and returns each signal output divided by the voxel volume through
 * the same stencil as a per-site source, summed over cells in cell order. The
 * grid then steps with those sources: forward Euler adds dt * source to the
 * explicit candidate, the implicit kinds add it to the right-hand side
 * (signals.cpp:1295-1310). The species are committed only when the grid step
 * is accepted. */

/* RateOp (cpp/include/microsimulator/species.hpp:12-16, codes of rate_ops.def). */
enum {
  MSIM_RATE_CONSTANT = 0, MSIM_RATE_SPECIES, MSIM_RATE_POSITION_X, MSIM_RATE_POSITION_Y,
  MSIM_RATE_POSITION_Z, MSIM_RATE_LENGTH, MSIM_RATE_RADIUS, MSIM_RATE_GROWTH_RATE,
  MSIM_RATE_CELL_TYPE, MSIM_RATE_VOLUME, MSIM_RATE_SURFACE_AREA, MSIM_RATE_ADD,
  MSIM_RATE_SUBTRACT, MSIM_RATE_MULTIPLY, MSIM_RATE_DIVIDE, MSIM_RATE_POWER, MSIM_RATE_MINIMUM,
  MSIM_RATE_MAXIMUM, MSIM_RATE_NEGATE, MSIM_RATE_EXPONENTIAL, MSIM_RATE_LOGARITHM,
  MSIM_RATE_LESS, MSIM_RATE_LESS_EQUAL, MSIM_RATE_GREATER, MSIM_RATE_GREATER_EQUAL,
  MSIM_RATE_EQUAL, MSIM_RATE_SELECT, MSIM_RATE_SIGNAL, MSIM_RATE_VOLUME_CHANGE_RATE
};

/* RateInstruction (species.hpp:18-24): inputs are earlier instructions'
 * values (first, second, third), a species or signal index (first), or the
 * constant `value`. */
typedef struct {
  uint32_t operation;
  uint32_t first;
  uint32_t second;
  uint32_t third;
  float value;
} msim_rate_instruction;

/* CoupledRatePlan (coupled_rates.hpp; checks of CoupledRatePlan::validate,
 * cpp/core/coupled_rates.cpp:60-146). */
typedef struct {
  uint32_t species_count;
  uint32_t signal_count;
  uint32_t instruction_count;
  const msim_rate_instruction* instructions;
  const uint32_t* species_outputs; /* species_count instruction indices */
  const uint32_t* signal_outputs;  /* signal_count instruction indices */
} msim_rate_plan;

/* The cells of a WorldState (world_state.hpp: geometry and attributes), one
 * array entry per cell. Species levels are a separate count x species_count
 * array, cell-major (WorldState::species_state). */
typedef struct {
  uint32_t count;
  const float* position_x;
  const float* position_y;
  const float* position_z;
  const float* lengths;
  const float* radii;
  const float* growth_rates;
  const int32_t* cell_types;
} msim_cells;

/* ComputeBackend::advance_coupled: one coupled step of dt. origin: the grid's
 * origin (SignalGridSpec::origin). levels (S*N) and species
 * (count*species_count) in place; previous_lengths: each cell's length before
 * this step's growth. The reference's checks and messages, in its order
 * (validate_coupled_step, cpu_coupled.cpp:90-123). report as
 * msim_hip_advance_signal_grid's. A refused or failed step leaves levels and
 * species unchanged. */
int msim_hip_advance_coupled(msim_ctx* ctx, const msim_signal_grid* grid, const float* origin,
                             float* levels, const msim_cells* cells, float* species,
                             const msim_rate_plan* plan, const float* previous_lengths,
                             float dt, msim_signal_report* report);

/* Device-resident coupled steps (no counterpart in the reference, which holds
 * the world state on the host): the same step as msim_hip_advance_coupled,
 * many times over cells that stay in place.
 *  coupled_upload   -- grid spec, levels, cells, species and plan, checked as
 *                      msim_hip_advance_coupled checks them.
 *  coupled_run      -- `steps` coupled steps of dt; the first step's previous
 *                      lengths are previous_lengths (NULL: the cells' own
 *                      lengths), later steps' the cells' lengths. Stops at the
 *                      first refused step, whose starting state is kept.
 *  coupled_download -- levels (S*N) and species (count*species_count); either
 *                      may be NULL. */
int msim_hip_coupled_upload(msim_ctx* ctx, const msim_signal_grid* grid, const float* origin,
                            const float* levels, const msim_cells* cells, const float* species,
                            const msim_rate_plan* plan);
int msim_hip_coupled_run(msim_ctx* ctx, const float* previous_lengths, float dt, uint32_t steps,
                         msim_signal_report* report);
int msim_hip_coupled_download(msim_ctx* ctx, float* levels, float* species);

/* Device time (seconds) of the kernels of the last coupled_run. */
double msim_hip_coupled_seconds(msim_ctx* ctx);

/* colony_volume_fraction (python/src/microsimulator/flow.py:282-305): each
 * cell's biomass volume pi r^2 (L + 2r) (biomass.py:6-14) spread over the
 * voxel integrals of a tent kernel of radius averaging_radius around it,
 * renormalised over the face-connected fluid component of its strongest voxel
 * (_deposition_weights, flow.py:224-279), divided by the voxel volume: density
 * (N doubles, site order). Cells outside the grid's extent are omitted. Only
 * positions, lengths and radii of `cells` are read (growth_rates and
 * cell_types may be NULL). The voxels' sums are FP64 atomics: they agree with
 * the reference's cell-ordered sums to FP64 rounding. Any number of layers. */
int msim_hip_colony_volume_fraction(msim_ctx* ctx, const msim_grid* grid, const float* origin,
                                    const msim_cells* cells, double averaging_radius,
                                    double* density);

/* colony_mobility (flow.py:334-382): base / (1 + base * drag) at each fluid
 * site (0 at solids), drag = C f^2 / (1 - f)^3 with f the column's mean
 * volume fraction over its fluid voxels, capped at max_volume_fraction.
 * base_count 0: base[0] is one value for every site; otherwise base holds
 * base_count values, one per site. mobility: N doubles. */
int msim_hip_colony_mobility(msim_ctx* ctx, const msim_grid* grid, const float* origin,
                             const msim_cells* cells, const double* base, uint32_t base_count,
                             double drag_coefficient, double max_volume_fraction,
                             double averaging_radius, double* mobility);

/* colony_flow_coupling (python/src/microsimulator/flow.py:384-424): the
 * colony's mobility as msim_hip_colony_mobility computes it, and the in-plane
 * body force per column for drag proportional to the fluid/biomass relative
 * velocity: each cell's deposited biomass per column times its in-plane
 * velocity, summed into the column's momentum, times drag / mass (mass the
 * column's deposited volume; zero force where it has none). velocities: 2
 * floats per cell (x, y). force: 2 doubles per column (x-major, y-minor), the
 * pairs DepthAveragedFlowParameters::body_force takes. Checks in the
 * reference's order: the radius, then per cell a finite velocity before the
 * deposit's own checks, then the mobility's. */
int msim_hip_colony_flow_coupling(msim_ctx* ctx, const msim_grid* grid, const float* origin,
                                  const msim_cells* cells, const float* velocities,
                                  const double* base, uint32_t base_count,
                                  double drag_coefficient, double max_volume_fraction,
                                  double averaging_radius, double* mobility, double* force);

/* Device time (seconds) of the last deposit's kernels. */
double msim_hip_colony_seconds(msim_ctx* ctx);

/* ---------------------------------------------------------------------------
 * Learned smoother of the LEARNED solver (no counterpart in the reference,
 * whose only preconditioner is Jacobi, cpp/cpu/cpu_flow.cpp:65-69).
 *
 * The finest level of the multilevel cycle replaces its damped-Jacobi
 * post-smoother z = x + 0.8 D^-1 r' by z = x + 0.8 D^-1 r' + s N(r'), where
 * N is a three-layer 3x3 convolutional network (4 -> 16 -> 16 -> 1 channels,
 * softsign, no biases) applied on 16 x 64 tiles with a 3-site halo, s is the
 * tile's largest |D^-1 r'| and the inputs are D^-1 r' / s times (1, the
 * site's x and y couplings and boundary term over the diagonal). Weights
 * are MSIM_LEARNED_WEIGHTS floats: W1[9][4][16], W2[9][16][16], W3[9][16]
 * ([tap][input][output], tap = 3 (dx + 1) + (dy + 1)). They run as bfloat16
 * with FP32 accumulation on the matrix cores; training keeps FP32 weights.
 * ------------------------------------------------------------------------- */
#define MSIM_LEARNED_WEIGHTS 3024

/* Sets / reads the smoother's weights (FP32). */
int msim_hip_learned_set_weights(msim_ctx* ctx, const float* weights, size_t count);
int msim_hip_learned_get_weights(msim_ctx* ctx, float* weights, size_t count);

/* One application of the LEARNED solver's cycle to a host residual of N
 * floats (needs weights and a prepared system). For checking. */
int msim_hip_learned_precondition(msim_ctx* ctx, const float* r, float* z);

/* The finest level's operator of the prepared system, N floats each (any
 * pointer may be NULL): coupling to the lower x and y neighbour, boundary
 * term, inverse diagonal (0 at inactive sites). */
int msim_hip_learned_operator(msim_ctx* ctx, float* wx, float* wy, float* bc, float* dinv);

/* Harvests training pairs from the prepared system. A flexible-CG solve of
 * the system (multilevel cycle; with the learned smoother when use_learned
 * is nonzero) gives one residual r per iteration; up to max_pairs of them,
 * spread evenly over the solve, are kept. For each, an accurate solve gives
 * e with A e = r, and the pair is taken after the cycle's pre-smoother and
 * coarse correction x: r' = r - A x, e' = e - x.
 * residuals and errors: max_pairs * N floats; *produced: pairs written. */
int msim_hip_learned_harvest(msim_ctx* ctx, uint32_t max_pairs, uint32_t use_learned,
                             float* residuals, float* errors, uint32_t* produced);

/* A training set: `geometries` operators of one nx x ny shape, as
 * operators[(g * 4 + k) * N + site] with k = wx, wy, bc, dinv; `pairs`
 * residual/error pairs (pairs * N floats each) and the geometry of each. */
typedef struct {
  uint32_t nx, ny;
  uint32_t geometries, pairs;
  const float* operators;
  const uint32_t* pair_geometry;
  const float* residuals;
  const float* errors;
} msim_train_set;

/* Adam on the mean over a batch of (e' - z)^T A (e' - z) / e'^T A e',
 * z = 0.8 D^-1 r' + s N(r'). Batches walk seeded shuffles of the set. */
typedef struct {
  uint32_t steps;
  uint32_t batch;
  float learning_rate;
  float beta1, beta2, epsilon;
  uint64_t seed;
} msim_train_params;

/* Copies a training set to the device (replacing any earlier one). */
int msim_hip_learned_upload(msim_ctx* ctx, const msim_train_set* set);

/* Trains the current weights in place on the uploaded set, entirely on the
 * device. losses: `steps` floats (batch loss before each step) or NULL;
 * seconds: device time, or NULL. */
int msim_hip_learned_train(msim_ctx* ctx, const msim_train_params* params, float* losses,
                           double* seconds);

/* Loss of one uploaded pair and its gradient with respect to the weights
 * (MSIM_LEARNED_WEIGHTS doubles), from the device's forward and backward
 * kernels. For checking them against a CPU restatement. */
int msim_hip_learned_gradients(msim_ctx* ctx, uint32_t pair, double* loss, double* gradient);

/* z = 0.8 D^-1 r' + s N(r') for one uploaded pair (N floats). */
int msim_hip_learned_apply(msim_ctx* ctx, uint32_t pair, float* z);

/* ---------------------------------------------------------------------------
 * Batched sweep (no counterpart in the reference, which runs one job per
 * invocation, docs/architecture/0005-batch-execution.md:30): many independent
 * one-layer devices of uneven sizes advanced together. Each instance has its
 * own geometry, mobility, occupancy, inlet pressure (Dirichlet at x = 0, 0 at
 * x = nx, flow along x, no scaling) and solute diffusivity; the solute has
 * fixed reservoirs beyond x = 0 (`reservoir`) and x = nx (0), no-flux walls
 * in y and Monod uptake vmax phi c / (K + c).
 *
 * The pressure (FP32 Jacobi-PCG corrections with FP64 dot products inside
 * FP64 iterative refinement of an FP64 pressure, warm-started, to an FP64
 * true relative residual) is solved by msim_hip_batch_solve_flow: one launch,
 * one workgroup per instance. A step is one launch over work items (blocks of
 * rows, spread over the chip): each checks its rows' FP64 pressure residual,
 * rebuilds their face velocities and advances them one FP32 forward-Euler
 * transport step whose sites compute exactly as msim_hip_advance_signal_grid's
 * do. Runs of steps can be captured in a HIP graph. The launch shape
 * (workgroup size, sweep row width, sites per work item) is chosen by
 * msim_hip_batch_autotune, which times each candidate on the uploaded batch.
 * ------------------------------------------------------------------------- */
typedef struct {
  uint32_t nx, ny;
  float dx, dy, dz;
  const uint8_t* obstacles; /* nx*ny, or NULL */
  const float* mobility;    /* nx*ny, required */
  const float* phi;         /* nx*ny occupancy for Monod uptake, or NULL */
  const float* levels;      /* nx*ny initial solute levels, or NULL (zero) */
  float inlet_pressure;
  float diffusion;
  float reservoir;
} msim_batch_instance;

typedef struct {
  float relative_tolerance; /* per-instance FP64 ||b - Ax|| / ||b|| */
  uint32_t max_iterations;  /* FP32 CG iterations per instance per flow phase */
  float monod_vmax, monod_k;
} msim_batch_params;

/* Uploads the instances and assembles their operators (replacing any batch). */
int msim_hip_batch_upload(msim_ctx* ctx, uint32_t count, const msim_batch_instance* instances,
                          const msim_batch_params* params);

/* Solves every instance's pressure from its current one and rebuilds its
 * velocities (one launch). Per-instance outputs (count each, any may be
 * NULL): FP32 CG iterations, final FP64 true relative residual, flags (1 CG
 * breakdown, 2 not converged within max_iterations, 4 a negative or
 * non-finite level). */
int msim_hip_batch_solve_flow(msim_ctx* ctx, uint32_t* iterations, float* residuals,
                              uint32_t* flags);

/* Times `steps` steps of each launch shape from the current levels (which it
 * restores), keeps the fastest and returns its index in *shape. seconds:
 * per-step time of each candidate (msim_hip_batch_shapes of them), or NULL.
 * Needs a solved flow. */
int msim_hip_batch_autotune(msim_ctx* ctx, float dt, uint32_t steps, int* shape,
                            double* seconds);

/* Number of launch shapes the autotuner tries, and shape k's workgroup size,
 * sweep row width and sites per work item. */
int msim_hip_batch_shapes(void);
int msim_hip_batch_shape(int index, uint32_t* threads, uint32_t* row_width,
                         uint32_t* item_sites);

/* Per-instance flags as msim_hip_batch_solve_flow describes (count values). */
int msim_hip_batch_flags(msim_ctx* ctx, uint32_t* flags);

/* `steps` steps of dt from a solved flow. use_graph != 0: captured once in a
 * HIP graph and replayed. ledger: steps*count*4 doubles per step and instance
 * (mass before, mass after, uptake over the step -- FP64 sums over fluid
 * sites times the voxel volume -- and the FP64 true relative residual of the
 * pressure the step checked), or NULL. seconds: device time, or NULL. */
int msim_hip_batch_run(msim_ctx* ctx, float dt, uint32_t steps, uint32_t use_graph,
                       double* ledger, double* seconds);

/* One instance's pressure (nx*ny doubles), x faces ((nx+1)*ny), y faces
 * (nx*(ny+1)) and levels (nx*ny); any pointer may be NULL. */
int msim_hip_batch_download(msim_ctx* ctx, uint32_t index, double* pressure, float* x_faces,
                            float* y_faces, float* levels);

#ifdef __cplusplus
}
#endif

#endif /* MSIM_HIP_H */
