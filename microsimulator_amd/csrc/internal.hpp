// Internal host-side declarations of the HIP backend: device buffers, the
// context, and the launchers each .hip file provides. Errors inside the
// library are C++ exceptions (Error) that the C ABI turns into status codes;
// none crosses the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "msim_hip.h"

namespace msim {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& message) : std::runtime_error(message), code_(code) {}
  [[nodiscard]] int code() const noexcept { return code_; }

 private:
  int code_;
};

inline void check(hipError_t status, const char* what) {
  if (status != hipSuccess) {
    throw Error(MSIM_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(status));
  }
}

// Device memory owned by one context. Reallocated only when it must grow.
template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }

  void reserve(std::size_t count) {
    if (count <= capacity_) {
      count_ = count;
      return;
    }
    release();
    void* ptr = nullptr;
    check(hipMalloc(&ptr, (count == 0 ? 1 : count) * sizeof(T)), "hipMalloc");
    data_ = static_cast<T*>(ptr);
    capacity_ = count;
    count_ = count;
  }
  void release() {
    if (data_ != nullptr) static_cast<void>(hipFree(data_));
    data_ = nullptr;
    capacity_ = 0;
    count_ = 0;
  }
  [[nodiscard]] T* data() const noexcept { return data_; }
  [[nodiscard]] std::size_t size() const noexcept { return count_; }

 private:
  T* data_ = nullptr;
  std::size_t capacity_ = 0;
  std::size_t count_ = 0;
};

// The device-side state of one Krylov solve. One instance lives in device
// memory; the host reads it back between batches of iterations.
struct CgState {
  double rho;        // r . z of the current iterate
  double alpha;      // step length of the last iteration
  double beta;       // weight of the previous direction in the next one
  double rr;         // r . r after the last iteration
  double curvature;  // p . A p of the last iteration
  double threshold;  // (tolerance * ||b||)^2
  double rr0;        // ||b||^2 of this solve
  double true_rr;    // ||b - A x||^2 from the residual kernel
  double bb;         // ||b||^2 from the residual kernel
  double inlet_sum;  // velocity reconstruction: sum over open inlet faces
  double inlet_count;
  double peak;       // largest |velocity|
  double scaled_peak;
  float factor;      // scaling factor, device copy
  uint32_t nonfinite;
  uint32_t iterations;
  uint32_t converged;
  uint32_t failure;  // ConjugateGradientFailure (conjugate_gradient.hpp:91-98)
  uint32_t done;
  // Iteration bookkeeping of the Krylov kernels. Each field is written by one
  // kernel's block 0 and read by the other kernel, never by blocks of the
  // launch that writes it (halt aside: see k_cg_apply).
  uint32_t it_update;   // last iteration whose update kernel ran
  uint32_t it_apply;    // iteration of the last operator kernel
  uint32_t halt;        // iteration at which r.r met the threshold; 0 while running
  uint32_t pad;
  uint32_t ticket[8];
};

enum Ticket : int {
  kTicketInit = 0,
  kTicketApply = 1,
  kTicketUpdate = 2,
  kTicketResidual = 3,
  kTicketVelocity = 4,
  kTicketScale = 5,
  kTicketAdr = 6,
  kTicketMisc = 7,
};

constexpr int kMaxSignals = 8;

// Bits of the transport flags (AdrDevice::flags[0]).
constexpr uint32_t kFlagNegative = 1;  // a non-finite or (upwind) negative explicit update
constexpr uint32_t kFlagLevels = 2;    // a level negative or not finite when the step is judged
constexpr uint32_t kFlagCell = 4;      // a coupled cell's rate or species update not finite
constexpr uint32_t kFlagSources = 8;   // a coupled source rate not finite

// One level of the multilevel preconditioner: a 5-point operator on a
// (nx, ny) grid of aggregates, as couplings to the lower x and y neighbour
// (wx, wy), a boundary term (bc) and the inverse diagonal; r and e hold the
// level's residual and correction during a cycle.
struct MgLevel {
  uint32_t nx = 0, ny = 0;
  // xp: the iterate after pre-smoothing (read by the up pass); t: the
  // single-workgroup kernel's intermediate sweep.
  DeviceBuffer<float> wx, wy, bc, dinv, r, e, xp, t;
  [[nodiscard]] std::size_t cells() const { return static_cast<std::size_t>(nx) * ny; }
};

struct MgDevice {
  std::vector<std::unique_ptr<MgLevel>> levels;
  uint32_t tail = 0;          // first level the single-block kernel handles
  bool split_finest = false;  // the finest level runs on the tile kernels even when small
  bool learned_cycle = false; // the finest up pass is the learned one (its own tile grid)
  uint32_t coarse_n = 0;      // cells of the coarsest level
  DeviceBuffer<float> cinv;   // inverse of the coarsest operator (dense)
  DeviceBuffer<double> zr, zq, pq, rr;  // flexible-CG partial sums, two generations
  // Iterations the last inner flexible-CG solve of each (cycle form,
  // refinement) took: the next solve launches that many plus one before its
  // first host check (0: no record).
  uint32_t fcg_expect[3][8] = {};
  bool ready = false;
};

// The boundary of a depth-averaged pressure system on its grid of columns.
// With `open` null: both ends of `axis` open, the lower held at pressure one
// and the upper at zero (axis_flow_boundary, flow_system.hpp:185-201).
// Otherwise, per exterior in-plane face, whether it is open and the pressure
// held half a spacing outside it (port_flow_boundary, 238-253), in device
// arrays of slots [x_lower ny | x_upper ny | y_lower nx | y_upper nx].
struct FlowEdges {
  const uint8_t* open = nullptr;
  const float* pressure = nullptr;
  uint32_t axis = 0;
  uint32_t held_zero = 0;  // every open face held at pressure zero (a forced solve)
};

// Everything the prepared depth-averaged system keeps on the device.
struct FlowDevice {
  uint32_t nx = 0, ny = 0, axis = 0;
  float dx = 1, dy = 1, dz = 1;
  bool prepared = false;
  bool solved = false;
  // Counts flow solves and preparations: a transport that reads the field in
  // place records the count it was uploaded at, and is refused once it moves.
  uint64_t generation = 0;
  // Port-driven flow: the open exterior faces and the pressures held outside
  // them (FlowEdges slots) instead of the axis boundary.
  bool ported = false;
  DeviceBuffer<uint8_t> edge_open;
  DeviceBuffer<float> edge_pressure;
  DeviceBuffer<uint8_t> solid;     // original obstacle mask
  DeviceBuffer<float> mobility;    // uploaded per-site mobility
  DeviceBuffer<float> cond;        // column conductance (reduction)
  DeviceBuffer<float> diag, rhs, bc, cxl, cyl;
  DeviceBuffer<double> x, r, q, p0, p1, res;
  DeviceBuffer<double> xc;         // FP64 correction of the learned solver's fallback
  // A body-forced solve: the force per column (fx, fy), the body flux per
  // face, the driven field, and the pressure-driven solve's right-hand side
  // and pressure kept across the forced solve.
  DeviceBuffer<float> force, body_flux, driven, rhs_saved;
  DeviceBuffer<double> x_saved;
  bool forced = false;             // the last solve carried a body force
  DeviceBuffer<float> base_mobility;  // the nutrient loop's mobility without biomass
  DeviceBuffer<float> x32, r32, q32, p032, p132;
  DeviceBuffer<double> partials;
  DeviceBuffer<double> pq, rz, rr; // per-block Krylov partial sums, two generations each
  DeviceBuffer<CgState> state;
  DeviceBuffer<float> faces;       // x faces then y faces (z faces are zero)
  CgState* host_state = nullptr;   // pinned readback
  MgDevice mg;
  // A grid of several layers is solved on its grid of columns (the device
  // arrays above); the host keeps what lifting the column flux back onto the
  // voxel faces needs, and the lifted field.
  struct Layers {
    uint32_t nz = 1;
    float dz = 1.0f;                 // voxel spacing along z
    uint32_t bottom = 0;             // common floor of the fluid columns
    std::vector<uint8_t> voxels;     // original obstacle mask, nx*ny*nz
    std::vector<float> heights;      // per column
    std::vector<float> conductance;  // per column
    std::vector<float> faces;        // lifted x|y|z voxel faces after a solve
  } layers;
  // The device copies of the original mask and the column heights, for the
  // lift of a column flux onto the voxel faces.
  DeviceBuffer<uint8_t> voxels;
  DeviceBuffer<float> heights;
  DeviceBuffer<float> lifted;       // x|y|z voxel faces of a lifted solve
  // Port-driven flow: one lifted field per solve, the original grid's port
  // faces, their values in every field, the combination's weights and the
  // combined field's largest |velocity| (as float bits).
  DeviceBuffer<float> port_fields;
  DeviceBuffer<uint64_t> port_faces;
  DeviceBuffer<float> port_values;
  DeviceBuffer<double> port_weights;
  DeviceBuffer<uint32_t> port_peak;
  [[nodiscard]] std::size_t sites() const { return static_cast<std::size_t>(nx) * ny; }
  [[nodiscard]] std::size_t x_faces() const { return static_cast<std::size_t>(nx + 1) * ny; }
  [[nodiscard]] std::size_t y_faces() const { return static_cast<std::size_t>(nx) * (ny + 1); }
  // x|y|z faces of the original grid (nz layers).
  [[nodiscard]] std::size_t voxel_faces() const {
    const std::size_t z = layers.nz;
    return (x_faces() + y_faces()) * z + sites() * (z + 1);
  }
  [[nodiscard]] FlowEdges edges() const {
    if (!ported) return FlowEdges{nullptr, nullptr, axis};
    return FlowEdges{edge_open.data(), edge_pressure.data(), axis};
  }
};

// Transport constants passed to the kernels by value.
struct AdrParams {
  uint32_t nx, ny, signals;
  uint32_t nz;  // layers: the one-layer kernels read nx, ny only
  float dx, dy, dz;
  uint8_t kind[6];
  uint8_t has_field, has_affine, has_monod;
  uint8_t limited;  // van Leer correction (SignalAdvectionKind::van_leer)
  // Reservoir values of the fixed faces site by site where the spec has
  // boundary patches (device array [S][x_lower ny | x_upper ny | y_lower nx |
  // y_upper nx] on one layer; on several layers every face's sites, see
  // adr_site3.hpp); null: the uniform `boundary` values.
  const float* edge;
  float diffusion[kMaxSignals];
  float advection[kMaxSignals][3];
  float boundary[kMaxSignals][6];
  float vmax[kMaxSignals];
  float monod_k[kMaxSignals];
};

// Iterations the first batch of an implicit run's first step enqueues.
constexpr uint32_t kPredictedIterations = 15;

// The state of one implicit step's Jacobi solve on the device.
struct ImplicitState {
  float rhs_rms;    // RMS of the right-hand side
  float scale;      // the implicit operator's scale
  float threshold;  // the stopping rule's threshold, set after iterate 0
  float rms;        // residual RMS of the last iterate judged
  uint32_t halted;  // the rule was met, or the solve cannot go on
  uint32_t converged;
  uint32_t iteration;  // the last iterate judged
  uint32_t nonfinite;  // an iterate held a non-finite level
};

struct AdrDevice {
  AdrParams params{};
  // SignalIntegrationKind and SignalSolveParameters of the uploaded spec;
  // the report of the last implicit step.
  uint32_t integration = MSIM_FORWARD_EULER;
  uint32_t max_iterations = 10000;
  float absolute_tolerance = 1.0e-7f, relative_tolerance = 1.0e-5f;
  msim_signal_report last_report{1, 0, 0.0f};
  uint32_t predicted_iterations = kPredictedIterations;  // first batch of the next step
  uint32_t precision = 32;
  bool uploaded = false;
  bool flow_field = false;
  uint64_t flow_generation = 0;  // FlowDevice::generation of the field read in place
  uint64_t revision = 0;
  double greatest_outflow = 0.0;
  double max_loss[kMaxSignals] = {};
  double phi_max = 0.0;  // largest Monod occupancy of the uploaded (or grown) phi
  double host_mass[kMaxSignals] = {};  // mass of each signal now (FP64 runs)
  DeviceBuffer<uint8_t> solid;
  DeviceBuffer<float> xf, yf, zf, phi, source, loss, edge;
  bool has_edge = false;
  DeviceBuffer<uint8_t> levels_a, levels_b;  // S*N of float or double
  // Implicit steps: a third level buffer, the right-hand side, the lagged
  // Monod loss, per-block partials and the solve's state.
  DeviceBuffer<uint8_t> levels_c;
  DeviceBuffer<float> rhs, lagged, implicit_least;
  DeviceBuffer<double> implicit_sums;
  DeviceBuffer<ImplicitState> implicit_state;
  int current = 0;
  DeviceBuffer<double> partials, ledger, mass;
  DeviceBuffer<uint32_t> flags;              // [0]: error bits, [1]: first failed step
  DeviceBuffer<CgState> state;
  double seconds = 0.0;
  [[nodiscard]] std::size_t sites() const {
    return static_cast<std::size_t>(params.nx) * params.ny * (params.nz == 0 ? 1u : params.nz);
  }
};

// The learned smoother: weights, the packed bfloat16 fragments the kernels
// read, Adam moments, and the uploaded training set with its scratch.
struct LearnedDevice {
  bool loaded = false;
  uint32_t adam_steps = 0;  // Adam steps since the weights were set
  DeviceBuffer<float> w, m1, m2;
  DeviceBuffer<uint16_t> packed;
  uint32_t nx = 0, ny = 0, geometries = 0, pairs = 0;
  DeviceBuffer<float> ops, rp, ep;
  DeviceBuffer<uint32_t> geometry, order;
  DeviceBuffer<double> energy, lparts, grad;
  DeviceBuffer<float> z, d, gz, grads, losses;
};

// Coupled cell/grid exchange (coupled.hip): the cells in spatial order, their
// species in two generations, the rate plan, and the step's scratch: per cell
// the weights of its 2 x 2 stencil corners and its exchange rates, and the
// implicit kinds' sources.
struct CoupledDevice {
  bool uploaded = false;
  bool has_mask = false;  // the grid carries an obstacle mask (stencils renormalise)
  uint32_t cells = 0, species = 0, signals = 0, instructions = 0;
  float origin[3] = {};
  DeviceBuffer<float> position_x, position_y, position_z, lengths, radii, growth, previous;
  DeviceBuffer<int32_t> types;
  DeviceBuffer<float> species_a, species_b;
  int species_current = 0;
  DeviceBuffer<msim_rate_instruction> plan;
  DeviceBuffer<uint32_t> outputs;  // species outputs, then signal outputs
  DeviceBuffer<uint32_t> order;    // slot -> the cell's index in the caller's order
  DeviceBuffer<uint32_t> offsets;  // N + 1: the slots whose lower corner is site k
  DeviceBuffer<float> staging;      // host arrays on their way to or from the slots
  DeviceBuffer<uint32_t> keys, entries, sorted_keys;  // the upload's sort
  DeviceBuffer<float> weights, exchange, sources;
  DeviceBuffer<unsigned long long> failure;  // (cell << 32 | instruction) of the first failure
  DeviceBuffer<double> largest;
  DeviceBuffer<uint8_t> sort_scratch;
  double seconds = 0.0;
};

// Biomass deposition (colony.hip): the cells' positions and sizes, the mask,
// the deposited volume (FP64, one per site), the output field, and the least
// failing cell (and an internal bound check).
struct ColonyDevice {
  DeviceBuffer<float> position_x, position_y, position_z, lengths, radii;
  DeviceBuffer<float> velocity_x, velocity_y;  // in-plane cell velocities (flow coupling)
  DeviceBuffer<uint8_t> solid;
  DeviceBuffer<double> volume, out, base;
  DeviceBuffer<double> momentum, force;        // per column (x, y) (flow coupling)
  DeviceBuffer<unsigned int> failure;
  double seconds = 0.0;
};

struct BatchDevice;  // batch.hip

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t start = nullptr, stop = nullptr;
  std::vector<hipEvent_t> probes;  // per-launch timing events (profiled solves)
  std::string error;
  std::string name;
  FlowDevice flow;
  AdrDevice adr;
  LearnedDevice learned;
  CoupledDevice coupled;
  ColonyDevice colony;
  std::shared_ptr<BatchDevice> batch;
};

// Summed event time and launch count of the two Krylov kernels of a solve.
struct KernelTimes {
  double apply_seconds = 0.0, update_seconds = 0.0;
  uint32_t apply_launches = 0, update_launches = 0;
  double down_seconds = 0.0, up_seconds = 0.0, rest_seconds = 0.0;
  uint32_t cycles = 0;
};

// flow.hip
void flow_assemble(Context& ctx, const uint8_t* solid_host, const float* mobility_host);
// Conductance and assembly of one grid already on the device, with the
// boundary `edges`.
void launch_assembly(hipStream_t stream, const uint8_t* solid, const float* mobility, float* cond,
                     float* diag, float* rhs, float* bc, float* cxl, float* cyl, uint32_t nx,
                     uint32_t ny, float dx, float dy, float dz, FlowEdges edges);
// A ported system: holds each open face at the pressure of its slot in
// `pressure_host` (FlowEdges layout) and reassembles the right-hand side.
void flow_hold_pressures(Context& ctx, const float* pressure_host);
uint32_t stream_blocks(size_t n);
uint32_t krylov_blocks(size_t n);
void launch_residual(Context& ctx, const double* x, double* out, int mode);
void read_state(Context& ctx);
int failure_status(uint32_t failure);
const char* failure_text(uint32_t failure);

// mg.hip: the multilevel preconditioner and flexible CG.
void mg_prepare(Context& ctx);
// One application of the multilevel cycle (with the learned smoother when
// `learned`) to a host residual (FP32).
void mg_precondition(Context& ctx, const float* r_host, float* z_host, bool learned);
// FP32 flexible CG preconditioned by the multilevel cycle inside FP64
// iterative refinement; leaves the FP64 pressure in flow.x and the true
// residual in the solve state. Returns the Krylov iterations taken;
// *converged tells whether the FP64 residual met the tolerance.
uint32_t mg_solve(Context& ctx, const msim_flow_params& params, msim_flow_report& report,
                  KernelTimes* times, bool learned, bool* converged);
// Training pairs from the prepared system (msim_hip_learned_harvest).
uint32_t mg_harvest(Context& ctx, uint32_t max_pairs, bool learned, float* residuals,
                    float* errors);
void mg_operator(Context& ctx, float* wx, float* wy, float* bc, float* dinv);

// learned.hip: the learned smoother and its training.
void learned_set_weights(Context& ctx, const float* weights);
void learned_get_weights(Context& ctx, float* weights);
void learned_upload(Context& ctx, const msim_train_set& set);
void learned_train(Context& ctx, const msim_train_params& params, float* losses, double* seconds);
void learned_gradients(Context& ctx, uint32_t pair, double* loss, double* gradient);
void learned_apply(Context& ctx, uint32_t pair, float* z);
// The finest level's up pass with the learned post-smoother; with q, the
// flexible-CG dots z.r and z.q into zr / zq.
void launch_up_learned(Context& ctx, const MgLevel& L0, const float* r, const float* xp,
                       const float* ec, uint32_t nyc, float* z, const CgState* st, const float* q,
                       double* zr, double* zq);
// Workgroups of the learned up pass on an nx x ny level (one dot partial each).
uint32_t learned_blocks(uint32_t nx, uint32_t ny);
// r' = r - A x, e' = e - x on the finest level.
void learned_make_pairs(Context& ctx, const MgLevel& L0, const float* r, const double* e,
                        const float* x, float* rp, float* ep);
// FP64 Jacobi-PCG from x = 0 with at most `budget` iterations (the LEARNED
// solver's fallback). Returns the iterations taken; *converged as above.
uint32_t jacobi_solve(Context& ctx, double tolerance, uint32_t budget, uint32_t interval,
                      bool* converged, double* relative, KernelTimes* times);
void flow_solve(Context& ctx, const msim_flow_params& params, msim_flow_report& report);
void flow_apply(Context& ctx, const double* x, double* y);

// ports.hip: the lift onto voxel faces and port-driven flow.
// The solved field on the original grid's x|y|z faces (device `out`, of
// flow.voxel_faces() floats): one layer copies the face velocities (and
// zeroes the z faces when `with_z`); several layers lift the column flux
// (ShallowFlowReduction::lift, flow_system.hpp:424-470). `source`: the
// column field to lift in place of the solve's (the driven field).
void flow_lift(Context& ctx, float* out, bool with_z = true, const float* source = nullptr);
// Uploads the voxel mask and the column heights a lift of several layers reads.
void flow_upload_layers(Context& ctx);
// A checked port-driven flow (ShallowPortFlow, flow_system.hpp:1268-1311):
// the ports; the exterior slots of the grid of columns (FlowEdges layout)
// that a port opens and the port each belongs to; and per port the original
// grid's faces it carries flow through (x|y|z face indices in
// for_each_port_face order, 221-236) with the signed area of one face
// (make_port_faces, 1124-1130).
struct PortPlan {
  std::vector<msim_flow_port> ports;
  std::vector<uint8_t> open;
  std::vector<int32_t> owner;
  std::vector<uint64_t> faces;
  std::vector<std::size_t> offsets;  // ports.size() + 1 entries
  std::vector<double> signed_area;
};
void port_flow_solve(Context& ctx, const PortPlan& plan, const msim_port_flow_params& params,
                     float* x_faces, float* y_faces, float* z_faces, msim_port_flow_report& report,
                     msim_port_state* states);

// mg.hip: the multilevel operators of one grid, for callers that keep their
// own hierarchy (the batched sweep). Couplings are to the lower neighbour,
// already over the spacing squared; dinv is one over the diagonal (0 where
// the diagonal is 0). The coarse inverse is dense (nx ny x nx ny floats).
std::vector<float> mg_coarse_inverse(const std::vector<float>& wx, const std::vector<float>& wy,
                                     const std::vector<float>& bc, uint32_t nx, uint32_t ny);
void mg_level0_operator(hipStream_t stream, const float* cxl, const float* cyl, const float* bc,
                        float* wx, float* wy, float* bcl, float* dinv, uint32_t nx, uint32_t ny,
                        float isx, float isy);
void mg_coarse_operator(hipStream_t stream, const float* fwx, const float* fwy, const float* fbc,
                        uint32_t fnx, uint32_t fny, float* wx, float* wy, float* bc, float* dinv,
                        uint32_t nx, uint32_t ny);

// port_face_geometry (flow_system.hpp:203-219): the axis a face is normal
// to, whether it is that axis's upper face, and its in-plane axes.
struct PortGeometry {
  int component = 0;
  bool upper = false;
  int in_plane[2] = {0, 0};
};

inline PortGeometry port_geometry(uint32_t face) {
  if (face / 2 > 2) throw Error(MSIM_INVALID_ARGUMENT, "unknown flow port face");
  static constexpr int planes[3][2] = {{1, 2}, {0, 2}, {0, 1}};
  PortGeometry p;
  p.component = static_cast<int>(face / 2);
  p.upper = face % 2 == 1;
  p.in_plane[0] = planes[p.component][0];
  p.in_plane[1] = planes[p.component][1];
  return p;
}

// for_each_port_face (flow_system.hpp:221-236) on a grid of (nx, ny, nz):
// body(face index into x|y|z faces, index of the site inside the face), the
// first in-plane index outer.
template <typename Body>
void for_each_port_face(const uint32_t dims[3], const msim_flow_port& port, Body&& body) {
  const PortGeometry geo = port_geometry(port.face);
  const size_t nx = dims[0], ny = dims[1], nz = dims[2];
  const size_t offsets[3] = {0, (nx + 1) * ny * nz, (nx + 1) * ny * nz + nx * (ny + 1) * nz};
  for (uint32_t first = port.begin[0]; first < port.end[0]; ++first) {
    for (uint32_t second = port.begin[1]; second < port.end[1]; ++second) {
      uint32_t face[3] = {0, 0, 0};
      face[geo.component] = geo.upper ? dims[geo.component] : 0;
      face[geo.in_plane[0]] = first;
      face[geo.in_plane[1]] = second;
      uint32_t site[3] = {face[0], face[1], face[2]};
      if (geo.upper) --site[geo.component];
      size_t index;
      if (geo.component == 0) {
        index = (face[0] * ny + face[1]) * nz + face[2];
      } else if (geo.component == 1) {
        index = (face[0] * (ny + 1) + face[1]) * nz + face[2];
      } else {
        index = (face[0] * ny + face[1]) * (nz + 1) + face[2];
      }
      body(offsets[geo.component] + index, (site[0] * ny + site[1]) * nz + site[2]);
    }
  }
}

// resolved.hip: the resolved Stokes-Brinkman flow along params.axis of a grid
// already validated (mask: its obstacles, 0/1 per site); faces_out: every
// face velocity (x|y|z), scaled.
void resolved_flow(Context& ctx, const msim_grid& g, const std::vector<uint8_t>& mask,
                   const float* drag, const msim_resolved_params& params, float* faces_out,
                   msim_resolved_report& report);

// The depth-averaged pressure of the ports held at the given pressures, one
// per site (each column's on all its voxels), or left empty where shallow flow
// does not describe the geometry (shallow_port_pressure,
// flow_system.hpp:1318-1338).
using ShallowPressure =
    std::function<void(const std::vector<float>& pressures, std::vector<double>& known)>;

// resolved.hip: the resolved flow driven through ports already validated
// (validate_flow_ports); faces_out: the combined field (x|y|z).
void resolved_port_flow(Context& ctx, const msim_grid& g, const std::vector<uint8_t>& mask,
                        const float* drag, const msim_flow_port* ports, uint32_t count,
                        const msim_resolved_port_params& params, const ShallowPressure& shallow,
                        float* faces_out, msim_resolved_port_report& report,
                        msim_port_state* states);

// batch.hip: the batched sweep (msim_hip_batch_*).
void batch_upload(Context& ctx, uint32_t count, const msim_batch_instance* inst,
                  const msim_batch_params& params);
void batch_solve_flow(Context& ctx, uint32_t* iterations, float* residual, uint32_t* flags);
int batch_autotune(Context& ctx, float dt, uint32_t steps, double* seconds);
void batch_shape(int index, int* threads, int* row, int* chunk);
int batch_shape_count();
void batch_flags(Context& ctx, uint32_t* flags);
void batch_run(Context& ctx, float dt, uint32_t steps, bool graph, double* ledger,
               double* seconds);
void batch_download(Context& ctx, uint32_t index, double* pressure, float* x_faces,
                    float* y_faces, float* levels);

// adr.hip
// The kernels' parameters of a one-layer signal grid (velocity field and
// patches aside).
AdrParams transport_params(const msim_signal_grid& grid);
void adr_upload(Context& ctx, const msim_signal_grid& grid, const void* levels,
                uint32_t precision, bool use_flow_field);
void adr_run(Context& ctx, float dt, uint32_t steps, double* ledger);
void adr_download(Context& ctx, void* levels);
// The greatest outflow over the fluid sites of the field the uploaded
// transport reads (after that field changed), kept for the explicit bound.
double adr_refresh_outflow(Context& ctx);
// The settle and validation of SignalGrid::accept_step (signals.cpp:979-1011)
// on `count` FP32 levels in place, with the largest level from `parts`
// per-block maxima (0: no settle, validation only); a failure sets the
// transport flags with `step`.
void launch_settle_f32(Context& ctx, float* levels, size_t count, const double* largest,
                       uint32_t parts, uint32_t step);

// The explicit candidate of one FP32 forward-Euler step from `in` into `out`,
// not settled; a non-finite or (upwind) negative update sets the transport
// flags with `step`.
void fe_candidate_f32(Context& ctx, float dt, const float* in, float* out, uint32_t step,
                      const float* xf, const float* yf);

// The allocation-free FP32 forward-Euler step: bind_static_scratch takes the
// device address of adr.hip's static scratch array on `device` (which must be
// current) once per device; adr_step_static validates the explicit bound and
// advances `levels` (host, S*N) one step with everything staged in that array.
void bind_static_scratch(int device);
size_t static_scratch_bytes();
void adr_step_static(Context& ctx, const msim_signal_grid& grid, float dt, float* levels);

// capi.cpp: the largest loss rate per signal of a spec, and the explicit
// bound (SignalGrid::validate_step, signals.cpp:1066-1102).
void max_losses(const msim_signal_grid& s, double* loss_out);
void validate_explicit_bound(const AdrParams& P, uint32_t integration, double greatest_outflow,
                             const double* max_loss, float dt);

// adr3.hip: transport on a grid of several layers (FP32 levels), per site
// with the reference's three-axis arithmetic (adr_site3.hpp). The survey of
// an upload (greatest outflow over the fluid sites, mass per signal);
// `steps` forward-Euler steps of the uploaded levels; and the implicit
// step's set-up and sweep kernels for implicit.hip's iteration.
void adr3_survey(Context& ctx, const float* xf, const float* yf, const float* zf,
                 const float* levels, CgState* state, double* partials, double* mass);
void adr3_run(Context& ctx, float dt, uint32_t steps, const float* xf, const float* yf,
              const float* zf);
// The forward-Euler candidate c + dt * rate on a grid of several layers,
// written without judging it (a coupled step adds its sources first).
void adr3_candidate(Context& ctx, float dt, const float* in, float* out, uint32_t step,
                    const float* xf, const float* yf, const float* zf);
void implicit3_setup(Context& ctx, uint32_t blocks, const float* xf, const float* yf,
                     const float* zf, const float* sources, const float* old, float explicit_dt,
                     float dt);
void implicit3_sweep(Context& ctx, uint32_t blocks, const float* xf, const float* yf,
                     const float* zf, const float* cur, float* nxt, float half_dt,
                     float relaxation);

// stream.hip: GB/s (read + written) of `repeats` device copies of `bytes`.
double stream_copy_bandwidth(Context& ctx, size_t bytes, uint32_t repeats);

// nutrient.hip: the nutrient loop (msim_hip_nutrient_run) over the prepared
// flow and the uploaded FP64 transport.
void nutrient_run(Context& ctx, const msim_nutrient_params& params, uint32_t steps,
                  double* ledger, msim_nutrient_step* per_step);

// implicit.hip: `steps` implicit steps of dt on the uploaded FP32 levels.
// sources (device, S*N, or null): rates added to each step's right-hand side as
// dt * source; label: the solve's name in the not-converged message.
void implicit_run(Context& ctx, float dt, uint32_t steps, const float* xf, const float* yf,
                  const float* sources = nullptr, const char* label = "Implicit signal solve");

// coupled.hip: device-resident coupled steps (msim_hip_coupled_*). The grid
// must be uploaded (adr_upload, FP32) first.
void coupled_upload(Context& ctx, const float* origin, const msim_cells& cells,
                    const float* species, const msim_rate_plan& plan, bool has_mask);
void coupled_run(Context& ctx, const float* previous_lengths, float dt, uint32_t steps);
void coupled_download(Context& ctx, float* levels, float* species);

// colony.hip: the voxels one cell's deposit support can span (an upper bound
// from the averaging radius and the spacing), and the deposit itself: the
// density (volume fraction, FP64 per site) into density_host, or with
// mobility_host the colony mobility (base: 1 or N values). Returns the least
// cell whose support has no connected fluid, or 0xffffffff.
uint32_t colony_region_cap(const msim_grid& grid, double radius);
// velocities (2 floats per cell, in-plane x then y) and force_host (2 doubles
// per column): colony_flow_coupling's body force; both null otherwise.
uint32_t colony_deposit(Context& ctx, const msim_grid& grid, const float* origin,
                        const msim_cells& cells, double radius, const double* base,
                        uint32_t base_count, double drag_coefficient, double cap,
                        double* density_host, double* mobility_host,
                        const float* velocities = nullptr, double* force_host = nullptr);

}  // namespace msim
