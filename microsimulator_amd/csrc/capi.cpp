// The C ABI (include/msim_hip.h): argument checks on the host in the
// reference's order and wording, then the device work in flow.hip / adr.hip.
// No exception leaves this file: each entry point maps Error and the standard
// exceptions onto msim_status and keeps the message on the context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "internal.hpp"
#include "msim_hip.h"

struct msim_ctx {
  msim::Context impl;
};

namespace msim {
namespace {

[[noreturn]] void invalid(const std::string& message) {
  throw Error(MSIM_INVALID_ARGUMENT, message);
}

template <typename Body>
int guarded(msim_ctx* ctx, Body&& body) {
  if (ctx == nullptr) return MSIM_INVALID_ARGUMENT;
  try {
    check(hipSetDevice(ctx->impl.device), "hipSetDevice");
    body(ctx->impl);
    ctx->impl.error.clear();
    return MSIM_OK;
  } catch (const Error& e) {
    ctx->impl.error = e.what();
    return e.code();
  } catch (const std::bad_alloc&) {
    ctx->impl.error = "out of host memory";
    return MSIM_RUNTIME_ERROR;
  } catch (const std::exception& e) {
    ctx->impl.error = e.what();
    return MSIM_RUNTIME_ERROR;
  }
}

// SignalGridSpec::validate_lattice and the mask and boundary checks of
// SignalGridSpec::validate (cpp/core/signals.cpp:584-599, 626-632, 721-731).
void validate_lattice(const msim_grid& g) {
  if (g.nx == 0 || g.ny == 0 || g.nz == 0) invalid("signal grid dimensions must be positive");
  if (!std::isfinite(g.dx) || !std::isfinite(g.dy) || !std::isfinite(g.dz) || g.dx <= 0.0f ||
      g.dy <= 0.0f || g.dz <= 0.0f) {
    invalid("signal grid spacing must be finite and positive");
  }
  const double sites = static_cast<double>(g.nx) * g.ny * g.nz;
  if (sites > static_cast<double>(std::numeric_limits<uint32_t>::max())) {
    throw Error(MSIM_OVERFLOW, "signal grid level count exceeds the uint32 index space");
  }
  for (int f = 0; f < 6; ++f) {
    if (g.boundary[f] > MSIM_FIXED) invalid("unknown signal grid boundary kind");
  }
  for (int a = 0; a < 3; ++a) {
    if ((g.boundary[2 * a] == MSIM_PERIODIC) != (g.boundary[2 * a + 1] == MSIM_PERIODIC)) {
      invalid(std::string("signal grid periodic ") + "xyz"[a] + " boundaries must be paired");
    }
  }
  if (g.obstacles != nullptr) {
    const size_t n = static_cast<size_t>(sites);
    for (size_t i = 0; i < n; ++i) {
      if (g.obstacles[i] > 1) invalid("signal grid obstacle mask values must be 0 or 1");
    }
  }
}

// Inlet-to-outlet reachability through sites not in `solid`
// (validate_flow_grid, cpp/cpu/cpu_flow.cpp:288-336); site (x*ny + y)*nz + z.
void check_reachable(uint32_t nx, uint32_t ny, uint32_t nz, const std::vector<uint8_t>& solid,
                     uint32_t axis) {
  const size_t n = static_cast<size_t>(nx) * ny * nz;
  std::vector<uint8_t> visited(n, 0);
  std::vector<size_t> pending;
  pending.reserve(1024);
  const uint32_t dims[3] = {nx, ny, nz};
  const size_t step[3] = {static_cast<size_t>(ny) * nz, nz, 1};
  auto coord = [&](size_t site, int a) -> uint32_t {
    return static_cast<uint32_t>((site / step[a]) % dims[a]);
  };
  for (size_t site = 0; site < n; ++site) {
    if (solid[site] == 0 && coord(site, static_cast<int>(axis)) == 0) {
      visited[site] = 1;
      pending.push_back(site);
    }
  }
  if (pending.empty()) invalid("the flow inlet boundary is entirely blocked");
  while (!pending.empty()) {
    const size_t site = pending.back();
    pending.pop_back();
    if (coord(site, static_cast<int>(axis)) + 1 == dims[axis]) return;
    for (int a = 0; a < 3; ++a) {
      const uint32_t c = coord(site, a);
      if (c > 0) {
        const size_t nb = site - step[a];
        if (!visited[nb] && solid[nb] == 0) {
          visited[nb] = 1;
          pending.push_back(nb);
        }
      }
      if (c + 1 < dims[a]) {
        const size_t nb = site + step[a];
        if (!visited[nb] && solid[nb] == 0) {
          visited[nb] = 1;
          pending.push_back(nb);
        }
      }
    }
  }
  throw Error(MSIM_RUNTIME_ERROR, "the device carries no through-flow: the outlet is unreachable");
}

// The grid of columns (ShallowFlowReduction, flow_system.hpp:257-330):
// contiguous fluid voxels from a floor common to every column, one mobility
// per column; height spacing.z per fluid voxel, conductance height *
// mobility, or without mobilities height * (height / max height)^2; a
// column without conductance is solid.
struct Columns {
  std::vector<uint8_t> solid;  // per column
  std::vector<float> heights, conductance;
  uint32_t bottom = 0;
};

Columns reduce_columns(const msim_grid& g, const float* mobility, const std::vector<uint8_t>& mask) {
  const size_t n = static_cast<size_t>(g.nx) * g.ny;
  const uint32_t nz = g.nz;
  Columns c;
  c.heights.assign(n, 0.0f);
  c.conductance.assign(n, 0.0f);
  c.solid.assign(n, 1);
  uint32_t common_bottom = nz;
  float max_height = 0.0f;
  for (size_t column = 0; column < n; ++column) {
    bool started = false, ended = false;
    float value = 1.0f;
    for (uint32_t z = 0; z < nz; ++z) {
      const size_t site = column * nz + z;
      if (mobility != nullptr && (!std::isfinite(mobility[site]) || mobility[site] < 0.0f)) {
        invalid("flow mobility must be finite and non-negative");
      }
      if (mask[site] != 0) {
        if (started) ended = true;
        continue;
      }
      if (ended) invalid("shallow flow needs contiguous fluid columns; use resolved flow");
      const float next = mobility == nullptr ? 1.0f : mobility[site];
      if (!started) {
        if (common_bottom == nz) common_bottom = z;
        if (z != common_bottom) invalid("shallow flow needs a common planar floor; use resolved flow");
        value = next;
      } else if (std::abs(next - value) > 2.0e-6f * std::max(std::abs(value), std::abs(next))) {
        invalid("shallow-flow mobility must be constant through each depth column");
      }
      started = true;
      c.heights[column] += g.dz;
    }
    if (started) {
      c.solid[column] = 0;
      c.conductance[column] = c.heights[column] * value;
      max_height = std::max(max_height, c.heights[column]);
    }
  }
  if (mobility == nullptr && max_height > 0.0f) {
    for (size_t i = 0; i < n; ++i) {
      const float relative = c.heights[i] / max_height;
      c.conductance[i] *= relative * relative;
    }
  }
  for (size_t i = 0; i < n; ++i) {
    if (c.conductance[i] == 0.0f) c.solid[i] = 1;
  }
  c.bottom = common_bottom;
  return c;
}

// What the device solve on the grid of columns takes: mask and per-site
// mobility, the device conductance being dz times the mobility. One layer
// keeps the user's mobility; several layers pass each column's conductance
// with dz = 1, so the device's conductance is that value exactly, and keep
// what the lift reads.
void take_columns(const msim_grid& g, const float* mobility, std::vector<uint8_t>&& mask,
                  Columns&& c, std::vector<uint8_t>& solid, std::vector<float>& mob,
                  FlowDevice::Layers& layers) {
  const size_t n = static_cast<size_t>(g.nx) * g.ny;
  layers = FlowDevice::Layers{};
  layers.nz = g.nz;
  layers.dz = g.dz;
  if (g.nz == 1) {
    solid = std::move(mask);
    mob.assign(n, 1.0f);
    if (mobility != nullptr) std::memcpy(mob.data(), mobility, n * sizeof(float));
    layers.heights = std::move(c.heights);
    return;
  }
  solid = c.solid;
  mob = c.conductance;
  layers.bottom = c.bottom;
  layers.voxels = std::move(mask);
  layers.heights = std::move(c.heights);
  layers.conductance = std::move(c.conductance);
}

std::vector<uint8_t> obstacle_mask(const msim_grid& g) {
  const size_t voxels = static_cast<size_t>(g.nx) * g.ny * g.nz;
  std::vector<uint8_t> mask(voxels, 0);
  if (g.obstacles != nullptr) std::memcpy(mask.data(), g.obstacles, voxels);
  return mask;
}

// The checks reduce_shallow_flow and the flow system's constructor make
// before a solve (flow_system.hpp:1255-1266, 260-330; cpu_flow.cpp:288-336),
// and the reduction to a grid of columns.
void validate_flow(const msim_grid& g, const float* mobility, uint32_t axis,
                   std::vector<uint8_t>& solid, std::vector<float>& mob,
                   FlowDevice::Layers& layers) {
  validate_lattice(g);
  if (axis > 2) invalid("unknown flow axis");
  for (int a = 0; a < 3; ++a) {
    if (g.boundary[2 * a] == MSIM_PERIODIC || g.boundary[2 * a + 1] == MSIM_PERIODIC) {
      invalid("native flow solvers do not support periodic boundaries");
    }
  }
  if (g.boundary[2 * axis] != MSIM_FIXED || g.boundary[2 * axis + 1] != MSIM_FIXED) {
    invalid("native flow-axis boundaries must be fixed to identify inlet and outlet");
  }
  std::vector<uint8_t> mask = obstacle_mask(g);
  check_reachable(g.nx, g.ny, g.nz, mask, axis);
  if (axis == 2) invalid("shallow flow is depth-integrated along z; flow axis must be x or y");
  Columns c = reduce_columns(g, mobility, mask);
  // The grid of columns must still carry flow.
  check_reachable(g.nx, g.ny, 1, c.solid, axis);
  take_columns(g, mobility, std::move(mask), std::move(c), solid, mob, layers);
}

// validate_flow_ports (cpp/cpu/cpu_flow.cpp:193-286) on a grid of `dims`
// with boundary kinds `kinds` and obstacle mask `solid`.
void validate_ports(const uint32_t dims[3], const uint8_t kinds[6],
                    const std::vector<uint8_t>& solid, const msim_flow_port* ports,
                    uint32_t count) {
  for (int a = 0; a < 3; ++a) {
    if (kinds[2 * a] == MSIM_PERIODIC || kinds[2 * a + 1] == MSIM_PERIODIC) {
      invalid("native flow solvers do not support periodic boundaries");
    }
  }
  if (count == 0) invalid("flow needs at least one port");
  const size_t nx = dims[0], ny = dims[1], nz = dims[2];
  const size_t sites = nx * ny * nz;
  std::vector<uint8_t> claimed((nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1), 0);
  std::vector<uint8_t> visited(sites, 0);
  std::vector<size_t> pending;
  bool holds_pressure = false;
  for (uint32_t p = 0; p < count; ++p) {
    const msim_flow_port& port = ports[p];
    if (port.kind != MSIM_PORT_SPEED && port.kind != MSIM_PORT_PRESSURE) {
      invalid("unknown flow port kind");
    }
    if (!std::isfinite(port.value)) invalid("flow port values must be finite");
    const PortGeometry geo = port_geometry(port.face);
    if (kinds[port.face] != MSIM_FIXED) invalid("a flow port must open a fixed boundary");
    const uint32_t extent[2] = {dims[geo.in_plane[0]], dims[geo.in_plane[1]]};
    if (port.end[0] <= port.begin[0] || port.end[1] <= port.begin[1] ||
        port.end[0] > extent[0] || port.end[1] > extent[1]) {
      invalid("a flow port must span at least one site of its face");
    }
    bool open = false;
    for_each_port_face(dims, port, [&](size_t face, size_t site) {
      if (claimed[face] != 0) invalid("flow ports must not overlap");
      claimed[face] = 1;
      if (solid[site] != 0) return;
      open = true;
      if (port.kind == MSIM_PORT_PRESSURE && visited[site] == 0) {
        visited[site] = 1;
        pending.push_back(site);
      }
    });
    if (!open) invalid("a flow port is entirely blocked");
    holds_pressure = holds_pressure || port.kind == MSIM_PORT_PRESSURE;
  }
  if (!holds_pressure) invalid("flow needs at least one pressure port");
  // Something flows: a speed port, or a pressure other than the first
  // pressure port's.
  float first = 0.0f;
  for (uint32_t p = 0; p < count; ++p) {
    if (ports[p].kind == MSIM_PORT_PRESSURE) {
      first = ports[p].value;
      break;
    }
  }
  bool drives = false;
  for (uint32_t p = 0; p < count; ++p) {
    drives = drives || ports[p].kind == MSIM_PORT_SPEED || ports[p].value != first;
  }
  if (!drives) invalid("flow ports state no speed and hold one pressure, so nothing flows");
  // A speed port's flow has to leave, or arrive, through a pressure port.
  const size_t step[3] = {ny * nz, nz, 1};
  while (!pending.empty()) {
    const size_t site = pending.back();
    pending.pop_back();
    for (int a = 0; a < 3; ++a) {
      const size_t c = (site / step[a]) % dims[a];
      if (c > 0) {
        const size_t nb = site - step[a];
        if (visited[nb] == 0 && solid[nb] == 0) {
          visited[nb] = 1;
          pending.push_back(nb);
        }
      }
      if (c + 1 < dims[a]) {
        const size_t nb = site + step[a];
        if (visited[nb] == 0 && solid[nb] == 0) {
          visited[nb] = 1;
          pending.push_back(nb);
        }
      }
    }
  }
  for (uint32_t p = 0; p < count; ++p) {
    if (ports[p].kind != MSIM_PORT_SPEED) continue;
    bool reached = false;
    for_each_port_face(dims, ports[p], [&](size_t, size_t site) {
      reached = reached || visited[site] != 0;
    });
    if (!reached) {
      throw Error(MSIM_RUNTIME_ERROR, "a speed port carries no flow: no pressure port is reachable");
    }
  }
}

// ShallowPortFlow's construction (flow_system.hpp:1272-1282): the ports'
// checks on the grid, the reduction, each port on the grid of columns
// (reduced_port, 347-372: an x or y face, spanning every fluid voxel of each
// column it reaches), the ports' checks on the grid of columns, and the faces
// each port carries flow through.
PortPlan plan_ports(const msim_grid& g, const float* mobility, const msim_flow_port* ports,
                    uint32_t count, std::vector<uint8_t>& solid, std::vector<float>& mob,
                    FlowDevice::Layers& layers, bool check_reduced = true) {
  validate_lattice(g);
  if (count > 0 && ports == nullptr) invalid("ports must not be null");
  std::vector<uint8_t> mask = obstacle_mask(g);
  const uint32_t dims[3] = {g.nx, g.ny, g.nz};
  validate_ports(dims, g.boundary, mask, ports, count);
  Columns c = reduce_columns(g, mobility, mask);
  const uint32_t nz = g.nz;
  std::vector<msim_flow_port> reduced(ports, ports + count);
  for (uint32_t p = 0; p < count; ++p) {
    const PortGeometry geo = port_geometry(ports[p].face);
    if (geo.component == 2) invalid("a shallow-flow port cannot open a z face; use resolved flow");
    uint32_t site[3] = {0, 0, 0};
    if (geo.upper) site[geo.component] = dims[geo.component] - 1;
    for (uint32_t along = ports[p].begin[0]; along < ports[p].end[0]; ++along) {
      site[geo.in_plane[0]] = along;
      for (uint32_t z = c.bottom; z < nz; ++z) {
        if (mask[(static_cast<size_t>(site[0]) * g.ny + site[1]) * nz + z] != 0) break;
        if (z < ports[p].begin[1] || z >= ports[p].end[1]) {
          invalid("a shallow-flow port must span the full depth of its columns; use resolved flow");
        }
      }
    }
    reduced[p].begin[1] = 0;
    reduced[p].end[1] = 1;
  }
  // On one layer whose every fluid site carries conductance the grid of
  // columns is the grid itself, and the ports on it are the same ports: the
  // checks there would repeat the ones above exactly.
  if (check_reduced && (nz > 1 || c.solid != mask)) {
    const uint32_t columns[3] = {g.nx, g.ny, 1};
    const uint8_t column_kinds[6] = {g.boundary[0], g.boundary[1], g.boundary[2],
                                     g.boundary[3], MSIM_NO_FLUX,   MSIM_NO_FLUX};
    validate_ports(columns, column_kinds, c.solid, reduced.data(), count);
  }
  PortPlan plan;
  plan.ports.assign(ports, ports + count);
  const size_t slots = 2 * (static_cast<size_t>(g.nx) + g.ny);
  plan.open.assign(slots, 0);
  plan.owner.assign(slots, -1);
  plan.offsets.assign(count + 1, 0);
  plan.signed_area.assign(count, 0.0);
  const float spacing[3] = {g.dx, g.dy, g.dz};
  for (uint32_t p = 0; p < count; ++p) {
    const PortGeometry geo = port_geometry(ports[p].face);
    // port_flow_boundary (238-253) on the grid of columns.
    const size_t base = ports[p].face == 0   ? 0
                        : ports[p].face == 1 ? g.ny
                        : ports[p].face == 2 ? 2 * static_cast<size_t>(g.ny)
                                             : 2 * static_cast<size_t>(g.ny) + g.nx;
    for (uint32_t along = ports[p].begin[0]; along < ports[p].end[0]; ++along) {
      plan.open[base + along] = 1;
      plan.owner[base + along] = static_cast<int32_t>(p);
    }
    // open_port_faces (385-393): faces onto fluid voxels of columns with
    // conductance.
    for_each_port_face(dims, ports[p], [&](size_t face, size_t site) {
      if (mask[site] == 0 && c.conductance[site / nz] > 0.0f) plan.faces.push_back(face);
    });
    plan.offsets[p + 1] = plan.faces.size();
    const double area = static_cast<double>(spacing[geo.in_plane[0]]) * spacing[geo.in_plane[1]];
    plan.signed_area[p] = geo.upper ? -area : area;
  }
  take_columns(g, mobility, std::move(mask), std::move(c), solid, mob, layers);
  return plan;
}

// Prepares the context's depth-averaged system for shallow_port_pressure
// (flow_system.hpp:1318-1338): the reduction with the default mobility, the
// ports on the grid of columns and the system they bound. Returns false where
// the reference's construction refuses (its invalid_argument: shallow flow
// does not describe the geometry, a port in a z face or short of its columns'
// depth, or a boundary that opens onto no fluid). The reference checks the
// reduced ports no further, so neither does this.
bool prepare_shallow_ports(Context& c, const msim_grid& g, const msim_flow_port* ports,
                           uint32_t count, PortPlan& plan) {
  std::vector<uint8_t> solid;
  std::vector<float> mob;
  auto& f = c.flow;
  f.prepared = false;
  try {
    plan = plan_ports(g, nullptr, ports, count, solid, mob, f.layers, false);
  } catch (const Error& e) {
    if (e.code() == MSIM_INVALID_ARGUMENT) return false;
    throw;
  }
  // A column conducts when it is fluid in the grid of columns with a nonzero
  // conductance; slots [x_lower ny | x_upper ny | y_lower nx | y_upper nx].
  const size_t nx = g.nx, ny = g.ny;
  bool open = false;
  for (size_t slot = 0; slot < plan.open.size() && !open; ++slot) {
    if (plan.open[slot] == 0) continue;
    size_t column;
    if (slot < ny) {
      column = slot;
    } else if (slot < 2 * ny) {
      column = (nx - 1) * ny + (slot - ny);
    } else if (slot < 2 * ny + nx) {
      column = (slot - 2 * ny) * ny;
    } else {
      column = (slot - 2 * ny - nx) * ny + ny - 1;
    }
    open = solid[column] == 0 && mob[column] > 0.0f;
  }
  if (!open) return false;
  f.nx = g.nx;
  f.ny = g.ny;
  f.dx = g.dx;
  f.dy = g.dy;
  f.dz = g.nz == 1 ? g.dz : 1.0f;
  f.axis = 0;
  f.ported = true;
  const size_t slots = plan.open.size();
  f.edge_open.reserve(slots);
  f.edge_pressure.reserve(slots);
  check(hipMemcpyAsync(f.edge_open.data(), plan.open.data(), slots, hipMemcpyHostToDevice,
                       c.stream),
        "upload open port faces");
  check(hipMemsetAsync(f.edge_pressure.data(), 0, slots * sizeof(float), c.stream),
        "clear port pressures");
  flow_assemble(c, solid.data(), mob.data());
  return true;
}

// The depth-averaged pressure of the prepared system with each port held at
// pressures[port]: FP64 Jacobi-PCG to 1e-6 in at most 50000 iterations (the
// reference's solve_depth_averaged_pressure, cpu_flow.cpp:172-175), each
// column's pressure on every voxel of the column.
void shallow_port_pressure(Context& c, const PortPlan& plan, uint32_t nz,
                           const std::vector<float>& pressures, std::vector<double>& known) {
  auto& f = c.flow;
  std::vector<float> edge(plan.open.size(), 0.0f);
  for (size_t slot = 0; slot < edge.size(); ++slot) {
    edge[slot] = plan.owner[slot] >= 0 ? pressures[static_cast<size_t>(plan.owner[slot])] : 0.0f;
  }
  flow_hold_pressures(c, edge.data());
  bool converged = false;
  double relative = 0.0;
  jacobi_solve(c, static_cast<double>(1.0e-6f), 50000, 64, &converged, &relative, nullptr);
  if (!converged) {
    throw Error(MSIM_NOT_CONVERGED,
                "HIP depth-averaged flow conjugate gradient did not converge: relative " +
                    std::to_string(relative));
  }
  std::vector<double> columns(f.sites());
  check(hipMemcpyAsync(columns.data(), f.x.data(), columns.size() * sizeof(double),
                       hipMemcpyDeviceToHost, c.stream),
        "download depth-averaged pressure");
  check(hipStreamSynchronize(c.stream), "depth-averaged pressure");
  known.resize(columns.size() * nz);
  for (size_t site = 0; site < known.size(); ++site) known[site] = columns[site / nz];
}

// scale_velocity (flow_system.hpp:1368-1415) over the lifted voxel faces: the
// open inlet faces are the axis's lower faces of fluid voxels in columns with
// conductance (open_inlet_faces, 449-459).
void scale_layers(FlowDevice& f, const msim_flow_params& params, msim_flow_report& report,
                  const std::vector<float>* driven) {
  auto& v = f.layers.faces;
  const auto& L = f.layers;
  const uint32_t nx = f.nx, ny = f.ny, nz = L.nz;
  float peak = 0.0f;
  for (const float value : v) {
    if (!std::isfinite(value)) throw Error(MSIM_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
    peak = std::max(peak, std::abs(value));
  }
  const size_t nxf = static_cast<size_t>(nx + 1) * ny * nz;
  double inlet_sum = 0.0, driven_sum = 0.0;
  size_t inlet_count = 0;
  const uint32_t across = f.axis == 0 ? ny : nx;
  for (uint32_t t = 0; t < across; ++t) {
    const size_t column = f.axis == 0 ? t : static_cast<size_t>(t) * ny;
    for (uint32_t z = 0; z < nz; ++z) {
      if (L.voxels[column * nz + z] != 0 || !(L.conductance[column] > 0.0f)) continue;
      const size_t face = f.axis == 0 ? static_cast<size_t>(t) * nz + z
                                      : nxf + static_cast<size_t>(t) * (ny + 1) * nz + z;
      inlet_sum += v[face];
      if (driven != nullptr) driven_sum += (*driven)[face];
      ++inlet_count;
    }
  }
  if (inlet_count == 0) throw Error(MSIM_RUNTIME_ERROR, "flow solve has no open inlet faces");
  const float solved_mean = static_cast<float>(inlet_sum / static_cast<double>(inlet_count));
  if (peak == 0.0f || solved_mean <= 1.0e-9f * peak) {
    throw Error(MSIM_RUNTIME_ERROR, "the device carries no through-flow: the outlet is unreachable");
  }
  const float driven_mean = static_cast<float>(driven_sum / static_cast<double>(inlet_count));
  const float factor =
      params.scale != 0 ? (params.mean_inlet_speed - driven_mean) / solved_mean : 1.0f;
  float scaled_peak = 0.0f;
  for (size_t i = 0; i < v.size(); ++i) {
    float& value = v[i];
    if (params.scale != 0) value = value * factor + (driven != nullptr ? (*driven)[i] : 0.0f);
    if (!std::isfinite(value)) throw Error(MSIM_RUNTIME_ERROR, "non-finite combined flow velocity");
    scaled_peak = std::max(scaled_peak, std::abs(value));
  }
  report.mean_inlet_speed = params.scale != 0 ? params.mean_inlet_speed : solved_mean;
  report.max_speed = scaled_peak;
  report.solved_mean = solved_mean;
  report.factor = factor;
}

void validate_flow_params(const msim_flow_params& p) {
  if (p.scale != 0 && (!std::isfinite(p.mean_inlet_speed) ||
                       (p.mean_inlet_speed == 0.0f && p.body_force == nullptr))) {
    invalid("depth-averaged mean inlet speed must be finite and nonzero without a body force");
  }
  if (p.body_force != nullptr && p.scale == 0) {
    invalid("a body-forced flow is returned scaled: set scale");
  }
  if (!std::isfinite(p.relative_tolerance) || p.relative_tolerance < 1.0e-15) {
    invalid("depth-averaged relative tolerance must be finite and at least 1e-15");
  }
  if (p.max_iterations == 0) invalid("depth-averaged iteration limit must be positive");
  if (p.solver > MSIM_SOLVER_MULTIGRID) invalid("unknown pressure solver");
}

// The transport fields of SignalGridSpec::validate (signals.cpp:609-700) for
// what this backend reads, plus the Monod inputs.
void validate_signal_grid(const msim_signal_grid& s, bool device_field) {
  validate_lattice(s.grid);
  if (s.signal_count == 0) invalid("signal grid must contain at least one signal");
  if (s.integration > MSIM_BACKWARD_EULER) {
    invalid("unknown signal integration kind");  // signals.cpp:718-725
  }
  if (s.advection_scheme != MSIM_UPWIND && s.advection_scheme != MSIM_VAN_LEER) {
    invalid("unknown signal advection scheme");  // signals.cpp:726-732
  }
  // SignalSolveParameters::validate (signals.cpp:482-491).
  if (s.solver_max_iterations == 0) invalid("signal solver iteration limit must be positive");
  if (!std::isfinite(s.solver_absolute_tolerance) || s.solver_absolute_tolerance < 0.0f ||
      !std::isfinite(s.solver_relative_tolerance) || s.solver_relative_tolerance < 0.0f ||
      (s.solver_absolute_tolerance == 0.0f && s.solver_relative_tolerance == 0.0f)) {
    invalid("signal solver tolerances must be finite and non-negative");
  }
  if (s.signal_count > kMaxSignals) {
    throw Error(MSIM_UNSUPPORTED, "the HIP transport step takes at most 8 signals");
  }
  if (s.diffusion == nullptr) invalid("signal grid transport arrays must match signal count");
  for (uint32_t k = 0; k < s.signal_count; ++k) {
    if (!std::isfinite(s.diffusion[k]) || s.diffusion[k] < 0.0f) {
      invalid("signal diffusion must be finite and non-negative");
    }
  }
  if (s.advection != nullptr) {
    for (uint32_t k = 0; k < 3 * s.signal_count; ++k) {
      if (!std::isfinite(s.advection[k])) invalid("signal advection must be finite");
    }
  }
  for (uint32_t k = 0; k < s.signal_count; ++k) {
    for (int f = 0; f < 6; ++f) {
      if (s.grid.boundary[f] != MSIM_FIXED) continue;
      const float v = s.boundary_values != nullptr ? s.boundary_values[6 * k + f] : 0.0f;
      if (!std::isfinite(v) || v < 0.0f) {
        invalid("fixed signal grid boundary values must be finite and non-negative");
      }
    }
  }
  // Reservoir patches (GridBoundary::validate and the face extents,
  // signals.cpp:449-479, 734-744).
  if (s.patch_count > 0 && s.patches == nullptr) invalid("signal grid boundary patches are missing");
  for (uint32_t k = 0; k < s.patch_count; ++k) {
    const msim_boundary_patch& pt = s.patches[k];
    if (pt.face > 5) invalid("unknown signal grid boundary face");
    if (s.grid.boundary[pt.face] != MSIM_FIXED) {
      invalid("non-fixed signal grid boundaries cannot have values");
    }
    if (pt.end[0] <= pt.begin[0] || pt.end[1] <= pt.begin[1]) {
      invalid("signal grid boundary patches must hold at least one site");
    }
    if (pt.values == nullptr) invalid("fixed signal grid boundary values must match signal count");
    for (uint32_t j = 0; j < s.signal_count; ++j) {
      if (!std::isfinite(pt.values[j]) || pt.values[j] < 0.0f) {
        invalid("fixed signal grid boundary values must be finite and non-negative");
      }
    }
    const uint32_t axis = pt.face / 2;
    const uint32_t extent[3][2] = {{s.grid.ny, s.grid.nz}, {s.grid.nx, s.grid.nz},
                                   {s.grid.nx, s.grid.ny}};
    if (pt.end[0] > extent[axis][0] || pt.end[1] > extent[axis][1]) {
      invalid("signal grid boundary patches must lie on their face");
    }
  }
  const uint32_t nx = s.grid.nx, ny = s.grid.ny, nz = s.grid.nz;
  const size_t n = static_cast<size_t>(nx) * ny * nz;
  auto solid = [&](size_t i) { return s.grid.obstacles != nullptr && s.grid.obstacles[i] != 0; };
  // SignalGridAffineReaction::validate (signals.cpp:493-497): both arrays or
  // neither.
  if ((s.loss_rates == nullptr) != (s.source_rates == nullptr)) {
    invalid("signal grid affine reaction arrays must match the grid level count");
  }
  if (s.loss_rates != nullptr) {
    for (size_t i = 0; i < s.signal_count * n; ++i) {
      const float src = s.source_rates[i];
      if (!std::isfinite(src) || src < 0.0f) {
        invalid("signal grid affine source rates must be finite and non-negative");
      }
      if (!std::isfinite(s.loss_rates[i]) || s.loss_rates[i] < 0.0f) {
        invalid("signal grid affine loss rates must be finite and non-negative");
      }
      if (solid(i % n) && (src != 0.0f || s.loss_rates[i] != 0.0f)) {
        invalid("signal grid affine reaction must be zero at obstacle sites");
      }
    }
  }
  if (s.phi != nullptr) {
    if (s.monod_vmax == nullptr || s.monod_k == nullptr) invalid("Monod uptake needs vmax and K");
    for (uint32_t k = 0; k < s.signal_count; ++k) {
      if (!std::isfinite(s.monod_vmax[k]) || s.monod_vmax[k] < 0.0f ||
          !std::isfinite(s.monod_k[k]) || s.monod_k[k] <= 0.0f) {
        invalid("Monod vmax must be finite and non-negative and K finite and positive");
      }
    }
    for (size_t i = 0; i < n; ++i) {
      if (!std::isfinite(s.phi[i]) || s.phi[i] < 0.0f) {
        invalid("Monod occupancy must be finite and non-negative");
      }
    }
  }
  const bool field = device_field || s.x_faces != nullptr;
  if (field && s.advection != nullptr) {
    for (uint32_t k = 0; k < 3 * s.signal_count; ++k) {
      if (s.advection[k] != 0.0f) {
        invalid("signal grid velocity field requires zero constant advection vectors");
      }
    }
  }
  if (!device_field && s.x_faces != nullptr) {
    if (s.y_faces == nullptr || (nz > 1 && s.z_faces == nullptr)) {
      invalid("signal grid velocity field must cover every lattice face");
    }
    const size_t nxf = static_cast<size_t>(nx + 1) * ny * nz;
    const size_t nyf = static_cast<size_t>(nx) * (ny + 1) * nz;
    const size_t nzf = static_cast<size_t>(nx) * ny * (nz + 1);
    for (size_t i = 0; i < nxf; ++i) {
      if (!std::isfinite(s.x_faces[i])) invalid("signal grid velocity field values must be finite");
    }
    for (size_t i = 0; i < nyf; ++i) {
      if (!std::isfinite(s.y_faces[i])) invalid("signal grid velocity field values must be finite");
    }
    if (s.z_faces != nullptr) {
      for (size_t i = 0; i < nzf; ++i) {
        if (!std::isfinite(s.z_faces[i])) {
          invalid("signal grid velocity field values must be finite");
        }
      }
    }
    auto xface = [&](uint32_t fx, uint32_t y, uint32_t z) {
      return s.x_faces[(static_cast<size_t>(fx) * ny + y) * nz + z];
    };
    auto yface = [&](uint32_t x, uint32_t fy, uint32_t z) {
      return s.y_faces[(static_cast<size_t>(x) * (ny + 1) + fy) * nz + z];
    };
    auto zface = [&](uint32_t x, uint32_t y, uint32_t fz) {
      return s.z_faces != nullptr ? s.z_faces[(static_cast<size_t>(x) * ny + y) * (nz + 1) + fz]
                                  : 0.0f;
    };
    auto site = [&](uint32_t x, uint32_t y, uint32_t z) {
      return (static_cast<size_t>(x) * ny + y) * nz + z;
    };
    // Closed faces carry no velocity (signals.cpp:654-676).
    const uint32_t dims[3] = {nx, ny, nz};
    for (uint32_t x = 0; x < nx; ++x) {
      for (uint32_t y = 0; y < ny; ++y) {
        for (uint32_t z = 0; z < nz; ++z) {
          const uint32_t c[3] = {x, y, z};
          const bool here = solid(site(x, y, z));
          for (int a = 0; a < 3; ++a) {
            if (dims[a] == 1) continue;
            uint32_t lo[3] = {x, y, z}, hi[3] = {x, y, z};
            bool clo, cup;
            if (c[a] == 0) {
              lo[a] = dims[a] - 1;
              clo = s.grid.boundary[2 * a] == MSIM_NO_FLUX ||
                    (s.grid.boundary[2 * a] == MSIM_PERIODIC && solid(site(lo[0], lo[1], lo[2])));
            } else {
              lo[a] = c[a] - 1;
              clo = solid(site(lo[0], lo[1], lo[2]));
            }
            if (c[a] + 1 == dims[a]) {
              hi[a] = 0;
              cup = s.grid.boundary[2 * a + 1] == MSIM_NO_FLUX ||
                    (s.grid.boundary[2 * a + 1] == MSIM_PERIODIC && solid(site(hi[0], hi[1], hi[2])));
            } else {
              hi[a] = c[a] + 1;
              cup = solid(site(hi[0], hi[1], hi[2]));
            }
            const float lower = a == 0 ? xface(x, y, z) : a == 1 ? yface(x, y, z) : zface(x, y, z);
            const float upper =
                a == 0 ? xface(x + 1, y, z) : a == 1 ? yface(x, y + 1, z) : zface(x, y, z + 1);
            if (((clo || here) && lower != 0.0f) || ((cup || here) && upper != 0.0f)) {
              invalid("signal grid velocity field must be zero on closed faces");
            }
          }
        }
      }
    }
    // The two sides of a periodic boundary are one face (signals.cpp:678-713).
    if (s.grid.boundary[0] == MSIM_PERIODIC && nx > 1) {
      for (uint32_t y = 0; y < ny; ++y)
        for (uint32_t z = 0; z < nz; ++z)
          if (xface(0, y, z) != xface(nx, y, z))
            invalid("signal grid velocity field periodic faces must hold equal values");
    }
    if (s.grid.boundary[2] == MSIM_PERIODIC && ny > 1) {
      for (uint32_t x = 0; x < nx; ++x)
        for (uint32_t z = 0; z < nz; ++z)
          if (yface(x, 0, z) != yface(x, ny, z))
            invalid("signal grid velocity field periodic faces must hold equal values");
    }
    if (s.grid.boundary[4] == MSIM_PERIODIC && nz > 1) {
      for (uint32_t x = 0; x < nx; ++x)
        for (uint32_t y = 0; y < ny; ++y)
          if (zface(x, y, 0) != zface(x, y, nz))
            invalid("signal grid velocity field periodic faces must hold equal values");
    }
  }
}

// validate_signal_levels (signals.cpp:754-784).
template <typename T>
void validate_levels(const msim_signal_grid& s, const T* levels) {
  const size_t n = static_cast<size_t>(s.grid.nx) * s.grid.ny * s.grid.nz;
  int fault = 0;
  for (size_t i = 0; i < s.signal_count * n; ++i) {
    const T v = levels[i];
    if (!std::isfinite(v) || v < 0) {
      fault = 1;
      break;
    }
    if (s.grid.obstacles != nullptr && v != 0 && s.grid.obstacles[i % n] != 0) fault = 2;
  }
  if (fault == 1) invalid("signal grid levels must be finite and non-negative");
  if (fault == 2) invalid("signal grid levels must be zero at obstacle sites");
}

}  // namespace

// Largest loss rate per signal, for the explicit bound: the affine loss's
// maximum (max_reaction_loss, signals.cpp:322-330) plus vmax*max(phi)/K.
void max_losses(const msim_signal_grid& s, double* loss_out) {
  const size_t n = static_cast<size_t>(s.grid.nx) * s.grid.ny * s.grid.nz;
  float phi_max = 0.0f;
  if (s.phi != nullptr) {
    for (size_t i = 0; i < n; ++i) phi_max = std::max(phi_max, s.phi[i]);
  }
  for (uint32_t k = 0; k < s.signal_count; ++k) {
    double loss = 0.0;
    if (s.loss_rates != nullptr) {
      loss = static_cast<double>(*std::max_element(s.loss_rates + k * n, s.loss_rates + (k + 1) * n));
    }
    if (s.phi != nullptr) {
      loss += static_cast<double>(s.monod_vmax[k]) * phi_max / static_cast<double>(s.monod_k[k]);
    }
    loss_out[k] = loss;
  }
}

// SignalGrid::validate_step (signals.cpp:1066-1102).
void validate_explicit_bound(const AdrParams& P, uint32_t integration, double greatest_outflow,
                             const double* max_loss, float dt) {
  if (!std::isfinite(dt) || dt < 0.0f) invalid("time step must be finite and non-negative");
  // The explicit bound binds forward Euler only (signals.cpp:1070-1072).
  if (integration != MSIM_FORWARD_EULER) return;
  const uint32_t dims[3] = {P.nx, P.ny, P.nz == 0 ? 1u : P.nz};
  const float spacing[3] = {P.dx, P.dy, P.dz};
  for (uint32_t k = 0; k < P.signals; ++k) {
    double inverse_square_sum = 0.0;
    double courant_sum = greatest_outflow;
    for (int a = 0; a < 3; ++a) {
      if (dims[a] == 1) continue;
      const double inverse_spacing = 1.0 / static_cast<double>(spacing[a]);
      inverse_square_sum += inverse_spacing * inverse_spacing;
      if (P.has_field == 0) {
        courant_sum += std::abs(static_cast<double>(P.advection[k][a])) * inverse_spacing;
      }
    }
    // A limited face flux carries up to twice the upwind one.
    if (P.limited != 0) courant_sum *= 2.0;
    const double factor =
        static_cast<double>(dt) *
        ((2.0 * static_cast<double>(P.diffusion[k]) * inverse_square_sum) + courant_sum +
         max_loss[k]);
    if (!std::isfinite(factor) || factor > 1.0) {
      invalid("signal grid time step violates the explicit stability bound");
    }
  }
}

namespace {

void record_losses(Context& ctx, const msim_signal_grid& s) {
  max_losses(s, ctx.adr.max_loss);
  const size_t n = static_cast<size_t>(s.grid.nx) * s.grid.ny * s.grid.nz;
  ctx.adr.phi_max = 0.0;
  if (s.phi != nullptr) {
    for (size_t i = 0; i < n; ++i) ctx.adr.phi_max = std::max(ctx.adr.phi_max, double(s.phi[i]));
  }
}

// A transport uploaded over the device's flow field reads it in place: any
// later flow solve or preparation (another grid, speed, mobility or mask)
// leaves it reading a field its explicit bound and mask were not taken from.
void check_flow_field(const Context& c) {
  if (c.adr.flow_field &&
      (!c.flow.solved || c.flow.generation != c.adr.flow_generation ||
       c.flow.layers.nz != 1 || c.flow.nx != c.adr.params.nx || c.flow.ny != c.adr.params.ny)) {
    invalid("the flow field on the device is no longer the transport's");
  }
}

void validate_step(const Context& ctx, float dt) {
  validate_explicit_bound(ctx.adr.params, ctx.adr.integration, ctx.adr.greatest_outflow,
                          ctx.adr.max_loss, dt);
}

// The world state's invariants (WorldState::validate, world_state.cpp:14-35)
// for what a coupled step reads, and its species levels.
void validate_cells(const msim_cells& cells, uint32_t species_count, const float* species) {
  if (cells.count > 0 &&
      (cells.position_x == nullptr || cells.position_y == nullptr ||
       cells.position_z == nullptr || cells.lengths == nullptr || cells.radii == nullptr ||
       cells.growth_rates == nullptr || cells.cell_types == nullptr)) {
    invalid("cell arrays must not be null");
  }
  for (uint32_t c = 0; c < cells.count; ++c) {
    const float v[6] = {cells.position_x[c], cells.position_y[c], cells.position_z[c],
                        cells.lengths[c],    cells.radii[c],      cells.growth_rates[c]};
    for (float x : v) {
      if (!std::isfinite(x)) invalid("cell fields must be finite");
    }
    if (cells.lengths[c] < 0.0f) invalid("cell length must be non-negative");
    if (cells.radii[c] <= 0.0f) invalid("cell radius must be positive");
  }
  const size_t levels = static_cast<size_t>(cells.count) * species_count;
  if (levels > 0 && species == nullptr) invalid("species levels must not be null");
  for (size_t i = 0; i < levels; ++i) {
    if (!std::isfinite(species[i])) invalid("cell species levels must be finite");
  }
}

// CoupledRatePlan::validate (coupled_rates.cpp:60-146).
void validate_rate_plan(const msim_rate_plan& plan) {
  if (plan.signal_count == 0) invalid("a coupled rate plan needs at least one signal");
  if (plan.instruction_count == 0) invalid("a coupled rate plan needs instructions");
  if (plan.instructions == nullptr || plan.signal_outputs == nullptr ||
      (plan.species_count > 0 && plan.species_outputs == nullptr)) {
    invalid("coupled rate plan arrays must not be null");
  }
  for (uint32_t index = 0; index < plan.instruction_count; ++index) {
    const msim_rate_instruction& in = plan.instructions[index];
    const char* bad = nullptr;
    switch (in.operation) {
      case MSIM_RATE_CONSTANT:
        if (!std::isfinite(in.value)) invalid("coupled rate constants must be finite");
        break;
      case MSIM_RATE_SPECIES:
        if (in.first >= plan.species_count) {
          invalid("coupled rate instruction uses an unknown species");
        }
        break;
      case MSIM_RATE_SIGNAL:
        if (in.first >= plan.signal_count) invalid("coupled rate instruction uses an unknown signal");
        break;
      case MSIM_RATE_POSITION_X: case MSIM_RATE_POSITION_Y: case MSIM_RATE_POSITION_Z:
      case MSIM_RATE_LENGTH: case MSIM_RATE_RADIUS: case MSIM_RATE_GROWTH_RATE:
      case MSIM_RATE_CELL_TYPE: case MSIM_RATE_VOLUME: case MSIM_RATE_VOLUME_CHANGE_RATE:
      case MSIM_RATE_SURFACE_AREA:
        break;
      case MSIM_RATE_NEGATE: case MSIM_RATE_EXPONENTIAL: case MSIM_RATE_LOGARITHM:
        if (in.first >= index) bad = "unary";
        break;
      case MSIM_RATE_ADD: case MSIM_RATE_SUBTRACT: case MSIM_RATE_MULTIPLY: case MSIM_RATE_DIVIDE:
      case MSIM_RATE_POWER: case MSIM_RATE_MINIMUM: case MSIM_RATE_MAXIMUM: case MSIM_RATE_LESS:
      case MSIM_RATE_LESS_EQUAL: case MSIM_RATE_GREATER: case MSIM_RATE_GREATER_EQUAL:
      case MSIM_RATE_EQUAL:
        if (in.first >= index) bad = "first";
        else if (in.second >= index) bad = "second";
        break;
      case MSIM_RATE_SELECT:
        if (in.first >= index) bad = "condition";
        else if (in.second >= index) bad = "true";
        else if (in.third >= index) bad = "false";
        break;
      default:
        invalid("coupled rate plan uses an unknown operation");
    }
    if (bad != nullptr) {
      invalid(std::string("coupled rate ") + bad + " input must refer to an earlier instruction");
    }
  }
  for (uint32_t k = 0; k < plan.species_count; ++k) {
    if (plan.species_outputs[k] >= plan.instruction_count) {
      invalid("coupled species output uses an unknown instruction");
    }
  }
  for (uint32_t k = 0; k < plan.signal_count; ++k) {
    if (plan.signal_outputs[k] >= plan.instruction_count) {
      invalid("coupled signal output uses an unknown instruction");
    }
  }
}

// interpolation_axis's bound (signals.cpp:98-127): false outside the grid;
// else the lower site.
bool inside_axis(float position, float origin, float spacing, uint32_t dimension,
                 uint32_t* lower) {
  *lower = 0;
  if (dimension == 1) return true;
  const double coordinate =
      (static_cast<double>(position) - static_cast<double>(origin)) / static_cast<double>(spacing);
  if (coordinate < 0.0 || coordinate > static_cast<double>(dimension - 1)) return false;
  *lower = static_cast<uint32_t>(std::floor(coordinate));
  return true;
}

// The stencil checks of validate_coupled_step (cpu_coupled.cpp:112-122), in
// cell order: each position inside the grid along x, y and z, and not with
// every corner of positive weight solid (signal_grid_stencil, signals.cpp:357-420).
void validate_positions(const msim_signal_grid& g, const float* origin, const msim_cells& cells) {
  const uint32_t nx = g.grid.nx, ny = g.grid.ny, nz = g.grid.nz;
  for (uint32_t c = 0; c < cells.count; ++c) {
    uint32_t x0, y0, z0;
    if (!inside_axis(cells.position_x[c], origin[0], g.grid.dx, nx, &x0)) {
      invalid("signal sample is outside the x grid bound");
    }
    if (!inside_axis(cells.position_y[c], origin[1], g.grid.dy, ny, &y0)) {
      invalid("signal sample is outside the y grid bound");
    }
    if (!inside_axis(cells.position_z[c], origin[2], g.grid.dz, nz, &z0)) {
      invalid("signal sample is outside the z grid bound");
    }
    if (g.grid.obstacles == nullptr) continue;
    // Weights along an axis: 1 - f and f with f the fraction past the lower
    // site (one site, weight 1, on the grid's upper edge or an axis of one).
    const auto weights = [&](float position, float o, float spacing, uint32_t dimension,
                             uint32_t lower, uint32_t count, float* w) {
      w[0] = 1.0f;
      w[1] = 0.0f;
      if (dimension == 1 || count == 1 || lower == dimension - 1) return;
      const double coordinate =
          (static_cast<double>(position) - static_cast<double>(o)) / static_cast<double>(spacing);
      const float fraction = static_cast<float>(coordinate - static_cast<double>(lower));
      w[0] = 1.0f - fraction;
      w[1] = fraction;
    };
    float wx[2], wy[2], wz[2];
    const uint32_t nxc = (nx == 1 || x0 == nx - 1) ? 1 : 2, nyc = (ny == 1 || y0 == ny - 1) ? 1 : 2;
    const uint32_t nzc = (nz == 1 || z0 == nz - 1) ? 1 : 2;
    weights(cells.position_x[c], origin[0], g.grid.dx, nx, x0, nxc, wx);
    weights(cells.position_y[c], origin[1], g.grid.dy, ny, y0, nyc, wy);
    weights(cells.position_z[c], origin[2], g.grid.dz, nz, z0, nzc, wz);
    bool fluid = false;
    for (uint32_t a = 0; a < nxc; ++a) {
      for (uint32_t b = 0; b < nyc; ++b) {
        for (uint32_t d = 0; d < nzc; ++d) {
          const size_t site = (static_cast<size_t>(x0 + a) * ny + y0 + b) * nz + z0 + d;
          if (g.grid.obstacles[site] == 0 && wx[a] * wy[b] * wz[d] > 0.0f) fluid = true;
        }
      }
    }
    if (!fluid) invalid("signal sample position is inside a grid obstacle");
  }
}

// Everything validate_coupled_step checks (cpu_coupled.cpp:90-123) but the
// time step, the explicit bound and the previous lengths, in its order, with
// the grid's own checks (those of constructing it) first.
void validate_coupled_inputs(const msim_signal_grid& g, const float* origin, const float* levels,
                             const msim_cells& cells, const float* species,
                             const msim_rate_plan& plan) {
  for (int a = 0; a < 3; ++a) {
    if (!std::isfinite(origin[a])) invalid("signal grid origin must be finite");
  }
  validate_cells(cells, plan.species_count, species);
  validate_signal_grid(g, false);
  validate_levels(g, levels);
}

// The plan-dependent checks of validate_coupled_step (cpu_coupled.cpp:97-104)
// and this backend's limit on the plan.
void validate_coupled_plan(const msim_signal_grid& g, const msim_rate_plan& plan) {
  validate_rate_plan(plan);
  if (plan.signal_count != g.signal_count) {
    invalid("coupled rate plan and signal grid counts disagree");
  }
}

void validate_previous(const msim_cells& cells, const float* previous) {
  if (cells.count > 0 && previous == nullptr) invalid("previous lengths must not be null");
  for (uint32_t c = 0; c < cells.count; ++c) {
    if (!std::isfinite(previous[c]) || previous[c] < 0.0f) {
      invalid("previous cell lengths must be finite and non-negative");
    }
  }
}

void check_coupled_limits(const msim_rate_plan& plan, const msim_cells& cells) {
  if (static_cast<uint64_t>(plan.instruction_count) + plan.signal_count + plan.species_count >
      256) {
    throw Error(MSIM_UNSUPPORTED,
                "the HIP coupled step takes at most 256 rate instructions, signals and species "
                "together");
  }
  if (cells.count > (1u << 29)) {
    throw Error(MSIM_OVERFLOW, "the HIP coupled step takes at most 2^29 cells");
  }
}

void upload_coupled(Context& c, const msim_signal_grid& g, const float* origin,
                    const float* levels, const msim_cells& cells, const float* species,
                    const msim_rate_plan& plan) {
  check_coupled_limits(plan, cells);
  adr_upload(c, g, levels, 32, false);
  record_losses(c, g);
  coupled_upload(c, origin, cells, species, plan, g.grid.obstacles != nullptr);
}


// The host's part of colony_volume_fraction's checks (flow.py:282-305,
// 224-241; biomass.py:12-13): the radius, and per cell in order a finite
// position and, for a cell inside the grid's extent, a valid length and
// radius. Returns the first refused cell (cells.count when none), its message
// and whether its size (not its position) was refused.
uint32_t colony_host_checks(const msim_grid& g, const float* origin, const msim_cells& cells,
                            double radius, std::string* message, bool* size_refused,
                            const float* velocities = nullptr) {
  *size_refused = false;
  if (!std::isfinite(radius) || radius <= 0.0) invalid("averaging radius must be finite and positive");
  validate_lattice(g);
  for (int a = 0; a < 3; ++a) {
    if (!std::isfinite(origin[a])) invalid("signal grid origin must be finite");
  }
  if (cells.count > 0 && (cells.position_x == nullptr || cells.position_y == nullptr ||
                          cells.position_z == nullptr || cells.lengths == nullptr ||
                          cells.radii == nullptr)) {
    invalid("cell arrays must not be null");
  }
  const uint32_t dims[3] = {g.nx, g.ny, g.nz};
  const double spacing[3] = {g.dx, g.dy, g.dz};
  for (uint32_t c = 0; c < cells.count; ++c) {
    // colony_flow_coupling checks each cell's velocity first (flow.py:409-411).
    if (velocities != nullptr &&
        (!std::isfinite(velocities[2 * c]) || !std::isfinite(velocities[2 * c + 1]))) {
      *message = "in-plane cell velocity must be finite";
      return c;
    }
    const double p[3] = {cells.position_x[c], cells.position_y[c], cells.position_z[c]};
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) {
      *message = "deposited positions must be finite";
      return c;
    }
    bool outside = false;
    for (int a = 0; a < 3; ++a) {
      const double o = origin[a], h = spacing[a];
      if (p[a] < o - h / 2.0 || p[a] >= o + (static_cast<double>(dims[a]) - 0.5) * h) {
        outside = true;
      }
    }
    const double length = cells.lengths[c], r = cells.radii[c];
    if (!outside && (!std::isfinite(length) || length < 0 || !std::isfinite(r) || r <= 0)) {
      *message = "length must be finite and nonnegative; radius finite and positive";
      *size_refused = true;
      return c;
    }
  }
  return cells.count;
}

// The deposit with the reference's first failure: the least cell the host or
// the device refuses, the host's message winning for its own cell only when
// the device found none before it (a cell's support is judged before its size).
void colony_run(Context& ctx, const msim_grid& g, const float* origin, const msim_cells& cells,
                double radius, const double* base, uint32_t base_count, double drag, double cap,
                double* density, double* mobility, const std::string& pending,
                const float* velocities = nullptr, double* force = nullptr) {
  std::string message;
  bool size_refused = false;
  const uint32_t refused =
      colony_host_checks(g, origin, cells, radius, &message, &size_refused, velocities);
  const uint32_t axis = [&] {
    uint32_t total = 0;
    const uint32_t dims[3] = {g.nx, g.ny, g.nz};
    const double spacing[3] = {g.dx, g.dy, g.dz};
    for (int a = 0; a < 3; ++a) {
      total += static_cast<uint32_t>(std::min(std::floor(2.0 * radius / spacing[a]) + 4.0,
                                              static_cast<double>(dims[a])));
    }
    return total;
  }();
  const uint32_t region = colony_region_cap(g, radius);
  if (region > 7000 || (static_cast<size_t>(axis) + region) * 8 + region > 65536) {
    throw Error(MSIM_UNSUPPORTED,
                "the HIP deposit takes kernel supports of at most 64 KiB of LDS (about 7000 "
                "voxels): shorten the averaging radius relative to the spacing");
  }
  msim_cells head = cells;
  // A refused size is judged after the cell's support, so that cell's
  // support is still deposited (and may fail first); a refused position is not.
  head.count = refused + (size_refused ? 1u : 0u);
  // With a refused mobility argument only the deposit's own failures can
  // come first: deposit into the density, then report.
  const bool mobility_ok = pending.empty();
  std::vector<double> scratch;
  double* density_out = density;
  if (mobility != nullptr && !mobility_ok) {
    scratch.resize(static_cast<size_t>(g.nx) * g.ny * g.nz);
    density_out = scratch.data();
  }
  const uint32_t failed =
      colony_deposit(ctx, g, origin, head, radius, base, base_count, drag, cap, density_out,
                     mobility_ok ? mobility : nullptr, velocities,
                     mobility_ok ? force : nullptr);
  if (failed != 0xffffffffu) invalid("biomass deposition has no connected fluid support");
  if (refused < cells.count) invalid(message);
  if (!mobility_ok) invalid(pending);
}

// _darcy_coefficients' checks (flow.py:359-374) and the drag's (91-96), which
// the reference makes after the deposit: the first failing one's message, or
// an empty string.
std::string darcy_pending(const msim_grid& grid, const double* base, uint32_t base_count,
                          double drag_coefficient, double max_volume_fraction) {
  const size_t n = static_cast<size_t>(grid.nx) * grid.ny * grid.nz;
  if (!(max_volume_fraction > 0.0 && max_volume_fraction < 1.0)) {
    return "maximum volume fraction must lie strictly between zero and one";
  }
  if (base_count == 0 && (!std::isfinite(base[0]) || base[0] <= 0.0)) {
    return "base mobility must be finite and positive";
  }
  if (base_count != 0 && base_count != n) return "base mobility must hold one value per grid site";
  if (base_count != 0 &&
      !std::all_of(base, base + n, [](double b) { return std::isfinite(b) && b >= 0.0; })) {
    return "base mobility values must be finite and non-negative";
  }
  if (!std::isfinite(drag_coefficient) || drag_coefficient < 0.0) {
    return "drag coefficient must be finite and non-negative";
  }
  return std::string();
}
}  // namespace
}  // namespace msim

using msim::Context;
using msim::Error;

extern "C" {

int msim_hip_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return 0;
  return count;
}

int msim_hip_create(int device, msim_ctx** out) {
  if (out == nullptr) return MSIM_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    return MSIM_NO_DEVICE;
  }
  auto* ctx = new (std::nothrow) msim_ctx;
  if (ctx == nullptr) return MSIM_RUNTIME_ERROR;
  ctx->impl.device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->impl.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->impl.start) != hipSuccess ||
      hipEventCreate(&ctx->impl.stop) != hipSuccess) {
    msim_hip_destroy(ctx);
    return MSIM_DEVICE_ERROR;
  }
  // The allocation-free entry point's scratch address on this device.
  try {
    msim::bind_static_scratch(device);
  } catch (const Error& e) {
    msim_hip_destroy(ctx);
    return e.code();
  }
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->impl.name = prop.name;
    if (ctx->impl.name.empty()) ctx->impl.name = prop.gcnArchName;
  }
  *out = ctx;
  return MSIM_OK;
}

void msim_hip_destroy(msim_ctx* ctx) {
  if (ctx == nullptr) return;
  static_cast<void>(hipSetDevice(ctx->impl.device));
  if (ctx->impl.stream != nullptr) static_cast<void>(hipStreamSynchronize(ctx->impl.stream));
  if (ctx->impl.flow.host_state != nullptr) static_cast<void>(hipHostFree(ctx->impl.flow.host_state));
  ctx->impl.flow.host_state = nullptr;
  for (hipEvent_t e : ctx->impl.probes) static_cast<void>(hipEventDestroy(e));
  if (ctx->impl.start != nullptr) static_cast<void>(hipEventDestroy(ctx->impl.start));
  if (ctx->impl.stop != nullptr) static_cast<void>(hipEventDestroy(ctx->impl.stop));
  if (ctx->impl.stream != nullptr) static_cast<void>(hipStreamDestroy(ctx->impl.stream));
  delete ctx;
}

const char* msim_hip_last_error(const msim_ctx* ctx) {
  return ctx == nullptr ? "null context" : ctx->impl.error.c_str();
}

int msim_hip_device_name(msim_ctx* ctx, char* name, size_t length) {
  return msim::guarded(ctx, [&](Context& c) {
    if (name == nullptr || length == 0) msim::invalid("name buffer must not be empty");
    std::snprintf(name, length, "%s", c.name.c_str());
  });
}

int msim_hip_flow_prepare(msim_ctx* ctx, const msim_grid* grid, const float* mobility,
                          uint32_t axis) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr) msim::invalid("grid must not be null");
    std::vector<uint8_t> solid;
    std::vector<float> mob;
    auto& f = c.flow;
    msim::validate_flow(*grid, mobility, axis, solid, mob, f.layers);
    f.nx = grid->nx;
    f.ny = grid->ny;
    f.dx = grid->dx;
    f.dy = grid->dy;
    // Several layers: the conductance goes up as the mobility with dz = 1.
    f.dz = grid->nz == 1 ? grid->dz : 1.0f;
    f.axis = axis;
    f.ported = false;
    f.prepared = false;
    msim::flow_assemble(c, solid.data(), mob.data());
    msim::flow_upload_layers(c);
  });
}

int msim_hip_flow_solve(msim_ctx* ctx, const msim_flow_params* params, msim_flow_report* report) {
  return msim::guarded(ctx, [&](Context& c) {
    if (params == nullptr || report == nullptr) msim::invalid("parameters must not be null");
    if (!c.flow.prepared) msim::invalid("no flow system has been prepared");
    if (c.flow.ported) msim::invalid("the prepared flow is driven through ports");
    msim::validate_flow_params(*params);
    if (params->axis != c.flow.axis) msim::invalid("the flow axis differs from the prepared one");
    auto& f = c.flow;
    if (f.layers.nz == 1) {
      msim::flow_solve(c, *params, *report);
      return;
    }
    // Several layers: the device solves the grid of columns, reconstructs
    // the column flux (unscaled, dz = 1) and lifts it onto the voxel faces;
    // the host scales the lifted field to the mean inlet speed.
    msim_flow_params column = *params;
    column.scale = 0;
    msim::flow_solve(c, column, *report);
    f.lifted.reserve(f.voxel_faces());
    msim::flow_lift(c, f.lifted.data());
    f.layers.faces.resize(f.voxel_faces());
    msim::check(hipMemcpyAsync(f.layers.faces.data(), f.lifted.data(),
                               f.layers.faces.size() * sizeof(float), hipMemcpyDeviceToHost,
                               c.stream),
                "download lifted faces");
    std::vector<float> driven;
    if (f.forced) {
      // The driven column flux, lifted as the pressure-driven one is.
      driven.resize(f.voxel_faces());
      msim::flow_lift(c, f.lifted.data(), true, f.driven.data());
      msim::check(hipMemcpyAsync(driven.data(), f.lifted.data(), driven.size() * sizeof(float),
                                 hipMemcpyDeviceToHost, c.stream),
                  "download driven faces");
    }
    msim::check(hipStreamSynchronize(c.stream), "download lifted faces");
    msim::scale_layers(f, *params, *report, f.forced ? &driven : nullptr);
  });
}

int msim_hip_flow_download(msim_ctx* ctx, double* pressure, float* x_faces, float* y_faces,
                           float* z_faces) {
  return msim::guarded(ctx, [&](Context& c) {
    auto& f = c.flow;
    if (!f.solved) msim::invalid("no flow solve to download");
    if (pressure != nullptr) {
      msim::check(hipMemcpyAsync(pressure, f.x.data(), f.sites() * sizeof(double),
                                 hipMemcpyDeviceToHost, c.stream),
                  "download pressure");
    }
    if (f.layers.nz > 1) {
      // The lifted voxel faces of a grid of several layers.
      msim::check(hipStreamSynchronize(c.stream), "download");
      const auto& v = f.layers.faces;
      const size_t nz = f.layers.nz;
      const size_t nxf = f.x_faces() * nz, nyf = f.y_faces() * nz;
      if (x_faces != nullptr) std::memcpy(x_faces, v.data(), nxf * sizeof(float));
      if (y_faces != nullptr) std::memcpy(y_faces, v.data() + nxf, nyf * sizeof(float));
      if (z_faces != nullptr) {
        std::memcpy(z_faces, v.data() + nxf + nyf, (v.size() - nxf - nyf) * sizeof(float));
      }
      return;
    }
    if (x_faces != nullptr) {
      msim::check(hipMemcpyAsync(x_faces, f.faces.data(), f.x_faces() * sizeof(float),
                                 hipMemcpyDeviceToHost, c.stream),
                  "download x faces");
    }
    if (y_faces != nullptr) {
      msim::check(hipMemcpyAsync(y_faces, f.faces.data() + f.x_faces(),
                                 f.y_faces() * sizeof(float), hipMemcpyDeviceToHost, c.stream),
                  "download y faces");
    }
    msim::check(hipStreamSynchronize(c.stream), "download");
    if (z_faces != nullptr) std::memset(z_faces, 0, 2 * f.sites() * sizeof(float));
  });
}

int msim_hip_solve_depth_averaged_flow(msim_ctx* ctx, const msim_grid* grid,
                                       const float* mobility, const msim_flow_params* params,
                                       float* x_faces, float* y_faces, float* z_faces,
                                       msim_flow_report* report) {
  if (params == nullptr) return MSIM_INVALID_ARGUMENT;
  int status = msim_hip_flow_prepare(ctx, grid, mobility, params->axis);
  if (status != MSIM_OK) return status;
  status = msim_hip_flow_solve(ctx, params, report);
  if (status != MSIM_OK) return status;
  return msim_hip_flow_download(ctx, nullptr, x_faces, y_faces, z_faces);
}

int msim_hip_solve_resolved_flow(msim_ctx* ctx, const msim_grid* grid, const float* drag,
                                 const msim_resolved_params* params, float* x_faces,
                                 float* y_faces, float* z_faces, msim_resolved_report* report) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || params == nullptr) msim::invalid("grid and parameters must not be null");
    // ResolvedFlowParameters::validate (cpu_flow.cpp:152-168).
    if (!std::isfinite(params->mean_inlet_speed) || params->mean_inlet_speed == 0.0f) {
      msim::invalid("resolved-flow mean inlet speed must be finite and nonzero");
    }
    const auto tolerance = [](float t, const char* name) {
      if (!std::isfinite(t) || t < std::numeric_limits<float>::epsilon()) {
        msim::invalid(std::string(name) + " must be finite and at least float machine epsilon");
      }
    };
    tolerance(params->relative_tolerance, "resolved-flow outer relative tolerance");
    tolerance(params->inner_relative_tolerance, "resolved-flow inner relative tolerance");
    if (params->max_outer_iterations == 0 || params->max_inner_iterations == 0) {
      msim::invalid("resolved-flow iteration limits must be positive");
    }
    if (params->axis > 2) msim::invalid("unknown resolved-flow axis");
    // validate_flow_grid (cpu_flow.cpp:288-336).
    msim::validate_lattice(*grid);
    for (int a = 0; a < 3; ++a) {
      if (grid->boundary[2 * a] == MSIM_PERIODIC || grid->boundary[2 * a + 1] == MSIM_PERIODIC) {
        msim::invalid("native flow solvers do not support periodic boundaries");
      }
    }
    const uint32_t axis = params->axis;
    if (grid->boundary[2 * axis] != MSIM_FIXED || grid->boundary[2 * axis + 1] != MSIM_FIXED) {
      msim::invalid("native flow-axis boundaries must be fixed to identify inlet and outlet");
    }
    const std::vector<uint8_t> mask = msim::obstacle_mask(*grid);
    msim::check_reachable(grid->nx, grid->ny, grid->nz, mask, axis);
    const size_t nfx = static_cast<size_t>(grid->nx + 1) * grid->ny * grid->nz;
    const size_t nfy = static_cast<size_t>(grid->nx) * (grid->ny + 1) * grid->nz;
    const size_t nfz = static_cast<size_t>(grid->nx) * grid->ny * (grid->nz + 1);
    std::vector<float> faces(nfx + nfy + nfz);
    msim_resolved_report rep{};
    msim::resolved_flow(c, *grid, mask, drag, *params, faces.data(), rep);
    if (x_faces != nullptr) std::memcpy(x_faces, faces.data(), nfx * sizeof(float));
    if (y_faces != nullptr) std::memcpy(y_faces, faces.data() + nfx, nfy * sizeof(float));
    if (z_faces != nullptr) std::memcpy(z_faces, faces.data() + nfx + nfy, nfz * sizeof(float));
    if (report != nullptr) *report = rep;
  });
}

int msim_hip_solve_resolved_port_flow(msim_ctx* ctx, const msim_grid* grid, const float* drag,
                                      const msim_flow_port* ports, uint32_t port_count,
                                      const msim_resolved_port_params* params, float* x_faces,
                                      float* y_faces, float* z_faces,
                                      msim_resolved_port_report* report,
                                      msim_port_state* states) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || params == nullptr) msim::invalid("grid and parameters must not be null");
    // ResolvedPortFlowParameters::validate (cpu_flow.cpp:179-185).
    const auto tolerance = [](float t, const char* name) {
      if (!std::isfinite(t) || t < std::numeric_limits<float>::epsilon()) {
        msim::invalid(std::string(name) + " must be finite and at least float machine epsilon");
      }
    };
    tolerance(params->relative_tolerance, "resolved-flow outer relative tolerance");
    tolerance(params->inner_relative_tolerance, "resolved-flow inner relative tolerance");
    if (params->max_outer_iterations == 0 || params->max_inner_iterations == 0) {
      msim::invalid("resolved-flow iteration limits must be positive");
    }
    // validate_flow_ports (cpu_flow.cpp:194-286).
    msim::validate_lattice(*grid);
    if (port_count > 0 && ports == nullptr) msim::invalid("ports must not be null");
    const std::vector<uint8_t> mask = msim::obstacle_mask(*grid);
    const uint32_t dims[3] = {grid->nx, grid->ny, grid->nz};
    msim::validate_ports(dims, grid->boundary, mask, ports, port_count);
    msim::PortPlan plan;
    const bool shallow = msim::prepare_shallow_ports(c, *grid, ports, port_count, plan);
    const uint32_t nz = grid->nz;
    const msim::ShallowPressure known = [&](const std::vector<float>& pressures,
                                            std::vector<double>& out) {
      out.clear();
      if (shallow) msim::shallow_port_pressure(c, plan, nz, pressures, out);
    };
    const size_t nfx = static_cast<size_t>(grid->nx + 1) * grid->ny * grid->nz;
    const size_t nfy = static_cast<size_t>(grid->nx) * (grid->ny + 1) * grid->nz;
    const size_t nfz = static_cast<size_t>(grid->nx) * grid->ny * (grid->nz + 1);
    std::vector<float> faces(nfx + nfy + nfz);
    msim_resolved_port_report rep{};
    msim::resolved_port_flow(c, *grid, mask, drag, ports, port_count, *params, known,
                             faces.data(), rep, states);
    if (x_faces != nullptr) std::memcpy(x_faces, faces.data(), nfx * sizeof(float));
    if (y_faces != nullptr) std::memcpy(y_faces, faces.data() + nfx, nfy * sizeof(float));
    if (z_faces != nullptr) std::memcpy(z_faces, faces.data() + nfx + nfy, nfz * sizeof(float));
    if (report != nullptr) *report = rep;
  });
}

int msim_hip_solve_depth_averaged_port_flow(msim_ctx* ctx, const msim_grid* grid,
                                            const float* mobility, const msim_flow_port* ports,
                                            uint32_t port_count,
                                            const msim_port_flow_params* params, float* x_faces,
                                            float* y_faces, float* z_faces,
                                            msim_port_flow_report* report,
                                            msim_port_state* states) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || params == nullptr || report == nullptr) {
      msim::invalid("grid, parameters and report must not be null");
    }
    // DepthAveragedPortFlowParameters::validate (cpu_flow.cpp:187-192).
    if (!std::isfinite(params->relative_tolerance) ||
        params->relative_tolerance < std::numeric_limits<float>::epsilon()) {
      msim::invalid(
          "depth-averaged relative tolerance must be finite and at least float machine epsilon");
    }
    if (params->max_iterations == 0) msim::invalid("depth-averaged iteration limit must be positive");
    if (params->solver > MSIM_SOLVER_MULTIGRID) msim::invalid("unknown pressure solver");
    std::vector<uint8_t> solid;
    std::vector<float> mob;
    auto& f = c.flow;
    f.prepared = false;
    const msim::PortPlan plan =
        msim::plan_ports(*grid, mobility, ports, port_count, solid, mob, f.layers);
    f.nx = grid->nx;
    f.ny = grid->ny;
    f.dx = grid->dx;
    f.dy = grid->dy;
    f.dz = grid->nz == 1 ? grid->dz : 1.0f;
    f.axis = 0;
    f.ported = true;
    const size_t slots = plan.open.size();
    f.edge_open.reserve(slots);
    f.edge_pressure.reserve(slots);
    msim::check(hipMemcpyAsync(f.edge_open.data(), plan.open.data(), slots,
                               hipMemcpyHostToDevice, c.stream),
                "upload open port faces");
    msim::check(hipMemsetAsync(f.edge_pressure.data(), 0, slots * sizeof(float), c.stream),
                "clear port pressures");
    msim::flow_assemble(c, solid.data(), mob.data());
    msim::flow_upload_layers(c);
    msim::port_flow_solve(c, plan, *params, x_faces, y_faces, z_faces, *report, states);
  });
}

int msim_hip_flow_apply(msim_ctx* ctx, const double* x, double* y) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.flow.prepared) msim::invalid("no flow system has been prepared");
    if (x == nullptr || y == nullptr) msim::invalid("vectors must not be null");
    msim::flow_apply(c, x, y);
  });
}

int msim_hip_flow_precondition(msim_ctx* ctx, const float* r, float* z) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.flow.prepared) msim::invalid("no flow system has been prepared");
    if (r == nullptr || z == nullptr) msim::invalid("vectors must not be null");
    msim::mg_precondition(c, r, z, false);
  });
}

int msim_hip_learned_set_weights(msim_ctx* ctx, const float* weights, size_t count) {
  return msim::guarded(ctx, [&](Context& c) {
    if (weights == nullptr) msim::invalid("weights must not be null");
    if (count != MSIM_LEARNED_WEIGHTS) {
      msim::invalid("the learned smoother has " + std::to_string(MSIM_LEARNED_WEIGHTS) +
                    " weights, got " + std::to_string(count));
    }
    for (size_t k = 0; k < count; ++k) {
      if (!std::isfinite(weights[k])) msim::invalid("weights must be finite");
    }
    msim::learned_set_weights(c, weights);
  });
}

int msim_hip_learned_get_weights(msim_ctx* ctx, float* weights, size_t count) {
  return msim::guarded(ctx, [&](Context& c) {
    if (weights == nullptr || count != MSIM_LEARNED_WEIGHTS) {
      msim::invalid("weights must hold MSIM_LEARNED_WEIGHTS floats");
    }
    msim::learned_get_weights(c, weights);
  });
}

int msim_hip_learned_precondition(msim_ctx* ctx, const float* r, float* z) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.flow.prepared) msim::invalid("no flow system has been prepared");
    if (r == nullptr || z == nullptr) msim::invalid("vectors must not be null");
    if (!c.learned.loaded) msim::invalid("the learned smoother has no weights: set them first");
    msim::mg_precondition(c, r, z, true);
  });
}

int msim_hip_learned_operator(msim_ctx* ctx, float* wx, float* wy, float* bc, float* dinv) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.flow.prepared) msim::invalid("no flow system has been prepared");
    msim::mg_operator(c, wx, wy, bc, dinv);
  });
}

int msim_hip_learned_harvest(msim_ctx* ctx, uint32_t max_pairs, uint32_t use_learned,
                             float* residuals, float* errors, uint32_t* produced) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.flow.prepared) msim::invalid("no flow system has been prepared");
    if (residuals == nullptr || errors == nullptr || produced == nullptr) {
      msim::invalid("output arrays must not be null");
    }
    if (max_pairs == 0) msim::invalid("max_pairs must be positive");
    if (use_learned != 0 && !c.learned.loaded) {
      msim::invalid("the learned smoother has no weights: set them first");
    }
    *produced = 0;
    *produced = msim::mg_harvest(c, max_pairs, use_learned != 0, residuals, errors);
  });
}

int msim_hip_learned_upload(msim_ctx* ctx, const msim_train_set* set) {
  return msim::guarded(ctx, [&](Context& c) {
    if (set == nullptr) msim::invalid("the training set must not be null");
    if (set->nx == 0 || set->ny == 0 || set->geometries == 0 || set->pairs == 0) {
      msim::invalid("the training set is empty");
    }
    if (static_cast<uint64_t>(set->nx) * set->ny > (uint64_t{1} << 31)) {
      throw msim::Error(MSIM_OVERFLOW, "training grids are limited to 2^31 sites");
    }
    if (set->operators == nullptr || set->pair_geometry == nullptr ||
        set->residuals == nullptr || set->errors == nullptr) {
      msim::invalid("training set arrays must not be null");
    }
    if (set->pairs > 65535) msim::invalid("at most 65535 pairs per training set");
    for (uint32_t p = 0; p < set->pairs; ++p) {
      if (set->pair_geometry[p] >= set->geometries) {
        msim::invalid("pair " + std::to_string(p) + " names geometry " +
                      std::to_string(set->pair_geometry[p]) + " of " +
                      std::to_string(set->geometries));
      }
    }
    msim::learned_upload(c, *set);
  });
}

int msim_hip_learned_train(msim_ctx* ctx, const msim_train_params* params, float* losses,
                           double* seconds) {
  return msim::guarded(ctx, [&](Context& c) {
    if (params == nullptr) msim::invalid("parameters must not be null");
    if (params->steps == 0 || params->batch == 0 || params->batch > 65535) {
      msim::invalid("steps and batch (1..65535) must be positive");
    }
    if (!(params->learning_rate > 0.0f) || !(params->beta1 >= 0.0f && params->beta1 < 1.0f) ||
        !(params->beta2 >= 0.0f && params->beta2 < 1.0f) || !(params->epsilon > 0.0f)) {
      msim::invalid("Adam needs learning_rate > 0, 0 <= beta < 1 and epsilon > 0");
    }
    msim::learned_train(c, *params, losses, seconds);
  });
}

int msim_hip_learned_gradients(msim_ctx* ctx, uint32_t pair, double* loss, double* gradient) {
  return msim::guarded(ctx, [&](Context& c) {
    if (loss == nullptr || gradient == nullptr) msim::invalid("outputs must not be null");
    msim::learned_gradients(c, pair, loss, gradient);
  });
}

int msim_hip_learned_apply(msim_ctx* ctx, uint32_t pair, float* z) {
  return msim::guarded(ctx, [&](Context& c) {
    if (z == nullptr) msim::invalid("output must not be null");
    msim::learned_apply(c, pair, z);
  });
}

int msim_hip_flow_arrays(msim_ctx* ctx, float* conductance, float* diagonal, float* rhs,
                         float* boundary_conductance) {
  return msim::guarded(ctx, [&](Context& c) {
    auto& f = c.flow;
    if (!f.prepared) msim::invalid("no flow system has been prepared");
    const size_t bytes = f.sites() * sizeof(float);
    const std::pair<float*, float*> copies[] = {{conductance, f.cond.data()},
                                                {diagonal, f.diag.data()},
                                                {rhs, f.rhs.data()},
                                                {boundary_conductance, f.bc.data()}};
    for (const auto& [dst, src] : copies) {
      if (dst != nullptr) {
        msim::check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream), "download");
      }
    }
    msim::check(hipStreamSynchronize(c.stream), "download");
  });
}

int msim_hip_advance_signal_grid(msim_ctx* ctx, const msim_signal_grid* grid, float dt,
                                 float* levels, msim_signal_report* report) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || levels == nullptr) msim::invalid("grid and levels must not be null");
    const bool refresh = grid->revision == 0 || grid->revision != c.adr.revision ||
                         !c.adr.uploaded || c.adr.flow_field;
    if (refresh) msim::validate_signal_grid(*grid, false);
    msim::validate_levels(*grid, levels);
    msim::adr_upload(c, *grid, levels, 32, false);
    if (refresh) msim::record_losses(c, *grid);
    msim::validate_step(c, dt);
    if (report != nullptr) *report = msim_signal_report{1, 0, 0.0f};
    if (dt == 0.0f) return;
    msim::adr_run(c, dt, 1, nullptr);
    if (report != nullptr) *report = c.adr.last_report;
    msim::adr_download(c, levels);
  });
}

int msim_hip_advance_signal_grid_static(msim_ctx* ctx, const msim_signal_grid* grid, float dt,
                                        float* levels, msim_signal_report* report) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || levels == nullptr) msim::invalid("grid and levels must not be null");
    msim::validate_signal_grid(*grid, false);
    msim::validate_levels(*grid, levels);
    if (report != nullptr) *report = msim_signal_report{1, 0, 0.0f};
    msim::adr_step_static(c, *grid, dt, levels);
  });
}

size_t msim_hip_static_scratch_bytes(void) { return msim::static_scratch_bytes(); }

int msim_hip_nutrient_run(msim_ctx* ctx, const msim_nutrient_params* params, uint32_t steps,
                          double* ledger, msim_nutrient_step* per_step) {
  return msim::guarded(ctx, [&](Context& c) {
    if (params == nullptr || params->base_mobility == nullptr) {
      msim::invalid("parameters and the base mobility must not be null");
    }
    auto& f = c.flow;
    auto& a = c.adr;
    if (!f.prepared || f.ported || f.layers.nz != 1) {
      msim::invalid("the nutrient loop needs a prepared axis flow on a grid of one layer");
    }
    if (!a.uploaded || a.precision != 64 || !a.flow_field || a.params.has_monod == 0 ||
        a.params.has_affine != 0 || a.integration != MSIM_FORWARD_EULER ||
        a.params.nx != f.nx || a.params.ny != f.ny) {
      msim::invalid(
          "the nutrient loop needs FP64 forward-Euler transport with Monod uptake and no affine "
          "reaction, uploaded over the prepared flow's field");
    }
    msim::validate_flow_params(params->flow);
    if (params->flow.scale == 0) msim::invalid("the nutrient loop takes the scaled field");
    if (params->flow.axis != f.axis) msim::invalid("the flow axis differs from the prepared one");
    if (!std::isfinite(params->dt) || params->dt <= 0.0f) {
      msim::invalid("time step must be finite and positive");
    }
    if (!std::isfinite(params->drag_coefficient) || params->drag_coefficient < 0.0f) {
      msim::invalid("drag coefficient must be finite and non-negative");
    }
    if (!std::isfinite(params->max_volume_fraction) || params->max_volume_fraction < 0.0f ||
        params->max_volume_fraction >= 1.0f) {
      msim::invalid("max volume fraction must be finite and in [0, 1)");
    }
    if (!std::isfinite(params->yield) || params->yield < 0.0f) {
      msim::invalid("yield must be finite and non-negative");
    }
    const size_t n = f.sites();
    for (size_t i = 0; i < n; ++i) {
      if (!std::isfinite(params->base_mobility[i]) || params->base_mobility[i] < 0.0f) {
        msim::invalid("flow mobility values must be finite and non-negative");
      }
    }
    msim::nutrient_run(c, *params, steps, ledger, per_step);
  });
}

int msim_hip_copy_bandwidth(msim_ctx* ctx, size_t bytes, uint32_t repeats, double* gb_per_s) {
  return msim::guarded(ctx, [&](Context& c) {
    if (gb_per_s == nullptr || repeats == 0 || bytes < 16) {
      msim::invalid("copy bandwidth needs an output, repeats and at least 16 bytes");
    }
    *gb_per_s = msim::stream_copy_bandwidth(c, bytes, repeats);
  });
}

int msim_hip_adr_occupancy(msim_ctx* ctx, float* phi) {
  return msim::guarded(ctx, [&](Context& c) {
    if (phi == nullptr) msim::invalid("phi must not be null");
    if (!c.adr.uploaded || c.adr.params.has_monod == 0) {
      msim::invalid("no transport with Monod uptake on the device");
    }
    msim::check(hipMemcpyAsync(phi, c.adr.phi.data(), c.adr.sites() * sizeof(float),
                               hipMemcpyDeviceToHost, c.stream),
                "download occupancy");
    msim::check(hipStreamSynchronize(c.stream), "download occupancy");
  });
}

int msim_hip_adr_upload(msim_ctx* ctx, const msim_signal_grid* grid, const void* levels,
                        uint32_t precision, uint32_t use_flow_field) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || levels == nullptr) msim::invalid("grid and levels must not be null");
    if (precision != 32 && precision != 64) msim::invalid("precision must be 32 or 64");
    if (precision == 64 && grid->integration != MSIM_FORWARD_EULER) {
      throw Error(MSIM_UNSUPPORTED, "implicit signal steps hold FP32 levels");
    }
    if (precision == 64 && grid->grid.nz != 1) {
      throw Error(MSIM_UNSUPPORTED, "FP64 levels take a grid of one layer (nz = 1)");
    }
    if (use_flow_field != 0) {
      if (!c.flow.solved) msim::invalid("no flow field on the device");
      if (c.flow.layers.nz != 1) msim::invalid("the flow field on the device is of a layered grid");
      if (c.flow.nx != grid->grid.nx || c.flow.ny != grid->grid.ny || grid->grid.nz != 1) {
        msim::invalid("the flow field is for another grid");
      }
    }
    msim::validate_signal_grid(*grid, use_flow_field != 0);
    if (precision == 64) {
      msim::validate_levels(*grid, static_cast<const double*>(levels));
    } else {
      msim::validate_levels(*grid, static_cast<const float*>(levels));
    }
    msim::adr_upload(c, *grid, levels, precision, use_flow_field != 0);
    msim::record_losses(c, *grid);
  });
}

int msim_hip_adr_run(msim_ctx* ctx, float dt, uint32_t steps, double* ledger) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.adr.uploaded) msim::invalid("no signal grid on the device");
    // A transport uploaded over the device's flow field reads it in place:
    // a later flow on another grid would leave it reading the wrong field.
    msim::check_flow_field(c);
    msim::validate_step(c, dt);
    if (dt == 0.0f || steps == 0) return;
    msim::adr_run(c, dt, steps, ledger);
  });
}

int msim_hip_adr_download(msim_ctx* ctx, void* levels) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.adr.uploaded) msim::invalid("no signal grid on the device");
    if (levels == nullptr) msim::invalid("levels must not be null");
    msim::adr_download(c, levels);
  });
}

double msim_hip_adr_seconds(msim_ctx* ctx) {
  return ctx == nullptr ? 0.0 : ctx->impl.adr.seconds;
}

int msim_hip_adr_report(msim_ctx* ctx, msim_signal_report* report) {
  return msim::guarded(ctx, [&](Context& c) {
    if (report == nullptr) msim::invalid("report must not be null");
    *report = c.adr.last_report;
  });
}

int msim_hip_advance_coupled(msim_ctx* ctx, const msim_signal_grid* grid, const float* origin,
                             float* levels, const msim_cells* cells, float* species,
                             const msim_rate_plan* plan, const float* previous_lengths,
                             float dt, msim_signal_report* report) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || origin == nullptr || levels == nullptr || cells == nullptr ||
        plan == nullptr) {
      msim::invalid("grid, origin, levels, cells and plan must not be null");
    }
    // validate_coupled_step (cpu_coupled.cpp:90-123), in its order.
    if (!std::isfinite(dt) || dt < 0.0f) {
      msim::invalid("coupled time step must be finite and non-negative");
    }
    msim::validate_coupled_inputs(*grid, origin, levels, *cells, species, *plan);
    c.coupled.uploaded = false;
    msim::adr_upload(c, *grid, levels, 32, false);
    msim::record_losses(c, *grid);
    msim::validate_step(c, dt);
    msim::validate_coupled_plan(*grid, *plan);
    msim::validate_previous(*cells, previous_lengths);
    msim::validate_positions(*grid, origin, *cells);
    msim::check_coupled_limits(*plan, *cells);
    msim::coupled_upload(c, origin, *cells, species, *plan, grid->grid.obstacles != nullptr);
    msim::coupled_run(c, previous_lengths, dt, 1);
    if (report != nullptr) *report = c.adr.last_report;
    msim::coupled_download(c, levels, species);
  });
}

int msim_hip_coupled_upload(msim_ctx* ctx, const msim_signal_grid* grid, const float* origin,
                            const float* levels, const msim_cells* cells, const float* species,
                            const msim_rate_plan* plan) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || origin == nullptr || levels == nullptr || cells == nullptr ||
        plan == nullptr) {
      msim::invalid("grid, origin, levels, cells and plan must not be null");
    }
    c.coupled.uploaded = false;
    msim::validate_coupled_inputs(*grid, origin, levels, *cells, species, *plan);
    msim::validate_coupled_plan(*grid, *plan);
    msim::validate_positions(*grid, origin, *cells);
    msim::upload_coupled(c, *grid, origin, levels, *cells, species, *plan);
  });
}

int msim_hip_coupled_run(msim_ctx* ctx, const float* previous_lengths, float dt, uint32_t steps,
                         msim_signal_report* report) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.coupled.uploaded || !c.adr.uploaded) msim::invalid("no coupled state on the device");
    msim::check_flow_field(c);
    if (!std::isfinite(dt) || dt < 0.0f) {
      msim::invalid("coupled time step must be finite and non-negative");
    }
    msim::validate_step(c, dt);
    if (previous_lengths != nullptr) {
      for (uint32_t k = 0; k < c.coupled.cells; ++k) {
        if (!std::isfinite(previous_lengths[k]) || previous_lengths[k] < 0.0f) {
          msim::invalid("previous cell lengths must be finite and non-negative");
        }
      }
    }
    if (steps > 0) msim::coupled_run(c, previous_lengths, dt, steps);
    if (report != nullptr) *report = c.adr.last_report;
  });
}

int msim_hip_coupled_download(msim_ctx* ctx, float* levels, float* species) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!c.coupled.uploaded || !c.adr.uploaded) msim::invalid("no coupled state on the device");
    msim::coupled_download(c, levels, species);
  });
}

int msim_hip_colony_volume_fraction(msim_ctx* ctx, const msim_grid* grid, const float* origin,
                                    const msim_cells* cells, double averaging_radius,
                                    double* density) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || origin == nullptr || cells == nullptr || density == nullptr) {
      msim::invalid("grid, origin, cells and density must not be null");
    }
    msim::colony_run(c, *grid, origin, *cells, averaging_radius, nullptr, 0, 0.0, 0.0, density,
                     nullptr, std::string());
  });
}

int msim_hip_colony_mobility(msim_ctx* ctx, const msim_grid* grid, const float* origin,
                             const msim_cells* cells, const double* base, uint32_t base_count,
                             double drag_coefficient, double max_volume_fraction,
                             double averaging_radius, double* mobility) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || origin == nullptr || cells == nullptr || base == nullptr ||
        mobility == nullptr) {
      msim::invalid("grid, origin, cells, base and mobility must not be null");
    }
    const std::string pending =
        msim::darcy_pending(*grid, base, base_count, drag_coefficient, max_volume_fraction);
    msim::colony_run(c, *grid, origin, *cells, averaging_radius, base,
                     base_count == 0 ? 1u : base_count, drag_coefficient, max_volume_fraction,
                     nullptr, mobility, pending);
  });
}

int msim_hip_colony_flow_coupling(msim_ctx* ctx, const msim_grid* grid, const float* origin,
                                  const msim_cells* cells, const float* velocities,
                                  const double* base, uint32_t base_count,
                                  double drag_coefficient, double max_volume_fraction,
                                  double averaging_radius, double* mobility, double* force) {
  return msim::guarded(ctx, [&](Context& c) {
    if (grid == nullptr || origin == nullptr || cells == nullptr || base == nullptr ||
        mobility == nullptr || force == nullptr ||
        (cells->count > 0 && velocities == nullptr)) {
      msim::invalid("grid, origin, cells, velocities, base, mobility and force must not be null");
    }
    const std::string pending =
        msim::darcy_pending(*grid, base, base_count, drag_coefficient, max_volume_fraction);
    msim::colony_run(c, *grid, origin, *cells, averaging_radius, base,
                     base_count == 0 ? 1u : base_count, drag_coefficient, max_volume_fraction,
                     nullptr, mobility, pending, velocities, force);
  });
}

double msim_hip_colony_seconds(msim_ctx* ctx) {
  return ctx == nullptr ? 0.0 : ctx->impl.colony.seconds;
}

double msim_hip_coupled_seconds(msim_ctx* ctx) {
  return ctx == nullptr ? 0.0 : ctx->impl.coupled.seconds;
}

}  // extern "C"

int msim_hip_batch_upload(msim_ctx* ctx, uint32_t count, const msim_batch_instance* instances,
                          const msim_batch_params* params) {
  return msim::guarded(ctx, [&](Context& c) {
    if (instances == nullptr || params == nullptr) msim::invalid("arguments must not be null");
    if (count == 0 || count > 65535) msim::invalid("a batch holds 1 to 65535 instances");
    if (!(params->relative_tolerance > 0.0f) || params->max_iterations == 0) {
      msim::invalid("the batch needs a positive tolerance and iteration cap");
    }
    if (!(params->monod_vmax >= 0.0f) || !(params->monod_k > 0.0f)) {
      msim::invalid("Monod uptake needs vmax >= 0 and K > 0");
    }
    uint64_t total = 0;
    for (uint32_t k = 0; k < count; ++k) {
      const msim_batch_instance& in = instances[k];
      const std::string where = "batch instance " + std::to_string(k) + ": ";
      if (in.nx < 2 || in.ny < 1) msim::invalid(where + "needs nx >= 2 and ny >= 1");
      const float h[3] = {in.dx, in.dy, in.dz};
      for (float v : h) {
        if (!(v > 0.0f) || !std::isfinite(v)) msim::invalid(where + "spacing must be positive");
      }
      if (in.mobility == nullptr) msim::invalid(where + "mobility must not be null");
      if (!std::isfinite(in.inlet_pressure) || !(in.diffusion >= 0.0f) ||
          !std::isfinite(in.diffusion) || !(in.reservoir >= 0.0f)) {
        msim::invalid(where + "inlet pressure, diffusion and reservoir must be finite (>= 0)");
      }
      const size_t n = static_cast<size_t>(in.nx) * in.ny;
      for (size_t i = 0; i < n; ++i) {
        if (!(in.mobility[i] >= 0.0f) || !std::isfinite(in.mobility[i])) {
          msim::invalid(where + "mobility must be finite and non-negative");
        }
      }
      total += n;
    }
    if (total > (uint64_t{1} << 31)) {
      throw msim::Error(MSIM_OVERFLOW, "a batch is limited to 2^31 sites");
    }
    msim::batch_upload(c, count, instances, *params);
  });
}

int msim_hip_batch_solve_flow(msim_ctx* ctx, uint32_t* iterations, float* residuals,
                              uint32_t* flags) {
  return msim::guarded(ctx, [&](Context& c) { msim::batch_solve_flow(c, iterations, residuals, flags); });
}

int msim_hip_batch_autotune(msim_ctx* ctx, float dt, uint32_t steps, int* shape,
                            double* seconds) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!(dt >= 0.0f) || !std::isfinite(dt) || steps == 0) {
      msim::invalid("autotuning needs dt >= 0 and steps > 0");
    }
    const int k = msim::batch_autotune(c, dt, steps, seconds);
    if (shape != nullptr) *shape = k;
  });
}

int msim_hip_batch_shapes(void) { return msim::batch_shape_count(); }

int msim_hip_batch_shape(int index, uint32_t* threads, uint32_t* row_width,
                         uint32_t* item_sites) {
  if (index < 0 || index >= msim::batch_shape_count() || threads == nullptr ||
      row_width == nullptr || item_sites == nullptr) {
    return MSIM_INVALID_ARGUMENT;
  }
  int t = 0, r = 0, k = 0;
  msim::batch_shape(index, &t, &r, &k);
  *threads = static_cast<uint32_t>(t);
  *row_width = static_cast<uint32_t>(r);
  *item_sites = static_cast<uint32_t>(k);
  return MSIM_OK;
}

int msim_hip_batch_flags(msim_ctx* ctx, uint32_t* flags) {
  return msim::guarded(ctx, [&](Context& c) {
    if (flags == nullptr) msim::invalid("flags must not be null");
    msim::batch_flags(c, flags);
  });
}

int msim_hip_batch_run(msim_ctx* ctx, float dt, uint32_t steps, uint32_t use_graph,
                       double* ledger, double* seconds) {
  return msim::guarded(ctx, [&](Context& c) {
    if (!(dt >= 0.0f) || !std::isfinite(dt)) msim::invalid("dt must be finite and >= 0");
    if (steps == 0) msim::invalid("steps must be positive");
    msim::batch_run(c, dt, steps, use_graph != 0, ledger, seconds);
  });
}

int msim_hip_batch_download(msim_ctx* ctx, uint32_t index, double* pressure, float* x_faces,
                            float* y_faces, float* levels) {
  return msim::guarded(ctx, [&](Context& c) {
    msim::batch_download(c, index, pressure, x_faces, y_faces, levels);
  });
}
