// Port-driven depth-averaged flow on the device, and the lift of a solved
// column flux onto the faces of a grid of several layers.
//
// A port-driven flow is a superposition (superpose_port_flows,
// cpp/core/flow_system.hpp:1139-1252): every solve shares one operator (all
// ports' faces open), only the pressures held outside the ports change. The
// device assembles the operator once, reassembles the right-hand side for
// each solve, keeps every solve's lifted field resident (one slot of x|y|z
// voxel faces per solve), gathers the values on the ports' faces for the
// host's small dense system, and combines the slots in FP64 into the final
// FP32 field in one pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kLiftThreads = 256;

// The x and y voxel faces of a lifted field (ShallowFlowReduction::lift,
// flow_system.hpp:424-444): a face between two fluid voxels (or an exterior
// face of a fluid voxel) carries its column face's flux over the smaller of
// the two columns' heights; a face next to a solid voxel, or between columns
// without height, carries nothing. flux: x|y faces of the grid of columns.
__global__ __launch_bounds__(kLiftThreads) void k_lift_faces(
    const float* __restrict__ flux, const uint8_t* __restrict__ voxels,
    const float* __restrict__ heights, float* __restrict__ out, uint32_t nx, uint32_t ny,
    uint32_t nz) {
  const size_t nxf = static_cast<size_t>(nx + 1) * ny * nz;
  const size_t nyf = static_cast<size_t>(nx) * (ny + 1) * nz;
  const size_t i = static_cast<size_t>(blockIdx.x) * kLiftThreads + threadIdx.x;
  if (i >= nxf + nyf) return;
  const bool is_x = i < nxf;
  const size_t local = is_x ? i : i - nxf;
  const uint32_t z = static_cast<uint32_t>(local % nz);
  const size_t t = local / nz;
  const uint32_t width = is_x ? ny : ny + 1;
  const uint32_t a = static_cast<uint32_t>(t / width);  // x (face x on an x face)
  const uint32_t b = static_cast<uint32_t>(t % width);  // y (face y on a y face)
  const uint32_t along = is_x ? a : b;
  const uint32_t dim = is_x ? nx : ny;
  const bool has_lower = along > 0;
  const bool has_upper = along < dim;
  const size_t up_col = static_cast<size_t>(a) * ny + b;  // read only when has_upper
  const size_t low_col = is_x ? up_col - ny : up_col - 1; // read only when has_lower
  float v = 0.0f;
  const bool blocked = (has_lower && voxels[low_col * nz + z] != 0) ||
                       (has_upper && voxels[up_col * nz + z] != 0);
  if (!blocked) {
    const size_t a_col = has_lower ? low_col : up_col;
    const size_t b_col = has_upper ? up_col : a_col;
    const float ha = heights[a_col], hb = heights[b_col];
    const float height = hb < ha ? hb : ha;
    const size_t column_face = is_x ? t : static_cast<size_t>(nx + 1) * ny + t;
    if (height > 0.0f) v = flux[column_face] / height;
  }
  out[i] = v;
}

// The z voxel faces of a lifted field (flow_system.hpp:445-469), one thread
// per column: the horizontal divergence of each fluid voxel from the common
// floor up, its mean over the column, and the vertical faces by continuity
// with the mean removed, so that floor and roof stay closed. Every other z
// face is zero. Reads the x and y faces k_lift_faces wrote.
__global__ __launch_bounds__(kLiftThreads) void k_lift_columns(
    const uint8_t* __restrict__ voxels, float* __restrict__ out, uint32_t nx, uint32_t ny,
    uint32_t nz, uint32_t bottom, float dx, float dy, float dz) {
  const size_t column = static_cast<size_t>(blockIdx.x) * kLiftThreads + threadIdx.x;
  if (column >= static_cast<size_t>(nx) * ny) return;
  const uint32_t x = static_cast<uint32_t>(column / ny);
  const uint32_t y = static_cast<uint32_t>(column % ny);
  const size_t nxf = static_cast<size_t>(nx + 1) * ny * nz;
  const size_t nyf = static_cast<size_t>(nx) * (ny + 1) * nz;
  float* zf = out + nxf + nyf + column * (nz + 1);
  for (uint32_t fz = 0; fz <= nz; ++fz) zf[fz] = 0.0f;
  const float* xf0 = out + (static_cast<size_t>(x) * ny + y) * nz;
  const float* xf1 = out + (static_cast<size_t>(x + 1) * ny + y) * nz;
  const float* yf0 = out + nxf + (static_cast<size_t>(x) * (ny + 1) + y) * nz;
  const float* yf1 = yf0 + nz;
  // horizontal(z): float differences over float spacings, summed in double.
  auto horizontal = [&](uint32_t z) {
    const double dxv = (xf1[z] - xf0[z]) / dx;
    const double dyv = (yf1[z] - yf0[z]) / dy;
    return dxv + dyv;
  };
  uint32_t count = 0;
  double sum = 0.0;
  for (uint32_t z = bottom; z < nz; ++z) {
    if (voxels[column * nz + z] != 0) break;
    sum += horizontal(z);
    ++count;
  }
  if (count == 0) return;
  const double mean = sum / static_cast<double>(count);
  double vertical = 0.0;
  for (uint32_t k = 0; k + 1 < count; ++k) {
    vertical -= static_cast<double>(dz) * (horizontal(bottom + k) - mean);
    zf[bottom + k + 1] = static_cast<float>(vertical);
  }
}

// values[j * P + i] = fields[j * F + faces[i]]: every solve's velocity on the
// ports' open faces.
__global__ void k_port_gather(const float* __restrict__ fields, size_t F,
                              const uint64_t* __restrict__ faces, size_t P, uint32_t J,
                              float* __restrict__ values) {
  const size_t total = P * J;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += stride) {
    const size_t j = k / P, i = k % P;
    values[k] = fields[j * F + faces[i]];
  }
}

// The combined field (superpose_port_flows, flow_system.hpp:1168-1240): per
// face, from 0.0, each solve's velocity in FP64 times its weight, added in
// solve order, then rounded to FP32. Each block leaves the largest |velocity|
// it wrote as float bits in peak[block] (positive floats order as their
// bits, and a NaN or an infinity compares above every finite value).
__global__ __launch_bounds__(256) void k_port_combine(const float* __restrict__ fields, size_t F,
                                                       uint32_t J,
                                                       const double* __restrict__ weights,
                                                       float* __restrict__ out,
                                                       uint32_t* __restrict__ peak) {
  __shared__ uint32_t s_peak[256 / kWave];
  uint32_t largest = 0;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t f = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < F;
       f += stride) {
    double v = 0.0;
    for (uint32_t j = 0; j < J; ++j) v += weights[j] * static_cast<double>(fields[j * F + f]);
    const float r = static_cast<float>(v);
    out[f] = r;
    largest = max(largest, __float_as_uint(fabsf(r)));
  }
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    largest = max(largest, static_cast<uint32_t>(__shfl_xor(static_cast<int>(largest), offset,
                                                            kWave)));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) s_peak[threadIdx.x / kWave] = largest;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < 256 / kWave; ++w) largest = max(largest, s_peak[w]);
    peak[blockIdx.x] = largest;
  }
}

uint32_t grid_blocks(size_t n, uint32_t threads, uint32_t cap) {
  const size_t blocks = (n + threads - 1) / threads;
  return static_cast<uint32_t>(std::max<size_t>(1, std::min<size_t>(blocks, cap)));
}

}  // namespace

void flow_upload_layers(Context& ctx) {
  auto& f = ctx.flow;
  const auto& L = f.layers;
  if (L.nz <= 1) return;
  f.voxels.reserve(L.voxels.size());
  f.heights.reserve(L.heights.size());
  check(hipMemcpyAsync(f.voxels.data(), L.voxels.data(), L.voxels.size(), hipMemcpyHostToDevice,
                       ctx.stream),
        "upload voxel mask");
  check(hipMemcpyAsync(f.heights.data(), L.heights.data(), L.heights.size() * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload column heights");
  check(hipStreamSynchronize(ctx.stream), "upload layers");
}

void flow_lift(Context& ctx, float* out, bool with_z, const float* source) {
  auto& f = ctx.flow;
  const uint32_t nz = f.layers.nz;
  const float* faces = source != nullptr ? source : f.faces.data();
  if (nz == 1) {
    // One layer: the face velocities are the field; no vertical faces flow.
    const size_t planar = f.x_faces() + f.y_faces();
    check(hipMemcpyAsync(out, faces, planar * sizeof(float), hipMemcpyDeviceToDevice,
                         ctx.stream),
          "copy face velocities");
    if (with_z) {
      check(hipMemsetAsync(out + planar, 0, 2 * f.sites() * sizeof(float), ctx.stream),
            "clear z faces");
    }
    return;
  }
  const size_t planar = (f.x_faces() + f.y_faces()) * nz;
  k_lift_faces<<<grid_blocks(planar, kLiftThreads, 0xffffffffu), kLiftThreads, 0, ctx.stream>>>(
      faces, f.voxels.data(), f.heights.data(), out, f.nx, f.ny, nz);
  check(hipGetLastError(), "launch k_lift_faces");
  k_lift_columns<<<grid_blocks(f.sites(), kLiftThreads, 0xffffffffu), kLiftThreads, 0,
                   ctx.stream>>>(f.voxels.data(), out, f.nx, f.ny, nz, f.layers.bottom, f.dx,
                                 f.dy, f.layers.dz);
  check(hipGetLastError(), "launch k_lift_columns");
}

void port_flow_solve(Context& ctx, const PortPlan& plan, const msim_port_flow_params& params,
                     float* x_faces, float* y_faces, float* z_faces, msim_port_flow_report& report,
                     msim_port_state* states) {
  auto& f = ctx.flow;
  std::memset(&report, 0, sizeof(report));
  const uint32_t count = static_cast<uint32_t>(plan.ports.size());
  std::vector<uint32_t> driven;
  std::vector<float> held(count, 0.0f);
  bool holds_pressure = false;
  for (uint32_t p = 0; p < count; ++p) {
    if (plan.ports[p].kind == MSIM_PORT_SPEED) {
      driven.push_back(p);
    } else {
      held[p] = plan.ports[p].value;
      holds_pressure = holds_pressure || plan.ports[p].value != 0.0f;
    }
  }
  const uint32_t first_driven = holds_pressure ? 1u : 0u;
  const uint32_t J = first_driven + static_cast<uint32_t>(driven.size());
  // A slot holds every face of the field; on one layer the vertical faces
  // are zero in every solve (a port cannot open a z face), so a slot holds
  // the x and y faces only.
  const size_t nz = f.layers.nz;
  const size_t F = nz == 1 ? f.x_faces() + f.y_faces() : f.voxel_faces();
  f.port_fields.reserve(static_cast<size_t>(J) * F);

  msim_flow_params solve{};
  solve.mean_inlet_speed = 1.0f;
  solve.relative_tolerance = static_cast<double>(params.relative_tolerance);
  solve.max_iterations = params.max_iterations;
  solve.solver = params.solver;
  solve.scale = 0;
  solve.check_interval = params.check_interval;
  solve.inner_tolerance = params.inner_tolerance;
  solve.fallback_iterations = params.fallback_iterations;
  double seconds = 0.0;
  std::vector<float> edge(plan.open.size(), 0.0f);
  // One pressure solve holding each port at pressures[port], lifted into
  // field slot `slot`.
  auto run = [&](const std::vector<float>& pressures, uint32_t slot) {
    for (size_t s = 0; s < edge.size(); ++s) {
      edge[s] = plan.owner[s] >= 0 ? pressures[static_cast<size_t>(plan.owner[s])] : 0.0f;
    }
    flow_hold_pressures(ctx, edge.data());
    msim_flow_report r{};
    flow_solve(ctx, solve, r);
    flow_lift(ctx, f.port_fields.data() + static_cast<size_t>(slot) * F, nz > 1);
    ++report.solves;
    report.iterations += r.iterations;
    report.relative_residual =
        std::max(report.relative_residual, static_cast<float>(r.relative_residual));
    report.true_relative_residual = std::max(report.true_relative_residual, r.true_relative_residual);
    seconds += r.seconds;
  };
  if (holds_pressure) run(held, 0);
  std::vector<float> unit(count, 0.0f);
  for (size_t j = 0; j < driven.size(); ++j) {
    std::fill(unit.begin(), unit.end(), 0.0f);
    unit[driven[j]] = 1.0f;
    run(unit, first_driven + static_cast<uint32_t>(j));
  }

  // Every solve's velocity on the ports' open faces.
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  const size_t P = plan.faces.size();
  std::vector<float> values(static_cast<size_t>(J) * P);
  if (P > 0) {
    f.port_faces.reserve(P);
    f.port_values.reserve(values.size());
    check(hipMemcpyAsync(f.port_faces.data(), plan.faces.data(), P * sizeof(uint64_t),
                         hipMemcpyHostToDevice, ctx.stream),
          "upload port faces");
    k_port_gather<<<grid_blocks(values.size(), 256, 1024), 256, 0, ctx.stream>>>(
        f.port_fields.data(), F, f.port_faces.data(), P, J, f.port_values.data());
    check(hipGetLastError(), "launch k_port_gather");
    check(hipMemcpyAsync(values.data(), f.port_values.data(), values.size() * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download port values");
  }
  check(hipStreamSynchronize(ctx.stream), "gather port values");

  // The flux through port q of slot j's field (PortFaces::flux,
  // flow_system.hpp:1116-1122): its faces' FP64 velocities summed in order,
  // times the signed area of one face.
  auto slot_flux = [&](uint32_t q, uint32_t j) {
    double sum = 0.0;
    for (size_t i = plan.offsets[q]; i < plan.offsets[q + 1]; ++i) {
      sum += static_cast<double>(values[static_cast<size_t>(j) * P + i]);
    }
    return sum * plan.signed_area[q];
  };
  auto open_area = [&](uint32_t q) {
    return std::abs(plan.signed_area[q]) *
           static_cast<double>(plan.offsets[q + 1] - plan.offsets[q]);
  };
  // The combination's weights in slot order; combined(q) is port q's flux of
  // the combination of the slots so far, computed face by face as the
  // device combines them.
  std::vector<double> weights;
  auto combined_flux = [&](uint32_t q) {
    double sum = 0.0;
    for (size_t i = plan.offsets[q]; i < plan.offsets[q + 1]; ++i) {
      double v = 0.0;
      for (size_t j = 0; j < weights.size(); ++j) {
        v += weights[j] * static_cast<double>(values[j * P + i]);
      }
      sum += v;
    }
    return sum * plan.signed_area[q];
  };
  if (holds_pressure) weights.push_back(1.0);
  std::vector<float> pressures = held;
  const size_t k = driven.size();
  if (k > 0) {
    // response[m][j]: the flux through speed port m when speed port j alone
    // is held at pressure one; the last column, each speed port's stated
    // flux less what the held pressures already carry through it.
    std::vector<std::vector<double>> response(k, std::vector<double>(k + 1, 0.0));
    for (size_t j = 0; j < k; ++j) {
      for (size_t m = 0; m < k; ++m) {
        response[m][j] = slot_flux(driven[m], first_driven + static_cast<uint32_t>(j));
      }
    }
    for (size_t m = 0; m < k; ++m) {
      const double base = holds_pressure ? combined_flux(driven[m]) : 0.0;
      response[m][k] =
          static_cast<double>(plan.ports[driven[m]].value) * open_area(driven[m]) - base;
    }
    // Gaussian elimination with partial pivoting on the augmented matrix.
    for (size_t column = 0; column < k; ++column) {
      size_t pivot = column;
      for (size_t row = column + 1; row < k; ++row) {
        if (std::abs(response[row][column]) > std::abs(response[pivot][column])) pivot = row;
      }
      if (response[pivot][column] == 0.0) {
        throw Error(MSIM_RUNTIME_ERROR,
                    "a speed port carries no flow: no pressure port is reachable");
      }
      std::swap(response[column], response[pivot]);
      for (size_t row = 0; row < k; ++row) {
        if (row == column) continue;
        const double factor = response[row][column] / response[column][column];
        for (size_t entry = column; entry <= k; ++entry) {
          response[row][entry] -= factor * response[column][entry];
        }
      }
    }
    for (size_t j = 0; j < k; ++j) {
      const double pressure = response[j][k] / response[j][j];
      pressures[driven[j]] = static_cast<float>(pressure);
      weights.push_back(pressure);
    }
  }

  // The combined field on the device.
  const uint32_t combine_blocks = grid_blocks(F, 256, 4096);
  f.port_weights.reserve(weights.size());
  f.port_peak.reserve(combine_blocks);
  f.lifted.reserve(F);
  check(hipMemcpyAsync(f.port_weights.data(), weights.data(), weights.size() * sizeof(double),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload weights");
  k_port_combine<<<combine_blocks, 256, 0, ctx.stream>>>(
      f.port_fields.data(), F, J, f.port_weights.data(), f.lifted.data(), f.port_peak.data());
  check(hipGetLastError(), "launch k_port_combine");
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  std::vector<uint32_t> block_peaks(combine_blocks);
  check(hipMemcpyAsync(block_peaks.data(), f.port_peak.data(), combine_blocks * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, ctx.stream),
        "download peak");
  check(hipStreamSynchronize(ctx.stream), "combine");
  const uint32_t peak_bits = *std::max_element(block_peaks.begin(), block_peaks.end());
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  seconds += ms * 1.0e-3;
  if (peak_bits >= 0x7f800000u) {
    throw Error(MSIM_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
  }
  float max_speed = 0.0f;
  std::memcpy(&max_speed, &peak_bits, sizeof(float));
  const size_t nxf = f.x_faces() * nz, nyf = f.y_faces() * nz;
  const float* out = f.lifted.data();
  if (x_faces != nullptr) {
    check(hipMemcpyAsync(x_faces, out, nxf * sizeof(float), hipMemcpyDeviceToHost, ctx.stream),
          "download x faces");
  }
  if (y_faces != nullptr) {
    check(hipMemcpyAsync(y_faces, out + nxf, nyf * sizeof(float), hipMemcpyDeviceToHost,
                         ctx.stream),
          "download y faces");
  }
  if (z_faces != nullptr && nz == 1) {
    std::memset(z_faces, 0, 2 * f.sites() * sizeof(float));
  } else if (z_faces != nullptr) {
    check(hipMemcpyAsync(z_faces, out + nxf + nyf, (F - nxf - nyf) * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download z faces");
  }
  check(hipStreamSynchronize(ctx.stream), "download port flow");
  if (states != nullptr) {
    for (uint32_t q = 0; q < count; ++q) {
      const double flux = combined_flux(q);
      states[q].pressure = pressures[q];
      states[q].flux = static_cast<float>(flux);
      states[q].mean_speed = static_cast<float>(flux / open_area(q));
    }
  }
  report.max_speed = max_speed;
  report.seconds = seconds;
  // The flow's faces hold the last solve, not the combination.
  f.prepared = true;
  f.solved = false;
}

}  // namespace msim
