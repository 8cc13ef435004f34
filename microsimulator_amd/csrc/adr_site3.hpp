// Per-site arithmetic of the FP32 transport step on a grid of several layers
// (nz > 1): the rate of one signal at one fluid site over the three axes in
// the reference's float order (transport_rates_into, cpp/core/signals.cpp:
// 1145-1291; face_closure 213-261, face_velocities 157-177, boundary_value
// 181-192), and the implicit operator's diagonal (signal_operator_diagonal,
// 263-298). A grid of one layer skips its z axis exactly as the reference
// does, so on nz = 1 these give the one-layer kernels' bits; those kernels
// (adr_site.hpp) stay the fast path there.
//
// Storage: site (x*ny + y)*nz + z; x faces (fx*ny + y)*nz + z, y faces
// (x*(ny+1) + fy)*nz + z, z faces (x*ny + y)*(nz+1) + fz. With reservoir
// patches, P.edge holds per signal the value beyond every boundary site of
// the six faces: [x_lower ny*nz | x_upper ny*nz | y_lower nx*nz |
// y_upper nx*nz | z_lower nx*ny | z_upper nx*ny], a face's sites in its
// (first, second) in-plane order (GridBoundary::values_at, 439-447).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "adr_site.hpp"
#include "internal.hpp"

namespace msim {

struct Site3 {
  uint32_t c[3];  // x, y, z
  size_t i;       // flat site
};

__device__ inline size_t flat3(const AdrParams& P, uint32_t x, uint32_t y, uint32_t z) {
  return (static_cast<size_t>(x) * P.ny + y) * P.nz + z;
}

__device__ inline Site3 site3(const AdrParams& P, size_t i) {
  Site3 s;
  const size_t plane = static_cast<size_t>(P.ny) * P.nz;
  s.c[0] = static_cast<uint32_t>(i / plane);
  s.c[1] = static_cast<uint32_t>((i % plane) / P.nz);
  s.c[2] = static_cast<uint32_t>(i % P.nz);
  s.i = i;
  return s;
}

// The site with coordinate `axis` replaced by `along`.
__device__ inline size_t along3(const AdrParams& P, const Site3& s, int axis, uint32_t along) {
  uint32_t c[3] = {s.c[0], s.c[1], s.c[2]};
  c[axis] = along;
  return flat3(P, c[0], c[1], c[2]);
}

// The exterior value beyond a boundary site by the face's kind
// (boundary_value, signals.cpp:181-192).
__device__ inline float exterior_kind(uint8_t kind, float current, float periodic,
                                      float reservoir) {
  if (kind == MSIM_NO_FLUX) return current;
  if (kind == MSIM_PERIODIC) return periodic;
  return reservoir;
}

// The value held beyond boundary face `face` (0..5) for signal s at the
// boundary site with in-plane indices (first, second).
template <bool kPatched>
__device__ inline float reservoir3(const AdrParams& P, int face, int s, uint32_t first,
                                   uint32_t second) {
  if constexpr (!kPatched) {
    return P.boundary[s][face];
  } else {
    const size_t yz = static_cast<size_t>(P.ny) * P.nz, xz = static_cast<size_t>(P.nx) * P.nz;
    const size_t xy = static_cast<size_t>(P.nx) * P.ny;
    const size_t row = 2 * (yz + xz + xy);
    const size_t offset[6] = {0, yz, 2 * yz, 2 * yz + xz, 2 * (yz + xz), 2 * (yz + xz) + xy};
    const uint32_t second_extent = face < 4 ? P.nz : P.ny;
    return P.edge[static_cast<size_t>(s) * row + offset[face] +
                  static_cast<size_t>(first) * second_extent + second];
  }
}

struct Geometry3 {
  bool clo[3], cup[3];   // face_closure
  float ulo[3], uup[3];  // face_velocities of the signal
};

__device__ inline Geometry3 geometry3(const AdrParams& P, const uint8_t* __restrict__ solid,
                                      const float* __restrict__ xf, const float* __restrict__ yf,
                                      const float* __restrict__ zf, const Site3& s, int signal) {
  const uint32_t dims[3] = {P.nx, P.ny, P.nz};
  Geometry3 g;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const uint32_t at = s.c[a];
    if (at == 0) {
      const uint8_t kind = P.kind[2 * a];
      g.clo[a] = kind == MSIM_NO_FLUX ||
                 (kind == MSIM_PERIODIC && solid[along3(P, s, a, dims[a] - 1)] != 0);
    } else {
      g.clo[a] = solid[along3(P, s, a, at - 1)] != 0;
    }
    if (at + 1 == dims[a]) {
      const uint8_t kind = P.kind[2 * a + 1];
      g.cup[a] = kind == MSIM_NO_FLUX || (kind == MSIM_PERIODIC && solid[along3(P, s, a, 0)] != 0);
    } else {
      g.cup[a] = solid[along3(P, s, a, at + 1)] != 0;
    }
  }
  if (P.has_field != 0) {
    const uint32_t x = s.c[0], y = s.c[1], z = s.c[2];
    const uint32_t ny = P.ny, nz = P.nz;
    g.ulo[0] = xf[(static_cast<size_t>(x) * ny + y) * nz + z];
    g.uup[0] = xf[(static_cast<size_t>(x + 1) * ny + y) * nz + z];
    g.ulo[1] = yf[(static_cast<size_t>(x) * (ny + 1) + y) * nz + z];
    g.uup[1] = yf[(static_cast<size_t>(x) * (ny + 1) + y + 1) * nz + z];
    g.ulo[2] = zf[(static_cast<size_t>(x) * ny + y) * (nz + 1) + z];
    g.uup[2] = zf[(static_cast<size_t>(x) * ny + y) * (nz + 1) + z + 1];
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) g.ulo[a] = g.uup[a] = P.advection[signal][a];
  }
  return g;
}

// The in-plane indices (first, second) of a site on the faces normal to axis a.
__device__ inline void in_plane3(const Site3& s, int a, uint32_t& first, uint32_t& second) {
  first = a == 0 ? s.c[1] : s.c[0];
  second = a == 2 ? s.c[1] : s.c[2];
}

// transport_rates_into for signal `signal` at site s from its levels lv.
template <bool kLimited, bool kPatched>
__device__ inline float transport_rate3_f32(const AdrParams& P, const uint8_t* __restrict__ solid,
                                            const Geometry3& g, const float* __restrict__ lv,
                                            const Site3& s, int signal) {
  const uint32_t dims[3] = {P.nx, P.ny, P.nz};
  const float spacing[3] = {P.dx, P.dy, P.dz};
  const float current = lv[s.i];
  float lower[3], upper[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    uint32_t first, second;
    in_plane3(s, a, first, second);
    const uint32_t at = s.c[a];
    if (at == 0) {
      const float periodic = lv[along3(P, s, a, dims[a] - 1)];
      lower[a] = exterior_kind(P.kind[2 * a], current, periodic,
                               reservoir3<kPatched>(P, 2 * a, signal, first, second));
    } else {
      lower[a] = lv[along3(P, s, a, at - 1)];
    }
    if (at + 1 == dims[a]) {
      const float periodic = lv[along3(P, s, a, 0)];
      upper[a] = exterior_kind(P.kind[2 * a + 1], current, periodic,
                               reservoir3<kPatched>(P, 2 * a + 1, signal, first, second));
    } else {
      upper[a] = lv[along3(P, s, a, at + 1)];
    }
  }
  const float diffusion = P.diffusion[signal];
  float rate = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (dims[a] == 1) continue;
    if (g.clo[a]) lower[a] = current;
    if (g.cup[a]) upper[a] = current;
    const float inverse_spacing = 1.0f / spacing[a];
    rate += diffusion * (lower[a] - (2.0f * current) + upper[a]) * inverse_spacing *
            inverse_spacing;
    float lower_flux = g.ulo[a] >= 0.0f ? g.ulo[a] * lower[a] : g.ulo[a] * current;
    float upper_flux = g.uup[a] >= 0.0f ? g.uup[a] * current : g.uup[a] * upper[a];
    if (g.clo[a]) lower_flux = 0.0f;
    if (g.cup[a]) upper_flux = 0.0f;
    rate -= (upper_flux - lower_flux) * inverse_spacing;
    if constexpr (kLimited) {
      // The limited correction (signals.cpp:1227-1286): a face takes it only
      // where its upwind, downwind and second-upwind sites are fluid sites
      // of the grid joined by open faces.
      const bool periodic = P.kind[2 * a] == MSIM_PERIODIC && dims[a] >= 3;
      const uint32_t at = s.c[a];
      auto far = [&](int offset, float& value) {
        const int64_t count = dims[a];
        const int64_t to = static_cast<int64_t>(at) + offset;
        value = 0.0f;
        if (!periodic && (to < 0 || to >= count)) return false;
        const size_t site = along3(P, s, a, static_cast<uint32_t>((to + count) % count));
        if (solid[site] != 0) return false;
        value = lv[site];
        return true;
      };
      const bool lower_site = !g.clo[a] && (at > 0 || periodic);
      const bool upper_site = !g.cup[a] && (at + 1 < dims[a] || periodic);
      float lower_correction = 0.0f, upper_correction = 0.0f;
      if (lower_site) {
        if (g.ulo[a] >= 0.0f) {
          float beyond;
          if (far(-2, beyond)) {
            lower_correction = g.ulo[a] * van_leer_increment(lower[a], current, beyond);
          }
        } else if (upper_site) {
          lower_correction = g.ulo[a] * van_leer_increment(current, lower[a], upper[a]);
        }
      }
      if (upper_site) {
        if (g.uup[a] >= 0.0f) {
          if (lower_site) {
            upper_correction = g.uup[a] * van_leer_increment(current, upper[a], lower[a]);
          }
        } else {
          float beyond;
          if (far(2, beyond)) {
            upper_correction = g.uup[a] * van_leer_increment(upper[a], current, beyond);
          }
        }
      }
      rate -= (upper_correction - lower_correction) * inverse_spacing;
    }
  }
  return rate;
}

// signal_operator_diagonal over three axes.
__device__ inline float operator_diagonal3_f32(const AdrParams& P, const Geometry3& g, int signal,
                                               bool reaction, float loss) {
  const uint32_t dims[3] = {P.nx, P.ny, P.nz};
  const float spacing[3] = {P.dx, P.dy, P.dz};
  float diagonal = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (dims[a] == 1) continue;
    const float ih = 1.0f / spacing[a];
    const float diffusion_scale = P.diffusion[signal] * ih * ih;
    diagonal -= 2.0f * diffusion_scale;
    if (g.clo[a]) diagonal += diffusion_scale;
    if (g.cup[a]) diagonal += diffusion_scale;
    if (!g.cup[a] && g.uup[a] > 0.0f) diagonal -= g.uup[a] * ih;
    if (!g.clo[a] && g.ulo[a] < 0.0f) diagonal += g.ulo[a] * ih;
  }
  if (reaction) diagonal -= loss;
  return diagonal;
}

}  // namespace msim
