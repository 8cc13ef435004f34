// Per-site arithmetic of the explicit FP32 transport step, shared by the
// single-grid kernel (adr.hip) and the batched sweep (batch.hip) so both
// compute every site in the same float operations and order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace msim {

// The reservoir value of signal s beyond the boundary site at position
// `along` of a fixed face (y on an x face, x on a y face): the face's uniform
// value, or with kPatched GridBoundary::values_at (signals.cpp:439-447)
// resolved per site on the host.
template <bool kPatched = false>
__device__ inline float reservoir(const AdrParams& P, int face, int s, uint32_t along) {
  if constexpr (!kPatched) {
    return P.boundary[s][face];
  } else {
    const uint32_t offset = (face & 1) * (face < 2 ? P.ny : P.nx) + (face < 2 ? 0u : 2u * P.ny) +
                            static_cast<uint32_t>(s) * (2u * P.ny + 2u * P.nx);
    return P.edge[offset + along];
  }
}

// The exterior value beyond a boundary site (boundary_value, signals.cpp:181-192).
template <bool kPatched = false>
__device__ inline float exterior(const AdrParams& P, int face, int s, float current,
                                 float periodic, uint32_t along) {
  const uint8_t kind = P.kind[face];
  if (kind == MSIM_NO_FLUX) return current;
  if (kind == MSIM_PERIODIC) return periodic;
  return reservoir<kPatched>(P, face, s, along);
}

struct SiteGeometry {
  bool clo[2], cup[2];      // face_closure (signals.cpp:213-261), x then y
  float ulo[2], uup[2];     // face velocities (signals.cpp:157-177)
};

__device__ inline bool solid_at(const uint8_t* solid, size_t i) { return solid[i] != 0; }

__device__ inline SiteGeometry geometry(const AdrParams& P, const uint8_t* __restrict__ solid,
                                        const float* __restrict__ xf,
                                        const float* __restrict__ yf, uint32_t x, uint32_t y,
                                        int s) {
  const uint32_t nx = P.nx, ny = P.ny;
  const size_t i = static_cast<size_t>(x) * ny + y;
  SiteGeometry g;
  // x axis
  if (x == 0) {
    g.clo[0] = P.kind[0] == MSIM_NO_FLUX ||
               (P.kind[0] == MSIM_PERIODIC && solid_at(solid, static_cast<size_t>(nx - 1) * ny + y));
  } else {
    g.clo[0] = solid_at(solid, i - ny);
  }
  if (x + 1 == nx) {
    g.cup[0] = P.kind[1] == MSIM_NO_FLUX || (P.kind[1] == MSIM_PERIODIC && solid_at(solid, y));
  } else {
    g.cup[0] = solid_at(solid, i + ny);
  }
  // y axis
  if (y == 0) {
    g.clo[1] = P.kind[2] == MSIM_NO_FLUX ||
               (P.kind[2] == MSIM_PERIODIC && solid_at(solid, static_cast<size_t>(x) * ny + ny - 1));
  } else {
    g.clo[1] = solid_at(solid, i - 1);
  }
  if (y + 1 == ny) {
    g.cup[1] = P.kind[3] == MSIM_NO_FLUX ||
               (P.kind[3] == MSIM_PERIODIC && solid_at(solid, static_cast<size_t>(x) * ny));
  } else {
    g.cup[1] = solid_at(solid, i + 1);
  }
  if (P.has_field != 0) {
    g.ulo[0] = xf[i];
    g.uup[0] = xf[i + ny];
    const size_t yfi = static_cast<size_t>(x) * (ny + 1) + y;
    g.ulo[1] = yf[yfi];
    g.uup[1] = yf[yfi + 1];
  } else {
    g.ulo[0] = g.uup[0] = P.advection[s][0];
    g.ulo[1] = g.uup[1] = P.advection[s][1];
  }
  return g;
}

// van_leer_increment (signals.cpp:202-206).
template <typename T>
__device__ inline T van_leer_increment(T upwind, T downwind, T far) {
  const T behind = upwind - far;
  const T ahead = downwind - upwind;
  return behind * ahead > T(0) ? behind * ahead / (behind + ahead) : T(0);
}

// What the limited (van Leer) correction of one site needs along each axis:
// whether the lower and upper neighbours are joined fluid sites of the grid
// (lower_site / upper_site, signals.cpp:1250-1253) and the level two sites
// away on each side where that site is on the grid and fluid (far(-2) and
// far(2), signals.cpp:1238-1249).
template <typename T>
struct Limiter {
  bool lower_site[2], upper_site[2];
  bool far_lo_ok[2], far_hi_ok[2];
  T far_lo[2], far_hi[2];
};

// The transport rate of signal s at one site from its neighbour values
// (transport_rates_into, cpp/core/signals.cpp:1145-1291, with the van Leer
// correction when kLimited) in the reference's float order. lower/upper hold
// the x and y neighbour values with exterior values already resolved
// (exterior()); a closed face replaces its neighbour by the site's own value
// here.
template <bool kLimited = false>
__device__ inline float transport_rate_f32(const AdrParams& P, const SiteGeometry& g,
                                           float current, float lower0, float upper0,
                                           float lower1, float upper1, int s,
                                           const Limiter<float>* lim = nullptr) {
  const float ih[2] = {1.0f / P.dx, 1.0f / P.dy};
  const uint32_t dims[2] = {P.nx, P.ny};
  float lower[2] = {lower0, lower1}, upper[2] = {upper0, upper1};
  const float diffusion = P.diffusion[s];
  float rate = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    if (dims[a] == 1) continue;
    if (g.clo[a]) lower[a] = current;
    if (g.cup[a]) upper[a] = current;
    const float inverse_spacing = ih[a];
    rate += diffusion * (lower[a] - (2.0f * current) + upper[a]) * inverse_spacing *
            inverse_spacing;
    float lower_flux = g.ulo[a] >= 0.0f ? g.ulo[a] * lower[a] : g.ulo[a] * current;
    float upper_flux = g.uup[a] >= 0.0f ? g.uup[a] * current : g.uup[a] * upper[a];
    if (g.clo[a]) lower_flux = 0.0f;
    if (g.cup[a]) upper_flux = 0.0f;
    rate -= (upper_flux - lower_flux) * inverse_spacing;
    if constexpr (kLimited) {
      float lower_correction = 0.0f, upper_correction = 0.0f;
      if (lim->lower_site[a]) {
        if (g.ulo[a] >= 0.0f) {
          if (lim->far_lo_ok[a]) {
            lower_correction =
                g.ulo[a] * van_leer_increment(lower[a], current, lim->far_lo[a]);
          }
        } else if (lim->upper_site[a]) {
          lower_correction = g.ulo[a] * van_leer_increment(current, lower[a], upper[a]);
        }
      }
      if (lim->upper_site[a]) {
        if (g.uup[a] >= 0.0f) {
          if (lim->lower_site[a]) {
            upper_correction = g.uup[a] * van_leer_increment(current, upper[a], lower[a]);
          }
        } else if (lim->far_hi_ok[a]) {
          upper_correction = g.uup[a] * van_leer_increment(upper[a], current, lim->far_hi[a]);
        }
      }
      rate -= (upper_correction - lower_correction) * inverse_spacing;
    }
  }
  return rate;
}

// Forward-Euler update of signal s at one site: the transport rate, the
// affine reaction (signals.cpp:305-320) and forward Euler (1109-1134), with
// Monod uptake as the affine loss vmax*phi/(K+c) it equals under forward
// Euler.
template <bool kLimited = false>
__device__ inline float fe_update_f32(const AdrParams& P, const SiteGeometry& g, float current,
                                      float lower0, float upper0, float lower1, float upper1,
                                      int s, float dt, float source, float loss, float phi,
                                      const Limiter<float>* lim = nullptr) {
  float rate = transport_rate_f32<kLimited>(P, g, current, lower0, upper0, lower1, upper1, s, lim);
  if (P.has_affine != 0) rate += source - (loss * current);
  if (P.has_monod != 0) {
    const float uptake = P.vmax[s] * phi / (P.monod_k[s] + current);
    rate += 0.0f - (uptake * current);
  }
  return current + (dt * rate);
}

// The diagonal of the implicit operator for signal s at a fluid site
// (signal_operator_diagonal, signals.cpp:263-298): per axis of extent above
// one, -2 D/h^2, D/h^2 back for each closed face, and the outflow through
// each open face; less the affine loss.
__device__ inline float operator_diagonal_f32(const AdrParams& P, const SiteGeometry& g, int s,
                                              bool reaction, float loss) {
  const float ih[2] = {1.0f / P.dx, 1.0f / P.dy};
  const uint32_t dims[2] = {P.nx, P.ny};
  float diagonal = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    if (dims[a] == 1) continue;
    const float diffusion_scale = P.diffusion[s] * ih[a] * ih[a];
    diagonal -= 2.0f * diffusion_scale;
    if (g.clo[a]) diagonal += diffusion_scale;
    if (g.cup[a]) diagonal += diffusion_scale;
    if (!g.cup[a] && g.uup[a] > 0.0f) diagonal -= g.uup[a] * ih[a];
    if (!g.clo[a] && g.ulo[a] < 0.0f) diagonal += g.ulo[a] * ih[a];
  }
  if (reaction) diagonal -= loss;
  return diagonal;
}

// Which neighbours the limited correction of the site (x, y) may use, and
// the levels two sites away, read from one signal's levels lv and the mask
// (signals.cpp:1233-1253): a face takes the correction only where its upwind,
// downwind and second-upwind sites are fluid sites joined by open faces; a
// periodic axis wraps only when it holds at least three sites.
__device__ inline void site_limiter(const AdrParams& P, const uint8_t* __restrict__ solid,
                                    const SiteGeometry& g, const float* __restrict__ lv,
                                    uint32_t x, uint32_t y, Limiter<float>& lim) {
  const uint32_t nx = P.nx, ny = P.ny;
  const bool xp3 = P.kind[0] == MSIM_PERIODIC && nx >= 3;
  const bool yp3 = P.kind[2] == MSIM_PERIODIC && ny >= 3;
  lim.lower_site[0] = !g.clo[0] && (x > 0 || xp3);
  lim.upper_site[0] = !g.cup[0] && (x + 1 < nx || xp3);
  lim.lower_site[1] = !g.clo[1] && (y > 0 || yp3);
  lim.upper_site[1] = !g.cup[1] && (y + 1 < ny || yp3);
  auto far = [&](int axis, int offset, float& value) {
    const int64_t count = axis == 0 ? nx : ny;
    const bool periodic = axis == 0 ? xp3 : yp3;
    const int64_t along = static_cast<int64_t>(axis == 0 ? x : y) + offset;
    value = 0.0f;
    if (!periodic && (along < 0 || along >= count)) return false;
    const uint32_t at = static_cast<uint32_t>((along + count) % count);
    const size_t site = axis == 0 ? static_cast<size_t>(at) * ny + y : static_cast<size_t>(x) * ny + at;
    if (solid[site] != 0) return false;
    value = lv[site];
    return true;
  };
  lim.far_lo_ok[0] = far(0, -2, lim.far_lo[0]);
  lim.far_hi_ok[0] = far(0, 2, lim.far_hi[0]);
  lim.far_lo_ok[1] = far(1, -2, lim.far_lo[1]);
  lim.far_hi_ok[1] = far(1, 2, lim.far_hi[1]);
}

// The same update reading the neighbours from the level array lv (one
// signal's levels) at site (x, y).
__device__ inline float fe_candidate_f32(const AdrParams& P, const SiteGeometry& g,
                                         const float* __restrict__ lv, uint32_t x, uint32_t y,
                                         int s, float dt, float source, float loss,
                                         const float* __restrict__ phi) {
  const size_t i = static_cast<size_t>(x) * P.ny + y;
  const float current = lv[i];
  const float lower0 =
      x == 0 ? exterior(P, 0, s, current, lv[static_cast<size_t>(P.nx - 1) * P.ny + y], y)
             : lv[i - P.ny];
  const float upper0 = x + 1 == P.nx ? exterior(P, 1, s, current, lv[y], y) : lv[i + P.ny];
  const float lower1 =
      y == 0 ? exterior(P, 2, s, current, lv[static_cast<size_t>(x) * P.ny + P.ny - 1], x)
             : lv[i - 1];
  const float upper1 =
      y + 1 == P.ny ? exterior(P, 3, s, current, lv[static_cast<size_t>(x) * P.ny], x) : lv[i + 1];
  return fe_update_f32(P, g, current, lower0, upper0, lower1, upper1, s, dt, source, loss,
                       P.has_monod != 0 ? phi[i] : 0.0f);
}

}  // namespace msim
