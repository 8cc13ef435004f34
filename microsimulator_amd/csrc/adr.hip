// Explicit advection-diffusion-reaction step on the device, one layer.
//
// One marching kernel, k_adr_step, in two arithmetic forms: a block owns a
// strip of columns and walks x, each thread advancing every signal of its
// site, so levels, faces, mask and occupancy are read from memory once.
//  * FP32 levels: the reference's per-site rate in its own float order
//    (cpp/core/signals.cpp:1145-1291 transport, 305-320 affine reaction,
//    1109-1134 forward Euler), Monod uptake as the affine loss
//    vmax*phi/(K+c) it equals under forward Euler.
//  * FP64 levels in face-flux form: every face flux is one expression of the
//    two sites it joins, so what leaves one site enters the other bit for
//    bit; the step's mass, boundary supply and uptake are tallied in FP64 for
//    the conservation ledger.
// A site whose update is negative or non-finite raises a flag (the reference
// throws, signals.cpp:1124-1130); the step is then refused as a whole.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "adr_site.hpp"
#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kAdrBlock = 256;

// The ledger of a step lives as per-block partial sums (mass after, boundary
// supply rate, uptake rate, for each signal) in two generations by step
// parity. Blocks j < 3S of the next step add partial j of the previous step
// (in a fixed order) into its ledger row, so no grid-wide fence is needed; a
// last small launch adds the final step's.
struct LedgerParts {
  double* parts;   // [2][3 * kMaxSignals][blocks]
  double* ledger;  // [steps][S][4]
  uint32_t blocks;
};

template <int S, int NT>
__device__ inline void fold_ledger(const AdrParams& P, LedgerParts L, uint32_t step, float dt,
                                   uint32_t j, double* lds) {
  // Partial j of `step` into ledger row `step`.
  const double* v =
      L.parts + (static_cast<size_t>(step & 1) * (3 * kMaxSignals) + j) * L.blocks;
  double total = 0.0;
  for (uint32_t b = threadIdx.x; b < L.blocks; b += NT) total += v[b];
  total = block_sum<NT>(total, lds);
  if (threadIdx.x == 0) {
    const double volume =
        static_cast<double>(P.dx) * static_cast<double>(P.dy) * static_cast<double>(P.dz);
    const uint32_t s = j / 3, what = j % 3;
    double* row = L.ledger + (static_cast<size_t>(step) * S + s) * 4;
    row[1 + what] = what == 0 ? total * volume : static_cast<double>(dt) * total * volume;
  }
}

// Sums of K values over a block in one pass: every wave reduces its K values,
// then thread 0 adds the waves in order. Only thread 0's result is meaningful.
template <int NT, int K>
__device__ inline void block_sums(double (&v)[K], double* lds) {
  constexpr int kWaves = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * kWaves + wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double t = 0.0;
      for (int w = 0; w < kWaves; ++w) t += lds[k * kWaves + w];
      v[k] = t;
    }
  }
}

// What one thread holds of one x-row of its column: the S levels and the
// mask; for sites the block updates also the faces, occupancy and (with
// affine reactions) the rates.
template <typename T, int S, bool kAffine>
struct AdrRow {
  T c[S];
  float src[kAffine ? S : 1], los[kAffine ? S : 1];
  float xf, ylo, yhi, phi;
  bool solid;
};

// One forward-Euler step for S signals, marching: a block owns a strip of
// TY - 2H columns and `rows` x-rows and walks x, so each level is read from
// memory once and its x neighbours come from registers, its y neighbours from
// a row in LDS. H = 1 for the upwind scheme, 2 under van Leer (whose
// correction reads the site two away). The strip's first and last H threads
// only load the columns just outside it (or, on a periodic y boundary, the
// partner columns), so no thread holds more than its own column. The next
// row's loads are issued one row ahead.
//  * T = float: the reference's per-site rate in its float order
//    (fe_update_f32), so every site matches its CPU step bit for bit.
//  * T = double: face-flux form, each face flux (with its limited
//    correction) one expression of the sites around it, so what leaves one
//    site enters the other bit for bit; the step's mass, boundary supply and
//    uptake go to the ledger partials.
// A site whose update is non-finite, or negative under upwind, raises a flag
// (the reference throws, signals.cpp:1119-1131); the step is then refused as
// a whole. Under van Leer each block also leaves the largest level it wrote,
// for the settle that follows.
// Least waves per SIMD asked of the step kernel: four for the FP64 upwind
// step without an affine reaction, which fits 128 VGPRs without spills once
// its ledger sums live in LDS (three waves otherwise, at 131); 1 (no demand)
// for the forms that would spill.
template <typename T, int S, bool kAffine, bool kLimited>
constexpr int step_waves() {
  return sizeof(T) == 8 && !kAffine && !kLimited && S <= 3 ? 4 : 1;
}

template <typename T, int S, int TY, bool kAffine, bool kLimited, bool kPatched>
__global__ __launch_bounds__(TY) __attribute__((amdgpu_waves_per_eu(step_waves<T, S, kAffine, kLimited>()))) void k_adr_step(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const float* __restrict__ phi, const float* __restrict__ source,
    const float* __restrict__ loss, const T* __restrict__ in, T* __restrict__ out, float dt,
    uint32_t* flags, uint32_t step, uint32_t rows, LedgerParts L, double* largest) {
  constexpr bool kF64 = sizeof(T) == 8;
  constexpr int kHalo = kLimited ? 2 : 1;
  constexpr int kWidth = TY - 2 * kHalo;
  using Row = AdrRow<T, S, kAffine>;
  __shared__ T s_c[2][S][TY];
  __shared__ uint8_t s_s[2][TY];
  __shared__ double s_red[(kF64 ? 3 * S : 1) * (TY / kWave)];
  // An earlier step failed: the run stops there. (flags[1] holds the first
  // failed step; a failure of this step may be seen here half-published, and
  // then the previous step's ledger row is still folded, as it is valid.)
  if (flags[0] != 0 && flags[1] < step) return;
  if constexpr (kF64) {
    if (step > 0 && blockIdx.x < 3 * S) fold_ledger<S, TY>(P, L, step - 1, dt, blockIdx.x, s_red);
  }
  if (flags[0] != 0) return;  // this step already failed elsewhere
  const uint32_t nx = P.nx, ny = P.ny;
  const size_t n = static_cast<size_t>(nx) * ny;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t strips = (ny + kWidth - 1) / kWidth;
  const uint32_t x0 = (tile / strips) * rows, x1 = min(nx, x0 + rows);
  const uint32_t t = threadIdx.x;
  const bool xper = P.kind[0] == MSIM_PERIODIC, yper = P.kind[2] == MSIM_PERIODIC;
  // Column of this thread; columns off the grid stand for the periodic
  // partner columns.
  const int col = static_cast<int>((tile % strips) * kWidth + t) - kHalo;
  int src_col = col;
  if (col < 0) src_col = yper ? col + static_cast<int>(ny) : -1;
  if (col >= static_cast<int>(ny)) src_col = yper ? col - static_cast<int>(ny) : -1;
  if (src_col < 0 || src_col >= static_cast<int>(ny)) src_col = -1;
  const bool loads = src_col >= 0;
  const bool owns = t >= kHalo && t < kHalo + kWidth && col >= 0 && col < static_cast<int>(ny);
  const uint32_t y = static_cast<uint32_t>(src_col);
  const bool field = P.has_field != 0, monod = P.has_monod != 0;
  // Row xx of the walk (xx in [-H, nx + H - 1]); rows off the grid are the
  // periodic partners when the x boundary is periodic and absent otherwise.
  // `full` rows are the ones the block updates.
  auto load = [&](int xx, bool full) {
    Row w{};
    int row = xx;
    if (xx < 0) row = xper ? xx + static_cast<int>(nx) : -1;
    if (xx >= static_cast<int>(nx)) row = xper ? xx - static_cast<int>(nx) : -1;
    if (row >= static_cast<int>(nx)) row = -1;
    if (owns && field && xx >= 0 && xx <= static_cast<int>(nx)) {
      w.xf = xf[static_cast<size_t>(xx) * ny + y];  // lower x face of row xx
    }
    if (row < 0 || !loads) return w;
    const size_t i = static_cast<size_t>(row) * ny + y;
    w.solid = solid[i] != 0;
#pragma unroll
    for (int s = 0; s < S; ++s) w.c[s] = in[s * n + i];
    if (!full || !owns) return w;
    if (field) {
      const size_t yfi = static_cast<size_t>(row) * (ny + 1) + y;
      w.ylo = yf[yfi];
      w.yhi = yf[yfi + 1];
    }
    if (monod) w.phi = phi[i];
    if constexpr (kAffine) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        w.src[s] = source[s * n + i];
        w.los[s] = loss[s * n + i];
      }
    }
    return w;
  };
  // FP64 ledger sums (mass after, supply, uptake per signal), one slot per
  // thread in LDS: kept in registers they would take 24 VGPRs and hold the
  // step at three waves per SIMD instead of four.
  __shared__ double s_acc[kF64 ? 3 * S : 1][TY];
  if constexpr (kF64) {
#pragma unroll
    for (int k = 0; k < 3 * S; ++k) s_acc[k][threadIdx.x] = 0.0;
  }
  T top = T(0);  // largest level written (van Leer)
  bool bad = false;
  if (x0 < x1) {
    const int xa = static_cast<int>(x0), xb = static_cast<int>(x1);
    // Window of rows xx - 2 .. xx + 2 (the outer two under van Leer only).
    Row p2 = kLimited ? load(xa - 2, false) : Row{};
    Row prev = load(xa - 1, false);
    Row cur = load(xa, true);
    Row next = load(xa + 1, xa + 1 < xb);
    Row n2 = kLimited ? load(xa + 2, xa + 2 < xb) : Row{};
    for (int xx = xa; xx < xb; ++xx) {
      constexpr int kAhead = kLimited ? 3 : 2;
      const Row ahead =
          xx + kAhead < xb + kHalo ? load(xx + kAhead, xx + kAhead < xb) : Row{};
      const int buf = xx & 1;
#pragma unroll
      for (int s = 0; s < S; ++s) s_c[buf][s][t] = cur.c[s];
      s_s[buf][t] = cur.solid ? 1 : 0;
      __syncthreads();
      if (owns) {
        const size_t i = static_cast<size_t>(xx) * ny + y;
        if (cur.solid) {
#pragma unroll
          for (int s = 0; s < S; ++s) {
            out[s * n + i] = cur.c[s];
            if constexpr (kLimited) top = max(top, cur.c[s]);
          }
        } else {
          const bool xlo_edge = xx == 0, xhi_edge = xx + 1 == static_cast<int>(nx);
          const bool ylo_edge = y == 0, yhi_edge = y + 1 == ny;
          const bool ylo_solid = s_s[buf][t - 1] != 0, yhi_solid = s_s[buf][t + 1] != 0;
          // face_closure (signals.cpp:213-261).
          SiteGeometry g;
          g.clo[0] = xlo_edge ? (P.kind[0] == MSIM_NO_FLUX ||
                                 (P.kind[0] == MSIM_PERIODIC && prev.solid))
                              : prev.solid;
          g.cup[0] = xhi_edge ? (P.kind[1] == MSIM_NO_FLUX ||
                                 (P.kind[1] == MSIM_PERIODIC && next.solid))
                              : next.solid;
          g.clo[1] = ylo_edge ? (P.kind[2] == MSIM_NO_FLUX ||
                                 (P.kind[2] == MSIM_PERIODIC && ylo_solid))
                              : ylo_solid;
          g.cup[1] = yhi_edge ? (P.kind[3] == MSIM_NO_FLUX ||
                                 (P.kind[3] == MSIM_PERIODIC && yhi_solid))
                              : yhi_solid;
          // Which sites the limited correction may use (signals.cpp:1235-1253).
          [[maybe_unused]] bool lower_site[2], upper_site[2], far_lo_ok[2], far_hi_ok[2];
          if constexpr (kLimited) {
            const bool xp3 = xper && nx >= 3, yp3 = yper && ny >= 3;
            lower_site[0] = !g.clo[0] && (!xlo_edge || xp3);
            upper_site[0] = !g.cup[0] && (!xhi_edge || xp3);
            lower_site[1] = !g.clo[1] && (!ylo_edge || yp3);
            upper_site[1] = !g.cup[1] && (!yhi_edge || yp3);
            far_lo_ok[0] = (xx >= 2 || xp3) && !p2.solid;
            far_hi_ok[0] = (xx + 2 < static_cast<int>(nx) || xp3) && !n2.solid;
            far_lo_ok[1] = (y >= 2 || yp3) && s_s[buf][t - 2] == 0;
            far_hi_ok[1] = (y + 2 < ny || yp3) && s_s[buf][t + 2] == 0;
          }
#pragma unroll
          for (int s = 0; s < S; ++s) {
            if (field) {
              g.ulo[0] = cur.xf;
              g.uup[0] = next.xf;
              g.ulo[1] = cur.ylo;
              g.uup[1] = cur.yhi;
            } else {
              g.ulo[0] = g.uup[0] = P.advection[s][0];
              g.ulo[1] = g.uup[1] = P.advection[s][1];
            }
            const T c = cur.c[s];
            const T ylo_v = s_c[buf][s][t - 1];
            const T yhi_v = s_c[buf][s][t + 1];
            [[maybe_unused]] T far_lo[2] = {T(0), T(0)}, far_hi[2] = {T(0), T(0)};
            if constexpr (kLimited) {
              far_lo[0] = p2.c[s];
              far_hi[0] = n2.c[s];
              far_lo[1] = s_c[buf][s][t - 2];
              far_hi[1] = s_c[buf][s][t + 2];
            }
            const float src = kAffine ? cur.src[kAffine ? s : 0] : 0.0f;
            const float los = kAffine ? cur.los[kAffine ? s : 0] : 0.0f;
            if constexpr (!kF64) {
              const float l0 = xlo_edge ? exterior<kPatched>(P, 0, s, c, prev.c[s], y) : prev.c[s];
              const float u0 = xhi_edge ? exterior<kPatched>(P, 1, s, c, next.c[s], y) : next.c[s];
              const float l1 = ylo_edge ? exterior<kPatched>(P, 2, s, c, ylo_v, static_cast<uint32_t>(xx)) : ylo_v;
              const float u1 = yhi_edge ? exterior<kPatched>(P, 3, s, c, yhi_v, static_cast<uint32_t>(xx)) : yhi_v;
              float v;
              if constexpr (kLimited) {
                Limiter<float> lim;
#pragma unroll
                for (int a2 = 0; a2 < 2; ++a2) {
                  lim.lower_site[a2] = lower_site[a2];
                  lim.upper_site[a2] = upper_site[a2];
                  lim.far_lo_ok[a2] = far_lo_ok[a2];
                  lim.far_hi_ok[a2] = far_hi_ok[a2];
                  lim.far_lo[a2] = far_lo[a2];
                  lim.far_hi[a2] = far_hi[a2];
                }
                v = fe_update_f32<true>(P, g, c, l0, u0, l1, u1, s, dt, src, los, cur.phi, &lim);
                if (!isfinite(v)) bad = true;
                top = max(top, v);
              } else {
                v = fe_update_f32(P, g, c, l0, u0, l1, u1, s, dt, src, los, cur.phi);
                if (!isfinite(v) || v < 0.0f) bad = true;
              }
              out[s * n + i] = v;
            } else {
              const double ih[2] = {1.0 / static_cast<double>(P.dx),
                                    1.0 / static_cast<double>(P.dy)};
              const uint32_t dims[2] = {nx, ny};
              const bool lo_edge[2] = {xlo_edge, ylo_edge}, hi_edge[2] = {xhi_edge, yhi_edge};
              const double nlo[2] = {prev.c[s], ylo_v}, nhi[2] = {next.c[s], yhi_v};
              const double D = static_cast<double>(P.diffusion[s]);
              double rate = 0.0, supply = 0.0;
#pragma unroll
              for (int a = 0; a < 2; ++a) {
                if (dims[a] == 1) continue;
                double fl = 0.0, fu = 0.0;
                const double ul = static_cast<double>(g.ulo[a]);
                const double uu = static_cast<double>(g.uup[a]);
                if (lo_edge[a] && P.kind[2 * a] == MSIM_FIXED) {
                  const double cb = static_cast<double>(
                      reservoir<kPatched>(P, 2 * a, s, a == 0 ? y : static_cast<uint32_t>(xx)));
                  fl = (ul >= 0.0 ? ul * cb : ul * c) - D * (c - cb) * ih[a];
                  supply += fl * ih[a];
                } else if (!g.clo[a]) {
                  const double cl = nlo[a];
                  fl = (ul >= 0.0 ? ul * cl : ul * c) - D * (c - cl) * ih[a];
                  if constexpr (kLimited) {
                    if (lower_site[a]) {
                      if (ul >= 0.0) {
                        if (far_lo_ok[a])
                          fl += ul * van_leer_increment<double>(cl, c, far_lo[a]);
                      } else if (upper_site[a]) {
                        fl += ul * van_leer_increment<double>(c, cl, nhi[a]);
                      }
                    }
                  }
                }
                if (hi_edge[a] && P.kind[2 * a + 1] == MSIM_FIXED) {
                  const double cb = static_cast<double>(
                      reservoir<kPatched>(P, 2 * a + 1, s, a == 0 ? y : static_cast<uint32_t>(xx)));
                  fu = (uu >= 0.0 ? uu * c : uu * cb) - D * (cb - c) * ih[a];
                  supply -= fu * ih[a];
                } else if (!g.cup[a]) {
                  const double cu = nhi[a];
                  fu = (uu >= 0.0 ? uu * c : uu * cu) - D * (cu - c) * ih[a];
                  if constexpr (kLimited) {
                    if (upper_site[a]) {
                      if (uu >= 0.0) {
                        if (lower_site[a]) fu += uu * van_leer_increment<double>(c, cu, nlo[a]);
                      } else if (far_hi_ok[a]) {
                        fu += uu * van_leer_increment<double>(cu, c, far_hi[a]);
                      }
                    }
                  }
                }
                rate -= (fu - fl) * ih[a];
              }
              double lossr = 0.0;
              if constexpr (kAffine) {
                rate += static_cast<double>(src);
                supply += static_cast<double>(src);
                lossr += static_cast<double>(los);
              }
              if (monod) {
                lossr += static_cast<double>(P.vmax[s]) * static_cast<double>(cur.phi) /
                         (static_cast<double>(P.monod_k[s]) + c);
              }
              rate -= lossr * c;
              const double v = c + static_cast<double>(dt) * rate;
              if (!isfinite(v) || (v < 0.0 && !kLimited)) bad = true;
              if constexpr (kLimited) top = max(top, v);
              out[s * n + i] = v;
              s_acc[3 * s + 0][t] += v;
              s_acc[3 * s + 1][t] += supply;
              s_acc[3 * s + 2][t] += lossr * c;
            }
          }
        }
      }
      if constexpr (kLimited) p2 = prev;
      prev = cur;
      cur = next;
      if constexpr (kLimited) {
        next = n2;
        n2 = ahead;
      } else {
        next = ahead;
      }
    }
  }
  if (bad) {
    atomicOr(&flags[0], kFlagNegative);
    atomicMin(&flags[1], step);
  }
  if constexpr (kLimited) {
    const double m = block_max<TY>(static_cast<double>(top), s_red);
    if (threadIdx.x == 0) largest[blockIdx.x] = m;
    __syncthreads();
  }
  if constexpr (kF64) {
    // Solid sites hold their (zero) level; the mass counts only what moved.
    double acc[3 * S];
#pragma unroll
    for (int k = 0; k < 3 * S; ++k) acc[k] = s_acc[k][threadIdx.x];
    block_sums<TY, 3 * S>(acc, s_red);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < 3 * S; ++k) {
        L.parts[(static_cast<size_t>(step & 1) * (3 * kMaxSignals) + k) * L.blocks + blockIdx.x] =
            acc[k];
      }
    }
  }
}

// settle_step_levels then validate_signal_levels (signals.cpp:979-1001,
// 754-784) after a van Leer step, in place over the S*N levels it wrote:
// levels below zero by no more than 16 epsilon (of the level type) times the
// largest level become zero; a level still negative or not finite refuses
// the step. Every block takes the largest level from the step's block
// maxima.
template <typename T>
__global__ __launch_bounds__(kAdrBlock) void k_adr_settle(T* __restrict__ levels, size_t count,
                                                          const double* __restrict__ largest,
                                                          uint32_t parts, uint32_t* flags,
                                                          uint32_t step) {
  __shared__ double s_red[kAdrBlock / kWave];
  __shared__ double s_top;
  if (flags[0] != 0) return;
  double m = 0.0;
  for (uint32_t b = threadIdx.x; b < parts; b += kAdrBlock) m = fmax(m, largest[b]);
  m = block_max<kAdrBlock>(m, s_red);
  if (threadIdx.x == 0) s_top = m;
  __syncthreads();
  const T top = static_cast<T>(s_top);
  const T eps = sizeof(T) == 8 ? static_cast<T>(DBL_EPSILON) : static_cast<T>(FLT_EPSILON);
  const T rounding = T(-16) * eps * top;
  bool bad = false;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kAdrBlock + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * kAdrBlock) {
    const T v = levels[i];
    if (v < T(0)) {
      if (v >= rounding) {
        levels[i] = T(0);
      } else {
        bad = true;
      }
    } else if (!isfinite(v)) {
      bad = true;
    }
  }
  if (bad) {
    atomicOr(&flags[0], kFlagLevels);
    atomicMin(&flags[1], step);
  }
}

// Adds the final step's ledger partials.
template <int S>
__global__ __launch_bounds__(kAdrBlock) void k_adr_ledger_tail(AdrParams P, LedgerParts L,
                                                               uint32_t step, float dt,
                                                               const uint32_t* flags) {
  __shared__ double s_red[kAdrBlock / kWave];
  if (flags[0] != 0) return;
  fold_ledger<S, kAdrBlock>(P, L, step, dt, blockIdx.x, s_red);
}

// Sum of levels per signal (FP64), and the greatest outflow rate over fluid
// sites (SignalGrid::transport_cache, signals.cpp:870-896) for the explicit
// bound.
template <typename T>
__global__ __launch_bounds__(kAdrBlock) void k_adr_survey(
    AdrParams P, const uint8_t* __restrict__ solid, const float* __restrict__ xf,
    const float* __restrict__ yf, const T* __restrict__ levels, double* partials, CgState* st,
    double* mass) {
  __shared__ double s_red[kAdrBlock / kWave];
  __shared__ bool s_last;
  const size_t n = static_cast<size_t>(P.nx) * P.ny;
  double acc[kMaxSignals + 1];
#pragma unroll
  for (int k = 0; k <= kMaxSignals; ++k) acc[k] = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kAdrBlock;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kAdrBlock + threadIdx.x; i < n; i += stride) {
    for (uint32_t s = 0; s < P.signals && s < kMaxSignals; ++s)
      acc[s] += static_cast<double>(levels[s * n + i]);
    if (P.has_field != 0 && solid[i] == 0) {
      const uint32_t x = static_cast<uint32_t>(i / P.ny), y = static_cast<uint32_t>(i % P.ny);
      double outflow = 0.0;
      if (P.nx > 1) {
        outflow += (fmax(static_cast<double>(xf[i + P.ny]), 0.0) -
                    fmin(static_cast<double>(xf[i]), 0.0)) *
                   (1.0 / static_cast<double>(P.dx));
      }
      if (P.ny > 1) {
        const size_t yfi = static_cast<size_t>(x) * (P.ny + 1) + y;
        outflow += (fmax(static_cast<double>(yf[yfi + 1]), 0.0) -
                    fmin(static_cast<double>(yf[yfi]), 0.0)) *
                   (1.0 / static_cast<double>(P.dy));
      }
      acc[kMaxSignals] = fmax(acc[kMaxSignals], outflow);
    }
  }
  double part[kMaxSignals + 1];
#pragma unroll
  for (int k = 0; k < kMaxSignals; ++k) part[k] = block_sum<kAdrBlock>(acc[k], s_red);
  part[kMaxSignals] = block_max<kAdrBlock>(acc[kMaxSignals], s_red);
  if (publish_partials<kMaxSignals + 1>(part, partials, &st->ticket[kTicketMisc], blockIdx.x,
                                        gridDim.x, &s_last)) {
    double total[kMaxSignals];
    sum_partials<kAdrBlock, kMaxSignals>(partials, gridDim.x, total, s_red);
    double outflow[1];
    max_partials<kAdrBlock, 1>(partials + static_cast<size_t>(kMaxSignals) * gridDim.x, gridDim.x,
                               outflow, s_red);
    if (threadIdx.x == 0) {
      st->ticket[kTicketMisc] = 0;
      const double volume =
          static_cast<double>(P.dx) * static_cast<double>(P.dy) * static_cast<double>(P.dz);
      for (int s = 0; s < kMaxSignals; ++s) mass[s] = total[s] * volume;
      st->peak = outflow[0];
    }
  }
}

// Launch shape of the marching step: TY threads per block (64, 128 or 256)
// for strips of TY - 2 columns, and 16 x-rows per block (fewer on small grids,
// for about 2048 blocks): on a 4096 x 2048 grid, 8 to 32 rows per block give
// the same speed within 10%, 16 the best for FP64 levels.
struct AdrShape {
  int ty;
  uint32_t rows, blocks;
};

AdrShape adr_shape(uint32_t nx, uint32_t ny, int halo, bool patched) {
  AdrShape s{};
  // The strip width (TY - 2 halo columns) that covers the grid's width with
  // the fewest threads, the wider on a tie; 128 columns with reservoir
  // patches.
  uint64_t best = ~0ull;
  for (int ty : {256, 128, 64}) {
    if (patched && ty != 128) continue;
    const uint32_t width = static_cast<uint32_t>(ty - 2 * halo);
    const uint64_t threads = static_cast<uint64_t>((ny + width - 1) / width) * ty;
    if (threads < best) {
      best = threads;
      s.ty = ty;
    }
  }
  const uint32_t width = static_cast<uint32_t>(s.ty - 2 * halo);
  const uint32_t strips = (ny + width - 1) / width;
  uint32_t rows = 16;
  while (rows > 4 && static_cast<uint64_t>(strips) * ((nx + rows - 1) / rows) < 2048) rows /= 2;
  s.rows = rows;
  s.blocks = strips * ((nx + rows - 1) / rows);
  return s;
}

// Blocks of a step launch: at least 3S, since blocks j < 3S fold the
// previous step's ledger partial j (blocks past the grid's strips only add
// zeros).
uint32_t step_blocks(const AdrShape& sh) {
  return std::max<uint32_t>(sh.blocks, 3 * kMaxSignals);
}

uint32_t site_blocks(size_t n) { return static_cast<uint32_t>((n + kAdrBlock - 1) / kAdrBlock); }

}  // namespace

// The kernels' parameters of a signal grid (one layer): shape, spacing,
// boundary kinds and values, diffusion, Monod constants, the scheme and
// which optional arrays it has. The velocity field and patches are set by
// the caller.
AdrParams transport_params(const msim_signal_grid& grid) {
  AdrParams P;
  std::memset(&P, 0, sizeof(P));
  P.nx = grid.grid.nx;
  P.ny = grid.grid.ny;
  P.nz = grid.grid.nz;
  P.signals = grid.signal_count;
  P.dx = grid.grid.dx;
  P.dy = grid.grid.dy;
  P.dz = grid.grid.dz;
  for (int f = 0; f < 6; ++f) P.kind[f] = grid.grid.boundary[f];
  for (uint32_t s = 0; s < grid.signal_count; ++s) {
    P.diffusion[s] = grid.diffusion[s];
    for (int f = 0; f < 6; ++f)
      P.boundary[s][f] = grid.boundary_values != nullptr ? grid.boundary_values[6 * s + f] : 0.0f;
    if (grid.monod_vmax != nullptr) {
      P.vmax[s] = grid.monod_vmax[s];
      P.monod_k[s] = grid.monod_k[s];
    }
    for (int a = 0; a < 3; ++a)
      P.advection[s][a] = grid.advection != nullptr ? grid.advection[3 * s + a] : 0.0f;
  }
  P.has_field = grid.x_faces != nullptr ? 1 : 0;
  P.limited = grid.advection_scheme == MSIM_VAN_LEER ? 1 : 0;
  P.has_affine = grid.loss_rates != nullptr ? 1 : 0;
  P.has_monod = grid.phi != nullptr ? 1 : 0;
  return P;
}

void adr_upload(Context& ctx, const msim_signal_grid& grid, const void* levels,
                uint32_t precision, bool use_flow_field) {
  // Coupled state belongs to the grid it was uploaded with.
  ctx.coupled.uploaded = false;
  auto& a = ctx.adr;
  auto& P = a.params;
  const uint32_t nz = grid.grid.nz;
  const size_t n = static_cast<size_t>(grid.grid.nx) * grid.grid.ny * nz;
  const uint32_t S = grid.signal_count;
  const bool same_spec = a.uploaded && grid.revision != 0 && grid.revision == a.revision &&
                         P.nx == grid.grid.nx && P.ny == grid.grid.ny && P.nz == nz &&
                         P.signals == S && !use_flow_field && !a.flow_field;
  P = transport_params(grid);
  a.integration = grid.integration;
  a.max_iterations = grid.solver_max_iterations;
  a.absolute_tolerance = grid.solver_absolute_tolerance;
  a.relative_tolerance = grid.solver_relative_tolerance;
  a.last_report = msim_signal_report{1, 0, 0.0f};
  a.predicted_iterations = kPredictedIterations;
  P.has_field = (use_flow_field || grid.x_faces != nullptr) ? 1 : 0;
  const bool patched = grid.patches != nullptr && grid.patch_count > 0;
  a.precision = precision;
  const size_t bytes = S * n * (precision == 64 ? sizeof(double) : sizeof(float));
  a.levels_a.reserve(bytes);
  a.levels_b.reserve(bytes);
  a.current = 0;
  a.flags.reserve(2);
  a.state.reserve(1);
  a.mass.reserve(kMaxSignals);
  // Ledger partials [2][3 * kMaxSignals][blocks], then the van Leer block
  // maxima [blocks]; the survey's partials reuse the front.
  const uint32_t step_parts =
      step_blocks(adr_shape(P.nx, P.ny, P.limited != 0 ? 2 : 1, patched));
  a.partials.reserve(std::max<size_t>(
      static_cast<size_t>(3 * kMaxSignals * 2 + 1) * step_parts,
      static_cast<size_t>(kMaxSignals + 1) * 2048));
  if (!same_spec) {
    a.has_edge = patched;
    if (patched && nz > 1) {
      // GridBoundary::values_at (signals.cpp:439-447) for every boundary site
      // of the six faces, in each face's (first, second) order: the last
      // patch holding the site, else the face's value.
      const uint32_t extent[6][2] = {{P.ny, nz}, {P.ny, nz}, {P.nx, nz},
                                     {P.nx, nz}, {P.nx, P.ny}, {P.nx, P.ny}};
      size_t offset[7] = {0};
      for (int f = 0; f < 6; ++f)
        offset[f + 1] = offset[f] + static_cast<size_t>(extent[f][0]) * extent[f][1];
      const size_t row = offset[6];
      std::vector<float> edge(S * row);
      for (uint32_t s = 0; s < S; ++s) {
        for (int f = 0; f < 6; ++f) {
          for (uint32_t first = 0; first < extent[f][0]; ++first) {
            for (uint32_t second = 0; second < extent[f][1]; ++second) {
              float v = P.boundary[s][f];
              for (uint32_t k = grid.patch_count; k-- > 0;) {
                const msim_boundary_patch& pt = grid.patches[k];
                if (static_cast<int>(pt.face) == f && first >= pt.begin[0] && first < pt.end[0] &&
                    second >= pt.begin[1] && second < pt.end[1]) {
                  v = pt.values[s];
                  break;
                }
              }
              edge[s * row + offset[f] + static_cast<size_t>(first) * extent[f][1] + second] = v;
            }
          }
        }
      }
      a.edge.reserve(edge.size());
      check(hipMemcpyAsync(a.edge.data(), edge.data(), edge.size() * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload boundary patches");
      check(hipStreamSynchronize(ctx.stream), "upload boundary patches");
    } else if (patched) {
      // GridBoundary::values_at (signals.cpp:439-447) for every boundary site
      // of the four in-plane faces (second in-plane index z = 0): the last
      // patch holding the site, else the face's value.
      const size_t row = 2 * static_cast<size_t>(P.ny) + 2 * static_cast<size_t>(P.nx);
      std::vector<float> edge(S * row);
      const uint32_t extent[4] = {P.ny, P.ny, P.nx, P.nx};
      const size_t offset[4] = {0, P.ny, 2 * static_cast<size_t>(P.ny),
                                2 * static_cast<size_t>(P.ny) + P.nx};
      for (uint32_t s = 0; s < S; ++s) {
        for (int f = 0; f < 4; ++f) {
          for (uint32_t along = 0; along < extent[f]; ++along) {
            float v = P.boundary[s][f];
            for (uint32_t k = grid.patch_count; k-- > 0;) {
              const msim_boundary_patch& pt = grid.patches[k];
              if (static_cast<int>(pt.face) == f && along >= pt.begin[0] && along < pt.end[0] &&
                  0 >= pt.begin[1] && 0 < pt.end[1]) {
                v = pt.values[s];
                break;
              }
            }
            edge[s * row + offset[f] + along] = v;
          }
        }
      }
      a.edge.reserve(edge.size());
      check(hipMemcpyAsync(a.edge.data(), edge.data(), edge.size() * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload boundary patches");
      check(hipStreamSynchronize(ctx.stream), "upload boundary patches");
    }
    a.solid.reserve(n);
    if (grid.grid.obstacles != nullptr) {
      check(hipMemcpyAsync(a.solid.data(), grid.grid.obstacles, n, hipMemcpyHostToDevice,
                           ctx.stream),
            "upload mask");
    } else {
      check(hipMemsetAsync(a.solid.data(), 0, n, ctx.stream), "clear mask");
    }
    if (use_flow_field) {
      a.xf.release();
      a.yf.release();
    } else if (grid.x_faces != nullptr) {
      a.xf.reserve(static_cast<size_t>(P.nx + 1) * P.ny * nz);
      a.yf.reserve(static_cast<size_t>(P.nx) * (P.ny + 1) * nz);
      check(hipMemcpyAsync(a.xf.data(), grid.x_faces, a.xf.size() * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload x faces");
      check(hipMemcpyAsync(a.yf.data(), grid.y_faces, a.yf.size() * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload y faces");
      if (nz > 1) {
        a.zf.reserve(static_cast<size_t>(P.nx) * P.ny * (nz + 1));
        check(hipMemcpyAsync(a.zf.data(), grid.z_faces, a.zf.size() * sizeof(float),
                             hipMemcpyHostToDevice, ctx.stream),
              "upload z faces");
      }
    }
    if (grid.phi != nullptr) {
      a.phi.reserve(n);
      check(hipMemcpyAsync(a.phi.data(), grid.phi, n * sizeof(float), hipMemcpyHostToDevice,
                           ctx.stream),
            "upload phi");
    }
    if (grid.loss_rates != nullptr) {
      a.source.reserve(S * n);
      a.loss.reserve(S * n);
      check(hipMemcpyAsync(a.loss.data(), grid.loss_rates, S * n * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload loss");
      check(hipMemcpyAsync(a.source.data(), grid.source_rates, S * n * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload source");
    }
  }
  P.edge = a.has_edge ? a.edge.data() : nullptr;
  a.flow_field = use_flow_field;
  a.flow_generation = use_flow_field ? ctx.flow.generation : 0;
  a.revision = grid.revision;
  check(hipMemcpyAsync(a.levels_a.data(), levels, bytes, hipMemcpyHostToDevice, ctx.stream),
        "upload levels");
  check(hipMemsetAsync(a.state.data(), 0, sizeof(CgState), ctx.stream), "clear state");
  const float* xf = use_flow_field ? ctx.flow.faces.data() : a.xf.data();
  const float* yf = use_flow_field ? ctx.flow.faces.data() + ctx.flow.x_faces() : a.yf.data();
  const uint32_t blocks = std::min<uint32_t>(site_blocks(n), 2048);
  if (nz > 1) {
    adr3_survey(ctx, xf, yf, a.zf.data(), reinterpret_cast<const float*>(a.levels_a.data()),
                a.state.data(), a.partials.data(), a.mass.data());
  } else if (precision == 64) {
    k_adr_survey<double><<<blocks, kAdrBlock, 0, ctx.stream>>>(
        P, a.solid.data(), xf, yf, reinterpret_cast<const double*>(a.levels_a.data()),
        a.partials.data(), a.state.data(), a.mass.data());
  } else {
    k_adr_survey<float><<<blocks, kAdrBlock, 0, ctx.stream>>>(
        P, a.solid.data(), xf, yf, reinterpret_cast<const float*>(a.levels_a.data()),
        a.partials.data(), a.state.data(), a.mass.data());
  }
  check(hipGetLastError(), "launch k_adr_survey");
  CgState host{};
  check(hipMemcpyAsync(&host, a.state.data(), sizeof(CgState), hipMemcpyDeviceToHost, ctx.stream),
        "read survey");
  check(hipStreamSynchronize(ctx.stream), "survey");
  a.greatest_outflow = P.has_field != 0 ? host.peak : 0.0;
  check(hipMemcpyAsync(a.host_mass, a.mass.data(), sizeof(a.host_mass), hipMemcpyDeviceToHost,
                       ctx.stream),
        "read mass");
  check(hipStreamSynchronize(ctx.stream), "read mass");
  a.uploaded = true;
}

// The device arrays one transport step reads and writes besides the levels:
// the context's buffers, or the static scratch of the allocation-free step.
struct AdrView {
  const AdrParams* P;
  const uint8_t* solid;
  const float *phi, *source, *loss;
  uint32_t* flags;
};

AdrView view_of(Context& ctx) {
  auto& a = ctx.adr;
  return AdrView{&a.params, a.solid.data(), a.phi.data(), a.source.data(), a.loss.data(),
                 a.flags.data()};
}

template <typename T, int S, int TY, bool kLimited, bool kPatched>
void launch_step(hipStream_t stream, const AdrView& v, const AdrShape& sh, const float* xf,
                 const float* yf, const uint8_t* src, uint8_t* dst, float dt, uint32_t k,
                 LedgerParts L, double* largest, bool settle) {
  const AdrParams& P = *v.P;
  if (P.has_affine != 0) {
    k_adr_step<T, S, TY, true, kLimited, kPatched><<<sh.blocks, TY, 0, stream>>>(
        P, v.solid, xf, yf, v.phi, v.source, v.loss, reinterpret_cast<const T*>(src),
        reinterpret_cast<T*>(dst), dt, v.flags, k, sh.rows, L, largest);
  } else {
    k_adr_step<T, S, TY, false, kLimited, kPatched><<<sh.blocks, TY, 0, stream>>>(
        P, v.solid, xf, yf, v.phi, v.source, v.loss, reinterpret_cast<const T*>(src),
        reinterpret_cast<T*>(dst), dt, v.flags, k, sh.rows, L, largest);
  }
  if (kLimited && settle) {
    const size_t count = static_cast<size_t>(S) * P.nx * P.ny;
    const uint32_t blocks = static_cast<uint32_t>(
        std::min<size_t>((count + kAdrBlock - 1) / kAdrBlock, 2048));
    k_adr_settle<T><<<blocks, kAdrBlock, 0, stream>>>(reinterpret_cast<T*>(dst), count, largest,
                                                      sh.blocks, v.flags, k);
  }
}

template <typename T, int S, bool kLimited>
void launch_typed(hipStream_t stream, const AdrView& v, const AdrShape& sh, const float* xf,
                  const float* yf, const uint8_t* src, uint8_t* dst, float dt, uint32_t k,
                  LedgerParts L, double* largest, bool settle = true) {
  // Specs with reservoir patches take one strip width (128 columns).
  if (v.P->edge != nullptr) {
    launch_step<T, S, 128, kLimited, true>(stream, v, sh, xf, yf, src, dst, dt, k, L, largest,
                                           settle);
    return;
  }
  switch (sh.ty) {
    case 64:
      launch_step<T, S, 64, kLimited, false>(stream, v, sh, xf, yf, src, dst, dt, k, L, largest,
                                             settle);
      break;
    case 128:
      launch_step<T, S, 128, kLimited, false>(stream, v, sh, xf, yf, src, dst, dt, k, L, largest,
                                              settle);
      break;
    default:
      launch_step<T, S, 256, kLimited, false>(stream, v, sh, xf, yf, src, dst, dt, k, L, largest,
                                              settle);
  }
}

// The explicit candidate of one FP32 step, unsettled.
template <int S>
void launch_candidate(Context& ctx, float dt, const float* in, float* out, uint32_t step,
                      const float* xf, const float* yf) {
  auto& a = ctx.adr;
  const bool limited = a.params.limited != 0;
  AdrShape sh = adr_shape(a.params.nx, a.params.ny, limited ? 2 : 1, a.params.edge != nullptr);
  sh.blocks = step_blocks(sh);
  LedgerParts L{a.partials.data(), a.ledger.data(), sh.blocks};
  double* largest = a.partials.data() + static_cast<size_t>(3 * kMaxSignals * 2) * sh.blocks;
  const auto* src = reinterpret_cast<const uint8_t*>(in);
  auto* dst = reinterpret_cast<uint8_t*>(out);
  const AdrView v = view_of(ctx);
  if (limited) {
    launch_typed<float, S, true>(ctx.stream, v, sh, xf, yf, src, dst, dt, step, L, largest, false);
  } else {
    launch_typed<float, S, false>(ctx.stream, v, sh, xf, yf, src, dst, dt, step, L, largest,
                                  false);
  }
}

template <int S>
void launch_steps(Context& ctx, float dt, uint32_t steps, const float* xf, const float* yf) {
  auto& a = ctx.adr;
  const auto& P = a.params;
  const bool limited = P.limited != 0;
  AdrShape sh = adr_shape(P.nx, P.ny, limited ? 2 : 1, P.edge != nullptr);
  sh.blocks = step_blocks(sh);
  LedgerParts L{a.partials.data(), a.ledger.data(), sh.blocks};
  double* largest = a.partials.data() + static_cast<size_t>(3 * kMaxSignals * 2) * sh.blocks;
  const AdrView v = view_of(ctx);
  for (uint32_t k = 0; k < steps; ++k) {
    const uint8_t* src = a.current == 0 ? a.levels_a.data() : a.levels_b.data();
    uint8_t* dst = a.current == 0 ? a.levels_b.data() : a.levels_a.data();
    if (a.precision == 64) {
      if (limited) {
        launch_typed<double, S, true>(ctx.stream, v, sh, xf, yf, src, dst, dt, k, L, largest);
      } else {
        launch_typed<double, S, false>(ctx.stream, v, sh, xf, yf, src, dst, dt, k, L, largest);
      }
    } else {
      if (limited) {
        launch_typed<float, S, true>(ctx.stream, v, sh, xf, yf, src, dst, dt, k, L, largest);
      } else {
        launch_typed<float, S, false>(ctx.stream, v, sh, xf, yf, src, dst, dt, k, L, largest);
      }
    }
    check(hipGetLastError(), "launch transport step");
    a.current ^= 1;
  }
  if (a.precision == 64 && steps > 0) {
    k_adr_ledger_tail<S><<<3 * S, kAdrBlock, 0, ctx.stream>>>(P, L, steps - 1, dt,
                                                             a.flags.data());
    check(hipGetLastError(), "launch ledger tail");
  }
}

void fe_candidate_f32(Context& ctx, float dt, const float* in, float* out, uint32_t step,
                      const float* xf, const float* yf) {
  switch (ctx.adr.params.signals) {
    case 1: launch_candidate<1>(ctx, dt, in, out, step, xf, yf); break;
    case 2: launch_candidate<2>(ctx, dt, in, out, step, xf, yf); break;
    case 3: launch_candidate<3>(ctx, dt, in, out, step, xf, yf); break;
    case 4: launch_candidate<4>(ctx, dt, in, out, step, xf, yf); break;
    case 5: launch_candidate<5>(ctx, dt, in, out, step, xf, yf); break;
    case 6: launch_candidate<6>(ctx, dt, in, out, step, xf, yf); break;
    case 7: launch_candidate<7>(ctx, dt, in, out, step, xf, yf); break;
    default: launch_candidate<8>(ctx, dt, in, out, step, xf, yf); break;
  }
  check(hipGetLastError(), "launch transport candidate");
}

void launch_settle_f32(Context& ctx, float* levels, size_t count, const double* largest,
                       uint32_t parts, uint32_t step) {
  const uint32_t blocks =
      static_cast<uint32_t>(std::min<size_t>((count + kAdrBlock - 1) / kAdrBlock, 2048));
  k_adr_settle<float><<<blocks, kAdrBlock, 0, ctx.stream>>>(levels, count, largest, parts,
                                                            ctx.adr.flags.data(), step);
  check(hipGetLastError(), "launch k_adr_settle");
}

void adr_run(Context& ctx, float dt, uint32_t steps, double* ledger) {
  auto& a = ctx.adr;
  const auto& P = a.params;
  const float* xf = a.flow_field ? ctx.flow.faces.data() : a.xf.data();
  const float* yf = a.flow_field ? ctx.flow.faces.data() + ctx.flow.x_faces() : a.yf.data();
  const uint32_t init_flags[2] = {0u, 0xffffffffu};
  check(hipMemcpyAsync(a.flags.data(), init_flags, sizeof(init_flags), hipMemcpyHostToDevice,
                       ctx.stream),
        "clear flags");
  if (a.integration != MSIM_FORWARD_EULER || P.nz > 1) {
    // Implicit steps (FP32 levels, each step's solve judged on the host),
    // and every step on a grid of several layers.
    check(hipEventRecord(ctx.start, ctx.stream), "event");
    if (a.integration != MSIM_FORWARD_EULER) {
      implicit_run(ctx, dt, steps, xf, yf);
    } else {
      adr3_run(ctx, dt, steps, xf, yf, a.zf.data());
    }
    check(hipEventRecord(ctx.stop, ctx.stream), "event");
    check(hipEventSynchronize(ctx.stop), "event sync");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
    a.seconds = ms * 1.0e-3;
    return;
  }
  if (a.precision == 64) a.ledger.reserve(static_cast<size_t>(steps) * P.signals * 4 + 1);
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  switch (P.signals) {
    case 1: launch_steps<1>(ctx, dt, steps, xf, yf); break;
    case 2: launch_steps<2>(ctx, dt, steps, xf, yf); break;
    case 3: launch_steps<3>(ctx, dt, steps, xf, yf); break;
    case 4: launch_steps<4>(ctx, dt, steps, xf, yf); break;
    case 5: launch_steps<5>(ctx, dt, steps, xf, yf); break;
    case 6: launch_steps<6>(ctx, dt, steps, xf, yf); break;
    case 7: launch_steps<7>(ctx, dt, steps, xf, yf); break;
    default: launch_steps<8>(ctx, dt, steps, xf, yf); break;
  }
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  uint32_t flags[2] = {0, 0};
  check(hipMemcpyAsync(flags, a.flags.data(), sizeof(flags), hipMemcpyDeviceToHost, ctx.stream),
        "read flags");
  std::vector<double> rows;
  if (a.precision == 64 && steps > 0) {
    rows.resize(static_cast<size_t>(steps) * P.signals * 4);
    check(hipMemcpyAsync(rows.data(), a.ledger.data(), rows.size() * sizeof(double),
                         hipMemcpyDeviceToHost, ctx.stream),
          "read ledger");
  }
  check(hipStreamSynchronize(ctx.stream), "transport steps");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  a.seconds = ms * 1.0e-3;
  if (flags[0] != 0) {
    // Levels stay those before the failed step: step k read buffer (k even ?
    // the starting one : the other).
    const uint32_t failed = flags[1];
    a.current = static_cast<int>((a.current ^ (steps & 1u)) ^ (failed & 1u));
    // The steps before it stand: their ledger rows carry the mass on, and go
    // to the caller.
    if (!rows.empty()) {
      const uint32_t kept = std::min(failed, steps);
      for (uint32_t k = 0; k < kept; ++k) {
        for (uint32_t s = 0; s < P.signals; ++s) {
          double* row = rows.data() + (static_cast<size_t>(k) * P.signals + s) * 4;
          row[0] = a.host_mass[s];
          a.host_mass[s] = row[1];
        }
      }
      if (ledger != nullptr && kept > 0) {
        std::memcpy(ledger, rows.data(),
                    static_cast<size_t>(kept) * P.signals * 4 * sizeof(double));
      }
    }
    if ((flags[0] & kFlagNegative) != 0) {
      throw Error(MSIM_DOMAIN_ERROR,
                  "signal grid update produced a non-finite or negative concentration (step " +
                      std::to_string(failed) + ")");
    }
    // The settle refused the step (accept_step, signals.cpp:1003-1011).
    throw Error(MSIM_INVALID_ARGUMENT, "signal grid levels must be finite and non-negative (step " +
                                           std::to_string(failed) + ")");
  }
  if (!rows.empty()) {
    // Mass before each step: the mass after the previous one.
    for (uint32_t k = 0; k < steps; ++k) {
      for (uint32_t s = 0; s < P.signals; ++s) {
        double* row = rows.data() + (static_cast<size_t>(k) * P.signals + s) * 4;
        row[0] = a.host_mass[s];
        a.host_mass[s] = row[1];
      }
    }
    if (ledger != nullptr) std::memcpy(ledger, rows.data(), rows.size() * sizeof(double));
  }
}

double adr_refresh_outflow(Context& ctx) {
  auto& a = ctx.adr;
  const auto& P = a.params;
  if (P.has_field == 0) return 0.0;
  const float* xf = a.flow_field ? ctx.flow.faces.data() : a.xf.data();
  const float* yf = a.flow_field ? ctx.flow.faces.data() + ctx.flow.x_faces() : a.yf.data();
  const uint8_t* levels = a.current == 0 ? a.levels_a.data() : a.levels_b.data();
  const size_t n = a.sites();
  const uint32_t blocks = std::min<uint32_t>(site_blocks(n), 2048);
  // The survey's mass goes to a.mass only; the ledger's running mass
  // (host_mass) is left to the steps' own rows.
  if (P.nz > 1) {
    adr3_survey(ctx, xf, yf, a.zf.data(), reinterpret_cast<const float*>(levels),
                a.state.data(), a.partials.data(), a.mass.data());
  } else if (a.precision == 64) {
    k_adr_survey<double><<<blocks, kAdrBlock, 0, ctx.stream>>>(
        P, a.solid.data(), xf, yf, reinterpret_cast<const double*>(levels), a.partials.data(),
        a.state.data(), a.mass.data());
  } else {
    k_adr_survey<float><<<blocks, kAdrBlock, 0, ctx.stream>>>(
        P, a.solid.data(), xf, yf, reinterpret_cast<const float*>(levels), a.partials.data(),
        a.state.data(), a.mass.data());
  }
  check(hipGetLastError(), "launch k_adr_survey");
  CgState host{};
  check(hipMemcpyAsync(&host, a.state.data(), sizeof(CgState), hipMemcpyDeviceToHost, ctx.stream),
        "read survey");
  check(hipStreamSynchronize(ctx.stream), "survey");
  a.greatest_outflow = host.peak;
  return host.peak;
}

void adr_download(Context& ctx, void* levels) {
  auto& a = ctx.adr;
  const size_t bytes =
      a.params.signals * a.sites() * (a.precision == 64 ? sizeof(double) : sizeof(float));
  const uint8_t* src = a.current == 0 ? a.levels_a.data() : a.levels_b.data();
  check(hipMemcpyAsync(levels, src, bytes, hipMemcpyDeviceToHost, ctx.stream), "download levels");
  check(hipStreamSynchronize(ctx.stream), "download levels");
}

// ---------------------------------------------------------------------------
// The allocation-free step (msim_hip_advance_signal_grid_static)
// ---------------------------------------------------------------------------
// Everything one FP32 forward-Euler step needs on the device besides the
// caller's host arrays lives in one file-scope array: the staged mask, faces,
// occupancy and reaction rates, both level buffers, the ledger and survey
// partials, the flags and the survey's state. Each device holds its own copy
// of the array for the life of the process; host code reaches it only through
// the address hipGetSymbolAddress gives on that device (the array's name is a
// host-side stand-in, not a device address).

namespace {

constexpr size_t kStaticScratchBytes = size_t(64) << 20;
constexpr int kStaticDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

void* g_scratch[kStaticDevices] = {};
std::mutex g_scratch_lock[kStaticDevices];

// Carves 256-byte aligned regions out of the scratch array, in order.
struct ScratchLayout {
  size_t used = 0;
  size_t take(size_t bytes) {
    const size_t at = used;
    used += (bytes + 255) & ~size_t(255);
    return at;
  }
};

__global__ void k_fill_bytes(uint8_t* __restrict__ out, size_t count, uint8_t value) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    out[i] = value;
  }
}

template <int S>
void launch_static_step(hipStream_t stream, const AdrView& v, const AdrShape& sh, const float* xf,
                        const float* yf, const uint8_t* src, uint8_t* dst, float dt,
                        LedgerParts L, double* largest) {
  if (v.P->limited != 0) {
    launch_typed<float, S, true>(stream, v, sh, xf, yf, src, dst, dt, 0, L, largest);
  } else {
    launch_typed<float, S, false>(stream, v, sh, xf, yf, src, dst, dt, 0, L, largest);
  }
}

}  // namespace

size_t static_scratch_bytes() { return kStaticScratchBytes; }

void bind_static_scratch(int device) {
  if (device < 0 || device >= kStaticDevices) return;
  std::lock_guard<std::mutex> hold(g_scratch_lock[device]);
  if (g_scratch[device] != nullptr) return;
  void* ptr = nullptr;
  check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  g_scratch[device] = ptr;
}

void adr_step_static(Context& ctx, const msim_signal_grid& grid, float dt, float* levels) {
  if (ctx.device < 0 || ctx.device >= kStaticDevices || g_scratch[ctx.device] == nullptr) {
    throw Error(MSIM_DEVICE_ERROR, "the static scratch of this device is not set up");
  }
  if (grid.integration != MSIM_FORWARD_EULER) {
    throw Error(MSIM_UNSUPPORTED, "the allocation-free transport step takes forward Euler only");
  }
  if (grid.patches != nullptr && grid.patch_count > 0) {
    throw Error(MSIM_UNSUPPORTED,
                "the allocation-free transport step takes no boundary patches");
  }
  if (grid.grid.nz != 1) {
    throw Error(MSIM_UNSUPPORTED, "the allocation-free transport step takes a grid of one layer");
  }
  AdrParams P = transport_params(grid);
  const uint32_t S = P.signals;
  const size_t n = static_cast<size_t>(P.nx) * P.ny;
  const AdrShape base = adr_shape(P.nx, P.ny, P.limited != 0 ? 2 : 1, false);
  AdrShape sh = base;
  sh.blocks = step_blocks(sh);
  const uint32_t survey_blocks = std::min<uint32_t>(site_blocks(n), 2048);
  ScratchLayout at;
  const size_t o_state = at.take(sizeof(CgState));
  const size_t o_flags = at.take(2 * sizeof(uint32_t));
  const size_t o_mass = at.take(kMaxSignals * sizeof(double));
  const size_t o_parts = at.take(
      std::max<size_t>(static_cast<size_t>(3 * kMaxSignals * 2 + 1) * sh.blocks,
                       static_cast<size_t>(kMaxSignals + 1) * survey_blocks) *
      sizeof(double));
  const size_t o_solid = at.take(n);
  const size_t nxf = static_cast<size_t>(P.nx + 1) * P.ny, nyf = static_cast<size_t>(P.nx) * (P.ny + 1);
  const size_t o_xf = P.has_field != 0 ? at.take(nxf * sizeof(float)) : 0;
  const size_t o_yf = P.has_field != 0 ? at.take(nyf * sizeof(float)) : 0;
  const size_t o_phi = P.has_monod != 0 ? at.take(n * sizeof(float)) : 0;
  const size_t o_loss = P.has_affine != 0 ? at.take(S * n * sizeof(float)) : 0;
  const size_t o_source = P.has_affine != 0 ? at.take(S * n * sizeof(float)) : 0;
  const size_t o_a = at.take(S * n * sizeof(float));
  const size_t o_b = at.take(S * n * sizeof(float));
  if (at.used > kStaticScratchBytes) {
    throw Error(MSIM_INVALID_ARGUMENT,
                "the allocation-free transport step needs " + std::to_string(at.used) +
                    " bytes of device scratch for this grid; its limit is " +
                    std::to_string(kStaticScratchBytes) + " bytes");
  }
  // One call at a time per device in this process: every call shares the
  // device's array.
  std::lock_guard<std::mutex> hold(g_scratch_lock[ctx.device]);
  uint8_t* base_ptr = static_cast<uint8_t*>(g_scratch[ctx.device]);
  auto* state = reinterpret_cast<CgState*>(base_ptr + o_state);
  auto* flags = reinterpret_cast<uint32_t*>(base_ptr + o_flags);
  auto* mass = reinterpret_cast<double*>(base_ptr + o_mass);
  auto* parts = reinterpret_cast<double*>(base_ptr + o_parts);
  uint8_t* solid = base_ptr + o_solid;
  float* xf = P.has_field != 0 ? reinterpret_cast<float*>(base_ptr + o_xf) : nullptr;
  float* yf = P.has_field != 0 ? reinterpret_cast<float*>(base_ptr + o_yf) : nullptr;
  float* phi = P.has_monod != 0 ? reinterpret_cast<float*>(base_ptr + o_phi) : nullptr;
  float* loss = P.has_affine != 0 ? reinterpret_cast<float*>(base_ptr + o_loss) : nullptr;
  float* source = P.has_affine != 0 ? reinterpret_cast<float*>(base_ptr + o_source) : nullptr;
  uint8_t* lv_a = base_ptr + o_a;
  uint8_t* lv_b = base_ptr + o_b;
  hipStream_t stream = ctx.stream;
  // The array keeps the last call's contents: the state (with its tickets),
  // the flags and every staged array are written afresh.
  const CgState zero_state{};
  const uint32_t init_flags[2] = {0u, 0xffffffffu};
  check(hipMemcpyAsync(state, &zero_state, sizeof(CgState), hipMemcpyHostToDevice, stream),
        "clear state");
  check(hipMemcpyAsync(flags, init_flags, sizeof(init_flags), hipMemcpyHostToDevice, stream),
        "clear flags");
  if (grid.grid.obstacles != nullptr) {
    check(hipMemcpyAsync(solid, grid.grid.obstacles, n, hipMemcpyHostToDevice, stream),
          "upload mask");
  } else {
    k_fill_bytes<<<std::min<uint32_t>(site_blocks(n), 1024), kAdrBlock, 0, stream>>>(solid, n, 0);
    check(hipGetLastError(), "launch k_fill_bytes");
  }
  if (P.has_field != 0) {
    check(hipMemcpyAsync(xf, grid.x_faces, nxf * sizeof(float), hipMemcpyHostToDevice, stream),
          "upload x faces");
    check(hipMemcpyAsync(yf, grid.y_faces, nyf * sizeof(float), hipMemcpyHostToDevice, stream),
          "upload y faces");
  }
  if (P.has_monod != 0) {
    check(hipMemcpyAsync(phi, grid.phi, n * sizeof(float), hipMemcpyHostToDevice, stream),
          "upload phi");
  }
  if (P.has_affine != 0) {
    check(hipMemcpyAsync(loss, grid.loss_rates, S * n * sizeof(float), hipMemcpyHostToDevice,
                         stream),
          "upload loss");
    check(hipMemcpyAsync(source, grid.source_rates, S * n * sizeof(float), hipMemcpyHostToDevice,
                         stream),
          "upload source");
  }
  check(hipMemcpyAsync(lv_a, levels, S * n * sizeof(float), hipMemcpyHostToDevice, stream),
        "upload levels");
  // The greatest outflow for the explicit bound (transport_cache).
  k_adr_survey<float><<<survey_blocks, kAdrBlock, 0, stream>>>(
      P, solid, xf, yf, reinterpret_cast<const float*>(lv_a), parts, state, mass);
  check(hipGetLastError(), "launch k_adr_survey");
  CgState host{};
  check(hipMemcpyAsync(&host, state, sizeof(CgState), hipMemcpyDeviceToHost, stream),
        "read survey");
  check(hipStreamSynchronize(stream), "survey");
  double max_loss[kMaxSignals] = {};
  max_losses(grid, max_loss);
  validate_explicit_bound(P, MSIM_FORWARD_EULER, P.has_field != 0 ? host.peak : 0.0, max_loss,
                          dt);
  if (dt == 0.0f) return;
  const AdrView v{&P, solid, phi, source, loss, flags};
  LedgerParts L{parts, nullptr, sh.blocks};
  double* largest = parts + static_cast<size_t>(3 * kMaxSignals * 2) * sh.blocks;
  switch (S) {
    case 1: launch_static_step<1>(stream, v, sh, xf, yf, lv_a, lv_b, dt, L, largest); break;
    case 2: launch_static_step<2>(stream, v, sh, xf, yf, lv_a, lv_b, dt, L, largest); break;
    case 3: launch_static_step<3>(stream, v, sh, xf, yf, lv_a, lv_b, dt, L, largest); break;
    case 4: launch_static_step<4>(stream, v, sh, xf, yf, lv_a, lv_b, dt, L, largest); break;
    case 5: launch_static_step<5>(stream, v, sh, xf, yf, lv_a, lv_b, dt, L, largest); break;
    case 6: launch_static_step<6>(stream, v, sh, xf, yf, lv_a, lv_b, dt, L, largest); break;
    case 7: launch_static_step<7>(stream, v, sh, xf, yf, lv_a, lv_b, dt, L, largest); break;
    default: launch_static_step<8>(stream, v, sh, xf, yf, lv_a, lv_b, dt, L, largest); break;
  }
  check(hipGetLastError(), "launch transport step");
  uint32_t result[2] = {0, 0};
  check(hipMemcpyAsync(result, flags, sizeof(result), hipMemcpyDeviceToHost, stream),
        "read flags");
  check(hipStreamSynchronize(stream), "transport step");
  if ((result[0] & kFlagNegative) != 0) {
    throw Error(MSIM_DOMAIN_ERROR,
                "signal grid update produced a non-finite or negative concentration (step 0)");
  }
  if (result[0] != 0) {
    throw Error(MSIM_INVALID_ARGUMENT, "signal grid levels must be finite and non-negative (step 0)");
  }
  check(hipMemcpyAsync(levels, lv_b, S * n * sizeof(float), hipMemcpyDeviceToHost, stream),
        "download levels");
  check(hipStreamSynchronize(stream), "download levels");
}

}  // namespace msim
