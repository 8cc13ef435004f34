// Multilevel preconditioner and flexible conjugate gradient for the
// depth-averaged pressure system (no counterpart in the reference, whose only
// preconditioner is Jacobi, cpp/cpu/cpu_flow.cpp:65-69).
//
// Hierarchy: 2x2 aggregation on the structured grid. With piecewise-constant
// prolongation P the Galerkin operator P^T A P is again a 5-point operator:
// the coupling between two aggregates is the sum of the fine couplings across
// their shared faces, the boundary term the sum over the aggregate. Coarsening
// stops at <= 256 cells, whose operator is inverted densely on the host.
//
// Cycle (V(2,2), damped Jacobi, omega = 0.8): on each level
//   x = omega D^-1 r;  x += omega D^-1 (r - A x);  r_c = P^T (r - A x);
//   e_c = cycle(r_c);  x += c P e_c;  twice x += omega D^-1 (r - A x),
// with c = 1.8 over-weighting the coarse correction (piecewise-constant
// aggregation under-corrects by about a factor of two). Each level is two
// fused passes: `down` (both pre-smoothing sweeps, residual, restriction;
// the pre-smoothed iterate is kept) and `up` (prolongation, both
// post-smoothing sweeps), each staging a 16 x 64 tile plus a two-site halo
// in LDS. A second sweep on each side takes the iterations at 4096 x 2048
// from 53 to about 33 for a quarter more work per cycle.
// Levels of <= 4096 cells run in one single-workgroup kernel (measured at
// 4096x2048: 16384 cost 130 us of coarse work per cycle, 4096 109 us).
//
// Outer solver: flexible CG (Notay's FCG(1): beta_k = -(z.q_{k-1}) /
// (p_{k-1}.q_{k-1})) in FP32 with FP64 sums, inside the FP64 iterative
// refinement of the mixed solver; the FP64 true residual decides convergence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "stencil.hpp"

namespace msim {
namespace {

constexpr int kTX = 16;                // fine rows (x) per tile
constexpr int kTY = 64;                // fine columns (y) per tile
constexpr int kHX = kTX + 2;           // with a one-site halo
constexpr int kHY = kTY + 2;
constexpr int kHX2 = kTX + 4;          // with a two-site halo
constexpr int kHY2 = kTY + 4;
constexpr int kMgThreads = (kTX / 2) * (kTY / 2);  // one thread per 2x2 aggregate
constexpr int kTailThreads = 1024;
constexpr uint32_t kCoarseCells = 256;  // densely inverted
constexpr uint32_t kTailCells = 4096;  // levels handled by the single-block kernel
constexpr uint32_t kMaxLevels = 24;
static_assert(4 * kCoarseCells <= kTailThreads, "the coarsest solve splits every row in parts");
constexpr float kOmega = 0.8f;
constexpr float kCorrection = 1.8f;

struct LevelView {
  const float* wx;
  const float* wy;
  const float* bc;
  const float* dinv;
  uint32_t nx, ny;
};

// Level 0 couplings from the assembled operator: wx = harm * 1/h^2.
__global__ void k_mg_level0(const float* __restrict__ cxl, const float* __restrict__ cyl,
                            const float* __restrict__ bc, float* __restrict__ wx,
                            float* __restrict__ wy, float* __restrict__ bcl, size_t n, float isx,
                            float isy) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  wx[i] = cxl[i] * isx;
  wy[i] = cyl[i] * isy;
  bcl[i] = bc[i];
}

__global__ void k_mg_dinv(const float* __restrict__ wx, const float* __restrict__ wy,
                          const float* __restrict__ bc, float* __restrict__ dinv, uint32_t nx,
                          uint32_t ny) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(nx) * ny) return;
  const uint32_t x = static_cast<uint32_t>(i / ny), y = static_cast<uint32_t>(i % ny);
  float d = bc[i] + wx[i] + wy[i];
  if (x + 1 < nx) d += wx[i + ny];
  if (y + 1 < ny) d += wy[i + 1];
  dinv[i] = d > 0.0f ? 1.0f / d : 0.0f;
}

// Galerkin coarse operator of 2x2 aggregation.
__global__ void k_mg_coarsen(LevelView f, float* __restrict__ wx, float* __restrict__ wy,
                             float* __restrict__ bc, uint32_t nxc, uint32_t nyc) {
  const size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<size_t>(nxc) * nyc) return;
  const uint32_t X = static_cast<uint32_t>(c / nyc), Y = static_cast<uint32_t>(c % nyc);
  const uint32_t fx = 2 * X, fy = 2 * Y;
  const bool x1 = fx + 1 < f.nx, y1 = fy + 1 < f.ny;
  const size_t i = static_cast<size_t>(fx) * f.ny + fy;
  float cx = 0.0f, cy = 0.0f;
  if (X > 0) cx = f.wx[i] + (y1 ? f.wx[i + 1] : 0.0f);
  if (Y > 0) cy = f.wy[i] + (x1 ? f.wy[i + f.ny] : 0.0f);
  float b = f.bc[i];
  if (y1) b += f.bc[i + 1];
  if (x1) b += f.bc[i + f.ny];
  if (x1 && y1) b += f.bc[i + f.ny + 1];
  wx[c] = cx;
  wy[c] = cy;
  bc[c] = b;
}

// A tile grown by 2 in LDS (kHX2 x kHY2, pitch kHY2): r, the inverse
// diagonal, the couplings to the lower x and y neighbours and the boundary
// term, each zero off the grid, and two iterates. Every sweep reads only
// LDS, so each global value of a tile is loaded once.
constexpr int kTile2 = kHX2 * kHY2;
struct TileLds {
  float r[kTile2], d[kTile2], wx[kTile2], wy[kTile2], bc[kTile2], x[kTile2], y[kTile2];
};

// Loads the tile's operator and r, and x = init(i, fx, fy) (zero off the grid).
template <class Init>
__device__ inline void stage_tile(TileLds& T, const LevelView& L, const float* __restrict__ r,
                                  uint32_t fx0, uint32_t fy0, Init&& init) {
  for (int idx = threadIdx.x; idx < kTile2; idx += kMgThreads) {
    const int64_t fx = static_cast<int64_t>(fx0) + idx / kHY2 - 2;
    const int64_t fy = static_cast<int64_t>(fy0) + idx % kHY2 - 2;
    float rv = 0.0f, dv = 0.0f, wxv = 0.0f, wyv = 0.0f, bv = 0.0f, xv = 0.0f;
    if (fx >= 0 && fx < L.nx && fy >= 0 && fy < L.ny) {
      const size_t i = static_cast<size_t>(fx) * L.ny + static_cast<size_t>(fy);
      rv = r[i];
      dv = L.dinv[i];
      wxv = L.wx[i];
      wyv = L.wy[i];
      bv = L.bc[i];
      xv = init(i, rv, dv, static_cast<uint32_t>(fx), static_cast<uint32_t>(fy));
    }
    T.r[idx] = rv;
    T.d[idx] = dv;
    T.wx[idx] = wxv;
    T.wy[idx] = wyv;
    T.bc[idx] = bv;
    T.x[idx] = xv;
  }
}

// (A s)_p at LDS position p: the boundary term, then the lower and upper x
// couplings, then the lower and upper y couplings.
__device__ inline float lds_ax(const TileLds& T, const float* s, int p) {
  const float xc = s[p];
  float ax = T.bc[p] * xc;
  ax += T.wx[p] * (xc - s[p - kHY2]);
  ax += T.wx[p + kHY2] * (xc - s[p + kHY2]);
  ax += T.wy[p] * (xc - s[p - 1]);
  ax += T.wy[p + 1] * (xc - s[p + 1]);
  return ax;
}

// One damped-Jacobi sweep out = in + omega D^-1 (r - A in) on the tile grown
// by 1; off the grid the inverse diagonal and the iterate are zero, so out is.
__device__ inline void lds_sweep(TileLds& T, const float* in, float* out) {
  for (int idx = threadIdx.x; idx < kHX * kHY; idx += kMgThreads) {
    const int p = (idx / kHY + 1) * kHY2 + idx % kHY + 1;
    out[p] = in[p] + kOmega * T.d[p] * (T.r[p] - lds_ax(T, in, p));
  }
}

__device__ inline bool cycle_stopped(const CgState* st) {
  return st->failure != 0 || st->halt != 0;
}

// Both pre-smoothing sweeps from zero, residual, restriction:
//   x1 = omega D^-1 r,  x2 = x1 + omega D^-1 (r - A x1),  r_c = P^T (r - A x2),
// x2 kept in xp for the up pass. x1 is staged on the tile grown by 2, x2 on
// the tile grown by 1; one thread per aggregate. One workgroup per tile,
// tiles in XCD-contiguous bands (t = (t / tiles_y, t % tiles_y)).
__global__ __launch_bounds__(kMgThreads) void k_mg_down(LevelView L, const float* __restrict__ r,
                                                        float* __restrict__ xp,
                                                        float* __restrict__ rc, uint32_t nxc,
                                                        uint32_t nyc, const CgState* st,
                                                        uint32_t tiles_y) {
  __shared__ TileLds T;
  if (cycle_stopped(st)) return;
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t fx0 = (t / tiles_y) * kTX, fy0 = (t % tiles_y) * kTY;
  stage_tile(T, L, r, fx0, fy0,
             [](size_t, float rv, float dv, uint32_t, uint32_t) { return kOmega * dv * rv; });
  __syncthreads();
  lds_sweep(T, T.x, T.y);
  __syncthreads();
  const int lx = threadIdx.x / (kTY / 2), ly = threadIdx.x % (kTY / 2);
  float sum = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const uint32_t fx = fx0 + 2 * lx + a, fy = fy0 + 2 * ly + b;
      if (fx < L.nx && fy < L.ny) {
        const size_t i = static_cast<size_t>(fx) * L.ny + fy;
        const int p = (2 * lx + a + 2) * kHY2 + 2 * ly + b + 2;
        sum += T.r[p] - lds_ax(T, T.y, p);
        xp[i] = T.y[p];
      }
    }
  }
  const uint32_t X = fx0 / 2 + lx, Y = fy0 / 2 + ly;
  if (X < nxc && Y < nyc) rc[static_cast<size_t>(X) * nyc + Y] = sum;
}

// Dot products the finest `up` pass leaves for flexible CG: z.r and z.q_prev.
struct UpDots {
  const float* q;  // previous A p (null: no dots)
  double* zr;      // [2][blocks]
  double* zq;
};

// Prolongation and both post-smoothing sweeps:
//   y = x_pre + c P e_c,  y1 = y + omega D^-1 (r - A y),
//   out = y1 + omega D^-1 (r - A y1).
// With `pre`, out = y instead (the cycle before post-smoothing). y is
// staged on the tile grown by 2, y1 on the tile grown by 1. Each workgroup
// walks a contiguous run of tiles (tile t = (t / tiles_y, t % tiles_y)) and
// each XCD a contiguous band, so tiles that share halo rows meet in one L2;
// the flexible-CG dots leave one partial per workgroup.
__global__ __launch_bounds__(kMgThreads) void k_mg_up(LevelView L, const float* __restrict__ r,
                                                      const float* __restrict__ xp,
                                                      const float* __restrict__ ec, uint32_t nyc,
                                                      float* __restrict__ out, const CgState* st,
                                                      UpDots dots, int pre, uint32_t tiles_y,
                                                      uint32_t tiles) {
  __shared__ TileLds T;
  __shared__ double s_red[kMgThreads / kWave];
  if (cycle_stopped(st)) return;
  const uint32_t per = (tiles + gridDim.x - 1) / gridDim.x;
  const uint32_t t0 = xcd_tile(blockIdx.x, gridDim.x) * per, t1 = min(tiles, t0 + per);
  double zr = 0.0, zq = 0.0;
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t fx0 = (t / tiles_y) * kTX, fy0 = (t % tiles_y) * kTY;
    stage_tile(T, L, r, fx0, fy0, [&](size_t i, float, float, uint32_t fx, uint32_t fy) {
      return xp[i] + kCorrection * ec[static_cast<size_t>(fx / 2) * nyc + fy / 2];
    });
    __syncthreads();
    if (pre == 0) lds_sweep(T, T.x, T.y);
    __syncthreads();
    // One site per thread in row order: unit-stride stores.
#pragma unroll
    for (int k = 0; k < (kTX * kTY) / kMgThreads; ++k) {
      const int idx = static_cast<int>(threadIdx.x) + k * kMgThreads;
      const int lxs = idx / kTY, lys = idx % kTY;
      const uint32_t fx = fx0 + lxs, fy = fy0 + lys;
      if (fx < L.nx && fy < L.ny) {
        const size_t i = static_cast<size_t>(fx) * L.ny + fy;
        const int p = (lxs + 2) * kHY2 + lys + 2;
        const float ri = T.r[p];
        const float z = pre != 0 ? T.x[p] : T.y[p] + kOmega * T.d[p] * (ri - lds_ax(T, T.y, p));
        out[i] = z;
        if (dots.q != nullptr) {
          zr += static_cast<double>(z) * static_cast<double>(ri);
          zq += static_cast<double>(z) * static_cast<double>(dots.q[i]);
        }
      }
    }
    __syncthreads();
  }
  if (dots.q != nullptr) {
    const uint32_t k = st->it_update + 1;
    const double a = block_sum<kMgThreads>(zr, s_red);
    const double b = block_sum<kMgThreads>(zq, s_red);
    if (threadIdx.x == 0) {
      dots.zr[(k & 1) * gridDim.x + blockIdx.x] = a;
      dots.zq[(k & 1) * gridDim.x + blockIdx.x] = b;
    }
  }
}

struct TailLevel {
  const float* wx;
  const float* wy;
  const float* bc;
  const float* dinv;
  float* r;
  float* e;
  float* xp;  // the pre-smoothed iterate, then the first post-smoothing sweep
  float* t;   // the first pre-smoothing sweep, then the prolonged iterate
  uint32_t nx, ny;
};

struct TailArgs {
  TailLevel lv[kMaxLevels];
  uint32_t count;      // levels from the first tail level to the coarsest
  const float* cinv;   // dense inverse of the coarsest operator, transposed
  uint32_t n;          // its size
  UpDots dots;         // when the tail starts at level 0
};

// (A x)_i on a tail level at site i = fx ny + fy, with x_j = xv(j); the
// terms in the order of the tile kernels.
template <typename X>
__device__ inline float tail_ax(const TailLevel& L, X xv, uint32_t i, uint32_t fx, uint32_t fy) {
  const float xc = xv(i);
  float ax = L.bc[i] * xc;
  if (fx > 0) ax += L.wx[i] * (xc - xv(i - L.ny));
  if (fx + 1 < L.nx) ax += L.wx[i + L.ny] * (xc - xv(i + L.ny));
  if (fy > 0) ax += L.wy[i] * (xc - xv(i - 1));
  if (fy + 1 < L.ny) ax += L.wy[i + 1] * (xc - xv(i + 1));
  return ax;
}

// out = x + omega D^-1 (r - A x) on every cell of a tail level, x_j = xv(j).
template <typename X>
__device__ inline void tail_sweep(const TailLevel& L, X xv, float* out) {
  const uint32_t cells = L.nx * L.ny;
  for (uint32_t i = threadIdx.x; i < cells; i += kTailThreads) {
    const uint32_t fx = i / L.ny, fy = i - fx * L.ny;
    out[i] = xv(i) + kOmega * L.dinv[i] * (L.r[i] - tail_ax(L, xv, i, fx, fy));
  }
}

// The cycle on the small levels in one workgroup: down passes, the dense
// coarsest solve, up passes, with workgroup barriers between the phases.
// The first sweep of each smoothing pair is not stored: the second sweep
// evaluates it at the sites its stencil reads (from zero going down: omega
// D^-1 r; going up: the prolonged iterate), so a level costs two phases each
// way. The coarsest inverse is stored transposed, so the threads of a wave
// read one row of it together.
__global__ __launch_bounds__(kTailThreads) void k_mg_tail(TailArgs T, const CgState* st) {
  __shared__ double s_red[kTailThreads / kWave];
  if (cycle_stopped(st)) return;
  for (uint32_t l = 0; l + 1 < T.count; ++l) {
    const TailLevel& F = T.lv[l];
    const TailLevel& Cl = T.lv[l + 1];
    const float* fr = F.r;
    const float* fd = F.dinv;
    tail_sweep(F, [fr, fd](uint32_t j) { return kOmega * fd[j] * fr[j]; }, F.xp);
    __syncthreads();
    const float* fxp = F.xp;
    const auto xp = [fxp](uint32_t j) { return fxp[j]; };
    const uint32_t cells = Cl.nx * Cl.ny;
    for (uint32_t c = threadIdx.x; c < cells; c += kTailThreads) {
      const uint32_t X = c / Cl.ny, Y = c - X * Cl.ny;
      float sum = 0.0f;
      for (int a = 0; a < 2; ++a) {
        for (int b = 0; b < 2; ++b) {
          const uint32_t fx = 2 * X + a, fy = 2 * Y + b;
          if (fx < F.nx && fy < F.ny) {
            const uint32_t i = fx * F.ny + fy;
            sum += F.r[i] - tail_ax(F, xp, i, fx, fy);
          }
        }
      }
      Cl.r[c] = sum;
    }
    __syncthreads();
  }
  double zr = 0.0, zq = 0.0;
  {
    // e = C^-1 r on the coarsest level (n <= kCoarseCells): the columns are
    // split into `parts` contiguous ranges, one thread per (row, range), the
    // rows of a wave adjacent, so each step of a thread's range is one
    // coalesced read of the transposed inverse; the ranges' sums are then
    // added in range order.
    const TailLevel& B = T.lv[T.count - 1];
    const bool dots = T.count == 1 && T.dots.q != nullptr;
    __shared__ float s_r[kCoarseCells];
    __shared__ float s_part[kTailThreads];
    const uint32_t n = T.n;
    for (uint32_t j = threadIdx.x; j < n; j += kTailThreads) s_r[j] = B.r[j];
    __syncthreads();
    const uint32_t parts = kTailThreads / n;  // n <= kCoarseCells <= kTailThreads / 4
    const uint32_t span = (n + parts - 1) / parts;
    if (threadIdx.x < parts * n) {
      const uint32_t row = threadIdx.x % n, part = threadIdx.x / n;
      const uint32_t j0 = part * span, j1 = min(n, j0 + span);
      float v = 0.0f;
      uint32_t j = j0;
      for (; j + 8 <= j1; j += 8) {
        float m[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) m[k] = T.cinv[static_cast<size_t>(j + k) * n + row];
#pragma unroll
        for (int k = 0; k < 8; ++k) v += m[k] * s_r[j + k];
      }
      for (; j < j1; ++j) v += T.cinv[static_cast<size_t>(j) * n + row] * s_r[j];
      s_part[part * n + row] = v;
    }
    __syncthreads();
    for (uint32_t row = threadIdx.x; row < n; row += kTailThreads) {
      float v = 0.0f;
      for (uint32_t part = 0; part < parts; ++part) v += s_part[part * n + row];
      B.e[row] = v;
      if (dots) {
        zr += static_cast<double>(v) * static_cast<double>(s_r[row]);
        zq += static_cast<double>(v) * static_cast<double>(T.dots.q[row]);
      }
    }
    __syncthreads();
  }
  for (int l = static_cast<int>(T.count) - 2; l >= 0; --l) {
    const TailLevel& F = T.lv[l];
    const TailLevel& Cl = T.lv[l + 1];
    const uint32_t cells = F.nx * F.ny;
    const bool dots = l == 0 && T.dots.q != nullptr;
    // The prolonged iterate x_pre + c P e_c, evaluated where it is read.
    const float* fxp = F.xp;
    const float* ce = Cl.e;
    const uint32_t ny = F.ny, cny = Cl.ny;
    const auto prolonged = [fxp, ce, ny, cny](uint32_t j) {
      const uint32_t jx = j / ny, jy = j - jx * ny;
      return fxp[j] + kCorrection * ce[(jx / 2) * cny + jy / 2];
    };
    tail_sweep(F, prolonged, F.t);
    __syncthreads();
    const float* ft = F.t;
    const auto smoothed = [ft](uint32_t j) { return ft[j]; };
    for (uint32_t i = threadIdx.x; i < cells; i += kTailThreads) {
      const uint32_t fx = i / F.ny, fy = i - fx * F.ny;
      const float ri = F.r[i];
      const float z = F.t[i] + kOmega * F.dinv[i] * (ri - tail_ax(F, smoothed, i, fx, fy));
      F.e[i] = z;
      if (dots) {
        zr += static_cast<double>(z) * static_cast<double>(ri);
        zq += static_cast<double>(z) * static_cast<double>(T.dots.q[i]);
      }
    }
    __syncthreads();
  }
  if (T.dots.q != nullptr) {
    const uint32_t k = st->it_update + 1;
    const double a = block_sum<kTailThreads>(zr, s_red);
    const double b = block_sum<kTailThreads>(zq, s_red);
    if (threadIdx.x == 0) {
      T.dots.zr[(k & 1)] = a;
      T.dots.zq[(k & 1)] = b;
    }
  }
}

// ----------------------------------------------------------------------------
// Flexible CG kernels (FP32 vectors, FP64 sums), iteration k = it_update + 1:
//   z_{k-1} = M(r_{k-1})                        (cycle kernels, with z.r, z.q_{k-1})
//   p_k = z_{k-1} + beta_k p_{k-1},  beta_k = -(z_{k-1}.q_{k-1}) / (p_{k-1}.q_{k-1})
//   q_k = A p_k,  x += alpha_{k-1} p_{k-1}      (k_fcg_apply, with p_k.q_k)
//   alpha_k = (z_{k-1}.r_{k-1}) / (p_k.q_k),  r_k = r_{k-1} - alpha_k q_k   (k_fcg_update)
// Scalars are re-added from per-block partials by every block in a fixed
// order, as in the Jacobi-PCG kernels; the stopping test of iteration k-1 is
// made by the operator kernel of iteration k.
// ----------------------------------------------------------------------------

struct FcgPartials {
  double* pq;  // [2][p1]
  double* rr;  // [2][p2]
  double* zr;  // [2][p0]
  double* zq;  // [2][p0]
  uint32_t p0, p1, p2;
};

template <typename T>
struct FcgRow {
  T z, po, x;
  float cx, cy, b;
  T lz, lpo;
  T hz, hpo;
  float hcy;
};

template <int TY>
__global__ __launch_bounds__(TY) void k_fcg_apply(
    const float* __restrict__ z, const float* __restrict__ p_old, float* __restrict__ p_new,
    float* __restrict__ q, float* __restrict__ xsol, const float* __restrict__ cxl,
    const float* __restrict__ cyl, const float* __restrict__ bc, uint32_t nx, uint32_t ny,
    uint32_t rows, float isx, float isy, CgState* st, FcgPartials parts) {
  using T = float;
  __shared__ T s_p[TY + 2];
  __shared__ float s_c[TY + 2];
  __shared__ double s_red[TY / kWave];
  __shared__ double s_b;
  if (st->failure != 0) return;
  const uint32_t k = st->it_update + 1;
  const uint32_t halt = st->halt;
  if (halt != 0 && halt + 1 < k) return;
  const Tile tl = tile_of<TY>(nx, ny, rows);
  const uint32_t t = threadIdx.x;
  const bool lead = blockIdx.x == 0 && t == 0;
  double alpha_prev = 0.0, beta = 0.0, rr_km1 = 0.0;
  bool finish = false;
  if (k >= 2) {
    const double pq_km1 = fixed_sum<TY>(parts.pq + ((k - 1) & 1) * parts.p1, parts.p1, s_red, &s_b);
    const double zr_km1 = fixed_sum<TY>(parts.zr + ((k - 1) & 1) * parts.p0, parts.p0, s_red, &s_b);
    rr_km1 = fixed_sum<TY>(parts.rr + ((k - 1) & 1) * parts.p2, parts.p2, s_red, &s_b);
    alpha_prev = zr_km1 / pq_km1;
    if (!isfinite(rr_km1)) {
      if (lead) st->failure = 3;
      return;
    }
    if (rr_km1 <= st->threshold) {
      finish = true;
    } else {
      const double zq_k = fixed_sum<TY>(parts.zq + (k & 1) * parts.p0, parts.p0, s_red, &s_b);
      beta = -zq_k / pq_km1;
      if (!isfinite(beta)) {
        if (lead) st->failure = 5;
        return;
      }
    }
  }
  if (lead) {
    st->it_apply = k;
    st->rr = rr_km1;
    st->alpha = alpha_prev;
    st->beta = beta;
    if (finish) st->halt = k - 1;
  }
  const T alpha = static_cast<T>(alpha_prev);
  if (finish) {
    if (tl.valid) {
      for (uint32_t xx = tl.x0; xx < tl.x1; ++xx) {
        const size_t i = static_cast<size_t>(xx) * ny + tl.y;
        xsol[i] = xsol[i] + alpha * p_old[i];
      }
    }
    return;
  }
  const T betaT = static_cast<T>(beta);
  auto load = [&](uint32_t xx, bool full) {
    FcgRow<T> w{};
    if (xx >= nx) return w;
    const size_t base = static_cast<size_t>(xx) * ny;
    if (tl.valid) {
      const size_t i = base + tl.y;
      w.z = z[i];
      w.po = p_old[i];
      w.cx = cxl[i];
      if (full) {
        w.x = xsol[i];
        w.cy = cyl[i];
        w.b = bc[i];
      }
    }
    if (full && tl.lo_halo) {
      const size_t i = base + tl.ylo;
      w.lz = z[i];
      w.lpo = p_old[i];
    }
    if (full && tl.hi_halo) {
      const size_t i = base + tl.yhi;
      w.hz = z[i];
      w.hpo = p_old[i];
      w.hcy = cyl[i];
    }
    return w;
  };
  double acc = 0.0;
  if (tl.x0 < tl.x1) {
    FcgRow<T> cur_row = load(tl.x0, true);
    FcgRow<T> next_row = load(tl.x0 + 1, tl.x0 + 1 < tl.x1);
    T prev = 0.0f;
    if (tl.x0 > 0) {
      const FcgRow<T> w = load(tl.x0 - 1, false);
      prev = w.z + betaT * w.po;
    }
    T cur = cur_row.z + betaT * cur_row.po;
    for (uint32_t xx = tl.x0; xx < tl.x1; ++xx) {
      const FcgRow<T> ahead = xx + 2 <= tl.x1 ? load(xx + 2, xx + 2 < tl.x1) : FcgRow<T>{};
      const T nxt = next_row.z + betaT * next_row.po;
      s_p[t + 1] = tl.valid ? cur : 0.0f;
      s_c[t + 1] = tl.valid ? cur_row.cy : 0.0f;
      if (t == 0) s_p[0] = tl.lo_halo ? cur_row.lz + betaT * cur_row.lpo : 0.0f;
      if (t == TY - 1) {
        s_p[TY + 1] = tl.hi_halo ? cur_row.hz + betaT * cur_row.hpo : 0.0f;
        s_c[TY + 1] = tl.hi_halo ? cur_row.hcy : 0.0f;
      }
      __syncthreads();
      if (tl.valid) {
        const size_t idx = static_cast<size_t>(xx) * ny + tl.y;
        T qv = cur_row.b * cur;
        qv += (cur_row.cx * isx) * (cur - prev);
        qv += (next_row.cx * isx) * (cur - nxt);
        qv += (s_c[t + 1] * isy) * (cur - s_p[t]);
        qv += (s_c[t + 2] * isy) * (cur - s_p[t + 2]);
        q[idx] = qv;
        p_new[idx] = cur;
        xsol[idx] = cur_row.x + alpha * cur_row.po;
        acc += static_cast<double>(cur) * static_cast<double>(qv);
      }
      __syncthreads();
      prev = cur;
      cur = nxt;
      cur_row = next_row;
      next_row = ahead;
    }
  }
  const double part = block_sum<TY>(acc, s_red);
  if (t == 0) parts.pq[(k & 1) * parts.p1 + blockIdx.x] = part;
}

constexpr int kUnroll = 4;

__global__ __launch_bounds__(kStream) void k_fcg_update(float* __restrict__ r,
                                                        const float* __restrict__ q, size_t n,
                                                        CgState* st, FcgPartials parts) {
  __shared__ double s_red[kStream / kWave];
  __shared__ double s_b;
  if (st->failure != 0) return;
  const uint32_t k = st->it_apply;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (st->halt != 0) {
    if (lead) st->it_update = k;
    return;
  }
  const double zr_k = fixed_sum<kStream>(parts.zr + (k & 1) * parts.p0, parts.p0, s_red, &s_b);
  const double pq_k = fixed_sum<kStream>(parts.pq + (k & 1) * parts.p1, parts.p1, s_red, &s_b);
  if (!isfinite(pq_k) || pq_k <= 0.0) {
    if (lead) st->failure = 1;
    return;
  }
  const double a = zr_k / pq_k;
  if (!isfinite(a)) {
    if (lead) st->failure = 2;
    return;
  }
  if (lead) {
    st->it_update = k;
    st->curvature = pq_k;
  }
  const float alpha = static_cast<float>(a);
  double rr = 0.0;
  const size_t chunk = static_cast<size_t>(kStream) * kUnroll;
  for (size_t base = static_cast<size_t>(blockIdx.x) * chunk + threadIdx.x; base < n;
       base += static_cast<size_t>(gridDim.x) * chunk) {
    float rv[kUnroll], qv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = base + static_cast<size_t>(u) * kStream;
      rv[u] = i < n ? r[i] : 0.0f;
      qv[u] = i < n ? q[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = base + static_cast<size_t>(u) * kStream;
      const float ri = rv[u] - alpha * qv[u];
      if (i < n) r[i] = ri;
      rr += static_cast<double>(ri) * static_cast<double>(ri);
    }
  }
  const double prr = block_sum<kStream>(rr, s_red);
  if (threadIdx.x == 0) parts.rr[(k & 1) * parts.p2 + blockIdx.x] = prr;
}

// r = scale * source, x = p = q = 0, r.r into the iteration-0 partials; the
// last block resets the bookkeeping and sets the threshold (tol ||r||)^2.
__global__ __launch_bounds__(kStream) void k_fcg_init(const double* __restrict__ source,
                                                      double scale, float* __restrict__ r,
                                                      float* __restrict__ xsol,
                                                      float* __restrict__ p_old,
                                                      float* __restrict__ q, size_t n,
                                                      double tolerance, CgState* st,
                                                      FcgPartials parts, double* partials) {
  __shared__ double s_red[kStream / kWave];
  __shared__ bool s_last;
  double rr = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kStream;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kStream + threadIdx.x; i < n; i += stride) {
    const float ri = static_cast<float>(source[i] * scale);
    r[i] = ri;
    xsol[i] = 0.0f;
    p_old[i] = 0.0f;
    q[i] = 0.0f;
    rr += static_cast<double>(ri) * static_cast<double>(ri);
  }
  double part[1] = {block_sum<kStream>(rr, s_red)};
  if (threadIdx.x == 0) parts.rr[blockIdx.x] = part[0];
  if (publish_partials<1>(part, partials, &st->ticket[kTicketInit], blockIdx.x, gridDim.x,
                          &s_last)) {
    double total[1];
    sum_partials<kStream, 1>(partials, gridDim.x, total, s_red);
    if (threadIdx.x == 0) {
      st->ticket[kTicketInit] = 0;
      st->rr0 = total[0];
      st->rr = total[0];
      const double bound = tolerance * sqrt(total[0]);
      st->threshold = bound * bound;
      st->alpha = 0.0;
      st->beta = 0.0;
      st->failure = 0;
      st->it_update = 0;
      st->it_apply = 0;
      st->halt = 0;
      st->converged = total[0] == 0.0 ? 1u : 0u;
      st->done = total[0] == 0.0 ? 1u : 0u;
    }
  }
}

__global__ void k_mg_refine(double* __restrict__ xsol, const float* __restrict__ e, double scale,
                            size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    xsol[i] += scale * static_cast<double>(e[i]);
}

// ----------------------------------------------------------------------------
// Host side
// ----------------------------------------------------------------------------

enum CycleMode { kCyclePlain, kCycleLearned, kCyclePreSmooth };

// Residual snapshots taken by a flexible-CG solve (one per iteration).
struct Snapshots {
  float* dst;
  uint32_t cap;
  uint32_t count;
};

// First level of the single-workgroup kernel: the finest level stays on the
// tile kernels when split_finest is set (the learned smoother lives there).
uint32_t active_tail(const MgDevice& mg) {
  if (mg.split_finest && mg.tail == 0 && mg.levels.size() > 1) return 1;
  return mg.tail;
}

LevelView view(const MgLevel& L) {
  return LevelView{L.wx.data(), L.wy.data(), L.bc.data(), L.dinv.data(), L.nx, L.ny};
}

dim3 tile_grid(const MgLevel& L) {
  return dim3((L.nx + kTX - 1) / kTX, (L.ny + kTY - 1) / kTY);
}

// Workgroups of the persistent up pass: about eight per CU, or one per tile
// on small levels.
uint32_t up_blocks(const MgLevel& L) {
  const dim3 g = tile_grid(L);
  return std::min<uint32_t>(g.x * g.y, 2048);
}

uint32_t fcg_rows(uint32_t nx, uint32_t ny, int ty, uint32_t* blocks) {
  const uint32_t nty = (ny + ty - 1) / ty;
  uint32_t rows = 32;
  while (rows > 4 && static_cast<uint64_t>(nty) * ((nx + rows - 1) / rows) < 1024) rows /= 2;
  *blocks = nty * ((nx + rows - 1) / rows);
  return rows;
}

int fcg_ty(uint32_t ny) { return ny <= 64 ? 64 : (ny <= 128 ? 128 : 256); }

// Inverse of the coarsest operator by Cholesky on its active cells (those
// with a nonzero diagonal); rows and columns of inactive cells stay zero. A
// pivot that vanishes (a fluid pocket with no open boundary) leaves its row
// and column zero as well: the residual is zero there.
std::vector<float> coarse_inverse(const std::vector<float>& wx, const std::vector<float>& wy,
                                  const std::vector<float>& bc, uint32_t nx, uint32_t ny) {
  const uint32_t n = nx * ny;
  std::vector<double> A(static_cast<size_t>(n) * n, 0.0);
  for (uint32_t x = 0; x < nx; ++x) {
    for (uint32_t y = 0; y < ny; ++y) {
      const uint32_t i = x * ny + y;
      A[static_cast<size_t>(i) * n + i] += bc[i];
      if (x > 0) {
        const uint32_t j = i - ny;
        const double w = wx[i];
        A[static_cast<size_t>(i) * n + i] += w;
        A[static_cast<size_t>(j) * n + j] += w;
        A[static_cast<size_t>(i) * n + j] -= w;
        A[static_cast<size_t>(j) * n + i] -= w;
      }
      if (y > 0) {
        const uint32_t j = i - 1;
        const double w = wy[i];
        A[static_cast<size_t>(i) * n + i] += w;
        A[static_cast<size_t>(j) * n + j] += w;
        A[static_cast<size_t>(i) * n + j] -= w;
        A[static_cast<size_t>(j) * n + i] -= w;
      }
    }
  }
  double scale = 0.0;
  for (uint32_t i = 0; i < n; ++i) scale = std::max(scale, A[static_cast<size_t>(i) * n + i]);
  std::vector<uint32_t> act;
  for (uint32_t i = 0; i < n; ++i)
    if (A[static_cast<size_t>(i) * n + i] > 0.0) act.push_back(i);
  const size_t m = act.size();
  std::vector<double> L(m * m, 0.0);
  std::vector<uint8_t> dead(m, 0);
  for (size_t j = 0; j < m; ++j) {
    double d = A[static_cast<size_t>(act[j]) * n + act[j]];
    for (size_t k = 0; k < j; ++k) d -= L[j * m + k] * L[j * m + k];
    if (d <= 1e-12 * scale) {
      dead[j] = 1;
      continue;
    }
    const double ljj = std::sqrt(d);
    L[j * m + j] = ljj;
    for (size_t i = j + 1; i < m; ++i) {
      double v = A[static_cast<size_t>(act[i]) * n + act[j]];
      for (size_t k = 0; k < j; ++k) v -= L[i * m + k] * L[j * m + k];
      L[i * m + j] = v / ljj;
    }
  }
  // Columns of the inverse: solve L L^T c = e_j.
  std::vector<float> inv(static_cast<size_t>(n) * n, 0.0f);
  std::vector<double> col(m);
  for (size_t j = 0; j < m; ++j) {
    if (dead[j]) continue;
    std::fill(col.begin(), col.end(), 0.0);
    col[j] = 1.0;
    for (size_t i = 0; i < m; ++i) {
      if (dead[i]) {
        col[i] = 0.0;
        continue;
      }
      double v = col[i];
      for (size_t k = 0; k < i; ++k) v -= L[i * m + k] * col[k];
      col[i] = v / L[i * m + i];
    }
    for (size_t ii = m; ii-- > 0;) {
      if (dead[ii]) {
        col[ii] = 0.0;
        continue;
      }
      double v = col[ii];
      for (size_t k = ii + 1; k < m; ++k) v -= L[k * m + ii] * col[k];
      col[ii] = v / L[ii * m + ii];
    }
    for (size_t i = 0; i < m; ++i)
      inv[static_cast<size_t>(act[i]) * n + act[j]] = static_cast<float>(col[i]);
  }
  return inv;
}

TailArgs tail_args(Context& ctx, float* r0, float* z0, const float* q0, FcgPartials parts) {
  auto& mg = ctx.flow.mg;
  const uint32_t tail = active_tail(mg);
  TailArgs T{};
  T.count = static_cast<uint32_t>(mg.levels.size()) - tail;
  for (uint32_t l = 0; l < T.count; ++l) {
    MgLevel& L = *mg.levels[tail + l];
    TailLevel& t = T.lv[l];
    t.wx = L.wx.data();
    t.wy = L.wy.data();
    t.bc = L.bc.data();
    t.dinv = L.dinv.data();
    t.r = L.r.data();
    t.e = L.e.data();
    t.xp = L.xp.data();
    t.t = L.t.data();
    t.nx = L.nx;
    t.ny = L.ny;
  }
  if (tail == 0) {
    T.lv[0].r = r0;
    T.lv[0].e = z0;
    if (q0 != nullptr) T.dots = UpDots{q0, parts.zr, parts.zq};
  }
  T.cinv = mg.cinv.data();
  T.n = mg.coarse_n;
  return T;
}

FcgPartials fcg_partials(Context& ctx) {
  auto& f = ctx.flow;
  auto& mg = f.mg;
  FcgPartials p{};
  p.pq = mg.pq.data();
  p.rr = mg.rr.data();
  p.zr = mg.zr.data();
  p.zq = mg.zq.data();
  if (mg.learned_cycle && active_tail(mg) > 0) {
    p.p0 = learned_blocks(mg.levels[0]->nx, mg.levels[0]->ny);
  } else if (active_tail(mg) == 0) {
    p.p0 = 1;
  } else {
    p.p0 = up_blocks(*mg.levels[0]);
  }
  fcg_rows(f.nx, f.ny, fcg_ty(f.ny), &p.p1);
  p.p2 = krylov_blocks(f.sites());
  return p;
}

// z = M(r): the cycle over all levels. With probes, events are recorded after
// the finest down pass (probes[0]) and before the finest up pass (probes[1]).
// kCycleLearned post-smooths the finest level with the learned smoother;
// kCyclePreSmooth stops the finest level before its post-smoother.
void launch_cycle(Context& ctx, float* r, float* z, const float* q, const hipEvent_t* probes,
                  CycleMode mode) {
  auto& mg = ctx.flow.mg;
  const uint32_t tail = active_tail(mg);
  const CgState* st = ctx.flow.state.data();
  const FcgPartials parts = fcg_partials(ctx);
  // Down passes on the large levels; level 0 reads the flexible-CG residual.
  for (uint32_t l = 0; l < tail; ++l) {
    MgLevel& L = *mg.levels[l];
    MgLevel& C = *mg.levels[l + 1];
    const float* rl = l == 0 ? r : L.r.data();
    const dim3 g = tile_grid(L);
    k_mg_down<<<g.x * g.y, kMgThreads, 0, ctx.stream>>>(view(L), rl, L.xp.data(), C.r.data(),
                                                        C.nx, C.ny, st, g.y);
    check(hipGetLastError(), "launch k_mg_down");
    if (l == 0 && probes != nullptr) check(hipEventRecord(probes[0], ctx.stream), "event");
  }
  if (tail == 0 && probes != nullptr) check(hipEventRecord(probes[0], ctx.stream), "event");
  k_mg_tail<<<1, kTailThreads, 0, ctx.stream>>>(tail_args(ctx, r, z, q, parts), st);
  check(hipGetLastError(), "launch k_mg_tail");
  if (tail == 0 && probes != nullptr) check(hipEventRecord(probes[1], ctx.stream), "event");
  for (int l = static_cast<int>(tail) - 1; l >= 0; --l) {
    MgLevel& L = *mg.levels[l];
    MgLevel& C = *mg.levels[l + 1];
    const float* rl = l == 0 ? r : L.r.data();
    float* out = l == 0 ? z : L.e.data();
    const UpDots dots = (l == 0 && q != nullptr) ? UpDots{q, parts.zr, parts.zq}
                                                 : UpDots{nullptr, nullptr, nullptr};
    if (l == 0 && probes != nullptr) check(hipEventRecord(probes[1], ctx.stream), "event");
    if (l == 0 && mode == kCycleLearned) {
      launch_up_learned(ctx, L, rl, L.xp.data(), C.e.data(), C.ny, out, st, dots.q, dots.zr,
                        dots.zq);
      continue;
    }
    const int pre = (l == 0 && mode == kCyclePreSmooth) ? 1 : 0;
    const dim3 g = tile_grid(L);
    k_mg_up<<<up_blocks(L), kMgThreads, 0, ctx.stream>>>(view(L), rl, L.xp.data(), C.e.data(),
                                                         C.ny, out, st, dots, pre, g.y,
                                                         g.x * g.y);
    check(hipGetLastError(), "launch k_mg_up");
  }
}

void launch_fcg_apply(Context& ctx, const float* z, const float* p_old, float* p_new, float* q,
                      float* x) {
  auto& f = ctx.flow;
  const int ty = fcg_ty(f.ny);
  uint32_t blocks = 0;
  const uint32_t rows = fcg_rows(f.nx, f.ny, ty, &blocks);
  const float isx = static_cast<float>(1.0 / (static_cast<double>(f.dx) * f.dx));
  const float isy = static_cast<float>(1.0 / (static_cast<double>(f.dy) * f.dy));
  const FcgPartials parts = fcg_partials(ctx);
  switch (ty) {
    case 64:
      k_fcg_apply<64><<<blocks, 64, 0, ctx.stream>>>(z, p_old, p_new, q, x, f.cxl.data(),
                                                     f.cyl.data(), f.bc.data(), f.nx, f.ny, rows,
                                                     isx, isy, f.state.data(), parts);
      break;
    case 128:
      k_fcg_apply<128><<<blocks, 128, 0, ctx.stream>>>(z, p_old, p_new, q, x, f.cxl.data(),
                                                       f.cyl.data(), f.bc.data(), f.nx, f.ny,
                                                       rows, isx, isy, f.state.data(), parts);
      break;
    default:
      k_fcg_apply<256><<<blocks, 256, 0, ctx.stream>>>(z, p_old, p_new, q, x, f.cxl.data(),
                                                       f.cyl.data(), f.bc.data(), f.nx, f.ny,
                                                       rows, isx, isy, f.state.data(), parts);
  }
  check(hipGetLastError(), "launch k_fcg_apply");
}

// Inner flexible-CG solve of A e = r / ||r|| to `tolerance` within `budget`
// iterations. Returns the iterations taken; *converged says whether the
// stopping test held.
uint32_t run_fcg(Context& ctx, double rnorm, double tolerance, uint32_t budget, uint32_t interval,
                 uint32_t first, bool* converged, KernelTimes* times, CycleMode mode,
                 Snapshots* snaps) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  float* r = f.r32.data();
  float* x = f.x32.data();
  float* q = f.q32.data();
  float* z = f.mg.levels[0]->e.data();
  float* buf[2] = {f.p032.data(), f.p132.data()};
  const FcgPartials parts = fcg_partials(ctx);
  k_fcg_init<<<krylov_blocks(n), kStream, 0, ctx.stream>>>(f.res.data(), 1.0 / rnorm, r, x,
                                                           buf[0], q, n, tolerance,
                                                           f.state.data(), parts,
                                                           f.partials.data());
  check(hipGetLastError(), "launch k_fcg_init");
  *converged = false;
  auto snapshot = [&]() {
    if (snaps == nullptr || snaps->count >= snaps->cap) return;
    check(hipMemcpyAsync(snaps->dst + static_cast<size_t>(snaps->count) * n, r, n * sizeof(float),
                         hipMemcpyDeviceToDevice, ctx.stream),
          "snapshot residual");
    ++snaps->count;
  };
  if (snaps != nullptr) {
    interval = 1;
    snapshot();
  }
  // The event probes cover one batch of `interval` iterations.
  if (snaps != nullptr || times != nullptr) first = interval;
  constexpr int kP = 6;  // events per iteration
  if (times != nullptr) {
    while (ctx.probes.size() < kP * static_cast<size_t>(interval) + kP) {
      hipEvent_t e;
      check(hipEventCreate(&e), "hipEventCreate");
      ctx.probes.push_back(e);
    }
  }
  read_state(ctx);
  if (f.host_state->done != 0) {
    *converged = true;
    return 0;
  }
  uint32_t launched = 0;
  bool tested_all = false;
  while (true) {
    const bool last_test = launched >= budget;
    const uint32_t batch =
        last_test ? 0 : std::min(launched == 0 ? std::max(first, interval) : interval,
                                 budget - launched);
    for (uint32_t j = 0; j < batch; ++j) {
      const uint32_t it = launched + j + 1;
      const hipEvent_t* pr = times != nullptr ? &ctx.probes[kP * j] : nullptr;
      if (pr != nullptr) check(hipEventRecord(pr[0], ctx.stream), "event");
      launch_cycle(ctx, r, z, q, pr != nullptr ? pr + 1 : nullptr, mode);
      if (pr != nullptr) check(hipEventRecord(pr[3], ctx.stream), "event");
      launch_fcg_apply(ctx, z, buf[(it + 1) % 2], buf[it % 2], q, x);
      if (pr != nullptr) check(hipEventRecord(pr[4], ctx.stream), "event");
      k_fcg_update<<<krylov_blocks(n), kStream, 0, ctx.stream>>>(r, q, n, f.state.data(), parts);
      check(hipGetLastError(), "launch k_fcg_update");
      if (pr != nullptr) check(hipEventRecord(pr[5], ctx.stream), "event");
    }
    if (last_test) {
      const uint32_t it = launched + 1;
      launch_fcg_apply(ctx, z, buf[(it + 1) % 2], buf[it % 2], q, x);
      tested_all = true;
    }
    const uint32_t before = launched;
    launched += batch;
    read_state(ctx);
    const CgState& s = *f.host_state;
    if (times != nullptr && batch > 0) {
      const uint32_t last_worked = s.halt != 0 ? s.halt : s.it_update;
      const uint32_t worked = std::min(batch, last_worked > before ? last_worked - before : 0u);
      for (uint32_t j = 0; j < worked; ++j) {
        const hipEvent_t* pr = &ctx.probes[kP * j];
        float t[5] = {};
        for (int e = 0; e < 5; ++e) {
          check(hipEventElapsedTime(&t[e], pr[e], pr[e + 1]), "event time");
        }
        times->down_seconds += t[0] * 1.0e-3;
        times->rest_seconds += t[1] * 1.0e-3;
        times->up_seconds += t[2] * 1.0e-3;
        times->apply_seconds += t[3] * 1.0e-3;
        times->update_seconds += t[4] * 1.0e-3;
        times->cycles += 1;
        times->apply_launches += 1;
        times->update_launches += 1;
      }
    }
    if (s.failure != 0) {
      throw Error(failure_status(s.failure),
                  std::string("HIP multilevel flexible conjugate gradient ") +
                      failure_text(s.failure));
    }
    if (s.halt != 0) {
      *converged = true;
      return s.halt;
    }
    if (batch > 0) snapshot();
    if (tested_all) return budget;
  }
}

}  // namespace

std::vector<float> mg_coarse_inverse(const std::vector<float>& wx, const std::vector<float>& wy,
                                     const std::vector<float>& bc, uint32_t nx, uint32_t ny) {
  return coarse_inverse(wx, wy, bc, nx, ny);
}

void mg_level0_operator(hipStream_t stream, const float* cxl, const float* cyl, const float* bc,
                        float* wx, float* wy, float* bcl, float* dinv, uint32_t nx, uint32_t ny,
                        float isx, float isy) {
  const size_t n = static_cast<size_t>(nx) * ny;
  const uint32_t blocks = static_cast<uint32_t>((n + 255) / 256);
  k_mg_level0<<<blocks, 256, 0, stream>>>(cxl, cyl, bc, wx, wy, bcl, n, isx, isy);
  check(hipGetLastError(), "launch k_mg_level0");
  k_mg_dinv<<<blocks, 256, 0, stream>>>(wx, wy, bcl, dinv, nx, ny);
  check(hipGetLastError(), "launch k_mg_dinv");
}

void mg_coarse_operator(hipStream_t stream, const float* fwx, const float* fwy, const float* fbc,
                        uint32_t fnx, uint32_t fny, float* wx, float* wy, float* bc, float* dinv,
                        uint32_t nx, uint32_t ny) {
  const size_t n = static_cast<size_t>(nx) * ny;
  const uint32_t blocks = static_cast<uint32_t>((n + 255) / 256);
  const LevelView f{fwx, fwy, fbc, nullptr, fnx, fny};
  k_mg_coarsen<<<blocks, 256, 0, stream>>>(f, wx, wy, bc, nx, ny);
  check(hipGetLastError(), "launch k_mg_coarsen");
  k_mg_dinv<<<blocks, 256, 0, stream>>>(wx, wy, bc, dinv, nx, ny);
  check(hipGetLastError(), "launch k_mg_dinv");
}

void mg_prepare(Context& ctx) {
  auto& f = ctx.flow;
  auto& mg = f.mg;
  // Levels of the same shapes keep their device buffers: a system that
  // changes only its coefficients (a new mobility on the same grid) rebuilds
  // the hierarchy without allocating.
  std::vector<std::unique_ptr<MgLevel>> kept = std::move(mg.levels);
  mg.levels.clear();
  auto take = [&](size_t l, uint32_t nx, uint32_t ny) {
    if (l < kept.size() && kept[l] && kept[l]->nx == nx && kept[l]->ny == ny) {
      return std::move(kept[l]);
    }
    auto level = std::make_unique<MgLevel>();
    level->nx = nx;
    level->ny = ny;
    return level;
  };
  auto level0 = take(0, f.nx, f.ny);
  const size_t n = f.sites();
  for (auto* b : {&level0->wx, &level0->wy, &level0->bc, &level0->dinv, &level0->r, &level0->e,
                  &level0->xp, &level0->t})
    b->reserve(n);
  const float isx = static_cast<float>(1.0 / (static_cast<double>(f.dx) * f.dx));
  const float isy = static_cast<float>(1.0 / (static_cast<double>(f.dy) * f.dy));
  const uint32_t blocks = static_cast<uint32_t>((n + 255) / 256);
  k_mg_level0<<<blocks, 256, 0, ctx.stream>>>(f.cxl.data(), f.cyl.data(), f.bc.data(),
                                              level0->wx.data(), level0->wy.data(),
                                              level0->bc.data(), n, isx, isy);
  check(hipGetLastError(), "launch k_mg_level0");
  k_mg_dinv<<<blocks, 256, 0, ctx.stream>>>(level0->wx.data(), level0->wy.data(),
                                            level0->bc.data(), level0->dinv.data(), f.nx, f.ny);
  check(hipGetLastError(), "launch k_mg_dinv");
  mg.levels.push_back(std::move(level0));
  while (mg.levels.back()->cells() > kCoarseCells && mg.levels.size() < kMaxLevels) {
    MgLevel& F = *mg.levels.back();
    auto C = take(mg.levels.size(), (F.nx + 1) / 2, (F.ny + 1) / 2);
    const size_t cn = C->cells();
    for (auto* b : {&C->wx, &C->wy, &C->bc, &C->dinv, &C->r, &C->e, &C->xp, &C->t})
      b->reserve(cn);
    const uint32_t cb = static_cast<uint32_t>((cn + 255) / 256);
    k_mg_coarsen<<<cb, 256, 0, ctx.stream>>>(view(F), C->wx.data(), C->wy.data(), C->bc.data(),
                                             C->nx, C->ny);
    check(hipGetLastError(), "launch k_mg_coarsen");
    k_mg_dinv<<<cb, 256, 0, ctx.stream>>>(C->wx.data(), C->wy.data(), C->bc.data(),
                                          C->dinv.data(), C->nx, C->ny);
    check(hipGetLastError(), "launch k_mg_dinv");
    mg.levels.push_back(std::move(C));
  }
  // The single-workgroup kernel takes every level from the first small one.
  mg.tail = static_cast<uint32_t>(mg.levels.size()) - 1;
  for (uint32_t l = 0; l < mg.levels.size(); ++l) {
    if (mg.levels[l]->cells() <= kTailCells) {
      mg.tail = l;
      break;
    }
  }
  // Dense inverse of the coarsest operator.
  MgLevel& B = *mg.levels.back();
  const size_t bn = B.cells();
  if (bn > kCoarseCells)
    throw Error(MSIM_INVALID_ARGUMENT, "multilevel: the grid needs more than 24 levels");
  std::vector<float> wx(bn), wy(bn), bc(bn);
  check(hipMemcpyAsync(wx.data(), B.wx.data(), bn * sizeof(float), hipMemcpyDeviceToHost,
                       ctx.stream),
        "download coarse operator");
  check(hipMemcpyAsync(wy.data(), B.wy.data(), bn * sizeof(float), hipMemcpyDeviceToHost,
                       ctx.stream),
        "download coarse operator");
  check(hipMemcpyAsync(bc.data(), B.bc.data(), bn * sizeof(float), hipMemcpyDeviceToHost,
                       ctx.stream),
        "download coarse operator");
  check(hipStreamSynchronize(ctx.stream), "coarse operator");
  const std::vector<float> inv = coarse_inverse(wx, wy, bc, B.nx, B.ny);
  // Stored transposed: entry (row, j) at j * n + row.
  std::vector<float> inv_t(inv.size());
  for (size_t row = 0; row < bn; ++row)
    for (size_t j = 0; j < bn; ++j) inv_t[j * bn + row] = inv[row * bn + j];
  mg.coarse_n = static_cast<uint32_t>(bn);
  mg.cinv.reserve(inv_t.size());
  check(hipMemcpyAsync(mg.cinv.data(), inv_t.data(), inv_t.size() * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload coarse inverse");
  check(hipStreamSynchronize(ctx.stream), "upload coarse inverse");
  // Flexible-CG partial sums, sized for the finest level on the tile kernels.
  const FcgPartials parts = fcg_partials(ctx);
  const size_t p0 = std::max<size_t>({static_cast<size_t>(up_blocks(*mg.levels[0])),
                                      learned_blocks(mg.levels[0]->nx, mg.levels[0]->ny), 1});
  mg.pq.reserve(2 * static_cast<size_t>(parts.p1));
  mg.rr.reserve(2 * static_cast<size_t>(parts.p2));
  mg.zr.reserve(2 * p0);
  mg.zq.reserve(2 * p0);
  check(hipStreamSynchronize(ctx.stream), "multilevel set-up");
  mg.ready = true;
}

void mg_precondition(Context& ctx, const float* r_host, float* z_host, bool learned) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  if (!f.mg.ready) mg_prepare(ctx);
  f.mg.split_finest = learned;
  f.mg.learned_cycle = learned;
  f.r32.reserve(n);
  f.p132.reserve(n);
  // A fresh state: the cycle kernels run unless a solve stopped or failed.
  check(hipMemsetAsync(f.state.data(), 0, sizeof(CgState), ctx.stream), "clear state");
  check(hipMemcpyAsync(f.r32.data(), r_host, n * sizeof(float), hipMemcpyHostToDevice,
                       ctx.stream),
        "upload residual");
  float* z = f.p132.data();
  launch_cycle(ctx, f.r32.data(), z, nullptr, nullptr, learned ? kCycleLearned : kCyclePlain);
  check(hipMemcpyAsync(z_host, z, n * sizeof(float), hipMemcpyDeviceToHost, ctx.stream),
        "download correction");
  check(hipStreamSynchronize(ctx.stream), "precondition");
  f.solved = false;
}

namespace {

uint32_t solve_refined(Context& ctx, const msim_flow_params& params, msim_flow_report& report,
                       KernelTimes* times, CycleMode mode, bool* converged, Snapshots* snaps) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  if (!f.mg.ready) mg_prepare(ctx);
  f.mg.split_finest = mode != kCyclePlain;
  f.mg.learned_cycle = mode == kCycleLearned;
  for (auto* b : {&f.x32, &f.r32, &f.q32, &f.p032, &f.p132}) b->reserve(n);
  const double inner = params.inner_tolerance > 0.0 ? params.inner_tolerance : 1.0e-5;
  const uint32_t interval = params.check_interval == 0 ? 8 : params.check_interval;
  check(hipMemsetAsync(f.x.data(), 0, n * sizeof(double), ctx.stream), "clear x");
  launch_residual(ctx, f.x.data(), f.res.data(), 1);
  read_state(ctx);
  const double bnorm = std::sqrt(f.host_state->bb);
  double rnorm = std::sqrt(f.host_state->true_rr);
  uint32_t iterations = 0, refinements = 0;
  bool done = bnorm == 0.0;
  while (!done) {
    if (rnorm <= params.relative_tolerance * bnorm) {
      done = true;
      break;
    }
    if (iterations >= params.max_iterations || refinements >= 64) break;
    bool inner_converged = false;
    uint32_t& expect = f.mg.fcg_expect[mode][std::min(refinements, 7u)];
    const uint32_t taken = run_fcg(ctx, rnorm, inner, params.max_iterations - iterations,
                                   interval, expect > 0 ? expect + 1 : interval,
                                   &inner_converged, times, mode, snaps);
    expect = inner_converged ? taken : 0;
    iterations += taken;
    ++refinements;
    k_mg_refine<<<stream_blocks(n), kStream, 0, ctx.stream>>>(f.x.data(), f.x32.data(), rnorm,
                                                              n);
    check(hipGetLastError(), "launch k_mg_refine");
    launch_residual(ctx, f.x.data(), f.res.data(), 1);
    read_state(ctx);
    const double next = std::sqrt(f.host_state->true_rr);
    if (!std::isfinite(next)) {
      throw Error(MSIM_CG_RESIDUAL, "HIP multilevel refinement produced a non-finite residual");
    }
    rnorm = next;
    if (!inner_converged && taken == 0) break;
  }
  report.refinements = refinements;
  report.relative_residual = bnorm > 0.0 ? rnorm / bnorm : 0.0;
  *converged = done;
  return iterations;
}

}  // namespace

uint32_t mg_solve(Context& ctx, const msim_flow_params& params, msim_flow_report& report,
                  KernelTimes* times, bool learned, bool* converged) {
  return solve_refined(ctx, params, report, times, learned ? kCycleLearned : kCyclePlain,
                       converged, nullptr);
}

uint32_t mg_harvest(Context& ctx, uint32_t max_pairs, bool learned, float* residuals,
                    float* errors) {
  auto& f = ctx.flow;
  auto& mg = f.mg;
  const size_t n = f.sites();
  if (!mg.ready) mg_prepare(ctx);
  if (mg.levels.size() < 2) {
    throw Error(MSIM_INVALID_ARGUMENT,
                "the grid is too small to harvest from: its hierarchy has a single level");
  }
  // Up to kSpread residuals per pair kept, so pairs cover the whole solve.
  constexpr uint32_t kSpread = 4;
  DeviceBuffer<float> snaps, saved, rp, ep;
  snaps.reserve(static_cast<size_t>(kSpread) * max_pairs * n);
  saved.reserve(n);
  rp.reserve(n);
  ep.reserve(n);
  check(hipMemcpyAsync(saved.data(), f.rhs.data(), n * sizeof(float), hipMemcpyDeviceToDevice,
                       ctx.stream),
        "save right-hand side");
  uint32_t count = 0;
  try {
    // Residuals of the solver's own iterations.
    msim_flow_params p{};
    p.relative_tolerance = 1.0e-8;
    p.max_iterations = 1000;
    msim_flow_report rep{};
    Snapshots sn{snaps.data(), kSpread * max_pairs, 0};
    bool ok = false;
    solve_refined(ctx, p, rep, nullptr, learned ? kCycleLearned : kCyclePlain, &ok, &sn);
    // Residuals spread evenly over the whole solve.
    count = std::min(sn.count, max_pairs);
    std::vector<uint32_t> pick(count);
    for (uint32_t k = 0; k < count; ++k) {
      pick[k] = count > 1 ? static_cast<uint32_t>(
                                (static_cast<uint64_t>(k) * (sn.count - 1) + (count - 1) / 2) /
                                (count - 1))
                          : 0;
    }
    // For each: e = A^-1 r, then the pair after the pre-smoother and the
    // coarse correction.
    p.relative_tolerance = 1.0e-10;
    p.max_iterations = 4000;
    for (uint32_t k = 0; k < count; ++k) {
      float* rk = snaps.data() + static_cast<size_t>(pick[k]) * n;
      check(hipMemcpyAsync(f.rhs.data(), rk, n * sizeof(float), hipMemcpyDeviceToDevice,
                           ctx.stream),
            "residual as right-hand side");
      solve_refined(ctx, p, rep, nullptr, kCyclePlain, &ok, nullptr);
      if (!ok) throw Error(MSIM_NOT_CONVERGED, "harvest: the error solve did not converge");
      mg.split_finest = true;
      mg.learned_cycle = false;
      check(hipMemsetAsync(f.state.data(), 0, sizeof(CgState), ctx.stream), "clear state");
      launch_cycle(ctx, rk, f.p132.data(), nullptr, nullptr, kCyclePreSmooth);
      learned_make_pairs(ctx, *mg.levels[0], rk, f.x.data(), f.p132.data(), rp.data(), ep.data());
      check(hipMemcpyAsync(residuals + static_cast<size_t>(k) * n, rp.data(), n * sizeof(float),
                           hipMemcpyDeviceToHost, ctx.stream),
            "download pair");
      check(hipMemcpyAsync(errors + static_cast<size_t>(k) * n, ep.data(), n * sizeof(float),
                           hipMemcpyDeviceToHost, ctx.stream),
            "download pair");
    }
  } catch (...) {
    static_cast<void>(hipMemcpyAsync(f.rhs.data(), saved.data(), n * sizeof(float),
                                     hipMemcpyDeviceToDevice, ctx.stream));
    static_cast<void>(hipStreamSynchronize(ctx.stream));
    f.solved = false;
    throw;
  }
  check(hipMemcpyAsync(f.rhs.data(), saved.data(), n * sizeof(float), hipMemcpyDeviceToDevice,
                       ctx.stream),
        "restore right-hand side");
  check(hipStreamSynchronize(ctx.stream), "harvest");
  f.solved = false;
  return count;
}

void mg_operator(Context& ctx, float* wx, float* wy, float* bc, float* dinv) {
  auto& f = ctx.flow;
  if (!f.mg.ready) mg_prepare(ctx);
  const MgLevel& L = *f.mg.levels[0];
  const size_t bytes = L.cells() * sizeof(float);
  const std::pair<float*, const float*> copies[4] = {
      {wx, L.wx.data()}, {wy, L.wy.data()}, {bc, L.bc.data()}, {dinv, L.dinv.data()}};
  for (const auto& c : copies) {
    if (c.first != nullptr) {
      check(hipMemcpyAsync(c.first, c.second, bytes, hipMemcpyDeviceToHost, ctx.stream),
            "download operator");
    }
  }
  check(hipStreamSynchronize(ctx.stream), "operator");
}

}  // namespace msim
