// Batched sweep: many independent depth-averaged devices of uneven sizes
// (a parameter sweep), each with its own geometry, mobility, occupancy,
// inlet pressure and solute diffusivity, advanced together. No counterpart
// in the reference, which runs one job per invocation
// (docs/architecture/0005-batch-execution.md:30); each instance follows the
// reference's single-grid algorithms:
//   pressure  -- the assembled operator (flow_system.hpp:522-695) solved by
//                FP32 Jacobi-PCG corrections (cpp/cpu/cpu_flow.cpp:65-140; FP64
//                dot products) inside FP64 iterative refinement of an FP64
//                pressure, warm-started, to an FP64 true relative residual;
//   velocity  -- face velocities and lift (flow_system.hpp:711-736, 424-445)
//                for the Dirichlet inlet pressure, unscaled;
//   transport -- one FP32 forward-Euler step of one solute with Monod uptake,
//                every site in the single-grid kernel's float operations
//                (adr_site.hpp), with FP64 mass tallies.
//
// A step runs over work items (instance, block of rows) of a few thousand
// sites, spread over the whole chip, in one launch. Each item
// recomputes its rows' FP64 pressure residual (the step's flow check), writes
// the face velocities it owns, and advances its sites, so no item waits on
// another. Instances up to 256 columns wide take a marching kernel: one
// thread per column walks the item's rows, neighbours from registers and an
// LDS row, each array read once and each x face computed once; wider ones a
// general sweep. Per-item FP64 partials (residual, masses, uptake) are added
// per instance in a fixed order on the host. A run of steps is captured once
// into a HIP graph and replayed.
//
// The pressure solve itself is one launch with one workgroup per instance
// (largest first): the refinement loop needs the instance's dot products,
// and a workgroup reduces them with barriers only. Each FP32 correction is
// conjugate gradient preconditioned by the instance's own multilevel cycle
// (2x2 aggregation, V(1,1) damped Jacobi, the coarse correction weighted by
// 1.8, the coarsest level of <= 64 cells inverted densely): the cycle of the
// single-grid multilevel form, run level by level inside the workgroup. It is
// linear and symmetric positive definite, so plain CG applies.
//
// The launch shapes (sites per item; workgroup size and sweep row width of
// the general sweep) are chosen by timing each candidate on the uploaded
// batch: the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "adr_site.hpp"
#include "common.hpp"
#include "internal.hpp"
#include "stencil.hpp"

namespace msim {
namespace {

struct BatchDesc {
  uint32_t nx, ny;
  uint64_t site, xface, yface;  // first site / x face / y face of the instance
  float dx, dy, dz;
  float inlet;      // pressure at x = 0 (0 at x = nx)
  float diffusion;  // of the solute
  float reservoir;  // solute level beyond x = 0 (0 beyond x = nx)
};

struct BatchItem {
  uint32_t inst, x0, x1, pad;
};

// The multilevel hierarchy of one instance: level l has nx[l] x ny[l] cells
// at offset off[l] of the level arrays (level 0 at the instance's own first
// site); the coarsest level's dense inverse is coarse_n x coarse_n floats, transposed, at
// offset cinv of the inverse pool.
constexpr uint32_t kBatchLevels = 16;
constexpr uint32_t kBatchCoarse = 64;
struct BatchLevels {
  uint32_t count, coarse_n;
  uint64_t cinv;
  uint64_t off[kBatchLevels];
  uint32_t nx[kBatchLevels], ny[kBatchLevels];
};

struct BatchData {
  const BatchDesc* desc;
  const uint32_t* order;
  const uint8_t* solid;
  const float* cond;
  const float* cxl;
  const float* cyl;
  const float* bc;
  const float* rhs;
  const float* phi;
  double* x64;  // pressure
  float* x;     // FP32 correction of one refinement
  float* r;
  float* p;
  float* q;
  float* xf;
  float* yf;
  float* la;
  float* lb;
  uint32_t* iterations;  // [instance]: FP32 CG iterations of the last flow solve
  float* residual;       // [instance]: FP64 true relative residual after it
  double* bnorm2;        // [instance]: ||b||^2
  uint32_t* flags;       // [instance]: 1 CG breakdown, 2 not converged, 4 bad level
  float* z;              // preconditioned residual
  const BatchLevels* levels;
  const float* lwx;      // level couplings to the lower x / y cell (over h^2),
  const float* lwy;      // boundary term and inverse diagonal
  const float* lbc;
  const float* ldinv;
  float* lr;             // level residual and correction (coarse levels)
  float* le;
  const float* cinv;     // dense inverses of the coarsest levels, each transposed
};

struct BatchParams {
  float tolerance;
  uint32_t max_iterations;
  float vmax, monod_k;
};

// FP64 sums over the workgroup, broadcast to every thread.
template <int NT, int K>
__device__ inline void block_sums(double (&v)[K], double* lds) {
  __shared__ double s_out[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double t = block_sum<NT>(v[k], lds);
    if (threadIdx.x == 0) s_out[k] = t;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = s_out[k];
  __syncthreads();
}

// The operator (flow_system.hpp:670-695) on an FP32 vector at site i.
__device__ inline float apply_f32(const float* __restrict__ v, const float* __restrict__ cxl,
                                  const float* __restrict__ cyl, const float* __restrict__ bc,
                                  uint32_t x, uint32_t y, uint32_t nx, uint32_t ny, size_t i,
                                  float isx, float isy) {
  const float c = v[i];
  float a = bc[i] * c;
  if (x > 0) a += (cxl[i] * isx) * (c - v[i - ny]);
  if (x + 1 < nx) a += (cxl[i + ny] * isx) * (c - v[i + ny]);
  if (y > 0) a += (cyl[i] * isy) * (c - v[i - 1]);
  if (y + 1 < ny) a += (cyl[i + 1] * isy) * (c - v[i + 1]);
  return a;
}

// b - A x in FP64 at site i, b = inlet * (the assembled Dirichlet-1 rhs).
__device__ inline double residual_f64(const double* __restrict__ x, const float* __restrict__ cxl,
                                      const float* __restrict__ cyl, const float* __restrict__ bc,
                                      const float* __restrict__ rhs, float inlet, uint32_t xx,
                                      uint32_t y, uint32_t nx, uint32_t ny, size_t i, double isx,
                                      double isy) {
  const double c = x[i];
  double a = static_cast<double>(bc[i]) * c;
  if (xx > 0) a += (static_cast<double>(cxl[i]) * isx) * (c - x[i - ny]);
  if (xx + 1 < nx) a += (static_cast<double>(cxl[i + ny]) * isx) * (c - x[i + ny]);
  if (y > 0) a += (static_cast<double>(cyl[i]) * isy) * (c - x[i - 1]);
  if (y + 1 < ny) a += (static_cast<double>(cyl[i + 1]) * isy) * (c - x[i + 1]);
  return static_cast<double>(inlet) * static_cast<double>(rhs[i]) - a;
}

// Face velocities (flow_system.hpp:711-736 with the lift of 424-445) from the
// FP64 pressure: x face fx (0..nx) of row y, y face fy (0..ny) of column x.
__device__ inline float face_x(const BatchDesc& I, const double* __restrict__ p,
                               const float* __restrict__ cond, const uint8_t* __restrict__ solid,
                               uint32_t fx, uint32_t y) {
  const uint32_t nx = I.nx, ny = I.ny;
  const bool has_lower = fx > 0, has_upper = fx < nx;
  const size_t upper = static_cast<size_t>(fx) * ny + y;
  const size_t lower = upper - ny;
  double value;
  if (has_lower && has_upper) {
    const float fm = harmonic(cond[lower], cond[upper]);
    value = -static_cast<double>(fm) * (p[upper] - p[lower]) / static_cast<double>(I.dx);
  } else {
    const size_t site = has_upper ? upper : lower;
    const float m = cond[site];
    float fc = 0.0f, fp = 0.0f;
    if (m != 0.0f) {
      fc = (has_upper ? 2.0f : -2.0f) * m / I.dx;
      fp = has_upper ? I.inlet : 0.0f;
    }
    value = static_cast<double>(fc) * (static_cast<double>(fp) - p[site]);
  }
  const float flux = static_cast<float>(value);
  if ((has_lower && solid[lower] != 0) || (has_upper && solid[upper] != 0)) return 0.0f;
  return flux / I.dz;
}

__device__ inline float face_y(const BatchDesc& I, const double* __restrict__ p,
                               const float* __restrict__ cond, const uint8_t* __restrict__ solid,
                               uint32_t x, uint32_t fy) {
  const uint32_t ny = I.ny;
  const bool has_lower = fy > 0, has_upper = fy < ny;
  const size_t upper = static_cast<size_t>(x) * ny + fy;
  const size_t lower = upper - 1;
  double value;
  if (has_lower && has_upper) {
    const float fm = harmonic(cond[lower], cond[upper]);
    value = -static_cast<double>(fm) * (p[upper] - p[lower]) / static_cast<double>(I.dy);
  } else {
    const size_t site = has_upper ? upper : lower;
    value = static_cast<double>(0.0f) * (static_cast<double>(0.0f) - p[site]);
  }
  const float flux = static_cast<float>(value);
  if ((has_lower && solid[lower] != 0) || (has_upper && solid[upper] != 0)) return 0.0f;
  return flux / I.dz;
}

// One level of an instance's hierarchy as the cycle sees it.
struct LevelRef {
  const float* wx;
  const float* wy;
  const float* bc;
  const float* dinv;
  float* r;
  float* e;
  uint32_t nx, ny;
};

__device__ inline LevelRef level_ref(const BatchData& D, const BatchLevels& V, uint32_t l,
                                     size_t site) {
  const size_t o = V.off[l];
  return LevelRef{D.lwx + o,
                  D.lwy + o,
                  D.lbc + o,
                  D.ldinv + o,
                  l == 0 ? D.r + site : D.lr + o,
                  l == 0 ? D.z + site : D.le + o,
                  V.nx[l],
                  V.ny[l]};
}

constexpr float kCycleOmega = 0.8f;
constexpr float kCycleCorrection = 1.8f;

// (A x)_i on a level, x at cell (gx, gy) given by xat, xc its value at i.
template <class X>
__device__ inline float level_ax(const LevelRef& L, size_t i, uint32_t x, uint32_t y, float xc,
                                 X&& xat) {
  float a = L.bc[i] * xc;
  if (x > 0) a += L.wx[i] * (xc - xat(x - 1, y));
  if (x + 1 < L.nx) a += L.wx[i + L.ny] * (xc - xat(x + 1, y));
  if (y > 0) a += L.wy[i] * (xc - xat(x, y - 1));
  if (y + 1 < L.ny) a += L.wy[i + 1] * (xc - xat(x, y + 1));
  return a;
}

// z = M r on one instance, by the whole workgroup: per level a down pass
// (pre-smoothing from zero, x = omega D^-1 r, its residual restricted to the
// next level), the coarsest level's dense inverse, then per level an up pass
// (x + 1.8 P e_c, post-smoothed). Returns r.z (FP64) in every thread.
template <int NT>
__device__ double batch_cycle(const BatchData& D, const BatchLevels& V, size_t site, double* lds) {
  const uint32_t K = V.count - 1;
  for (uint32_t l = 0; l < K; ++l) {
    const LevelRef F = level_ref(D, V, l, site);
    const LevelRef C = level_ref(D, V, l + 1, site);
    auto pre = [&](uint32_t gx, uint32_t gy) {
      const size_t j = static_cast<size_t>(gx) * F.ny + gy;
      return kCycleOmega * F.dinv[j] * F.r[j];
    };
    const size_t cells = static_cast<size_t>(C.nx) * C.ny;
    for (size_t c = threadIdx.x; c < cells; c += NT) {
      const uint32_t X = static_cast<uint32_t>(c / C.ny), Y = static_cast<uint32_t>(c % C.ny);
      float sum = 0.0f;
      for (uint32_t a = 0; a < 2; ++a) {
        for (uint32_t b = 0; b < 2; ++b) {
          const uint32_t fx = 2 * X + a, fy = 2 * Y + b;
          if (fx >= F.nx || fy >= F.ny) continue;
          const size_t i = static_cast<size_t>(fx) * F.ny + fy;
          const float xc = kCycleOmega * F.dinv[i] * F.r[i];
          sum += F.r[i] - level_ax(F, i, fx, fy, xc, pre);
        }
      }
      C.r[c] = sum;
    }
    __syncthreads();
  }
  double rz = 0.0;
  {
    const LevelRef B = level_ref(D, V, K, site);
    // The inverse is stored transposed; each row's columns are split into
    // NT / n contiguous ranges, one thread each (the rows of a wave adjacent,
    // so a step reads the inverse coalesced), added in range order.
    static_assert(NT >= static_cast<int>(kBatchCoarse), "one thread per row at least");
    __shared__ float s_r[kBatchCoarse];
    __shared__ float s_part[NT];
    const uint32_t n = V.coarse_n;
    const float* inv = D.cinv + V.cinv;
    for (uint32_t j = threadIdx.x; j < n; j += NT) s_r[j] = B.r[j];
    __syncthreads();
    const uint32_t parts = NT / n;
    const uint32_t span = (n + parts - 1) / parts;
    if (threadIdx.x < parts * n) {
      const uint32_t row = threadIdx.x % n, part = threadIdx.x / n;
      const uint32_t j1 = min(n, part * span + span);
      float v = 0.0f;
      for (uint32_t j = part * span; j < j1; ++j) v += inv[static_cast<size_t>(j) * n + row] * s_r[j];
      s_part[part * n + row] = v;
    }
    __syncthreads();
    for (uint32_t row = threadIdx.x; row < n; row += NT) {
      float v = 0.0f;
      for (uint32_t part = 0; part < parts; ++part) v += s_part[part * n + row];
      B.e[row] = v;
      if (K == 0) rz += static_cast<double>(v) * static_cast<double>(s_r[row]);
    }
    __syncthreads();
  }
  for (int l = static_cast<int>(K) - 1; l >= 0; --l) {
    const LevelRef F = level_ref(D, V, static_cast<uint32_t>(l), site);
    const LevelRef C = level_ref(D, V, static_cast<uint32_t>(l) + 1, site);
    auto corrected = [&](uint32_t gx, uint32_t gy) {
      const size_t j = static_cast<size_t>(gx) * F.ny + gy;
      return kCycleOmega * F.dinv[j] * F.r[j] +
             kCycleCorrection * C.e[static_cast<size_t>(gx / 2) * C.ny + gy / 2];
    };
    const size_t cells = static_cast<size_t>(F.nx) * F.ny;
    for (size_t i = threadIdx.x; i < cells; i += NT) {
      const uint32_t fx = static_cast<uint32_t>(i / F.ny), fy = static_cast<uint32_t>(i % F.ny);
      const float ri = F.r[i];
      const float xc = corrected(fx, fy);
      const float z = xc + kCycleOmega * F.dinv[i] * (ri - level_ax(F, i, fx, fy, xc, corrected));
      F.e[i] = z;
      if (l == 0) rz += static_cast<double>(z) * static_cast<double>(ri);
    }
    __syncthreads();
  }
  double out[1] = {rz};
  block_sums<NT, 1>(out, lds);
  return out[0];
}

// The pressure solve: one workgroup per instance (order[b], largest first),
// FP64 refinement around FP32 corrections by conjugate gradient with the
// instance's multilevel cycle, then every face velocity of the instance.
template <int NT, int TY>
__global__ __launch_bounds__(NT) void k_batch_flow(BatchData D, BatchParams P) {
  __shared__ double s_red[NT / kWave];
  constexpr int RX = NT / TY;
  const uint32_t inst = D.order[blockIdx.x];
  const BatchDesc I = D.desc[inst];
  const uint32_t nx = I.nx, ny = I.ny;
  const size_t o = I.site;
  const uint32_t tx = threadIdx.x / TY, ty = threadIdx.x % TY;
  const uint8_t* solid = D.solid + o;
  const float* cond = D.cond + o;
  const float* cxl = D.cxl + o;
  const float* cyl = D.cyl + o;
  const float* bc = D.bc + o;
  const float* rhs = D.rhs + o;
  double* x64 = D.x64 + o;
  float* xv = D.x + o;
  float* r = D.r + o;
  float* p = D.p + o;
  float* q = D.q + o;
  const float isx = 1.0f / (I.dx * I.dx), isy = 1.0f / (I.dy * I.dy);
  auto sweep = [&](auto&& f) {
    for (uint32_t x0 = 0; x0 < nx; x0 += RX) {
      const uint32_t x = x0 + tx;
      if (x >= nx) continue;
      for (uint32_t y = ty; y < ny; y += TY) f(x, y, static_cast<size_t>(x) * ny + y);
    }
  };
  constexpr uint32_t kRefinements = 16;
  constexpr double kInner = 1.0e-4;  // relative tolerance of each FP32 correction
  const BatchLevels& V = D.levels[inst];
  float* z = D.z + o;
  uint32_t it = 0, failed = 0;
  double rr64 = 0.0, bb = 0.0, threshold = 0.0;
  for (uint32_t ref = 0; ref <= kRefinements; ++ref) {
    // FP64 true residual; its FP32 copy starts the correction e = 0.
    double s2[2] = {0.0, 0.0};  // r.r (FP64), b.b
    sweep([&](uint32_t x, uint32_t y, size_t i) {
      const double res = residual_f64(x64, cxl, cyl, bc, rhs, I.inlet, x, y, nx, ny, i, isx, isy);
      const double b = static_cast<double>(I.inlet) * static_cast<double>(rhs[i]);
      r[i] = static_cast<float>(res);
      xv[i] = 0.0f;
      s2[0] += res * res;
      s2[1] += b * b;
    });
    block_sums<NT, 2>(s2, s_red);
    rr64 = s2[0];
    bb = s2[1];
    const double bound = static_cast<double>(P.tolerance) * sqrt(bb);
    threshold = bound * bound;
    if (rr64 <= threshold || ref == kRefinements || it >= P.max_iterations) break;
    double rz = batch_cycle<NT>(D, V, o, s_red);
    if (!(rz > 0.0) || !isfinite(rz)) {
      failed = 1;
      break;
    }
    sweep([&](uint32_t x, uint32_t y, size_t i) { p[i] = z[i]; });
    __syncthreads();
    double rr = rr64;
    const double inner = kInner * kInner * rr64;
    while (rr > inner && it < P.max_iterations) {
      double s1[1] = {0.0};
      sweep([&](uint32_t x, uint32_t y, size_t i) {
        const float qi = apply_f32(p, cxl, cyl, bc, x, y, nx, ny, i, isx, isy);
        q[i] = qi;
        s1[0] += static_cast<double>(p[i]) * static_cast<double>(qi);
      });
      block_sums<NT, 1>(s1, s_red);
      if (!(s1[0] > 0.0) || !isfinite(s1[0])) {
        failed = 1;
        break;
      }
      const float alpha = static_cast<float>(rz / s1[0]);
      double s3[1] = {0.0};
      sweep([&](uint32_t x, uint32_t y, size_t i) {
        xv[i] = xv[i] + alpha * p[i];
        const float ri = r[i] - alpha * q[i];
        r[i] = ri;
        s3[0] += static_cast<double>(ri) * static_cast<double>(ri);
      });
      block_sums<NT, 1>(s3, s_red);
      rr = s3[0];
      ++it;
      if (rr <= inner) break;
      const double rz_next = batch_cycle<NT>(D, V, o, s_red);
      if (!(rz_next > 0.0) || !isfinite(rz_next)) {
        failed = 1;
        break;
      }
      const float beta = static_cast<float>(rz_next / rz);
      rz = rz_next;
      sweep([&](uint32_t x, uint32_t y, size_t i) { p[i] = z[i] + beta * p[i]; });
      __syncthreads();
    }
    if (failed != 0) break;
    sweep([&](uint32_t x, uint32_t y, size_t i) { x64[i] += static_cast<double>(xv[i]); });
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    D.iterations[inst] = it;
    D.residual[inst] = bb > 0.0 ? static_cast<float>(sqrt(rr64 / bb)) : 0.0f;
    D.bnorm2[inst] = bb;
    if (failed != 0) D.flags[inst] |= 1u;
    if (rr64 > threshold && failed == 0) D.flags[inst] |= 2u;
  }
  // Face velocities.
  float* xf = D.xf + I.xface;
  float* yf = D.yf + I.yface;
  for (size_t f = threadIdx.x; f < static_cast<size_t>(nx + 1) * ny; f += NT)
    xf[f] = face_x(I, x64, cond, solid, static_cast<uint32_t>(f / ny), static_cast<uint32_t>(f % ny));
  for (size_t f = threadIdx.x; f < static_cast<size_t>(nx) * (ny + 1); f += NT)
    yf[f] = face_y(I, x64, cond, solid, static_cast<uint32_t>(f / (ny + 1)),
                   static_cast<uint32_t>(f % (ny + 1)));
}

// One step for work item blockIdx.x (rows [x0, x1) of an instance): the FP64
// residual of the rows, their face velocities, and the forward-Euler step of
// their sites. partials[(step * items + item) * 4 + k]: residual r.r, mass
// before, mass after, uptake rate (FP64, before the volume factor).
template <int NT, int TY>
__global__ __launch_bounds__(NT) void k_batch_step(BatchData D, BatchParams P,
                                                   const BatchItem* __restrict__ items,
                                                   double* __restrict__ partials, uint32_t step,
                                                   float dt, uint32_t base, uint32_t total) {
  __shared__ double s_red[NT / kWave];
  constexpr int RX = NT / TY;
  const BatchItem item = items[base + blockIdx.x];
  const uint32_t inst = item.inst;
  const BatchDesc I = D.desc[inst];
  const uint32_t nx = I.nx, ny = I.ny;
  const size_t o = I.site;
  const uint8_t* solid = D.solid + o;
  const float* cond = D.cond + o;
  const double* x64 = D.x64 + o;
  float* xf = D.xf + I.xface;
  float* yf = D.yf + I.yface;
  const float* in = ((step & 1) == 0 ? D.la : D.lb) + o;
  float* out = ((step & 1) == 0 ? D.lb : D.la) + o;
  const float* phi = D.phi != nullptr ? D.phi + o : nullptr;
  const double isx = 1.0 / (static_cast<double>(I.dx) * I.dx);
  const double isy = 1.0 / (static_cast<double>(I.dy) * I.dy);
  AdrParams A{};
  A.nx = nx;
  A.ny = ny;
  A.signals = 1;
  A.dx = I.dx;
  A.dy = I.dy;
  A.dz = I.dz;
  A.kind[0] = A.kind[1] = MSIM_FIXED;
  A.kind[2] = A.kind[3] = A.kind[4] = A.kind[5] = MSIM_NO_FLUX;
  A.has_field = 0;  // velocities are set per site below
  A.has_affine = 0;
  A.has_monod = phi != nullptr ? 1 : 0;
  A.diffusion[0] = I.diffusion;
  A.boundary[0][0] = I.reservoir;
  A.vmax[0] = P.vmax;
  A.monod_k[0] = P.monod_k;
  const uint32_t tx = threadIdx.x / TY, ty = threadIdx.x % TY;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t bad = 0;
  for (uint32_t xr = item.x0; xr < item.x1; xr += RX) {
    const uint32_t x = xr + tx;
    if (x >= item.x1) continue;
    for (uint32_t y = ty; y < ny; y += TY) {
      const size_t i = static_cast<size_t>(x) * ny + y;
      const double res = residual_f64(x64, D.cxl + o, D.cyl + o, D.bc + o, D.rhs + o, I.inlet, x,
                                      y, nx, ny, i, isx, isy);
      acc[0] += res * res;
      // This site's four faces; the lower x and y faces (and the upper ones
      // at the grid's end) are this item's to store.
      const float uxl = face_x(I, x64, cond, solid, x, y);
      const float uxu = face_x(I, x64, cond, solid, x + 1, y);
      const float uyl = face_y(I, x64, cond, solid, x, y);
      const float uyu = face_y(I, x64, cond, solid, x, y + 1);
      xf[i] = uxl;
      if (x + 1 == nx) xf[i + ny] = uxu;
      const size_t yi = static_cast<size_t>(x) * (ny + 1) + y;
      yf[yi] = uyl;
      if (y + 1 == ny) yf[yi + 1] = uyu;
      const float current = in[i];
      if (solid[i] != 0) {
        out[i] = current;
        continue;
      }
      SiteGeometry g = geometry(A, solid, xf, yf, x, y, 0);
      g.ulo[0] = uxl;
      g.uup[0] = uxu;
      g.ulo[1] = uyl;
      g.uup[1] = uyu;
      const float candidate = fe_candidate_f32(A, g, in, x, y, 0, dt, 0.0f, 0.0f, phi);
      if (!isfinite(candidate) || candidate < 0.0f) bad = 1;
      out[i] = candidate;
      acc[1] += static_cast<double>(current);
      acc[2] += static_cast<double>(candidate);
      if (phi != nullptr) {
        acc[3] += static_cast<double>(P.vmax) * static_cast<double>(phi[i]) *
                  static_cast<double>(current) / (static_cast<double>(P.monod_k) + current);
      }
    }
  }
  if (bad) atomicOr(&D.flags[inst], 4u);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double t = block_sum<NT>(acc[k], s_red);
    if (threadIdx.x == 0) {
      partials[(static_cast<size_t>(step) * total + base + blockIdx.x) * 4 + k] = t;
    }
  }
}

// The same step for instances at most 256 columns wide, marching, every
// width class in one launch. A workgroup of 256 threads holds 256 / W groups
// of W threads (W = 64, 128 or 256: the width class), each group one work
// item (rows [x0, x1) of an instance), one thread per column. A group walks
// x, so each array is read once; the x neighbours come from registers, the y
// neighbours from the group's segment of an LDS row, and each x face is
// computed once (the upper face of a row is the next row's lower face). The
// groups of a workgroup walk the same number of rows (the most any of them
// has; a group past its own rows idles), so the row barriers are uniform.
// Every site's arithmetic is the general kernel's, so the two agree bit for
// bit.
constexpr int kMarchThreads = 256;
constexpr int kMarchGroups = 4;  // groups per workgroup at the narrowest width

struct MarchBlock {
  uint32_t width;                 // W
  uint32_t rows;                  // rows every group walks
  uint32_t item[kMarchGroups];    // work items; 0xffffffff: none
};

__global__ __launch_bounds__(kMarchThreads) void k_batch_march(
    BatchData D, BatchParams P, const BatchItem* __restrict__ items,
    const MarchBlock* __restrict__ blocks, double* __restrict__ partials, uint32_t step, float dt,
    uint32_t total) {
  struct Row {
    double p;
    float cond, c, cxl, cyl, bc, rhs, phi;
    bool solid;
  };
  __shared__ double s_p[2][kMarchThreads];
  __shared__ float s_cond[2][kMarchThreads], s_c[2][kMarchThreads], s_cyl[2][kMarchThreads];
  __shared__ uint8_t s_s[2][kMarchThreads];
  __shared__ double s_red[4][kMarchThreads / kWave];
  const MarchBlock blk = blocks[blockIdx.x];
  const uint32_t W = blk.width;
  const uint32_t group = threadIdx.x / W;
  const uint32_t y = threadIdx.x % W;
  const uint32_t seg = group * W;  // the group's segment of the LDS rows
  const uint32_t item_index = blk.item[group];
  const bool has_item = item_index != 0xffffffffu;
  const BatchItem item = has_item ? items[item_index] : BatchItem{0, 0, 0, 0};
  const uint32_t inst = item.inst;
  const BatchDesc I = D.desc[inst];
  const uint32_t nx = I.nx, ny = I.ny;
  const size_t o = I.site;
  const bool valid = has_item && y < ny;
  const float* in = ((step & 1) == 0 ? D.la : D.lb) + o;
  float* out = ((step & 1) == 0 ? D.lb : D.la) + o;
  const bool monod = D.phi != nullptr;
  float* xf = D.xf + I.xface;
  float* yf = D.yf + I.yface;
  const double isx = 1.0 / (static_cast<double>(I.dx) * I.dx);
  const double isy = 1.0 / (static_cast<double>(I.dy) * I.dy);
  AdrParams A{};
  A.nx = nx;
  A.ny = ny;
  A.signals = 1;
  A.dx = I.dx;
  A.dy = I.dy;
  A.dz = I.dz;
  A.kind[0] = A.kind[1] = MSIM_FIXED;
  A.kind[2] = A.kind[3] = A.kind[4] = A.kind[5] = MSIM_NO_FLUX;
  A.has_field = 1;
  A.has_monod = monod ? 1 : 0;
  A.diffusion[0] = I.diffusion;
  A.boundary[0][0] = I.reservoir;
  A.vmax[0] = P.vmax;
  A.monod_k[0] = P.monod_k;
  const int x0 = static_cast<int>(item.x0), x1 = static_cast<int>(item.x1);
  auto load = [&](int xx, bool full) {
    Row w{};
    if (!valid || xx < 0 || xx >= static_cast<int>(nx) || xx > x1) return w;
    const size_t i = o + static_cast<size_t>(xx) * ny + y;
    w.p = D.x64[i];
    w.cond = D.cond[i];
    w.c = in[i - o];
    w.solid = D.solid[i] != 0;
    w.cxl = D.cxl[i];  // also the upper x coupling of the row before
    if (!full) return w;
    w.cyl = D.cyl[i];
    w.bc = D.bc[i];
    w.rhs = D.rhs[i];
    if (monod) w.phi = D.phi[i];
    return w;
  };
  // Face x fx of this column from the rows either side (face_x's arithmetic).
  auto face_x_of = [&](const Row& lo, const Row& up, int fx) -> float {
    const bool has_lower = fx > 0, has_upper = fx < static_cast<int>(nx);
    double value;
    if (has_lower && has_upper) {
      const float fm = harmonic(lo.cond, up.cond);
      value = -static_cast<double>(fm) * (up.p - lo.p) / static_cast<double>(I.dx);
    } else {
      const Row& site = has_upper ? up : lo;
      float fc = 0.0f, fp = 0.0f;
      if (site.cond != 0.0f) {
        fc = (has_upper ? 2.0f : -2.0f) * site.cond / I.dx;
        fp = has_upper ? I.inlet : 0.0f;
      }
      value = static_cast<double>(fc) * (static_cast<double>(fp) - site.p);
    }
    const float flux = static_cast<float>(value);
    if ((has_lower && lo.solid) || (has_upper && up.solid)) return 0.0f;
    return flux / I.dz;
  };
  // Face y fy of row buf from the group's LDS row (face_y's arithmetic).
  auto face_y_of = [&](int buf, uint32_t fy) -> float {
    const bool has_lower = fy > 0, has_upper = fy < ny;
    double value;
    if (has_lower && has_upper) {
      const float fm = harmonic(s_cond[buf][seg + fy - 1], s_cond[buf][seg + fy]);
      value = -static_cast<double>(fm) * (s_p[buf][seg + fy] - s_p[buf][seg + fy - 1]) /
              static_cast<double>(I.dy);
    } else {
      const double ps = has_upper ? s_p[buf][seg + fy] : s_p[buf][seg + fy - 1];
      value = static_cast<double>(0.0f) * (static_cast<double>(0.0f) - ps);
    }
    const float flux = static_cast<float>(value);
    if ((has_lower && s_s[buf][seg + fy - 1] != 0) || (has_upper && s_s[buf][seg + fy] != 0))
      return 0.0f;
    return flux / I.dz;
  };
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t bad = 0;
  Row prev = load(x0 - 1, false);
  Row cur = load(x0, true);
  Row next = load(x0 + 1, x0 + 1 < x1);
  float uxl = face_x_of(prev, cur, x0);
  for (uint32_t r = 0; r < blk.rows; ++r) {
    const int xx = x0 + static_cast<int>(r);
    const bool live = valid && xx < x1;
    const Row ahead = xx + 2 <= x1 ? load(xx + 2, xx + 2 < x1) : Row{};
    const int buf = static_cast<int>(r & 1);
    s_p[buf][threadIdx.x] = cur.p;
    s_cond[buf][threadIdx.x] = cur.cond;
    s_c[buf][threadIdx.x] = cur.c;
    s_cyl[buf][threadIdx.x] = cur.cyl;
    s_s[buf][threadIdx.x] = cur.solid ? 1 : 0;
    __syncthreads();
    const float uxu = face_x_of(cur, next, xx + 1);
    if (live) {
      const size_t i = static_cast<size_t>(xx) * ny + y;
      // residual_f64's arithmetic.
      const double c = cur.p;
      double a = static_cast<double>(cur.bc) * c;
      if (xx > 0) a += (static_cast<double>(cur.cxl) * isx) * (c - prev.p);
      if (xx + 1 < static_cast<int>(nx)) a += (static_cast<double>(next.cxl) * isx) * (c - next.p);
      if (y > 0) a += (static_cast<double>(cur.cyl) * isy) * (c - s_p[buf][seg + y - 1]);
      if (y + 1 < ny)
        a += (static_cast<double>(s_cyl[buf][seg + y + 1]) * isy) * (c - s_p[buf][seg + y + 1]);
      const double res = static_cast<double>(I.inlet) * static_cast<double>(cur.rhs) - a;
      acc[0] += res * res;
      const float uyl = face_y_of(buf, y);
      const float uyu = face_y_of(buf, y + 1);
      xf[i] = uxl;
      if (xx + 1 == static_cast<int>(nx)) xf[i + ny] = uxu;
      const size_t yi = static_cast<size_t>(xx) * (ny + 1) + y;
      yf[yi] = uyl;
      if (y + 1 == ny) yf[yi + 1] = uyu;
      const float current = cur.c;
      if (cur.solid) {
        out[i] = current;
      } else {
        const bool xlo_edge = xx == 0, xhi_edge = xx + 1 == static_cast<int>(nx);
        const bool ylo_edge = y == 0, yhi_edge = y + 1 == ny;
        SiteGeometry g;
        g.clo[0] = xlo_edge ? false : prev.solid;
        g.cup[0] = xhi_edge ? false : next.solid;
        g.clo[1] = ylo_edge ? true : s_s[buf][seg + y - 1] != 0;
        g.cup[1] = yhi_edge ? true : s_s[buf][seg + y + 1] != 0;
        g.ulo[0] = uxl;
        g.uup[0] = uxu;
        g.ulo[1] = uyl;
        g.uup[1] = uyu;
        const float l0 = xlo_edge ? exterior(A, 0, 0, current, 0.0f, y) : prev.c;
        const float u0 = xhi_edge ? exterior(A, 1, 0, current, 0.0f, y) : next.c;
        const float l1 = ylo_edge ? exterior(A, 2, 0, current, 0.0f, static_cast<uint32_t>(xx))
                                  : s_c[buf][seg + y - 1];
        const float u1 = yhi_edge ? exterior(A, 3, 0, current, 0.0f, static_cast<uint32_t>(xx))
                                  : s_c[buf][seg + y + 1];
        const float candidate =
            fe_update_f32(A, g, current, l0, u0, l1, u1, 0, dt, 0.0f, 0.0f, cur.phi);
        if (!isfinite(candidate) || candidate < 0.0f) bad = 1;
        out[i] = candidate;
        acc[1] += static_cast<double>(current);
        acc[2] += static_cast<double>(candidate);
        if (monod) {
          acc[3] += static_cast<double>(P.vmax) * static_cast<double>(cur.phi) *
                    static_cast<double>(current) / (static_cast<double>(P.monod_k) + current);
        }
      }
    }
    uxl = uxu;
    prev = cur;
    cur = next;
    next = ahead;
  }
  if (bad) atomicOr(&D.flags[inst], 4u);
  // Per group: wave sums, then the group's waves in order.
  const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) s_red[k][wave] = v;
  }
  __syncthreads();
  if (has_item && y == 0) {
    const uint32_t first = seg / kWave, waves = W / kWave;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double t = 0.0;
      for (uint32_t w = 0; w < waves; ++w) t += s_red[k][first + w];
      partials[(static_cast<size_t>(step) * total + item_index) * 4 + k] = t;
    }
  }
}

using FlowKernel = void (*)(BatchData, BatchParams);
using StepKernel = void (*)(BatchData, BatchParams, const BatchItem*, double*, uint32_t, float,
                            uint32_t, uint32_t);

// A launch shape of the step: sites per work item, and for instances wider
// than 256 columns (the general kernel) the workgroup size and sweep row
// width.
struct Shape {
  int threads, row, chunk;
  StepKernel step;
};

const Shape kShapes[] = {
    {256, 64, 512, k_batch_step<256, 64>},    {256, 64, 1024, k_batch_step<256, 64>},
    {256, 64, 2048, k_batch_step<256, 64>},   {256, 64, 4096, k_batch_step<256, 64>},
    {512, 64, 8192, k_batch_step<512, 64>},   {256, 128, 16384, k_batch_step<256, 128>},
};
constexpr int kShapeCount = static_cast<int>(sizeof(kShapes) / sizeof(kShapes[0]));
constexpr int kDefaultShape = 1;
// Instance classes by width: the marching kernel's groups of 64, 128 and 256
// threads for at most that many columns, then the general kernel.
constexpr int kClasses = 4;
constexpr uint32_t kClassWidth[kClasses - 1] = {64, 128, 256};

int class_of(uint32_t ny) {
  for (int c = 0; c < kClasses - 1; ++c) {
    if (ny <= kClassWidth[c]) return c;
  }
  return kClasses - 1;
}
constexpr FlowKernel kFlowKernel = k_batch_flow<1024, 128>;
constexpr int kFlowThreads = 1024;

}  // namespace

// Device state of an uploaded batch.
struct BatchDevice {
  uint32_t count = 0;
  uint64_t sites = 0, xfaces = 0, yfaces = 0;
  std::vector<BatchDesc> desc;
  std::vector<uint32_t> order;
  std::vector<std::vector<BatchItem>> items;  // per shape, grouped by instance class
  std::vector<std::array<uint32_t, 5>> class_begin;  // per shape: first item of each class, end
  std::vector<std::unique_ptr<DeviceBuffer<BatchItem>>> d_items;
  std::vector<std::vector<MarchBlock>> march;  // per shape: workgroups of the marching launch
  std::vector<std::unique_ptr<DeviceBuffer<MarchBlock>>> d_march;
  DeviceBuffer<BatchDesc> d_desc;
  DeviceBuffer<uint32_t> d_order, iterations, flags;
  DeviceBuffer<uint8_t> solid;
  DeviceBuffer<float> mobility, cond, cxl, cyl, bc, diag, rhs, phi, x, r, p, q, z, xf, yf, la, lb,
      residual, saved;
  std::vector<BatchLevels> levels;  // per instance
  DeviceBuffer<BatchLevels> d_levels;
  DeviceBuffer<float> lwx, lwy, lbc, ldinv, lr, le, cinv;
  DeviceBuffer<double> x64, bnorm2, partials;
  BatchParams params{};
  bool has_phi = false;
  bool solved = false;
  int shape = -1;   // chosen by the autotuner; -1: the default
  int current = 0;  // 0: levels in la, 1: in lb
  hipGraphExec_t graph = nullptr;  // the captured run and what it was captured with
  uint32_t graph_steps = 0;
  float graph_dt = 0.0f;
  int graph_shape = -1;
  BatchData graph_data{};
  const double* graph_partials = nullptr;
  ~BatchDevice() {
    if (graph != nullptr) static_cast<void>(hipGraphExecDestroy(graph));
  }
};

namespace {

BatchData batch_data(BatchDevice& B) {
  BatchData D{};
  D.desc = B.d_desc.data();
  D.order = B.d_order.data();
  D.solid = B.solid.data();
  D.cond = B.cond.data();
  D.cxl = B.cxl.data();
  D.cyl = B.cyl.data();
  D.bc = B.bc.data();
  D.rhs = B.rhs.data();
  D.phi = B.has_phi ? B.phi.data() : nullptr;
  D.x64 = B.x64.data();
  D.x = B.x.data();
  D.r = B.r.data();
  D.p = B.p.data();
  D.q = B.q.data();
  D.xf = B.xf.data();
  D.yf = B.yf.data();
  D.la = B.la.data();
  D.lb = B.lb.data();
  D.iterations = B.iterations.data();
  D.residual = B.residual.data();
  D.bnorm2 = B.bnorm2.data();
  D.flags = B.flags.data();
  D.z = B.z.data();
  D.levels = B.d_levels.data();
  D.lwx = B.lwx.data();
  D.lwy = B.lwy.data();
  D.lbc = B.lbc.data();
  D.ldinv = B.ldinv.data();
  D.lr = B.lr.data();
  D.le = B.le.data();
  D.cinv = B.cinv.data();
  return D;
}

int shape_of(const BatchDevice& B) { return B.shape >= 0 ? B.shape : kDefaultShape; }

void launch_step(Context& ctx, BatchDevice& B, const BatchData& D, int shape, uint32_t step,
                 float dt) {
  const Shape& s = kShapes[shape];
  const auto& begin = B.class_begin[shape];
  const uint32_t total = static_cast<uint32_t>(B.items[shape].size());
  // Every instance at most 256 columns wide: one launch.
  const uint32_t marching = static_cast<uint32_t>(B.march[shape].size());
  if (marching > 0) {
    k_batch_march<<<marching, kMarchThreads, 0, ctx.stream>>>(
        D, B.params, B.d_items[shape]->data(), B.d_march[shape]->data(), B.partials.data(), step,
        dt, total);
    check(hipGetLastError(), "launch batch step");
  }
  // Wider instances: the general sweep.
  const uint32_t base = begin[kClasses - 1], n = begin[kClasses] - begin[kClasses - 1];
  if (n > 0) {
    hipLaunchKernelGGL(s.step, dim3(n), dim3(static_cast<uint32_t>(s.threads)), 0, ctx.stream, D,
                       B.params, B.d_items[shape]->data(), B.partials.data(), step, dt, base,
                       total);
    check(hipGetLastError(), "launch batch step");
  }
}

// Runs `steps` steps from the current levels, in a graph or eagerly.
void run_steps(Context& ctx, BatchDevice& B, int shape, float dt, uint32_t steps, bool graph) {
  const BatchData D = batch_data(B);
  if (B.current != 0) {
    check(hipMemcpyAsync(B.la.data(), B.lb.data(), B.sites * sizeof(float),
                         hipMemcpyDeviceToDevice, ctx.stream),
          "levels to the first buffer");
    B.current = 0;
  }
  size_t most = 0;
  for (const auto& it : B.items) most = std::max(most, it.size());
  B.partials.reserve(static_cast<size_t>(steps) * most * 4);
  if (!graph) {
    for (uint32_t s = 0; s < steps; ++s) launch_step(ctx, B, D, shape, s, dt);
  } else {
    // A captured graph holds every pointer: capture again when any changed.
    if (B.graph == nullptr || B.graph_steps != steps || B.graph_dt != dt ||
        B.graph_shape != shape || std::memcmp(&B.graph_data, &D, sizeof(D)) != 0 ||
        B.graph_partials != B.partials.data()) {
      if (B.graph != nullptr) {
        check(hipGraphExecDestroy(B.graph), "destroy graph");
        B.graph = nullptr;
      }
      hipGraph_t g = nullptr;
      check(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal), "begin capture");
      for (uint32_t s = 0; s < steps; ++s) launch_step(ctx, B, D, shape, s, dt);
      check(hipStreamEndCapture(ctx.stream, &g), "end capture");
      const hipError_t e = hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0);
      static_cast<void>(hipGraphDestroy(g));
      check(e, "instantiate graph");
      B.graph_steps = steps;
      B.graph_dt = dt;
      B.graph_shape = shape;
      B.graph_data = D;
      B.graph_partials = B.partials.data();
    }
    check(hipGraphLaunch(B.graph, ctx.stream), "launch graph");
  }
  B.current = static_cast<int>(steps & 1);
}

}  // namespace

namespace {

// Each instance's multilevel hierarchy: level shapes and offsets, the level
// operators on the device (level 0 from the assembled couplings), and the
// dense inverse of every coarsest level, computed on the host.
void build_hierarchies(Context& ctx, BatchDevice& B) {
  B.levels.assign(B.count, BatchLevels{});
  uint64_t pool = B.sites, inverses = 0;
  for (uint32_t k = 0; k < B.count; ++k) {
    const BatchDesc& d = B.desc[k];
    BatchLevels& L = B.levels[k];
    uint32_t nx = d.nx, ny = d.ny;
    L.off[0] = d.site;
    L.nx[0] = nx;
    L.ny[0] = ny;
    L.count = 1;
    while (static_cast<uint64_t>(nx) * ny > kBatchCoarse) {
      if (L.count == kBatchLevels) throw Error(MSIM_OVERFLOW, "batch instance needs too many levels");
      nx = (nx + 1) / 2;
      ny = (ny + 1) / 2;
      L.off[L.count] = pool;
      L.nx[L.count] = nx;
      L.ny[L.count] = ny;
      pool += static_cast<uint64_t>(nx) * ny;
      ++L.count;
    }
    L.coarse_n = nx * ny;
    L.cinv = inverses;
    inverses += static_cast<uint64_t>(L.coarse_n) * L.coarse_n;
  }
  for (auto* b : {&B.lwx, &B.lwy, &B.lbc, &B.ldinv, &B.lr, &B.le}) b->reserve(pool);
  B.cinv.reserve(std::max<uint64_t>(inverses, 1));
  B.d_levels.reserve(B.count);
  check(hipMemcpyAsync(B.d_levels.data(), B.levels.data(), B.count * sizeof(BatchLevels),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload hierarchies");
  for (uint32_t k = 0; k < B.count; ++k) {
    const BatchDesc& d = B.desc[k];
    const BatchLevels& L = B.levels[k];
    const float isx = 1.0f / (d.dx * d.dx), isy = 1.0f / (d.dy * d.dy);
    mg_level0_operator(ctx.stream, B.cxl.data() + d.site, B.cyl.data() + d.site,
                       B.bc.data() + d.site, B.lwx.data() + d.site, B.lwy.data() + d.site,
                       B.lbc.data() + d.site, B.ldinv.data() + d.site, d.nx, d.ny, isx, isy);
    for (uint32_t l = 1; l < L.count; ++l) {
      const uint64_t f = L.off[l - 1], c = L.off[l];
      mg_coarse_operator(ctx.stream, B.lwx.data() + f, B.lwy.data() + f, B.lbc.data() + f,
                         L.nx[l - 1], L.ny[l - 1], B.lwx.data() + c, B.lwy.data() + c,
                         B.lbc.data() + c, B.ldinv.data() + c, L.nx[l], L.ny[l]);
    }
  }
  // The coarsest operators to the host, their inverses back.
  std::vector<std::vector<float>> wx(B.count), wy(B.count), bc(B.count);
  for (uint32_t k = 0; k < B.count; ++k) {
    const BatchLevels& L = B.levels[k];
    const uint64_t c = L.off[L.count - 1];
    for (auto* v : {&wx[k], &wy[k], &bc[k]}) v->resize(L.coarse_n);
    check(hipMemcpyAsync(wx[k].data(), B.lwx.data() + c, L.coarse_n * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download coarse operator");
    check(hipMemcpyAsync(wy[k].data(), B.lwy.data() + c, L.coarse_n * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download coarse operator");
    check(hipMemcpyAsync(bc[k].data(), B.lbc.data() + c, L.coarse_n * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download coarse operator");
  }
  check(hipStreamSynchronize(ctx.stream), "coarse operators");
  std::vector<float> all(inverses);
  for (uint32_t k = 0; k < B.count; ++k) {
    const BatchLevels& L = B.levels[k];
    const std::vector<float> inv =
        mg_coarse_inverse(wx[k], wy[k], bc[k], L.nx[L.count - 1], L.ny[L.count - 1]);
    // Transposed: entry (row, j) at j * n + row.
    const size_t n = L.coarse_n;
    for (size_t row = 0; row < n; ++row)
      for (size_t j = 0; j < n; ++j) all[L.cinv + j * n + row] = inv[row * n + j];
  }
  check(hipMemcpyAsync(B.cinv.data(), all.data(), all.size() * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload coarse inverses");
  check(hipStreamSynchronize(ctx.stream), "upload coarse inverses");
}

}  // namespace

BatchDevice& batch_of(Context& ctx) {
  if (ctx.batch == nullptr) ctx.batch = std::make_shared<BatchDevice>();
  return *ctx.batch;
}

void batch_upload(Context& ctx, uint32_t count, const msim_batch_instance* inst,
                  const msim_batch_params& params) {
  ctx.batch.reset();
  BatchDevice& B = batch_of(ctx);
  B.count = count;
  B.desc.resize(count);
  uint64_t site = 0, xface = 0, yface = 0;
  B.has_phi = false;
  for (uint32_t k = 0; k < count; ++k) {
    const msim_batch_instance& in = inst[k];
    BatchDesc& d = B.desc[k];
    d.nx = in.nx;
    d.ny = in.ny;
    d.site = site;
    d.xface = xface;
    d.yface = yface;
    d.dx = in.dx;
    d.dy = in.dy;
    d.dz = in.dz;
    d.inlet = in.inlet_pressure;
    d.diffusion = in.diffusion;
    d.reservoir = in.reservoir;
    site += static_cast<uint64_t>(in.nx) * in.ny;
    xface += static_cast<uint64_t>(in.nx + 1) * in.ny;
    yface += static_cast<uint64_t>(in.nx) * (in.ny + 1);
    if (in.phi != nullptr) B.has_phi = true;
  }
  B.sites = site;
  B.xfaces = xface;
  B.yfaces = yface;
  // Largest instances first: they are dispatched first.
  B.order.resize(count);
  std::iota(B.order.begin(), B.order.end(), 0u);
  std::stable_sort(B.order.begin(), B.order.end(), [&](uint32_t a, uint32_t b) {
    return static_cast<uint64_t>(B.desc[a].nx) * B.desc[a].ny >
           static_cast<uint64_t>(B.desc[b].nx) * B.desc[b].ny;
  });
  // Work items of each shape: blocks of whole rows of about `chunk` sites.
  B.items.assign(kShapeCount, {});
  B.class_begin.assign(kShapeCount, {});
  B.d_items.clear();
  B.march.clear();
  B.d_march.clear();
  for (int s = 0; s < kShapeCount; ++s) {
    for (int c = 0; c < kClasses; ++c) {
      B.class_begin[s][c] = static_cast<uint32_t>(B.items[s].size());
      for (uint32_t k : B.order) {
        const BatchDesc& d = B.desc[k];
        if (class_of(d.ny) != c) continue;
        const uint32_t rows =
            std::max<uint32_t>(1, static_cast<uint32_t>(kShapes[s].chunk) / d.ny);
        for (uint32_t x0 = 0; x0 < d.nx; x0 += rows)
          B.items[s].push_back(BatchItem{k, x0, std::min(d.nx, x0 + rows), 0});
      }
    }
    B.class_begin[s][kClasses] = static_cast<uint32_t>(B.items[s].size());
    if (B.items[s].size() > 0x7FFFFFFFu) throw Error(MSIM_OVERFLOW, "too many batch work items");
    // Workgroups of the marching launch: 256 / W consecutive items of a
    // class of width W, walking the most rows any of them has.
    std::vector<MarchBlock> blocks;
    for (int c = 0; c < kClasses - 1; ++c) {
      const uint32_t width = kClassWidth[c];
      const uint32_t per = static_cast<uint32_t>(kMarchThreads) / width;
      for (uint32_t j = B.class_begin[s][c]; j < B.class_begin[s][c + 1]; j += per) {
        MarchBlock b{};
        b.width = width;
        for (uint32_t g = 0; g < static_cast<uint32_t>(kMarchGroups); ++g) {
          const uint32_t it = j + g;
          if (g < per && it < B.class_begin[s][c + 1]) {
            b.item[g] = it;
            b.rows = std::max(b.rows, B.items[s][it].x1 - B.items[s][it].x0);
          } else {
            b.item[g] = 0xffffffffu;
          }
        }
        blocks.push_back(b);
      }
    }
    auto mbuf = std::make_unique<DeviceBuffer<MarchBlock>>();
    mbuf->reserve(std::max<size_t>(blocks.size(), 1));
    check(hipMemcpyAsync(mbuf->data(), blocks.data(), blocks.size() * sizeof(MarchBlock),
                         hipMemcpyHostToDevice, ctx.stream),
          "upload marching workgroups");
    B.d_march.push_back(std::move(mbuf));
    B.march.push_back(std::move(blocks));
    auto buf = std::make_unique<DeviceBuffer<BatchItem>>();
    buf->reserve(B.items[s].size());
    check(hipMemcpyAsync(buf->data(), B.items[s].data(), B.items[s].size() * sizeof(BatchItem),
                         hipMemcpyHostToDevice, ctx.stream),
          "upload work items");
    B.d_items.push_back(std::move(buf));
  }
  const size_t n = site;
  B.solid.reserve(n);
  for (auto* b : {&B.mobility, &B.cond, &B.cxl, &B.cyl, &B.bc, &B.diag, &B.rhs, &B.phi, &B.x,
                  &B.r, &B.p, &B.q, &B.z, &B.la, &B.lb, &B.saved})
    b->reserve(n);
  B.x64.reserve(n);
  B.xf.reserve(xface);
  B.yf.reserve(yface);
  B.d_desc.reserve(count);
  B.d_order.reserve(count);
  B.iterations.reserve(count);
  B.flags.reserve(count);
  B.residual.reserve(count);
  B.bnorm2.reserve(count);
  check(hipMemcpyAsync(B.d_desc.data(), B.desc.data(), count * sizeof(BatchDesc),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload batch");
  check(hipMemcpyAsync(B.d_order.data(), B.order.data(), count * sizeof(uint32_t),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload batch order");
  check(hipMemsetAsync(B.flags.data(), 0, count * sizeof(uint32_t), ctx.stream), "clear flags");
  check(hipMemsetAsync(B.x64.data(), 0, n * sizeof(double), ctx.stream), "clear pressure");
  for (uint32_t k = 0; k < count; ++k) {
    const msim_batch_instance& in = inst[k];
    const BatchDesc& d = B.desc[k];
    const size_t m = static_cast<size_t>(d.nx) * d.ny;
    if (in.obstacles != nullptr) {
      check(hipMemcpyAsync(B.solid.data() + d.site, in.obstacles, m, hipMemcpyHostToDevice,
                           ctx.stream),
            "upload obstacles");
    } else {
      check(hipMemsetAsync(B.solid.data() + d.site, 0, m, ctx.stream), "clear obstacles");
    }
    check(hipMemcpyAsync(B.mobility.data() + d.site, in.mobility, m * sizeof(float),
                         hipMemcpyHostToDevice, ctx.stream),
          "upload mobility");
    if (B.has_phi) {
      if (in.phi != nullptr) {
        check(hipMemcpyAsync(B.phi.data() + d.site, in.phi, m * sizeof(float),
                             hipMemcpyHostToDevice, ctx.stream),
              "upload phi");
      } else {
        check(hipMemsetAsync(B.phi.data() + d.site, 0, m * sizeof(float), ctx.stream),
              "clear phi");
      }
    }
    if (in.levels != nullptr) {
      check(hipMemcpyAsync(B.la.data() + d.site, in.levels, m * sizeof(float),
                           hipMemcpyHostToDevice, ctx.stream),
            "upload levels");
    } else {
      check(hipMemsetAsync(B.la.data() + d.site, 0, m * sizeof(float), ctx.stream),
            "clear levels");
    }
    launch_assembly(ctx.stream, B.solid.data() + d.site, B.mobility.data() + d.site,
                    B.cond.data() + d.site, B.diag.data() + d.site, B.rhs.data() + d.site,
                    B.bc.data() + d.site, B.cxl.data() + d.site, B.cyl.data() + d.site, d.nx,
                    d.ny, d.dx, d.dy, d.dz, FlowEdges{});
  }
  build_hierarchies(ctx, B);
  B.params.tolerance = params.relative_tolerance;
  B.params.max_iterations = params.max_iterations;
  B.params.vmax = params.monod_vmax;
  B.params.monod_k = params.monod_k;
  B.current = 0;
  B.solved = false;
  check(hipStreamSynchronize(ctx.stream), "batch upload");
}

// Solves every instance's pressure from its current one and rebuilds its
// velocities (one launch).
void batch_solve_flow(Context& ctx, uint32_t* iterations, float* residual, uint32_t* flags) {
  BatchDevice& B = batch_of(ctx);
  if (B.count == 0) throw Error(MSIM_INVALID_ARGUMENT, "no batch is uploaded");
  hipLaunchKernelGGL(kFlowKernel, dim3(B.count), dim3(kFlowThreads), 0, ctx.stream,
                     batch_data(B), B.params);
  check(hipGetLastError(), "launch k_batch_flow");
  B.solved = true;
  if (iterations != nullptr) {
    check(hipMemcpyAsync(iterations, B.iterations.data(), B.count * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download iterations");
  }
  if (residual != nullptr) {
    check(hipMemcpyAsync(residual, B.residual.data(), B.count * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download residuals");
  }
  if (flags != nullptr) {
    check(hipMemcpyAsync(flags, B.flags.data(), B.count * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download flags");
  }
  check(hipStreamSynchronize(ctx.stream), "batch flow");
}

int batch_autotune(Context& ctx, float dt, uint32_t steps, double* seconds) {
  BatchDevice& B = batch_of(ctx);
  if (B.count == 0) throw Error(MSIM_INVALID_ARGUMENT, "no batch is uploaded");
  if (!B.solved) throw Error(MSIM_INVALID_ARGUMENT, "solve the batch's flow first");
  // Every candidate starts from the same levels, and the batch is left as it was.
  float* levels = B.current == 0 ? B.la.data() : B.lb.data();
  check(hipMemcpyAsync(B.saved.data(), levels, B.sites * sizeof(float), hipMemcpyDeviceToDevice,
                       ctx.stream),
        "save levels");
  auto restore = [&]() {
    check(hipMemcpyAsync(B.la.data(), B.saved.data(), B.sites * sizeof(float),
                         hipMemcpyDeviceToDevice, ctx.stream),
          "restore levels");
    B.current = 0;
  };
  double best = 0.0;
  int choice = 0;
  for (int c = 0; c < kShapeCount; ++c) {
    restore();
    run_steps(ctx, B, c, dt, 1, false);  // warm
    restore();
    check(hipEventRecord(ctx.start, ctx.stream), "event");
    run_steps(ctx, B, c, dt, steps, false);
    check(hipEventRecord(ctx.stop, ctx.stream), "event");
    check(hipEventSynchronize(ctx.stop), "autotune");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
    const double t = ms * 1.0e-3 / steps;
    if (seconds != nullptr) seconds[c] = t;
    if (c == 0 || t < best) {
      best = t;
      choice = c;
    }
  }
  restore();
  check(hipMemsetAsync(B.flags.data(), 0, B.count * sizeof(uint32_t), ctx.stream), "clear flags");
  check(hipStreamSynchronize(ctx.stream), "autotune");
  B.shape = choice;
  return choice;
}

void batch_shape(int index, int* threads, int* row, int* chunk) {
  *threads = kShapes[index].threads;
  *row = kShapes[index].row;
  *chunk = kShapes[index].chunk;
}

int batch_shape_count() { return kShapeCount; }

void batch_run(Context& ctx, float dt, uint32_t steps, bool graph, double* ledger,
               double* seconds) {
  BatchDevice& B = batch_of(ctx);
  if (B.count == 0) throw Error(MSIM_INVALID_ARGUMENT, "no batch is uploaded");
  if (!B.solved) throw Error(MSIM_INVALID_ARGUMENT, "solve the batch's flow first");
  const int shape = shape_of(B);
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  run_steps(ctx, B, shape, dt, steps, graph);
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  check(hipEventSynchronize(ctx.stop), "batch run");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  if (seconds != nullptr) *seconds = ms * 1.0e-3;
  if (ledger == nullptr) return;
  // Per-instance rows from the per-item partials, added in item order.
  const auto& items = B.items[shape];
  std::vector<double> parts(static_cast<size_t>(steps) * items.size() * 4);
  std::vector<double> bb(B.count);
  check(hipMemcpyAsync(parts.data(), B.partials.data(), parts.size() * sizeof(double),
                       hipMemcpyDeviceToHost, ctx.stream),
        "download partials");
  check(hipMemcpyAsync(bb.data(), B.bnorm2.data(), B.count * sizeof(double),
                       hipMemcpyDeviceToHost, ctx.stream),
        "download norms");
  check(hipStreamSynchronize(ctx.stream), "ledger");
  std::fill(ledger, ledger + static_cast<size_t>(steps) * B.count * 4, 0.0);
  for (uint32_t s = 0; s < steps; ++s) {
    for (size_t j = 0; j < items.size(); ++j) {
      const double* pj = &parts[(static_cast<size_t>(s) * items.size() + j) * 4];
      double* row = ledger + (static_cast<size_t>(s) * B.count + items[j].inst) * 4;
      row[0] += pj[1];
      row[1] += pj[2];
      row[2] += pj[3];
      row[3] += pj[0];
    }
    for (uint32_t k = 0; k < B.count; ++k) {
      const BatchDesc& d = B.desc[k];
      const double volume = static_cast<double>(d.dx) * d.dy * d.dz;
      double* row = ledger + (static_cast<size_t>(s) * B.count + k) * 4;
      row[0] *= volume;
      row[1] *= volume;
      row[2] *= static_cast<double>(dt) * volume;
      row[3] = bb[k] > 0.0 ? std::sqrt(row[3] / bb[k]) : 0.0;
    }
  }
}

void batch_flags(Context& ctx, uint32_t* flags) {
  BatchDevice& B = batch_of(ctx);
  if (B.count == 0) throw Error(MSIM_INVALID_ARGUMENT, "no batch is uploaded");
  check(hipMemcpyAsync(flags, B.flags.data(), B.count * sizeof(uint32_t), hipMemcpyDeviceToHost,
                       ctx.stream),
        "download flags");
  check(hipStreamSynchronize(ctx.stream), "flags");
}

void batch_download(Context& ctx, uint32_t index, double* pressure, float* x_faces,
                    float* y_faces, float* levels) {
  BatchDevice& B = batch_of(ctx);
  if (index >= B.count) throw Error(MSIM_INVALID_ARGUMENT, "instance index out of range");
  const BatchDesc& d = B.desc[index];
  const size_t m = static_cast<size_t>(d.nx) * d.ny;
  if (pressure != nullptr) {
    check(hipMemcpyAsync(pressure, B.x64.data() + d.site, m * sizeof(double),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download pressure");
  }
  if (x_faces != nullptr) {
    check(hipMemcpyAsync(x_faces, B.xf.data() + d.xface,
                         static_cast<size_t>(d.nx + 1) * d.ny * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download faces");
  }
  if (y_faces != nullptr) {
    check(hipMemcpyAsync(y_faces, B.yf.data() + d.yface,
                         static_cast<size_t>(d.nx) * (d.ny + 1) * sizeof(float),
                         hipMemcpyDeviceToHost, ctx.stream),
          "download faces");
  }
  if (levels != nullptr) {
    const float* src = (B.current == 0 ? B.la.data() : B.lb.data()) + d.site;
    check(hipMemcpyAsync(levels, src, m * sizeof(float), hipMemcpyDeviceToHost, ctx.stream),
          "download levels");
  }
  check(hipStreamSynchronize(ctx.stream), "batch download");
}

}  // namespace msim
