// Depth-averaged pressure solve on the device: operator assembly, the
// Jacobi-preconditioned conjugate gradient in FP64 (and in FP32 inside FP64
// iterative refinement), the FP64 true residual, and face-velocity
// reconstruction with the reference's lift and scaling.
//
// Layout: one layer, site = x*ny + y with y contiguous (as the reference,
// cpp/core/flow_system.hpp:64-67). The operator is stored on the structured
// grid, not compacted: per site the conductance of its lower x face (cxl) and
// lower y face (cyl), its boundary conductance (bc) and its Jacobi diagonal.
// An inactive site (solid, or fluid without conductance) has every
// coefficient zero, so its row of A is zero, its preconditioned residual is
// zero, and it stays exactly zero through the iteration: the structured
// system is the reference's active system (flow_system.hpp:500-518) with
// zero rows added, and its sums gain only exact zeros.
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



// Reduction (flow_system.hpp:276-327) for one layer: the column conductance
// is the layer height times the mobility; solid sites carry none.
__global__ void k_conductance(const uint8_t* __restrict__ solid, const float* __restrict__ mobility,
                              float* __restrict__ cond, size_t n, float dz) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cond[i] = solid[i] != 0 ? 0.0f : dz * mobility[i];
}

// Whether the exterior face `side` (0 x lower, 1 x upper, 2 y lower, 3 y
// upper) at in-plane index `along` is open, and the pressure held half a
// spacing outside it.
__device__ __forceinline__ bool open_edge(const FlowEdges& e, uint32_t side, uint32_t along,
                                          uint32_t nx, uint32_t ny, float& pressure) {
  if (e.open == nullptr) {
    pressure = (side & 1u) != 0 || e.held_zero != 0 ? 0.0f : 1.0f;
    return (side >> 1) == e.axis;
  }
  const uint32_t slot =
      along + (side == 0 ? 0u : side == 1 ? ny : side == 2 ? 2u * ny : 2u * ny + nx);
  pressure = e.held_zero != 0 ? 0.0f : e.pressure[slot];
  return e.open[slot] != 0;
}

// ShallowFlowReduction::body_flux (flow_system.hpp:395-420) per face of the
// grid of columns (x faces, then y faces): the body force integrated across
// each half cell in series, harm(m_l, m_u) (f_l/2 + f_u/2) on an interior
// face, m f through an open exterior face, nothing elsewhere. force: (fx, fy)
// per column.
__global__ __launch_bounds__(256) void k_body_flux(const float* __restrict__ cond,
                                                   const float2* __restrict__ force,
                                                   float* __restrict__ flux, uint32_t nx,
                                                   uint32_t ny, FlowEdges edges) {
  const size_t nxf = static_cast<size_t>(nx + 1) * ny;
  const size_t total = nxf + static_cast<size_t>(nx) * (ny + 1);
  const size_t f = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= total) return;
  uint32_t along, dim, across;
  size_t lower, upper;
  const bool xface = f < nxf;
  if (xface) {
    along = static_cast<uint32_t>(f / ny);
    across = static_cast<uint32_t>(f % ny);
    dim = nx;
    upper = static_cast<size_t>(along) * ny + across;
    lower = upper - ny;
  } else {
    const size_t g = f - nxf;
    across = static_cast<uint32_t>(g / (ny + 1));
    along = static_cast<uint32_t>(g % (ny + 1));
    dim = ny;
    upper = static_cast<size_t>(across) * ny + along;
    lower = upper - 1;
  }
  float value = 0.0f;
  float held = 0.0f;
  if (along > 0 && along < dim) {
    const float fl = xface ? force[lower].x : force[lower].y;
    const float fu = xface ? force[upper].x : force[upper].y;
    value = harmonic(cond[lower], cond[upper]) * (0.5f * fl + 0.5f * fu);
  } else if (open_edge(edges, (xface ? 0u : 2u) + (along == 0 ? 0u : 1u), across, nx, ny,
                       held)) {
    const size_t site = along == 0 ? upper : lower;
    value = cond[site] * (xface ? force[site].x : force[site].y);
  }
  flux[f] = value;
}

// The right-hand side of the forced system (flow_system.hpp:569-576 with
// every boundary pressure zero): per site with conductance, the divergence
// of the body flux, x then y, in float.
__global__ __launch_bounds__(256) void k_forced_rhs(const float* __restrict__ cond,
                                                    const float* __restrict__ flux,
                                                    float* __restrict__ rhs, uint32_t nx,
                                                    uint32_t ny, float dx, float dy) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(nx) * ny) return;
  float r = 0.0f;
  if (cond[i] != 0.0f) {
    const uint32_t x = static_cast<uint32_t>(i / ny), y = static_cast<uint32_t>(i % ny);
    const size_t nxf = static_cast<size_t>(nx + 1) * ny;
    const size_t yf = nxf + static_cast<size_t>(x) * (ny + 1) + y;
    r += (flux[i] - flux[i + ny]) / dx;
    r += (flux[yf] - flux[yf + 1]) / dy;
  }
  rhs[i] = r;
}

// Assembly (flow_system.hpp:549-599): diagonal, right-hand side and boundary
// conductance in the reference's float order (x lower, x upper, y lower,
// y upper), and the lower-face conductances the operator uses. Each open
// exterior face adds 2m/h^2 toward the pressure held outside it.
__global__ void k_assemble(const float* __restrict__ cond, float* __restrict__ diag,
                           float* __restrict__ rhs, float* __restrict__ bc,
                           float* __restrict__ cxl, float* __restrict__ cyl, uint32_t nx,
                           uint32_t ny, float dx, float dy, FlowEdges edges) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(nx) * ny) return;
  const uint32_t x = static_cast<uint32_t>(i / ny);
  const uint32_t y = static_cast<uint32_t>(i % ny);
  const float m = cond[i];
  cxl[i] = x > 0 ? harmonic(m, cond[i - ny]) : 0.0f;
  cyl[i] = y > 0 ? harmonic(m, cond[i - 1]) : 0.0f;
  float d = 0.0f, r = 0.0f, b = 0.0f;
  if (m != 0.0f) {
    const float isx = 1.0f / (dx * dx);
    const float isy = 1.0f / (dy * dy);
    float held = 0.0f;
    // x lower
    if (x > 0) {
      d += harmonic(m, cond[i - ny]) * isx;
    } else if (open_edge(edges, 0, y, nx, ny, held)) {
      const float c = 2.0f * m * isx;
      d += c;
      b += c;
      r += c * held;
    }
    // x upper
    if (x + 1 < nx) {
      d += harmonic(m, cond[i + ny]) * isx;
    } else if (open_edge(edges, 1, y, nx, ny, held)) {
      const float c = 2.0f * m * isx;
      d += c;
      b += c;
      r += c * held;
    }
    // y lower
    if (y > 0) {
      d += harmonic(m, cond[i - 1]) * isy;
    } else if (open_edge(edges, 2, x, nx, ny, held)) {
      const float c = 2.0f * m * isy;
      d += c;
      b += c;
      r += c * held;
    }
    // y upper
    if (y + 1 < ny) {
      d += harmonic(m, cond[i + 1]) * isy;
    } else if (open_edge(edges, 3, x, nx, ny, held)) {
      const float c = 2.0f * m * isy;
      d += c;
      b += c;
      r += c * held;
    }
  }
  diag[i] = d;
  rhs[i] = r;
  bc[i] = b;
}

// ---------------------------------------------------------------------------
// Marching stencil kernels. A block of TY threads owns TY consecutive y of a
// strip of `rows` x-rows and walks along x: each thread keeps the operand of
// its column at x-1, x and x+1 in registers (so every row is read from memory
// once), and the y neighbours come from an LDS row staged with a one-site
// halo on each side. The loads of row x+2 are issued while row x is computed,
// so a row's memory latency is hidden behind the row before it. Tiles are
// visited in XCD-contiguous order, so the halo rows a tile shares with its
// neighbours meet in one L2.
// ---------------------------------------------------------------------------

// What one thread loads of one row for the CG operator kernel.
template <typename T>
struct CgRow {
  T r, po, x;
  float d, cx, cy, b;
  T lr, lpo;  // site below the tile (thread 0)
  float ld;
  T hr, hpo;  // site above the tile (thread TY-1)
  float hd, hcy;
};

// One CG iteration k's first half, with iteration k-1's solution update and
// the direction update folded in (k = it_update + 1):
//   p_k = M^-1 r_{k-1} + beta_k p_{k-1}   (conjugate_gradient.hpp:82, cpu_flow.cpp:92-97)
//   x  += alpha_{k-1} p_{k-1}             (cpu_flow.cpp:84-89, deferred one iteration)
//   q_k = A p_k                           (DepthAveragedFlowSystem::apply, flow_system.hpp:670-695)
// and this block's part of p_k . q_k. The scalars come from the partial sums
// the previous kernels left: every block adds them in the same fixed order, so
// every block holds the same bits and no grid-wide fence is needed.
//   rho_{k-1} = r.z after iteration k-1 (the set-up's for k = 1),
//   alpha_{k-1} = rho_{k-2} / (p_{k-1} . q_{k-1}), beta_k = rho_{k-1} / rho_{k-2}.
// The stopping test of iteration k-1 (r.r <= (tol ||b||)^2, the batched
// solve's, conjugate_gradient.hpp:126-166) is made here: when it holds the
// kernel only applies the deferred solution update and records the halt.
template <typename T, int TY>
__global__ __launch_bounds__(TY) void k_cg_apply(
    const T* __restrict__ r, const float* __restrict__ diag, const T* __restrict__ p_old,
    T* __restrict__ p_new, T* __restrict__ q, T* __restrict__ xsol, const float* __restrict__ cxl,
    const float* __restrict__ cyl, const float* __restrict__ bc, uint32_t nx, uint32_t ny,
    uint32_t rows, T isx, T isy, CgState* st, CgPartials parts) {
  __shared__ T s_p[TY + 2];
  __shared__ float s_c[TY + 2];
  __shared__ double s_red[TY / kWave];
  __shared__ double s_b;
  if (st->failure != 0) return;
  const uint32_t k = st->it_update + 1;
  const uint32_t halt = st->halt;
  // A launch after the one that saw the halt does nothing. (The launch that
  // records the halt reads either 0 or k-1 here and goes on either way.)
  if (halt != 0 && halt + 1 < k) return;
  const Tile tl = tile_of<TY>(nx, ny, rows);
  const uint32_t t = threadIdx.x;
  const bool lead = blockIdx.x == 0 && t == 0;
  const double rho_km1 = fixed_sum<TY>(parts.rz + ((k - 1) & 1) * parts.p2, parts.p2, s_red, &s_b);
  const double rr_km1 = fixed_sum<TY>(parts.rr + ((k - 1) & 1) * parts.p2, parts.p2, s_red, &s_b);
  double alpha_prev = 0.0, beta = 0.0;
  bool finish = false;
  if (k >= 2) {
    const double rho_km2 = fixed_sum<TY>(parts.rz + (k & 1) * parts.p2, parts.p2, s_red, &s_b);
    const double pq_km1 = fixed_sum<TY>(parts.pq + ((k - 1) & 1) * parts.p1, parts.p1, s_red, &s_b);
    alpha_prev = rho_km2 / pq_km1;
    if (!isfinite(rr_km1)) {
      if (lead) st->failure = 3;
      return;
    }
    if (rr_km1 <= st->threshold) {
      finish = true;
    } else {
      if (!isfinite(rho_km1) || rho_km2 == 0.0) {
        if (lead) st->failure = 4;
        return;
      }
      beta = rho_km1 / rho_km2;
      if (!isfinite(beta)) {
        if (lead) st->failure = 5;
        return;
      }
    }
  }
  if (lead) {
    st->it_apply = k;
    st->rr = rr_km1;
    st->rho = rho_km1;
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
  const T zero = static_cast<T>(0);
  // Own rows load everything; the halo rows x0-1 and x1 only the operand
  // (and x1 its lower-x conductance, the upper one of the strip's last row).
  auto load = [&](uint32_t xx, bool full) {
    CgRow<T> w{};
    if (xx >= nx) return w;
    const size_t base = static_cast<size_t>(xx) * ny;
    if (tl.valid) {
      const size_t i = base + tl.y;
      w.r = r[i];
      w.po = p_old[i];
      w.d = diag[i];
      w.cx = cxl[i];
      if (full) {
        w.x = xsol[i];
        w.cy = cyl[i];
        w.b = bc[i];
      }
    }
    if (full && tl.lo_halo) {
      const size_t i = base + tl.ylo;
      w.lr = r[i];
      w.lpo = p_old[i];
      w.ld = diag[i];
    }
    if (full && tl.hi_halo) {
      const size_t i = base + tl.yhi;
      w.hr = r[i];
      w.hpo = p_old[i];
      w.hd = diag[i];
      w.hcy = cyl[i];
    }
    return w;
  };
  auto op = [&](T rv, float d, T po) -> T {
    const T z = d > 0.0f ? rv / static_cast<T>(d) : zero;
    return z + betaT * po;
  };
  double acc = 0.0;
  if (tl.x0 < tl.x1) {
    CgRow<T> cur_row = load(tl.x0, true);
    CgRow<T> next_row = load(tl.x0 + 1, tl.x0 + 1 < tl.x1);
    T prev = zero;
    if (tl.x0 > 0) {
      const CgRow<T> w = load(tl.x0 - 1, false);
      prev = op(w.r, w.d, w.po);
    }
    T cur = op(cur_row.r, cur_row.d, cur_row.po);
    for (uint32_t xx = tl.x0; xx < tl.x1; ++xx) {
      const CgRow<T> ahead = xx + 2 <= tl.x1 ? load(xx + 2, xx + 2 < tl.x1) : CgRow<T>{};
      const T nxt = op(next_row.r, next_row.d, next_row.po);
      const float cxhi = next_row.cx;
      s_p[t + 1] = tl.valid ? cur : zero;
      s_c[t + 1] = tl.valid ? cur_row.cy : 0.0f;
      if (t == 0) s_p[0] = tl.lo_halo ? op(cur_row.lr, cur_row.ld, cur_row.lpo) : zero;
      if (t == TY - 1) {
        s_p[TY + 1] = tl.hi_halo ? op(cur_row.hr, cur_row.hd, cur_row.hpo) : zero;
        s_c[TY + 1] = tl.hi_halo ? cur_row.hcy : 0.0f;
      }
      __syncthreads();
      if (tl.valid) {
        const size_t idx = static_cast<size_t>(xx) * ny + tl.y;
        const T pl = s_p[t];
        const T pu = s_p[t + 2];
        const float cylo = s_c[t + 1];
        const float cyhi = s_c[t + 2];
        T qv = static_cast<T>(cur_row.b) * cur;
        qv += (static_cast<T>(cur_row.cx) * isx) * (cur - prev);
        qv += (static_cast<T>(cxhi) * isx) * (cur - nxt);
        qv += (static_cast<T>(cylo) * isy) * (cur - pl);
        qv += (static_cast<T>(cyhi) * isy) * (cur - pu);
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

// Streaming kernels walk their range in chunks of kUnroll * kStream
// elements: each thread issues its kUnroll loads before it uses any of them.
constexpr int kUnroll = 4;

// Iteration k's second half: alpha_k = rho_{k-1} / (p_k . q_k) with the
// curvature and step failure codes (conjugate_gradient.hpp:58-66), then
// r -= alpha_k q (cpu_flow.cpp:84-89) and this block's parts of r.r and
// r.M^-1 r (cpu_flow.cpp:65-69).
template <typename T>
__global__ __launch_bounds__(kStream) void k_cg_update(T* __restrict__ r, const T* __restrict__ q,
                                                       const float* __restrict__ diag, size_t n,
                                                       CgState* st, CgPartials parts) {
  __shared__ double s_red[kStream / kWave];
  __shared__ double s_b;
  if (st->failure != 0) return;
  const uint32_t k = st->it_apply;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (st->halt != 0) {
    if (lead) st->it_update = k;
    return;
  }
  const double rho_km1 =
      fixed_sum<kStream>(parts.rz + ((k - 1) & 1) * parts.p2, parts.p2, s_red, &s_b);
  const double curvature = fixed_sum<kStream>(parts.pq + (k & 1) * parts.p1, parts.p1, s_red, &s_b);
  if (!isfinite(curvature) || curvature <= 0.0) {
    if (lead) st->failure = 1;
    return;
  }
  const double a = rho_km1 / curvature;
  if (!isfinite(a)) {
    if (lead) st->failure = 2;
    return;
  }
  if (lead) {
    st->it_update = k;
    st->curvature = curvature;
  }
  const T alpha = static_cast<T>(a);
  double rr = 0.0, rz = 0.0;
  const size_t chunk = static_cast<size_t>(kStream) * kUnroll;
  for (size_t base = static_cast<size_t>(blockIdx.x) * chunk + threadIdx.x; base < n;
       base += static_cast<size_t>(gridDim.x) * chunk) {
    T rv[kUnroll], qv[kUnroll];
    float dv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = base + static_cast<size_t>(u) * kStream;
      const bool in = i < n;
      rv[u] = in ? r[i] : static_cast<T>(0);
      qv[u] = in ? q[i] : static_cast<T>(0);
      dv[u] = in ? diag[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = base + static_cast<size_t>(u) * kStream;
      const T ri = rv[u] - alpha * qv[u];
      if (i < n) r[i] = ri;
      const T z = dv[u] > 0.0f ? ri / static_cast<T>(dv[u]) : static_cast<T>(0);
      rr += static_cast<double>(ri) * static_cast<double>(ri);
      rz += static_cast<double>(ri) * static_cast<double>(z);
    }
  }
  const double prr = block_sum<kStream>(rr, s_red);
  const double prz = block_sum<kStream>(rz, s_red);
  if (threadIdx.x == 0) {
    parts.rr[(k & 1) * parts.p2 + blockIdx.x] = prr;
    parts.rz[(k & 1) * parts.p2 + blockIdx.x] = prz;
  }
}

// x = 0, r = b (or a scaled FP64 residual), p_old = 0; r.r and r.M^-1 r into
// the iteration-0 partials (cpu_flow.cpp:74-80; conjugate_gradient.hpp:49-55);
// the last block also resets the bookkeeping and sets the threshold.
template <typename T, typename S>
__global__ __launch_bounds__(kStream) void k_cg_init(const S* __restrict__ source, double scale,
                                                     T* __restrict__ r, T* __restrict__ xsol,
                                                     T* __restrict__ p_old,
                                                     const float* __restrict__ diag, size_t n,
                                                     double tolerance, CgState* st,
                                                     CgPartials parts, double* partials) {
  __shared__ double s_red[kStream / kWave];
  __shared__ bool s_last;
  double rr = 0.0, rz = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kStream;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kStream + threadIdx.x; i < n; i += stride) {
    const T ri = static_cast<T>(static_cast<double>(source[i]) * scale);
    r[i] = ri;
    xsol[i] = static_cast<T>(0);
    p_old[i] = static_cast<T>(0);
    const float d = diag[i];
    const T z = d > 0.0f ? ri / static_cast<T>(d) : static_cast<T>(0);
    rr += static_cast<double>(ri) * static_cast<double>(ri);
    rz += static_cast<double>(ri) * static_cast<double>(z);
  }
  double part[2] = {block_sum<kStream>(rr, s_red), block_sum<kStream>(rz, s_red)};
  if (threadIdx.x == 0) {
    parts.rr[blockIdx.x] = part[0];
    parts.rz[blockIdx.x] = part[1];
  }
  if (publish_partials<2>(part, partials, &st->ticket[kTicketInit], blockIdx.x, gridDim.x,
                          &s_last)) {
    double total[2];
    sum_partials<kStream, 2>(partials, gridDim.x, total, s_red);
    if (threadIdx.x == 0) {
      st->ticket[kTicketInit] = 0;
      st->rr0 = total[0];
      st->rr = total[0];
      st->rho = total[1];
      const double bound = tolerance * sqrt(total[0]);
      st->threshold = bound * bound;
      st->alpha = 0.0;
      st->beta = 0.0;
      st->iterations = 0;
      st->failure = 0;
      st->it_update = 0;
      st->it_apply = 0;
      st->halt = 0;
      st->converged = total[0] == 0.0 ? 1u : 0u;
      st->done = total[0] == 0.0 ? 1u : 0u;
    }
  }
}

// FP64 refinement update: x += scale * e.
__global__ void k_refine(double* __restrict__ xsol, const float* __restrict__ e, double scale,
                         size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    xsol[i] += scale * static_cast<double>(e[i]);
}

// FP64 correction: x += e.
__global__ void k_add_correction(double* __restrict__ xsol, const double* __restrict__ e,
                                 size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    xsol[i] += e[i];
}

struct ResRow {
  double x, lx, hx;
  float cx, cy, b, rhs, hcy;
};

// FP64 operator on x (flow_system.hpp:670-695): out = A x (mode 0) or
// out = b - A x (mode 1, the true residual), with ||out||^2 and ||b||^2.
template <int TY>
__global__ __launch_bounds__(TY) void k_residual(
    const double* __restrict__ xin, const float* __restrict__ rhs, double* __restrict__ out,
    const float* __restrict__ cxl, const float* __restrict__ cyl, const float* __restrict__ bc,
    uint32_t nx, uint32_t ny, uint32_t rows, double isx, double isy, int mode, CgState* st,
    double* partials) {
  __shared__ double s_p[TY + 2];
  __shared__ float s_c[TY + 2];
  __shared__ double s_red[TY / kWave];
  __shared__ bool s_last;
  const Tile tl = tile_of<TY>(nx, ny, rows);
  const uint32_t t = threadIdx.x;
  auto load = [&](uint32_t xx, bool full) {
    ResRow w{};
    if (xx >= nx) return w;
    const size_t base = static_cast<size_t>(xx) * ny;
    if (tl.valid) {
      const size_t i = base + tl.y;
      w.x = xin[i];
      w.cx = cxl[i];
      if (full) {
        w.cy = cyl[i];
        w.b = bc[i];
        w.rhs = rhs[i];
      }
    }
    if (full && tl.lo_halo) w.lx = xin[base + tl.ylo];
    if (full && tl.hi_halo) {
      w.hx = xin[base + tl.yhi];
      w.hcy = cyl[base + tl.yhi];
    }
    return w;
  };
  double acc = 0.0, bb = 0.0;
  if (tl.x0 < tl.x1) {
    ResRow cur_row = load(tl.x0, true);
    ResRow next_row = load(tl.x0 + 1, tl.x0 + 1 < tl.x1);
    double prev = tl.x0 > 0 ? load(tl.x0 - 1, false).x : 0.0;
    for (uint32_t xx = tl.x0; xx < tl.x1; ++xx) {
      const ResRow ahead = xx + 2 <= tl.x1 ? load(xx + 2, xx + 2 < tl.x1) : ResRow{};
      const double cur = cur_row.x;
      const double nxt = next_row.x;
      s_p[t + 1] = tl.valid ? cur : 0.0;
      s_c[t + 1] = tl.valid ? cur_row.cy : 0.0f;
      if (t == 0) s_p[0] = tl.lo_halo ? cur_row.lx : 0.0;
      if (t == TY - 1) {
        s_p[TY + 1] = tl.hi_halo ? cur_row.hx : 0.0;
        s_c[TY + 1] = tl.hi_halo ? cur_row.hcy : 0.0f;
      }
      __syncthreads();
      if (tl.valid) {
        const size_t idx = static_cast<size_t>(xx) * ny + tl.y;
        double qv = static_cast<double>(cur_row.b) * cur;
        qv += (static_cast<double>(cur_row.cx) * isx) * (cur - prev);
        qv += (static_cast<double>(next_row.cx) * isx) * (cur - nxt);
        qv += (static_cast<double>(s_c[t + 1]) * isy) * (cur - s_p[t]);
        qv += (static_cast<double>(s_c[t + 2]) * isy) * (cur - s_p[t + 2]);
        const double b = static_cast<double>(cur_row.rhs);
        const double v = mode == 0 ? qv : b - qv;
        out[idx] = v;
        acc += v * v;
        bb += b * b;
      }
      __syncthreads();
      prev = cur;
      cur_row = next_row;
      next_row = ahead;
    }
  }
  double part[2] = {block_sum<TY>(acc, s_red), block_sum<TY>(bb, s_red)};
  if (publish_partials<2>(part, partials, &st->ticket[kTicketResidual], blockIdx.x, gridDim.x,
                          &s_last)) {
    double total[2];
    sum_partials<TY, 2>(partials, gridDim.x, total, s_red);
    if (t == 0) {
      st->ticket[kTicketResidual] = 0;
      st->true_rr = total[0];
      st->bb = total[1];
    }
  }
}

// Face velocities from the pressure (flow_system.hpp:711-736) lifted to the
// one-layer grid (424-445): interior faces -harm(m_l, m_u) (p_u - p_l) / h,
// faces of the flow axis' ends 2m/h toward the boundary pressure held half a
// spacing outside, faces next to a solid site zero, all divided by the layer
// height. Accumulates the open-inlet sum and the peak for scale_velocity
// (1368-1415).
__global__ __launch_bounds__(kStream) void k_velocity(
    const double* __restrict__ p, const float* __restrict__ cond, const uint8_t* __restrict__ solid,
    float* __restrict__ faces, uint32_t nx, uint32_t ny, float dx, float dy, float dz,
    FlowEdges edges, const float* __restrict__ body_flux, CgState* st, double* partials) {
  __shared__ double s_red[kStream / kWave];
  __shared__ bool s_last;
  const size_t nxf = static_cast<size_t>(nx + 1) * ny;
  const size_t total = nxf + static_cast<size_t>(nx) * (ny + 1);
  double inlet = 0.0, count = 0.0, peak = 0.0, bad = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kStream;
  for (size_t f = static_cast<size_t>(blockIdx.x) * kStream + threadIdx.x; f < total;
       f += stride) {
    uint32_t comp, along, dim, across;
    size_t lower, upper;
    if (f < nxf) {
      comp = 0;
      const uint32_t fx = static_cast<uint32_t>(f / ny), y = static_cast<uint32_t>(f % ny);
      along = fx;
      across = y;
      dim = nx;
      upper = static_cast<size_t>(fx) * ny + y;
      lower = upper - ny;  // meaningful only when fx > 0
    } else {
      comp = 1;
      const size_t g = f - nxf;
      const uint32_t x = static_cast<uint32_t>(g / (ny + 1)), fy = static_cast<uint32_t>(g % (ny + 1));
      along = fy;
      across = x;
      dim = ny;
      upper = static_cast<size_t>(x) * ny + fy;
      lower = upper - 1;
    }
    const bool has_lower = along > 0;
    const bool has_upper = along < dim;
    const float sp = comp == 0 ? dx : dy;
    double value;
    if (has_lower && has_upper) {
      const float fm = harmonic(cond[lower], cond[upper]);
      value = -static_cast<double>(fm) * (p[upper] - p[lower]) / static_cast<double>(sp);
    } else {
      const size_t site = has_upper ? upper : lower;
      const float m = cond[site];
      float fc = 0.0f, fp = 0.0f, held = 0.0f;
      if (m != 0.0f && open_edge(edges, 2 * comp + (has_upper ? 0u : 1u), across, nx, ny, held)) {
        fc = (has_upper ? 2.0f : -2.0f) * m / sp;
        fp = held;
      }
      value = static_cast<double>(fc) * (static_cast<double>(fp) - p[site]);
    }
    float flux = static_cast<float>(value);
    // A forced solve's correction adds to the body flux (flow_system.hpp:1438-1439).
    if (body_flux != nullptr) flux = body_flux[f] + flux;
    float v = 0.0f;
    if (!((has_lower && solid[lower] != 0) || (has_upper && solid[upper] != 0))) v = flux / dz;
    faces[f] = v;
    if (!isfinite(v)) bad = 1.0;
    peak = fmax(peak, static_cast<double>(fabsf(v)));
    if (comp == edges.axis && along == 0 && solid[upper] == 0 && cond[upper] > 0.0f) {
      inlet += static_cast<double>(v);
      count += 1.0;
    }
  }
  double part[4] = {block_sum<kStream>(inlet, s_red), block_sum<kStream>(count, s_red),
                    block_max<kStream>(peak, s_red), block_max<kStream>(bad, s_red)};
  if (publish_partials<4>(part, partials, &st->ticket[kTicketVelocity], blockIdx.x, gridDim.x,
                          &s_last)) {
    double sums[2], maxes[2];
    double tmp[4];
    sum_partials<kStream, 2>(partials, gridDim.x, sums, s_red);
    max_partials<kStream, 2>(partials + 2 * static_cast<size_t>(gridDim.x), gridDim.x, maxes,
                             s_red);
    (void)tmp;
    if (threadIdx.x == 0) {
      st->ticket[kTicketVelocity] = 0;
      st->inlet_sum = sums[0];
      st->inlet_count = sums[1];
      st->peak = maxes[0];
      st->nonfinite = maxes[1] > 0.0 ? 1u : 0u;
    }
  }
}

// scaled = v * factor + driven (flow_system.hpp:1401-1407; driven is zero
// without a body force), the scaled peak, and whether any is not finite.
__global__ __launch_bounds__(kStream) void k_scale(float* __restrict__ faces, size_t total,
                                                   float factor, const float* __restrict__ driven,
                                                   CgState* st, double* partials) {
  __shared__ double s_red[kStream / kWave];
  __shared__ bool s_last;
  double peak = 0.0, bad = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kStream;
  for (size_t f = static_cast<size_t>(blockIdx.x) * kStream + threadIdx.x; f < total;
       f += stride) {
    const float v = faces[f] * factor + (driven != nullptr ? driven[f] : 0.0f);
    faces[f] = v;
    if (!isfinite(v)) bad = 1.0;
    peak = fmax(peak, static_cast<double>(fabsf(v)));
  }
  double part[2] = {block_max<kStream>(peak, s_red), block_max<kStream>(bad, s_red)};
  if (publish_partials<2>(part, partials, &st->ticket[kTicketScale], blockIdx.x, gridDim.x,
                          &s_last)) {
    double m[2];
    max_partials<kStream, 2>(partials, gridDim.x, m, s_red);
    if (threadIdx.x == 0) {
      st->ticket[kTicketScale] = 0;
      st->scaled_peak = m[0];
      st->nonfinite = m[1] > 0.0 ? 1u : 0u;
    }
  }
}

}  // namespace

// --------------------------------------------------------------------------
// Host orchestration
// --------------------------------------------------------------------------

uint32_t stream_blocks(size_t n) {
  const size_t per_block = static_cast<size_t>(kStream) * kUnroll;
  const size_t blocks = (n + per_block - 1) / per_block;
  return static_cast<uint32_t>(std::clamp<size_t>(blocks, 1, 2048));
}

// Blocks of the Krylov streaming kernels (set-up and update). Every operator
// block re-adds their partials, so they are kept few: 512 blocks of 256
// threads (two per CU) with 4-way unrolled loads keep HBM busy.
uint32_t krylov_blocks(size_t n) {
  const size_t per_block = static_cast<size_t>(kStream) * kUnroll;
  const size_t blocks = (n + per_block - 1) / per_block;
  return static_cast<uint32_t>(std::clamp<size_t>(blocks, 1, 512));
}

struct MarchShape {
  int ty;          // threads per block (64, 128 or 256)
  uint32_t rows;   // x-rows per block
  uint32_t blocks;
};

MarchShape march_shape(uint32_t nx, uint32_t ny) {
  MarchShape s{};
  s.ty = ny <= 64 ? 64 : (ny <= 128 ? 128 : 256);
  const uint32_t nty = (ny + s.ty - 1) / s.ty;
  // Aim for ~1024 blocks (four per CU) while keeping a strip long enough that
  // its two halo rows cost little; small grids get shorter strips.
  uint32_t rows = 32;
  while (rows > 4 && static_cast<uint64_t>(nty) * ((nx + rows - 1) / rows) < 1024) rows /= 2;
  s.rows = rows;
  s.blocks = nty * ((nx + rows - 1) / rows);
  return s;
}

CgPartials krylov_partials(Context& ctx) {
  auto& f = ctx.flow;
  CgPartials parts{};
  parts.pq = f.pq.data();
  parts.rz = f.rz.data();
  parts.rr = f.rr.data();
  parts.p1 = march_shape(f.nx, f.ny).blocks;
  parts.p2 = krylov_blocks(f.sites());
  return parts;
}

template <typename T>
void launch_apply(Context& ctx, const T* r, const T* p_old, T* p_new, T* q, T* x) {
  auto& f = ctx.flow;
  const MarchShape s = march_shape(f.nx, f.ny);
  const T isx = static_cast<T>(1.0 / (static_cast<double>(f.dx) * f.dx));
  const T isy = static_cast<T>(1.0 / (static_cast<double>(f.dy) * f.dy));
  const CgPartials parts = krylov_partials(ctx);
  switch (s.ty) {
    case 64:
      k_cg_apply<T, 64><<<s.blocks, 64, 0, ctx.stream>>>(r, f.diag.data(), p_old, p_new, q, x,
                                                         f.cxl.data(), f.cyl.data(), f.bc.data(),
                                                         f.nx, f.ny, s.rows, isx, isy,
                                                         f.state.data(), parts);
      break;
    case 128:
      k_cg_apply<T, 128><<<s.blocks, 128, 0, ctx.stream>>>(
          r, f.diag.data(), p_old, p_new, q, x, f.cxl.data(), f.cyl.data(), f.bc.data(), f.nx,
          f.ny, s.rows, isx, isy, f.state.data(), parts);
      break;
    default:
      k_cg_apply<T, 256><<<s.blocks, 256, 0, ctx.stream>>>(
          r, f.diag.data(), p_old, p_new, q, x, f.cxl.data(), f.cyl.data(), f.bc.data(), f.nx,
          f.ny, s.rows, isx, isy, f.state.data(), parts);
  }
  check(hipGetLastError(), "launch k_cg_apply");
}

template <typename T>
void launch_update(Context& ctx, T* r, const T* q) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  k_cg_update<T><<<krylov_blocks(n), kStream, 0, ctx.stream>>>(r, q, f.diag.data(), n,
                                                               f.state.data(),
                                                               krylov_partials(ctx));
  check(hipGetLastError(), "launch k_cg_update");
}

template <typename T, typename S>
void launch_init(Context& ctx, const S* source, double scale, T* r, T* x, T* p_old,
                 double tolerance) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  k_cg_init<T, S><<<krylov_blocks(n), kStream, 0, ctx.stream>>>(
      source, scale, r, x, p_old, f.diag.data(), n, tolerance, f.state.data(),
      krylov_partials(ctx), f.partials.data());
  check(hipGetLastError(), "launch k_cg_init");
}

void launch_residual(Context& ctx, const double* x, double* out, int mode) {
  auto& f = ctx.flow;
  const MarchShape s = march_shape(f.nx, f.ny);
  const double isx = 1.0 / (static_cast<double>(f.dx) * f.dx);
  const double isy = 1.0 / (static_cast<double>(f.dy) * f.dy);
  switch (s.ty) {
    case 64:
      k_residual<64><<<s.blocks, 64, 0, ctx.stream>>>(x, f.rhs.data(), out, f.cxl.data(),
                                                      f.cyl.data(), f.bc.data(), f.nx, f.ny,
                                                      s.rows, isx, isy, mode, f.state.data(),
                                                      f.partials.data());
      break;
    case 128:
      k_residual<128><<<s.blocks, 128, 0, ctx.stream>>>(x, f.rhs.data(), out, f.cxl.data(),
                                                        f.cyl.data(), f.bc.data(), f.nx, f.ny,
                                                        s.rows, isx, isy, mode, f.state.data(),
                                                        f.partials.data());
      break;
    default:
      k_residual<256><<<s.blocks, 256, 0, ctx.stream>>>(x, f.rhs.data(), out, f.cxl.data(),
                                                        f.cyl.data(), f.bc.data(), f.nx, f.ny,
                                                        s.rows, isx, isy, mode, f.state.data(),
                                                        f.partials.data());
  }
  check(hipGetLastError(), "launch k_residual");
}

void read_state(Context& ctx) {
  check(hipMemcpyAsync(ctx.flow.host_state, ctx.flow.state.data(), sizeof(CgState),
                       hipMemcpyDeviceToHost, ctx.stream),
        "read solve state");
  check(hipStreamSynchronize(ctx.stream), "synchronize solve");
}

int failure_status(uint32_t failure) {
  return failure == 0 ? MSIM_OK : MSIM_CG_CURVATURE + static_cast<int>(failure) - 1;
}

const char* failure_text(uint32_t failure) {
  switch (failure) {
    case 1: return "encountered non-positive curvature";
    case 2: return "produced a non-finite step";
    case 3: return "produced a non-finite residual";
    case 4: return "encountered a preconditioner breakdown";
    case 5: return "produced a non-finite direction";
    default: return "failed";
  }
}

// Runs Jacobi-PCG iterations of precision T on the prepared system from the
// state launch_init left, until the stopping test holds, a failure, or
// `budget` iterations (conjugate_gradient.hpp:126-166: the host looks at the
// device's state once per `interval` iterations). The operator kernel of
// iteration k+1 makes iteration k's stopping test; one more operator launch
// after the budget makes the last test. Returns the iterations taken (the
// iteration that met the test, or the budget). Throws on a CG failure.
template <typename T>
uint32_t run_cg(Context& ctx, T* r, T* x, T* q, T* pa, T* pb, uint32_t budget, uint32_t interval,
                bool* converged, double* relative, KernelTimes* times) {
  auto& f = ctx.flow;
  T* buf[2] = {pa, pb};
  *converged = false;
  if (times != nullptr) {
    while (ctx.probes.size() < 3 * static_cast<size_t>(interval) + 3) {
      hipEvent_t e;
      check(hipEventCreate(&e), "hipEventCreate");
      ctx.probes.push_back(e);
    }
  }
  read_state(ctx);
  if (f.host_state->done != 0) {  // r.r of the right-hand side is zero
    *converged = true;
    *relative = 0.0;
    return 0;
  }
  const double rr0 = f.host_state->rr0;
  uint32_t launched = 0;  // iterations launched
  bool tested_all = false;
  while (true) {
    // A final operator launch tests the last iteration of the budget.
    const bool last_test = launched >= budget;
    const uint32_t batch = last_test ? 0 : std::min(interval, budget - launched);
    for (uint32_t j = 0; j < batch; ++j) {
      const uint32_t it = launched + j + 1;  // 1-based iteration
      if (times != nullptr) check(hipEventRecord(ctx.probes[3 * j], ctx.stream), "event");
      launch_apply<T>(ctx, r, buf[(it + 1) % 2], buf[it % 2], q, x);
      if (times != nullptr) check(hipEventRecord(ctx.probes[3 * j + 1], ctx.stream), "event");
      launch_update<T>(ctx, r, q);
      if (times != nullptr) check(hipEventRecord(ctx.probes[3 * j + 2], ctx.stream), "event");
    }
    if (last_test) {
      const uint32_t it = launched + 1;
      launch_apply<T>(ctx, r, buf[(it + 1) % 2], buf[it % 2], q, x);
      tested_all = true;
    }
    const uint32_t before = launched;
    launched += batch;
    read_state(ctx);
    const CgState& s = *f.host_state;
    if (times != nullptr && batch > 0) {
      // Iterations of this batch that did work (later launches return at once).
      const uint32_t last_worked = s.halt != 0 ? s.halt : s.it_update;
      const uint32_t worked =
          std::min(batch, last_worked > before ? last_worked - before : 0u);
      for (uint32_t j = 0; j < worked; ++j) {
        float a = 0.0f, b = 0.0f;
        check(hipEventElapsedTime(&a, ctx.probes[3 * j], ctx.probes[3 * j + 1]), "event time");
        check(hipEventElapsedTime(&b, ctx.probes[3 * j + 1], ctx.probes[3 * j + 2]),
              "event time");
        times->apply_seconds += a * 1.0e-3;
        times->update_seconds += b * 1.0e-3;
        times->apply_launches += 1;
        times->update_launches += 1;
      }
    }
    if (s.failure != 0) {
      throw Error(failure_status(s.failure),
                  std::string("HIP depth-averaged flow conjugate gradient ") +
                      failure_text(s.failure));
    }
    *relative = rr0 > 0.0 ? std::sqrt(std::max(0.0, s.rr)) / std::sqrt(rr0) : 0.0;
    if (s.halt != 0) {
      *converged = true;
      return s.halt;
    }
    if (tested_all) return budget;
  }
}

uint32_t jacobi_solve(Context& ctx, double tolerance, uint32_t budget, uint32_t interval,
                      bool* converged, double* relative, KernelTimes* times) {
  auto& f = ctx.flow;
  launch_init<double, float>(ctx, f.rhs.data(), 1.0, f.r.data(), f.x.data(), f.p0.data(),
                             tolerance);
  const uint32_t iterations = run_cg<double>(ctx, f.r.data(), f.x.data(), f.q.data(),
                                             f.p0.data(), f.p1.data(), budget, interval,
                                             converged, relative, times);
  launch_residual(ctx, f.x.data(), f.res.data(), 1);
  read_state(ctx);
  return iterations;
}

// FP64 Jacobi-PCG continued from the solution x already in f.x (the
// learned solver's fallback): each pass solves A e = b - A x by the
// reference's iteration (conjugate_gradient.hpp:126-166) from e = 0, to the
// fraction of ||b - A x|| that would meet the target if the recursive residual
// held, adds e to x and recomputes the FP64 true residual, which alone decides
// convergence. A non-finite iterate restarts from x = 0.
uint32_t jacobi_continue(Context& ctx, double tolerance, uint32_t budget, uint32_t interval,
                         bool* converged, double* relative, KernelTimes* times) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  f.xc.reserve(n);
  launch_residual(ctx, f.x.data(), f.res.data(), 1);
  read_state(ctx);
  const double bnorm = std::sqrt(f.host_state->bb);
  double rnorm = std::sqrt(f.host_state->true_rr);
  if (!std::isfinite(rnorm) || rnorm > bnorm) {
    check(hipMemsetAsync(f.x.data(), 0, n * sizeof(double), ctx.stream), "clear x");
    launch_residual(ctx, f.x.data(), f.res.data(), 1);
    read_state(ctx);
    rnorm = std::sqrt(f.host_state->true_rr);
  }
  uint32_t iterations = 0;
  *converged = bnorm == 0.0 || rnorm <= tolerance * bnorm;
  for (uint32_t pass = 0; !*converged && iterations < budget && pass < 16; ++pass) {
    // Half the needed reduction as the recursive target: the true residual
    // of an FP64 solve stays within a small factor of the recursive one.
    const double inner = std::max(0.5 * tolerance * bnorm / rnorm, 1.0e-15);
    launch_init<double, double>(ctx, f.res.data(), 1.0, f.r.data(), f.xc.data(), f.p0.data(),
                                inner);
    bool inner_converged = false;
    double inner_relative = 0.0;
    const uint32_t taken = run_cg<double>(ctx, f.r.data(), f.xc.data(), f.q.data(), f.p0.data(),
                                          f.p1.data(), budget - iterations, interval,
                                          &inner_converged, &inner_relative, times);
    iterations += taken;
    k_add_correction<<<stream_blocks(n), kStream, 0, ctx.stream>>>(f.x.data(), f.xc.data(), n);
    check(hipGetLastError(), "launch k_add_correction");
    launch_residual(ctx, f.x.data(), f.res.data(), 1);
    read_state(ctx);
    rnorm = std::sqrt(f.host_state->true_rr);
    if (!std::isfinite(rnorm)) {
      throw Error(MSIM_CG_RESIDUAL, "HIP Jacobi-PCG fallback produced a non-finite residual");
    }
    *converged = rnorm <= tolerance * bnorm;
    if (!inner_converged && taken == 0) break;
  }
  *relative = bnorm > 0.0 ? rnorm / bnorm : 0.0;
  return iterations;
}

void launch_assembly(hipStream_t stream, const uint8_t* solid, const float* mobility, float* cond,
                     float* diag, float* rhs, float* bc, float* cxl, float* cyl, uint32_t nx,
                     uint32_t ny, float dx, float dy, float dz, FlowEdges edges) {
  const size_t n = static_cast<size_t>(nx) * ny;
  const uint32_t blocks = static_cast<uint32_t>((n + 255) / 256);
  k_conductance<<<blocks, 256, 0, stream>>>(solid, mobility, cond, n, dz);
  check(hipGetLastError(), "launch k_conductance");
  k_assemble<<<blocks, 256, 0, stream>>>(cond, diag, rhs, bc, cxl, cyl, nx, ny, dx, dy, edges);
  check(hipGetLastError(), "launch k_assemble");
}

void flow_hold_pressures(Context& ctx, const float* pressure_host) {
  auto& f = ctx.flow;
  if (!f.ported) throw Error(MSIM_INVALID_ARGUMENT, "the prepared flow has no ports");
  const size_t slots = 2 * (static_cast<size_t>(f.nx) + f.ny);
  check(hipMemcpyAsync(f.edge_pressure.data(), pressure_host, slots * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload port pressures");
  // Only the right-hand side changes: the operator and the multilevel
  // hierarchy built on it stay valid.
  const size_t n = f.sites();
  k_assemble<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, ctx.stream>>>(
      f.cond.data(), f.diag.data(), f.rhs.data(), f.bc.data(), f.cxl.data(), f.cyl.data(), f.nx,
      f.ny, f.dx, f.dy, f.edges());
  check(hipGetLastError(), "launch k_assemble");
  f.solved = false;
}

void flow_assemble(Context& ctx, const uint8_t* solid_host, const float* mobility_host) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  f.solid.reserve(n);
  f.mobility.reserve(n);
  f.cond.reserve(n);
  for (auto* b : {&f.diag, &f.rhs, &f.bc, &f.cxl, &f.cyl}) b->reserve(n);
  for (auto* b : {&f.x, &f.r, &f.q, &f.p0, &f.p1, &f.res}) b->reserve(n);
  f.faces.reserve(f.x_faces() + f.y_faces());
  // Partial sums: up to 4 values per block of any kernel here.
  f.partials.reserve(4 * static_cast<size_t>(std::max<uint32_t>(march_shape(f.nx, f.ny).blocks, 4096)));
  f.pq.reserve(2 * static_cast<size_t>(march_shape(f.nx, f.ny).blocks));
  f.rz.reserve(2 * static_cast<size_t>(krylov_blocks(n)));
  f.rr.reserve(2 * static_cast<size_t>(krylov_blocks(n)));
  f.state.reserve(1);
  if (f.host_state == nullptr) {
    void* ptr = nullptr;
    check(hipHostMalloc(&ptr, sizeof(CgState), hipHostMallocDefault), "hipHostMalloc");
    f.host_state = static_cast<CgState*>(ptr);
  }
  check(hipMemsetAsync(f.state.data(), 0, sizeof(CgState), ctx.stream), "clear state");
  check(hipMemcpyAsync(f.solid.data(), solid_host, n, hipMemcpyHostToDevice, ctx.stream),
        "upload mask");
  check(hipMemcpyAsync(f.mobility.data(), mobility_host, n * sizeof(float), hipMemcpyHostToDevice,
                       ctx.stream),
        "upload mobility");
  launch_assembly(ctx.stream, f.solid.data(), f.mobility.data(), f.cond.data(), f.diag.data(),
                  f.rhs.data(), f.bc.data(), f.cxl.data(), f.cyl.data(), f.nx, f.ny, f.dx, f.dy,
                  f.dz, f.edges());
  check(hipStreamSynchronize(ctx.stream), "assemble");
  f.prepared = true;
  f.solved = false;
  ++f.generation;
  f.mg.ready = false;
}

void flow_apply(Context& ctx, const double* x, double* y) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  check(hipMemcpyAsync(f.x.data(), x, n * sizeof(double), hipMemcpyHostToDevice, ctx.stream),
        "upload vector");
  launch_residual(ctx, f.x.data(), f.res.data(), 0);
  check(hipMemcpyAsync(y, f.res.data(), n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream),
        "download vector");
  check(hipStreamSynchronize(ctx.stream), "apply");
  f.solved = false;
}

namespace {

// One pressure solve of the prepared system's right-hand side by the chosen
// form, from x = 0, leaving the solution in f.x and its FP64 true residual in
// the state: iterations, both residuals, refinements and whether the learned
// solver fell back go to `report`.
void solve_pressure(Context& ctx, const msim_flow_params& params, msim_flow_report& report,
                    KernelTimes* times) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  const uint32_t interval = params.check_interval == 0 ? 64 : params.check_interval;
  uint32_t iterations = 0;
  bool converged = false;
  double relative = 0.0;
  if (params.solver == MSIM_SOLVER_JACOBI_FP64) {
    iterations = jacobi_solve(ctx, params.relative_tolerance, params.max_iterations, interval,
                              &converged, &relative, times);
    if (!converged) {
      throw Error(MSIM_NOT_CONVERGED,
                  "HIP depth-averaged flow conjugate gradient did not converge: relative " +
                      std::to_string(relative));
    }
  } else if (params.solver == MSIM_SOLVER_MIXED) {
    for (auto* b : {&f.x32, &f.r32, &f.q32, &f.p032, &f.p132}) b->reserve(n);
    const double inner = params.inner_tolerance > 0.0 ? params.inner_tolerance : 1.0e-4;
    // x = 0, res = b, ||b||.
    check(hipMemsetAsync(f.x.data(), 0, n * sizeof(double), ctx.stream), "clear x");
    launch_residual(ctx, f.x.data(), f.res.data(), 1);
    read_state(ctx);
    const double bnorm = std::sqrt(f.host_state->bb);
    double rnorm = std::sqrt(f.host_state->true_rr);
    uint32_t refinements = 0;
    if (bnorm == 0.0) {
      converged = true;
    }
    while (!converged) {
      if (rnorm <= params.relative_tolerance * bnorm) {
        converged = true;
        break;
      }
      if (iterations >= params.max_iterations || refinements >= 64) break;
      // Inner FP32 solve of A e = res / ||res|| to `inner`.
      launch_init<float, double>(ctx, f.res.data(), 1.0 / rnorm, f.r32.data(), f.x32.data(),
                                 f.p032.data(), inner);
      bool inner_converged = false;
      double inner_relative = 0.0;
      const uint32_t taken = run_cg<float>(ctx, f.r32.data(), f.x32.data(), f.q32.data(),
                                           f.p032.data(), f.p132.data(),
                                           params.max_iterations - iterations, interval,
                                           &inner_converged, &inner_relative, times);
      iterations += taken;
      ++refinements;
      k_refine<<<stream_blocks(n), kStream, 0, ctx.stream>>>(f.x.data(), f.x32.data(), rnorm, n);
      check(hipGetLastError(), "launch k_refine");
      launch_residual(ctx, f.x.data(), f.res.data(), 1);
      read_state(ctx);
      const double next = std::sqrt(f.host_state->true_rr);
      if (!std::isfinite(next)) {
        throw Error(MSIM_CG_RESIDUAL, "HIP mixed-precision refinement produced a non-finite residual");
      }
      rnorm = next;
      if (!inner_converged && taken == 0) break;
    }
    report.refinements = refinements;
    relative = bnorm > 0.0 ? rnorm / bnorm : 0.0;
    if (!converged) {
      throw Error(MSIM_NOT_CONVERGED,
                  "HIP mixed-precision depth-averaged flow did not converge: relative " +
                      std::to_string(relative));
    }
  } else if (params.solver == MSIM_SOLVER_MULTIGRID) {
    iterations = mg_solve(ctx, params, report, times, false, &converged);
    relative = report.relative_residual;
    if (!converged) {
      throw Error(MSIM_NOT_CONVERGED,
                  "HIP multilevel depth-averaged flow did not converge: relative " +
                      std::to_string(relative));
    }
  } else if (params.solver == MSIM_SOLVER_LEARNED) {
    if (!ctx.learned.loaded) {
      throw Error(MSIM_INVALID_ARGUMENT, "the learned smoother has no weights: set them first");
    }
    // Learned flexible CG up to the cap; FP64 Jacobi-PCG after, continued
    // from the flexible-CG iterate and judged by the FP64 true residual.
    const uint32_t cap = std::min(params.max_iterations, params.fallback_iterations == 0
                                                             ? 200u
                                                             : params.fallback_iterations);
    msim_flow_params capped = params;
    capped.max_iterations = cap;
    bool fell = false;
    try {
      iterations = mg_solve(ctx, capped, report, times, true, &converged);
      relative = report.relative_residual;
      fell = !converged;
    } catch (const Error& e) {
      if (e.code() < MSIM_CG_CURVATURE || e.code() > MSIM_CG_DIRECTION) throw;
      fell = true;
      iterations = cap;
    }
    if (fell) {
      report.fell_back = 1;
      uint32_t extra = 0;
      if (params.max_iterations > iterations) {
        extra = jacobi_continue(ctx, params.relative_tolerance,
                                params.max_iterations - iterations, interval, &converged,
                                &relative, times);
      }
      iterations += extra;
      if (!converged) {
        throw Error(MSIM_NOT_CONVERGED,
                    "HIP learned depth-averaged flow did not converge after the Jacobi-PCG "
                    "fallback: relative " +
                        std::to_string(relative));
      }
    }
  } else {
    throw Error(MSIM_UNSUPPORTED, "this pressure solver form is not available");
  }
  const CgState& s = *f.host_state;
  report.iterations = iterations;
  report.relative_residual = relative;
  report.true_relative_residual = s.bb > 0.0 ? std::sqrt(s.true_rr / s.bb) : 0.0;
}

// The host's checks of a body force (ShallowFlowReduction::body_flux,
// flow_system.hpp:398-407), made after the pressure-driven solve as the
// reference makes them.
void check_body_force(const FlowDevice& f, const msim_flow_params& params) {
  if (params.body_force_columns != f.sites()) {
    throw Error(MSIM_INVALID_ARGUMENT, "shallow flow needs one in-plane body force per column");
  }
  for (size_t i = 0; i < 2 * f.sites(); ++i) {
    if (!std::isfinite(params.body_force[i])) {
      throw Error(MSIM_INVALID_ARGUMENT, "shallow-flow body force must be finite");
    }
  }
}

// The forced solve (solve_shallow_axis_flow, flow_system.hpp:1433-1441): the
// body flux, the same operator with every boundary pressure zero and the
// flux's divergence as its right-hand side, solved by the same form; its
// velocity added to the body flux and lifted is the driven field (f.driven),
// whose inlet sum the state then holds. The pressure-driven solve's
// right-hand side and pressure are kept.
void solve_forced(Context& ctx, const msim_flow_params& params, msim_flow_report& report,
                  KernelTimes* times) {
  auto& f = ctx.flow;
  const size_t n = f.sites();
  const size_t nfaces = f.x_faces() + f.y_faces();
  f.force.reserve(2 * n);
  f.body_flux.reserve(nfaces);
  f.driven.reserve(nfaces);
  f.rhs_saved.reserve(n);
  f.x_saved.reserve(n);
  check(hipMemcpyAsync(f.force.data(), params.body_force, 2 * n * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload body force");
  k_body_flux<<<static_cast<uint32_t>((nfaces + 255) / 256), 256, 0, ctx.stream>>>(
      f.cond.data(), reinterpret_cast<const float2*>(f.force.data()), f.body_flux.data(), f.nx,
      f.ny, f.edges());
  check(hipGetLastError(), "launch k_body_flux");
  check(hipMemcpyAsync(f.rhs_saved.data(), f.rhs.data(), n * sizeof(float),
                       hipMemcpyDeviceToDevice, ctx.stream),
        "keep right-hand side");
  check(hipMemcpyAsync(f.x_saved.data(), f.x.data(), n * sizeof(double), hipMemcpyDeviceToDevice,
                       ctx.stream),
        "keep pressure");
  k_forced_rhs<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, ctx.stream>>>(
      f.cond.data(), f.body_flux.data(), f.rhs.data(), f.nx, f.ny, f.dx, f.dy);
  check(hipGetLastError(), "launch k_forced_rhs");
  auto restore = [&] {
    check(hipMemcpyAsync(f.rhs.data(), f.rhs_saved.data(), n * sizeof(float),
                         hipMemcpyDeviceToDevice, ctx.stream),
          "restore right-hand side");
  };
  try {
    solve_pressure(ctx, params, report, times);
  } catch (...) {
    restore();
    static_cast<void>(hipStreamSynchronize(ctx.stream));
    throw;
  }
  restore();
  FlowEdges held = f.edges();
  held.held_zero = 1;
  k_velocity<<<stream_blocks(nfaces), kStream, 0, ctx.stream>>>(
      f.x.data(), f.cond.data(), f.solid.data(), f.driven.data(), f.nx, f.ny, f.dx, f.dy, f.dz,
      held, f.body_flux.data(), f.state.data(), f.partials.data());
  check(hipGetLastError(), "launch k_velocity");
  check(hipMemcpyAsync(f.x.data(), f.x_saved.data(), n * sizeof(double), hipMemcpyDeviceToDevice,
                       ctx.stream),
        "restore pressure");
  read_state(ctx);
}

}  // namespace

void flow_solve(Context& ctx, const msim_flow_params& params, msim_flow_report& report) {
  auto& f = ctx.flow;
  // The field is rewritten from here on, whatever the outcome.
  ++f.generation;
  f.solved = false;
  f.forced = false;
  std::memset(&report, 0, sizeof(report));
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  KernelTimes kernel_times;
  KernelTimes* times = params.profile != 0 ? &kernel_times : nullptr;
  solve_pressure(ctx, params, report, times);
  const CgState& s = *f.host_state;

  // Velocity reconstruction, lift and scaling.
  const size_t nfaces = f.x_faces() + f.y_faces();
  k_velocity<<<stream_blocks(nfaces), kStream, 0, ctx.stream>>>(
      f.x.data(), f.cond.data(), f.solid.data(), f.faces.data(), f.nx, f.ny, f.dx, f.dy, f.dz,
      f.edges(), nullptr, f.state.data(), f.partials.data());
  check(hipGetLastError(), "launch k_velocity");
  read_state(ctx);
  // A solve of a port-driven flow returns the unscaled field of the held
  // pressures, which the caller combines with the other solves' fields.
  float solved_mean = 0.0f;
  const float peak = static_cast<float>(s.peak);
  const double inlet_count = s.inlet_count;
  if (!f.ported) {
    if (s.nonfinite != 0) {
      throw Error(MSIM_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
    }
    if (s.inlet_count == 0.0) throw Error(MSIM_RUNTIME_ERROR, "flow solve has no open inlet faces");
    solved_mean = static_cast<float>(s.inlet_sum / s.inlet_count);
    if (peak == 0.0f || solved_mean <= 1.0e-9f * peak) {
      throw Error(MSIM_RUNTIME_ERROR,
                  "the device carries no through-flow: the outlet is unreachable");
    }
  }
  double driven_sum = 0.0;
  if (params.body_force != nullptr && !f.ported) {
    check_body_force(f, params);
    msim_flow_report second{};
    solve_forced(ctx, params, second, times);
    driven_sum = s.inlet_sum;
    report.iterations += second.iterations;
    report.refinements += second.refinements;
    report.relative_residual = std::max(report.relative_residual, second.relative_residual);
    report.true_relative_residual =
        std::max(report.true_relative_residual, second.true_relative_residual);
    report.fell_back |= second.fell_back;
    f.forced = true;
  }
  float factor = 1.0f;
  float max_speed = peak;
  if (params.scale != 0 && !f.ported) {
    const float driven_mean = static_cast<float>(driven_sum / inlet_count);
    factor = (params.mean_inlet_speed - driven_mean) / solved_mean;
    k_scale<<<stream_blocks(nfaces), kStream, 0, ctx.stream>>>(
        f.faces.data(), nfaces, factor, f.forced ? f.driven.data() : nullptr, f.state.data(),
        f.partials.data());
    check(hipGetLastError(), "launch k_scale");
    read_state(ctx);
    if (f.host_state->nonfinite != 0) {
      throw Error(MSIM_RUNTIME_ERROR, "non-finite combined flow velocity");
    }
    max_speed = static_cast<float>(f.host_state->scaled_peak);
  }
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  check(hipEventSynchronize(ctx.stop), "event sync");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  report.mean_inlet_speed = params.scale != 0 ? params.mean_inlet_speed : solved_mean;
  report.max_speed = max_speed;
  report.solved_mean = solved_mean;
  report.factor = factor;
  report.seconds = ms * 1.0e-3;
  report.apply_seconds = kernel_times.apply_seconds;
  report.apply_launches = kernel_times.apply_launches;
  report.update_seconds = kernel_times.update_seconds;
  report.update_launches = kernel_times.update_launches;
  report.cycle_down_seconds = kernel_times.down_seconds;
  report.cycle_up_seconds = kernel_times.up_seconds;
  report.cycle_rest_seconds = kernel_times.rest_seconds;
  report.cycles = kernel_times.cycles;
  f.solved = true;
}

}  // namespace msim
