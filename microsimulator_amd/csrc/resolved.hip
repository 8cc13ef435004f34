// Resolved Stokes-Brinkman flow on the device (msim_hip_solve_resolved_flow):
// solve_resolved_flow_cpu (cpp/cpu/cpu_flow.cpp:379-467) over
// ResolvedFlowSystem (cpp/core/flow_system.hpp:755-1077) with the reference's
// flexible GMRES (cpp/core/flexible_gmres.hpp:33-108).
//
// Unknowns: the velocity on every face of the voxel grid (x|y|z faces, the
// reference's order) and the pressure at every site, FP64 throughout, as the
// reference computes them. The system [A G; cs D] (momentum with Brinkman
// drag and the no-slip wall rule, pressure gradient, scaled divergence) is
// applied by one kernel; the preconditioner is Jacobi-PCG on the momentum
// block (the float-rounded right-hand side and the float diagonal, as the
// reference's conjugate_gradient takes them) and the scaled pressure
// diagonal. Every scalar of both Krylov methods stays on the device: dot
// products leave fixed-order block partials that a one-workgroup kernel adds
// and turns into the step (CG's alpha and beta, the stopping test; GMRES's
// projections and norms), so the vectors never wait on the host. The host
// reads a few scalars per Arnoldi step (its Hessenberg column) and runs the
// Givens rotations and the small triangular solve, as the reference does.
//
// The system's arrays (active and existing faces, face drag, diagonal,
// forcing, pressure diagonal) are built on the host from the grid, in the
// reference's float arithmetic, and uploaded once per solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace msim {
namespace {

constexpr int kResBlock = 256;
constexpr uint32_t kRestart = 40;  // flexible_gmres.hpp:46

// The grid as the kernels see it.
struct ResView {
  uint32_t dims[3];
  uint64_t off[3];    // first face of each component
  uint64_t nf, n;     // faces, sites
  double h[3];        // spacing
  double inv[3];      // 1 / h^2, 0 on an axis of one site (no viscous term)
  double cs;          // continuity scale 1 / min spacing
  uint32_t chunk;     // entries per workgroup of the reducing kernels
  uint32_t cg_faces;  // faces per thread of the momentum CG's kernels (1, 2, 4 or 8)
  const uint8_t* active;
  // Per face, bit 2a + (side > 0): the momentum stencil reads the neighbour
  // face along axis a on that side (it is on the face lattice and, across
  // the face's own axis, exists); otherwise it mirrors (flow_system.hpp:
  // 887-905).
  const uint8_t* neighbours;
  const uint8_t* exists;
  const uint8_t* fluid;
  const float* face_drag;
  const float* diag;
  const float* pd;    // pressure diagonal
};

__device__ inline int face_component(const ResView& V, uint64_t f) {
  return f < V.off[1] ? 0 : (f < V.off[2] ? 1 : 2);
}

// Face f's coordinates in its component's face lattice (32-bit division:
// the C ABI refuses grids of 2^32 faces or more).
__device__ inline void face_xyz(const ResView& V, int c, uint64_t f, uint32_t fd[3],
                                uint32_t co[3]) {
  fd[0] = V.dims[0];
  fd[1] = V.dims[1];
  fd[2] = V.dims[2];
  ++fd[c];
  uint32_t local = static_cast<uint32_t>(f - V.off[c]);
  co[2] = local % fd[2];
  local /= fd[2];
  co[1] = local % fd[1];
  co[0] = local / fd[1];
}

__device__ inline uint64_t face_at(const ResView& V, int c, const uint32_t fd[3], uint32_t x,
                                   uint32_t y, uint32_t z) {
  return V.off[c] + (static_cast<uint64_t>(x) * fd[1] + y) * fd[2] + z;
}

__device__ inline uint64_t site_at(const ResView& V, uint32_t x, uint32_t y, uint32_t z) {
  return (static_cast<uint64_t>(x) * V.dims[1] + y) * V.dims[2] + z;
}

// apply_momentum (flow_system.hpp:881-910) at face f; u(i) is the input at
// face i. The neighbour bits decide every read up front, so the loads of all
// six neighbours issue together.
template <typename U>
__device__ inline double momentum_at(const ResView& V, const U& u, uint64_t f, int c,
                                     const uint32_t fd[3], const uint32_t co[3]) {
  const double uf = u(f);
  const uint32_t bits = V.neighbours[f];
  double value[6];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int slot = 2 * a + side;
      if ((bits >> slot) & 1u) {
        uint32_t nb[3] = {co[0], co[1], co[2]};
        nb[a] = side == 0 ? nb[a] - 1 : nb[a] + 1;
        value[slot] = u(face_at(V, c, fd, nb[0], nb[1], nb[2]));
      } else {
        value[slot] = a == c ? uf : -uf;
      }
    }
  }
  double result = static_cast<double>(V.face_drag[f]) * uf;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (V.dims[a] == 1) continue;
    result -= (value[2 * a] - uf) * V.inv[a];
    result -= (value[2 * a + 1] - uf) * V.inv[a];
  }
  return result;
}

// out = [A u + G p; cs D u] (cpu_flow.cpp:414-423; gradient and divergence,
// flow_system.hpp:912-956).
__global__ __launch_bounds__(kResBlock) void k_res_apply(ResView V, const double* __restrict__ in,
                                                         double* __restrict__ out) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kResBlock + threadIdx.x;
  if (k < V.nf) {
    const uint64_t f = k;
    if (V.active[f] == 0) {
      out[f] = 0.0 + 0.0;
      return;
    }
    const int c = face_component(V, f);
    uint32_t fd[3], co[3];
    face_xyz(V, c, f, fd, co);
    const double m = momentum_at(V, [in](uint64_t i) { return in[i]; }, f, c, fd, co);
    const double* p = in + V.nf;
    double lv = 0.0, uv = 0.0;
    if (co[c] > 0) {
      uint32_t s[3] = {co[0], co[1], co[2]};
      --s[c];
      const uint64_t site = site_at(V, s[0], s[1], s[2]);
      if (V.fluid[site] != 0) lv = p[site];
    }
    if (co[c] < V.dims[c]) {
      const uint64_t site = site_at(V, co[0], co[1], co[2]);
      if (V.fluid[site] != 0) uv = p[site];
    }
    out[f] = m + (uv - lv) / V.h[c];
    return;
  }
  if (k >= V.nf + V.n) return;
  const uint64_t site = k - V.nf;
  double d = 0.0;
  if (V.fluid[site] != 0) {
    const uint32_t z = static_cast<uint32_t>(site % V.dims[2]);
    const uint32_t y = static_cast<uint32_t>((site / V.dims[2]) % V.dims[1]);
    const uint32_t x = static_cast<uint32_t>(site / (static_cast<uint64_t>(V.dims[1]) * V.dims[2]));
    const uint32_t co[3] = {x, y, z};
    for (int c = 0; c < 3; ++c) {
      uint32_t fd[3] = {V.dims[0], V.dims[1], V.dims[2]};
      ++fd[c];
      uint32_t up[3] = {co[0], co[1], co[2]};
      ++up[c];
      const double hi = in[face_at(V, c, fd, up[0], up[1], up[2])];
      const double lo = in[face_at(V, c, fd, co[0], co[1], co[2])];
      d += (hi - lo) / V.h[c];
    }
  }
  out[k] = V.cs * d;
}

// Block partials of a . b over chunks of `chunk` entries.
__global__ __launch_bounds__(kResBlock) void k_res_dot(const double* __restrict__ a,
                                                       const double* __restrict__ b, uint64_t count,
                                                       uint32_t chunk,
                                                       double* __restrict__ partials) {
  __shared__ double s_red[kResBlock / kWave];
  const uint64_t begin = static_cast<uint64_t>(blockIdx.x) * chunk;
  const uint64_t end = begin + chunk < count ? begin + chunk : count;
  double v = 0.0;
  for (uint64_t i = begin + threadIdx.x; i < end; i += kResBlock) v += a[i] * b[i];
  v = block_sum<kResBlock>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// The solve's scalars, in device memory. The two kernels of a CG iteration
// each write their own stop flag and failure code and read only the other's,
// so no launch reads a field that it writes.
struct ResState {
  double rhs_norm, relative;
  // r.z after iteration k's update at rho[k & 1], recorded by the first
  // kernel of iteration k + 1 (for the second kernel and the next beta).
  double rho[2];
  uint32_t iterations, stop1, stop2, failure1, failure2, pad;
};

// The sum of `blocks` partials in a fixed order, the same in every workgroup
// (and the same bits as a one-workgroup sum of them).
__device__ inline double all_sum(const double* partials, uint32_t blocks, double* s_red,
                                 double* s_b) {
  double v = 0.0;
  for (uint32_t i = threadIdx.x; i < blocks; i += kResBlock) v += partials[i];
  v = block_sum<kResBlock>(v, s_red);
  if (threadIdx.x == 0) *s_b = v;
  __syncthreads();
  const double out = *s_b;
  __syncthreads();
  return out;
}

// all_sum of two partial arrays in one pass: the same sums, bit for bit, in
// one round of loads instead of two.
__device__ inline void all_sum2(const double* a, const double* b, uint32_t blocks,
                                double* s_red, double* s_b, double& sa, double& sb) {
  double va = 0.0, vb = 0.0;
  for (uint32_t i = threadIdx.x; i < blocks; i += kResBlock) {
    va += a[i];
    vb += b[i];
  }
  va = wave_sum(va);
  vb = wave_sum(vb);
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_red[wave] = va;
    s_red[kResBlock / kWave + wave] = vb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0.0, tb = 0.0;
#pragma unroll
    for (int w = 0; w < kResBlock / kWave; ++w) {
      ta += s_red[w];
      tb += s_red[kResBlock / kWave + w];
    }
    s_b[0] = ta;
    s_b[1] = tb;
  }
  __syncthreads();
  sa = s_b[0];
  sb = s_b[1];
  __syncthreads();
}

__device__ inline bool lead() { return blockIdx.x == 0 && threadIdx.x == 0; }

// The CG partials: r.r and r.z of the residual after iteration k in slot
// (k & 1), p.q of iteration k.
struct CgParts {
  double* base;
  uint32_t blocks;
  __device__ double* rr(uint32_t k) const { return base + (k & 1) * blocks; }
  __device__ double* rz(uint32_t k) const { return base + (2 + (k & 1)) * blocks; }
  __device__ double* pq() const { return base + 4 * blocks; }
};

// x = 0, r = double(float(rhs)), z = D^-1 r, and the partials of r.r and r.z
// (the reference's initialize, conjugate_gradient.hpp:49-55); resets the
// solve's flags.
__global__ __launch_bounds__(kResBlock) void k_cg_init(ResView V, const double* __restrict__ rhs,
                                                       double* __restrict__ x,
                                                       double* __restrict__ r,
                                                       double* __restrict__ z, CgParts P,
                                                       ResState* st) {
  __shared__ double s_red[kResBlock / kWave];
  // The CG kernels' layout: cg_faces faces per thread, kResBlock apart.
  const uint64_t chunk = static_cast<uint64_t>(V.cg_faces) * kResBlock;
  const uint64_t begin = static_cast<uint64_t>(blockIdx.x) * chunk;
  const uint64_t end = begin + chunk < V.nf ? begin + chunk : V.nf;
  double rr = 0.0, rz = 0.0;
  for (uint64_t i = begin + threadIdx.x; i < end; i += kResBlock) {
    const double ri = static_cast<double>(static_cast<float>(rhs[i]));
    const float d = V.diag[i];
    const double zi = d > 0.0f ? ri / d : 0.0;
    x[i] = 0.0;
    r[i] = ri;
    z[i] = zi;
    rr += ri * ri;
    rz += ri * zi;
  }
  rr = block_sum<kResBlock>(rr, s_red);
  rz = block_sum<kResBlock>(rz, s_red);
  if (threadIdx.x == 0) {
    P.rr(0)[blockIdx.x] = rr;
    P.rz(0)[blockIdx.x] = rz;
  }
  if (lead()) {
    st->stop1 = st->stop2 = 0;
    st->failure1 = st->failure2 = 0;
    st->iterations = 0;
    st->relative = 1.0;
  }
}

// The first half of CG iteration k. For k > 0: the relative residual of
// iteration k - 1's update and the stopping test, then beta = r.z / rho and
// the direction p = z + beta p_old (conjugate_gradient.hpp:66-84); for k = 0
// the right-hand side's norm and p = z. Then q = A p and the partials of p.q.
// Every workgroup re-adds the partials it needs in the same fixed order, so
// all hold the same scalars; workgroup 0 records them. A neighbour's p is
// formed from z and p_old where it is read, so p_old (another buffer) is
// never written here. `last`: test only (the iteration limit is reached).
// Each thread takes F faces, F * kResBlock apart, their loads interleaved.
template <int F>
__global__ __launch_bounds__(kResBlock) void k_cg_direction(
    ResView V, const double* __restrict__ z, const double* __restrict__ p_old,
    double* __restrict__ p, double* __restrict__ q, CgParts P, uint32_t k, uint32_t last,
    float tolerance, ResState* st) {
  __shared__ double s_red[2 * kResBlock / kWave];
  __shared__ double s_b[2];
  if (st->stop2 != 0) {
    if (lead()) st->stop1 = 1;
    return;
  }
  double beta = 0.0;
  double rr, rz;
  all_sum2(P.rr(k), P.rz(k), P.blocks, s_red, s_b, rr, rz);
  if (k == 0) {
    if (rr == 0.0) {
      if (lead()) {
        st->rhs_norm = 0.0;
        st->stop1 = 1;
      }
      return;
    }
    if (lead()) {
      st->rhs_norm = sqrt(rr);
      st->rho[0] = rz;
    }
  } else {
    const double relative = sqrt(rr > 0.0 ? rr : 0.0) / st->rhs_norm;
    if (!isfinite(relative)) {
      if (lead()) {
        st->failure1 = MSIM_CG_RESIDUAL;
        st->stop1 = 1;
      }
      return;
    }
    if (lead()) {
      st->iterations = k;
      st->relative = relative;
    }
    if (relative <= static_cast<double>(tolerance)) {
      if (lead()) st->stop1 = 1;
      return;
    }
    if (last != 0) return;
    const double rho = st->rho[(k + 1) & 1];
    if (!isfinite(rz) || rho == 0.0) {
      if (lead()) {
        st->failure1 = MSIM_CG_BREAKDOWN;
        st->stop1 = 1;
      }
      return;
    }
    beta = rz / rho;
    if (!isfinite(beta)) {
      if (lead()) {
        st->failure1 = MSIM_CG_DIRECTION;
        st->stop1 = 1;
      }
      return;
    }
    if (lead()) st->rho[k & 1] = rz;
  }
  const bool first = k == 0;
  const auto dir = [&](uint64_t i) { return first ? z[i] : z[i] + beta * p_old[i]; };
  const uint64_t begin = static_cast<uint64_t>(blockIdx.x) * (F * kResBlock);
  double pq = 0.0;
#pragma unroll
  for (int u = 0; u < F; ++u) {
    const uint64_t f = begin + threadIdx.x + static_cast<uint64_t>(u) * kResBlock;
    if (f < V.nf) {
      const double pf = dir(f);
      double v = 0.0;
      if (V.active[f] != 0) {
        const int c = face_component(V, f);
        uint32_t fd[3], co[3];
        face_xyz(V, c, f, fd, co);
        v = momentum_at(V, dir, f, c, fd, co);
      }
      p[f] = pf;
      q[f] = v;
      pq += pf * v;
    }
  }
  pq = block_sum<kResBlock>(pq, s_red);
  if (threadIdx.x == 0) P.pq()[blockIdx.x] = pq;
}

// The second half of CG iteration k: alpha = rho / p.q with the curvature and
// step checks (conjugate_gradient.hpp:57-65), x += alpha p, r -= alpha q,
// z = D^-1 r, and the partials of r.r and r.z for iteration k + 1.
template <int F>
__global__ __launch_bounds__(kResBlock) void k_cg_step(ResView V, double* __restrict__ x,
                                                       double* __restrict__ r,
                                                       double* __restrict__ z,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ q, CgParts P,
                                                       uint32_t k, ResState* st) {
  __shared__ double s_red[kResBlock / kWave];
  __shared__ double s_b;
  if (st->stop1 != 0) {
    if (lead()) st->stop2 = 1;
    return;
  }
  const double curvature = all_sum(P.pq(), P.blocks, s_red, &s_b);
  if (!isfinite(curvature) || curvature <= 0.0) {
    if (lead()) {
      st->failure2 = MSIM_CG_CURVATURE;
      st->stop2 = 1;
    }
    return;
  }
  const double alpha = st->rho[k & 1] / curvature;
  if (!isfinite(alpha)) {
    if (lead()) {
      st->failure2 = MSIM_CG_STEP;
      st->stop2 = 1;
    }
    return;
  }
  const uint64_t begin = static_cast<uint64_t>(blockIdx.x) * (F * kResBlock);
  double rr = 0.0, rz = 0.0;
#pragma unroll
  for (int u = 0; u < F; ++u) {
    const uint64_t i = begin + threadIdx.x + static_cast<uint64_t>(u) * kResBlock;
    if (i < V.nf) {
      x[i] += alpha * p[i];
      const double ri = r[i] - alpha * q[i];
      r[i] = ri;
      const float d = V.diag[i];
      const double zi = d > 0.0f ? ri / d : 0.0;
      z[i] = zi;
      rr += ri * ri;
      rz += ri * zi;
    }
  }
  rr = block_sum<kResBlock>(rr, s_red);
  rz = block_sum<kResBlock>(rz, s_red);
  if (threadIdx.x == 0) {
    P.rr(k + 1)[blockIdx.x] = rr;
    P.rz(k + 1)[blockIdx.x] = rz;
  }
}

// The pressure part of the preconditioner (cpu_flow.cpp:424-431):
// out = -in / (cs pd) where pd > 0, else 0.
__global__ __launch_bounds__(kResBlock) void k_res_pressure_precondition(
    ResView V, const double* __restrict__ in, double* __restrict__ out) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kResBlock + threadIdx.x;
  if (i >= V.n) return;
  const float pd = V.pd[i];
  out[V.nf + i] = pd > 0.0f ? -in[V.nf + i] / (V.cs * static_cast<double>(pd)) : 0.0;
}

// x += weights[0] d_0, then += weights[1] d_1, ...: the reference's sequence
// of axpy updates (flexible_gmres.hpp:97-104), in one pass.
struct UpdateWeights {
  double w[kRestart];
};

__global__ __launch_bounds__(kResBlock) void k_res_update(uint64_t n, double* __restrict__ x,
                                                          const double* __restrict__ directions,
                                                          UpdateWeights weights, uint32_t used) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kResBlock + threadIdx.x;
  if (i >= n) return;
  double v = x[i];
  for (uint32_t k = 0; k < used; ++k) v += weights.w[k] * directions[k * n + i];
  x[i] = v;
}

// The one-workgroup sum of a dot product's partials into `slot`.
__global__ __launch_bounds__(kResBlock) void k_res_sum(const double* partials, uint32_t blocks,
                                                       double* slot) {
  __shared__ double s_red[kResBlock / kWave];
  __shared__ double s_b;
  const double v = all_sum(partials, blocks, s_red, &s_b);
  if (threadIdx.x == 0) *slot = v;
}

// Arnoldi orthogonalisation by classical Gram-Schmidt applied twice (CGS2).
// The reference orthogonalises by modified Gram-Schmidt applied twice
// (flexible_gmres.hpp:70-76); both give w minus its projection on the basis
// and the same coefficients in exact arithmetic, and both are applied twice
// for the same reason. CGS2 forms all of a pass's dot products in one launch,
// so a step of j basis vectors takes six launches instead of 2j + 4.

// Partials of v_i . w for i < count over this workgroup's chunk, at
// partials[i * gridDim.x + block]. V holds the basis vectors, n apart.
__global__ __launch_bounds__(kResBlock) void k_res_multidot(uint64_t n, uint32_t chunk,
                                                            const double* __restrict__ V,
                                                            uint32_t count,
                                                            const double* __restrict__ w,
                                                            double* __restrict__ partials) {
  __shared__ double s_part[kRestart + 1][kResBlock / kWave];
  const uint64_t begin = static_cast<uint64_t>(blockIdx.x) * chunk;
  const uint64_t end = begin + chunk < n ? begin + chunk : n;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (uint32_t i = 0; i < count; ++i) {
    const double* vi = V + static_cast<size_t>(i) * n;
    double acc = 0.0;
    for (uint64_t e = begin + threadIdx.x; e < end; e += kResBlock) acc += vi[e] * w[e];
    acc = wave_sum(acc);
    if (lane == 0) s_part[i][wave] = acc;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < count; i += kResBlock) {
    double total = 0.0;
#pragma unroll
    for (int k = 0; k < kResBlock / kWave; ++k) total += s_part[i][k];
    partials[static_cast<size_t>(i) * gridDim.x + blockIdx.x] = total;
  }
}

// out[i] = the fixed-order sum of partials[i * blocks + 0 .. blocks): one
// workgroup per coefficient i.
__global__ __launch_bounds__(kResBlock) void k_res_multisum(const double* __restrict__ partials,
                                                            uint32_t blocks,
                                                            double* __restrict__ out) {
  __shared__ double s_red[kResBlock / kWave];
  __shared__ double s_b;
  const double v = all_sum(partials + static_cast<size_t>(blockIdx.x) * blocks, blocks, s_red,
                           &s_b);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// w -= sum_i h[i] v_i over this workgroup's chunk (in order of i), then the
// partials of the next pass's dot products with the updated w: v_i . w for
// i < count (multidot layout), or w . w when `norm` (at partials[block]).
__global__ __launch_bounds__(kResBlock) void k_res_cgs(uint64_t n, uint32_t chunk,
                                                       const double* __restrict__ V,
                                                       uint32_t count, const double* __restrict__ h,
                                                       double* w, uint32_t norm,
                                                       double* __restrict__ partials) {
  __shared__ double s_h[kRestart + 1];
  __shared__ double s_part[kRestart + 1][kResBlock / kWave];
  if (threadIdx.x < count) s_h[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  const uint64_t begin = static_cast<uint64_t>(blockIdx.x) * chunk;
  const uint64_t end = begin + chunk < n ? begin + chunk : n;
  for (uint64_t e = begin + threadIdx.x; e < end; e += kResBlock) {
    double we = w[e];
    for (uint32_t i = 0; i < count; ++i) we += -s_h[i] * V[static_cast<size_t>(i) * n + e];
    w[e] = we;
  }
  // Each thread reads back only the entries it wrote.
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const uint32_t dots = norm != 0 ? 1u : count;
  for (uint32_t i = 0; i < dots; ++i) {
    const double* vi = norm != 0 ? w : V + static_cast<size_t>(i) * n;
    double acc = 0.0;
    for (uint64_t e = begin + threadIdx.x; e < end; e += kResBlock) acc += vi[e] * w[e];
    acc = wave_sum(acc);
    if (lane == 0) s_part[i][wave] = acc;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < dots; i += kResBlock) {
    double total = 0.0;
#pragma unroll
    for (int k = 0; k < kResBlock / kWave; ++k) total += s_part[i][k];
    partials[static_cast<size_t>(i) * gridDim.x + blockIdx.x] = total;
  }
}

// out = 0 + s * in (a basis vector from w or the residual).
__global__ __launch_bounds__(kResBlock) void k_res_scaled(uint64_t n, double* __restrict__ out,
                                                          double s, const double* __restrict__ in) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kResBlock + threadIdx.x;
  if (i < n) out[i] = 0.0 + s * in[i];
}

// r = rhs + (-1) * w.
__global__ __launch_bounds__(kResBlock) void k_res_residual(uint64_t n, double* __restrict__ r,
                                                            const double* __restrict__ rhs,
                                                            const double* __restrict__ w) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kResBlock + threadIdx.x;
  if (i < n) r[i] = rhs[i] + -1.0 * w[i];
}

uint32_t grid_of(uint64_t n) { return static_cast<uint32_t>((n + kResBlock - 1) / kResBlock); }
uint32_t chunks_of(uint64_t n, uint32_t chunk) {
  return static_cast<uint32_t>(std::max<uint64_t>((n + chunk - 1) / chunk, 1));
}

// Entries per workgroup of the reducing kernels: one per thread until a
// launch would need more than 1024 workgroups, then whole multiples of the
// workgroup, so small systems spread over the chip and large ones keep the
// partials few.
uint32_t chunk_for(uint64_t n) {
  const uint64_t per = (n + kResBlock * 1024ull - 1) / (kResBlock * 1024ull);
  return kResBlock * static_cast<uint32_t>(std::max<uint64_t>(per, 1));
}

// The system, built on the host (flow_system.hpp:755-858, 958-1041).
struct ResSystem {
  uint32_t dims[3];
  float h[3];
  uint64_t off[3], cnt[3], nf, n;
  std::vector<uint8_t> fluid, active, exists, neighbours;
  std::vector<float> face_drag, diag, force, pd;
  uint32_t axis;
};

void face_coords_host(const ResSystem& s, uint64_t index, int* c, uint32_t co[3]) {
  int k = 0;
  while (k < 2 && index >= s.off[k] + s.cnt[k]) ++k;
  uint64_t local = index - s.off[k];
  uint32_t fd[3] = {s.dims[0], s.dims[1], s.dims[2]};
  ++fd[k];
  co[2] = static_cast<uint32_t>(local % fd[2]);
  local /= fd[2];
  co[1] = static_cast<uint32_t>(local % fd[1]);
  co[0] = static_cast<uint32_t>(local / fd[1]);
  *c = k;
}

uint64_t face_host(const ResSystem& s, int c, uint32_t x, uint32_t y, uint32_t z) {
  uint32_t fd[3] = {s.dims[0], s.dims[1], s.dims[2]};
  ++fd[c];
  return s.off[c] + (static_cast<uint64_t>(x) * fd[1] + y) * fd[2] + z;
}

uint64_t site_host(const ResSystem& s, uint32_t x, uint32_t y, uint32_t z) {
  return (static_cast<uint64_t>(x) * s.dims[1] + y) * s.dims[2] + z;
}

// The general constructor (flow_system.hpp:789-858). axis < 3: with
// axis_flow_boundary (186-201) and no known pressure, then the single-axis
// forcing (762-781). axis == 3: the faces `open` flags (port_flow_boundary,
// 238-253) opened; the forcing, which depends on the pressures held and the
// known pressure, is formed per solve (port_force).
ResSystem build_system(const msim_grid& g, const std::vector<uint8_t>& mask, const float* drag,
                       uint32_t axis, const std::vector<uint8_t>* open_faces = nullptr) {
  ResSystem s{};
  s.axis = axis;
  s.dims[0] = g.nx;
  s.dims[1] = g.ny;
  s.dims[2] = g.nz;
  s.h[0] = g.dx;
  s.h[1] = g.dy;
  s.h[2] = g.dz;
  s.n = static_cast<uint64_t>(g.nx) * g.ny * g.nz;
  s.cnt[0] = static_cast<uint64_t>(g.nx + 1) * g.ny * g.nz;
  s.cnt[1] = static_cast<uint64_t>(g.nx) * (g.ny + 1) * g.nz;
  s.cnt[2] = static_cast<uint64_t>(g.nx) * g.ny * (g.nz + 1);
  s.off[0] = 0;
  s.off[1] = s.cnt[0];
  s.off[2] = s.cnt[0] + s.cnt[1];
  s.nf = s.off[2] + s.cnt[2];
  s.fluid.assign(s.n, 1);
  std::vector<float> d(s.n, 0.0f);
  if (drag != nullptr) std::copy(drag, drag + s.n, d.begin());
  for (uint64_t site = 0; site < s.n; ++site) {
    s.fluid[site] = mask[site] != 0 ? 0 : 1;
    if (!std::isfinite(d[site]) || d[site] < 0.0f) {
      throw Error(MSIM_INVALID_ARGUMENT, "resolved-flow drag values must be finite and non-negative");
    }
    if (s.fluid[site] == 0) d[site] = 0.0f;
  }
  s.active.assign(s.nf, 0);
  s.exists.assign(s.nf, 0);
  s.face_drag.assign(s.nf, 0.0f);
  s.diag.assign(s.nf, 0.0f);
  s.force.assign(s.nf, 0.0f);
  for (uint64_t index = 0; index < s.nf; ++index) {
    int c;
    uint32_t co[3];
    face_coords_host(s, index, &c, co);
    int64_t lower = -1, upper = -1;
    if (co[c] > 0) {
      uint32_t t[3] = {co[0], co[1], co[2]};
      --t[c];
      lower = static_cast<int64_t>(site_host(s, t[0], t[1], t[2]));
    }
    if (co[c] < s.dims[c]) upper = static_cast<int64_t>(site_host(s, co[0], co[1], co[2]));
    const bool lower_fluid = lower >= 0 && s.fluid[lower] != 0;
    const bool upper_fluid = upper >= 0 && s.fluid[upper] != 0;
    s.exists[index] = lower_fluid || upper_fluid ? 1 : 0;
    const bool exterior = lower < 0 || upper < 0;
    const bool open = axis < 3 ? c == static_cast<int>(axis) &&
                                     (co[axis] == 0 || co[axis] == s.dims[axis])
                               : (*open_faces)[index] != 0;
    const bool act = exterior ? open && (lower_fluid || upper_fluid) : lower_fluid && upper_fluid;
    s.active[index] = act ? 1 : 0;
    if (!act) continue;
    float sum = 0.0f, count = 0.0f;
    if (lower_fluid) {
      sum += d[lower];
      count += 1.0f;
    }
    if (upper_fluid) {
      sum += d[upper];
      count += 1.0f;
    }
    s.face_drag[index] = sum / count;
    s.diag[index] = s.face_drag[index];
    for (int a = 0; a < 3; ++a) {
      if (s.dims[a] > 1) s.diag[index] += 2.0f / (s.h[a] * s.h[a]);
    }
  }
  if (axis < 3) {
    const float force = 1.0f / ((static_cast<float>(s.dims[axis]) + 1.0f) * s.h[axis]);
    bool open_inlet = false;
    for (uint64_t index = s.off[axis]; index < s.off[axis] + s.cnt[axis]; ++index) {
      if (s.active[index] == 0) continue;
      int c;
      uint32_t co[3];
      face_coords_host(s, index, &c, co);
      s.force[index] = force;
      open_inlet = open_inlet || co[axis] == 0;
    }
    if (!open_inlet) {
      throw Error(MSIM_INVALID_ARGUMENT, "the resolved-flow inlet boundary is entirely blocked");
    }
  }
  // The momentum stencil's neighbour bits of each active face.
  s.neighbours.assign(s.nf, 0);
  for (uint64_t index = 0; index < s.nf; ++index) {
    if (s.active[index] == 0) continue;
    int c;
    uint32_t co[3];
    face_coords_host(s, index, &c, co);
    uint32_t fd[3] = {s.dims[0], s.dims[1], s.dims[2]};
    ++fd[c];
    uint8_t bits = 0;
    for (int a = 0; a < 3; ++a) {
      for (int side = 0; side < 2; ++side) {
        const bool has = side == 0 ? co[a] > 0 : co[a] + 1 < fd[a];
        if (!has) continue;
        uint32_t nb[3] = {co[0], co[1], co[2]};
        nb[a] = side == 0 ? nb[a] - 1 : nb[a] + 1;
        if (a == c || s.exists[face_host(s, c, nb[0], nb[1], nb[2])] != 0) {
          bits |= static_cast<uint8_t>(1u << (2 * a + side));
        }
      }
    }
    s.neighbours[index] = bits;
  }
  // pressure_diagonal (958-978).
  s.pd.assign(s.n, 0.0f);
  for (uint64_t site = 0; site < s.n; ++site) {
    if (s.fluid[site] == 0) continue;
    const uint32_t co[3] = {static_cast<uint32_t>(site / (static_cast<uint64_t>(s.dims[1]) * s.dims[2])),
                            static_cast<uint32_t>((site / s.dims[2]) % s.dims[1]),
                            static_cast<uint32_t>(site % s.dims[2])};
    for (int a = 0; a < 3; ++a) {
      uint32_t up[3] = {co[0], co[1], co[2]};
      ++up[a];
      const uint64_t faces[2] = {face_host(s, a, co[0], co[1], co[2]),
                                 face_host(s, a, up[0], up[1], up[2])};
      for (uint64_t face : faces) {
        if (s.active[face] != 0 && s.diag[face] > 0.0f) {
          s.pd[site] += 1.0f / (s.diag[face] * s.h[a] * s.h[a]);
        }
      }
    }
  }
  return s;
}

// minimum_gap_voxels of a single-axis system (998-1041).
uint32_t min_gap(const ResSystem& s) {
  uint32_t shortest = 0;
  for (int axis = 0; axis < 3; ++axis) {
    if (axis == static_cast<int>(s.axis) || s.dims[axis] <= 1) continue;
    const int fa = (axis + 1) % 3, sa = (axis + 2) % 3;
    for (uint32_t first = 0; first < s.dims[fa]; ++first) {
      for (uint32_t second = 0; second < s.dims[sa]; ++second) {
        uint32_t co[3] = {0, 0, 0};
        co[fa] = first;
        co[sa] = second;
        const auto open_face = [&](uint32_t along) {
          uint32_t f[3] = {co[0], co[1], co[2]};
          f[axis] = along;
          return s.active[face_host(s, axis, f[0], f[1], f[2])] != 0;
        };
        uint32_t run = 0;
        bool through = false;
        const auto close = [&] {
          if (run != 0 && !through) shortest = shortest == 0 ? run : std::min(shortest, run);
          run = 0;
          through = false;
        };
        for (uint32_t along = 0; along < s.dims[axis]; ++along) {
          co[axis] = along;
          if (s.fluid[site_host(s, co[0], co[1], co[2])] != 0) {
            through = through || (along == 0 && open_face(0));
            ++run;
          } else {
            close();
          }
        }
        through = through || (run != 0 && open_face(s.dims[axis]));
        close();
      }
    }
  }
  return shortest;
}

// Device state of one solve.
struct ResDevice {
  ResView V{};
  uint64_t m = 0;
  DeviceBuffer<uint8_t> active, exists, fluid, neighbours;
  DeviceBuffer<float> face_drag, diag, pd;
  // cg_p holds two direction buffers (iterations alternate between them).
  DeviceBuffer<double> rhs, x, r, w, basis, directions, cg_r, cg_z, cg_p, cg_q, cg_parts,
      partials, slots;
  DeviceBuffer<ResState> state;
  // Iterations the last momentum solve took: the next solve's first batch.
  uint32_t cg_expect = 16;
};

template <typename T>
void upload(Context& ctx, DeviceBuffer<T>& buf, const std::vector<T>& host, const char* what) {
  buf.reserve(host.size());
  check(hipMemcpyAsync(buf.data(), host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice,
                       ctx.stream),
        what);
}

// dot(a, b) of `count` entries into the device scalar `slot`.
void device_dot(Context& ctx, ResDevice& D, const double* a, const double* b, uint64_t count,
                double* slot) {
  const uint32_t blocks = chunks_of(count, D.V.chunk);
  k_res_dot<<<blocks, kResBlock, 0, ctx.stream>>>(a, b, count, D.V.chunk, D.partials.data());
  k_res_sum<<<1, kResBlock, 0, ctx.stream>>>(D.partials.data(), blocks, slot);
}

double read_scalar(Context& ctx, const double* slot) {
  double v = 0.0;
  check(hipMemcpyAsync(&v, slot, sizeof(double), hipMemcpyDeviceToHost, ctx.stream), "read scalar");
  check(hipStreamSynchronize(ctx.stream), "resolved flow");
  return v;
}

// The momentum block by Jacobi-PCG from zero (conjugate_gradient,
// cpu_flow.cpp:53-103) on rhs (FP64, rounded to float as the reference
// rounds it): the solution into `out`. Adds the iterations to `inner`. Two
// launches per iteration. The host reads the state after a first batch as
// long as the last solve took (plus the launch that tests its last update),
// then after batches of 8: the solves of one flow take similar counts, so
// most end in one batch with few launches that find the solve stopped.
void momentum_cg(Context& ctx, ResDevice& D, const double* rhs, double* out, float tolerance,
                 uint32_t max_iterations, uint64_t& inner) {
  const ResView& V = D.V;
  const uint32_t blocks = chunks_of(V.nf, V.cg_faces * kResBlock);
  const CgParts P{D.cg_parts.data(), blocks};
  ResState* st = D.state.data();
  double* const p[2] = {D.cg_p.data(), D.cg_p.data() + V.nf};
  k_cg_init<<<blocks, kResBlock, 0, ctx.stream>>>(V, rhs, out, D.cg_r.data(), D.cg_z.data(), P,
                                                  st);
  check(hipGetLastError(), "launch momentum CG");
  auto direction = [&](uint32_t k, uint32_t last) {
    switch (V.cg_faces) {
#define MSIM_CG_DIRECTION(F)                                                               \
  case F:                                                                                  \
    k_cg_direction<F><<<blocks, kResBlock, 0, ctx.stream>>>(V, D.cg_z.data(), p[(k + 1) & 1], \
                                                            p[k & 1], D.cg_q.data(), P, k,  \
                                                            last, tolerance, st);           \
    break;
      MSIM_CG_DIRECTION(1)
      MSIM_CG_DIRECTION(2)
      MSIM_CG_DIRECTION(4)
      default: MSIM_CG_DIRECTION(8)
#undef MSIM_CG_DIRECTION
    }
  };
  auto step = [&](uint32_t k) {
    switch (V.cg_faces) {
#define MSIM_CG_STEP(F)                                                                     \
  case F:                                                                                   \
    k_cg_step<F><<<blocks, kResBlock, 0, ctx.stream>>>(V, out, D.cg_r.data(), D.cg_z.data(), \
                                                       p[k & 1], D.cg_q.data(), P, k, st);  \
    break;
      MSIM_CG_STEP(1)
      MSIM_CG_STEP(2)
      MSIM_CG_STEP(4)
      default: MSIM_CG_STEP(8)
#undef MSIM_CG_STEP
    }
  };
  uint32_t k = 0;
  uint32_t next = std::max(D.cg_expect + 1, 4u);
  ResState s{};
  for (;;) {
    const uint32_t batch = std::min(next, max_iterations - k);
    next = 8;
    for (uint32_t i = 0; i < batch; ++i, ++k) {
      direction(k, 0);
      step(k);
    }
    // The test of the last iteration's update.
    if (k >= max_iterations) direction(k, 1);
    check(hipGetLastError(), "launch momentum CG");
    check(hipMemcpyAsync(&s, st, sizeof(ResState), hipMemcpyDeviceToHost, ctx.stream),
          "read CG state");
    check(hipStreamSynchronize(ctx.stream), "momentum CG");
    if (s.stop1 != 0 || s.stop2 != 0 || k >= max_iterations) break;
  }
  inner += s.iterations;
  D.cg_expect = s.iterations;
  const std::string label = "resolved-flow momentum conjugate gradient ";
  switch (s.failure1 != 0 ? s.failure1 : s.failure2) {
    case MSIM_CG_CURVATURE:
      throw Error(MSIM_CG_CURVATURE, label + "encountered non-positive curvature");
    case MSIM_CG_STEP: throw Error(MSIM_CG_STEP, label + "produced a non-finite step");
    case MSIM_CG_RESIDUAL: throw Error(MSIM_CG_RESIDUAL, label + "produced a non-finite residual");
    case MSIM_CG_BREAKDOWN:
      throw Error(MSIM_CG_BREAKDOWN, label + "encountered a preconditioner breakdown");
    case MSIM_CG_DIRECTION:
      throw Error(MSIM_CG_DIRECTION, label + "produced a non-finite direction");
    default: break;
  }
  if (s.stop1 == 0 && s.stop2 == 0) {
    throw Error(MSIM_NOT_CONVERGED,
                label + "did not converge: relative " + std::to_string(s.relative));
  }
}

void apply_system(Context& ctx, ResDevice& D, const double* in, double* out) {
  k_res_apply<<<grid_of(D.m), kResBlock, 0, ctx.stream>>>(D.V, in, out);
  check(hipGetLastError(), "launch k_res_apply");
}

void precondition(Context& ctx, ResDevice& D, const double* in, double* out, float tolerance,
                  uint32_t max_iterations, uint64_t& inner) {
  momentum_cg(ctx, D, in, out, tolerance, max_iterations, inner);
  k_res_pressure_precondition<<<grid_of(D.V.n), kResBlock, 0, ctx.stream>>>(D.V, in, out);
  check(hipGetLastError(), "launch k_res_pressure_precondition");
}

// The device arrays of a built system, and the buffers of its solves.
void setup_device(Context& ctx, const ResSystem& S, ResDevice& D) {
  const uint64_t faces = S.nf;
  if (faces >= (1ull << 32)) {
    throw Error(MSIM_INVALID_ARGUMENT, "resolved flow needs fewer than 2^32 faces");
  }
  D.m = S.nf + S.n;
  ResView& V = D.V;
  for (int a = 0; a < 3; ++a) {
    V.dims[a] = S.dims[a];
    V.off[a] = S.off[a];
    V.h[a] = static_cast<double>(S.h[a]);
    V.inv[a] = 1.0 / (static_cast<double>(S.h[a]) * S.h[a]);
  }
  V.nf = S.nf;
  V.n = S.n;
  V.cs = 1.0 / std::min({S.h[0], S.h[1], S.h[2]});
  V.chunk = chunk_for(D.m);
  // Faces per thread of the momentum CG: the fewest that keep a launch at
  // or under 2048 workgroups.
  V.cg_faces = 1;
  while (V.cg_faces < 8 && (S.nf + V.cg_faces * kResBlock - 1) / (V.cg_faces * kResBlock) > 2048) {
    V.cg_faces *= 2;
  }
  upload(ctx, D.active, S.active, "upload resolved system");
  upload(ctx, D.exists, S.exists, "upload resolved system");
  upload(ctx, D.fluid, S.fluid, "upload resolved system");
  upload(ctx, D.neighbours, S.neighbours, "upload resolved system");
  upload(ctx, D.face_drag, S.face_drag, "upload resolved system");
  upload(ctx, D.diag, S.diag, "upload resolved system");
  upload(ctx, D.pd, S.pd, "upload resolved system");
  V.active = D.active.data();
  V.exists = D.exists.data();
  V.fluid = D.fluid.data();
  V.neighbours = D.neighbours.data();
  V.face_drag = D.face_drag.data();
  V.diag = D.diag.data();
  V.pd = D.pd.data();
  const uint64_t m = D.m;
  for (auto* b : {&D.rhs, &D.x, &D.r, &D.w}) b->reserve(m);
  D.basis.reserve(static_cast<size_t>(kRestart + 1) * m);
  D.directions.reserve(static_cast<size_t>(kRestart) * m);
  for (auto* b : {&D.cg_r, &D.cg_z, &D.cg_q}) b->reserve(S.nf);
  D.cg_p.reserve(2 * S.nf);
  D.cg_parts.reserve(5 * static_cast<size_t>(chunks_of(S.nf, V.cg_faces * kResBlock)));
  D.partials.reserve((kRestart + 1) * static_cast<size_t>(chunks_of(m, V.chunk)));
  // slots: [0] the last norm, [1 + 2 * i + pass] the Arnoldi projections.
  D.slots.reserve(2 + 2 * kRestart + 2);
  D.state.reserve(1);
}

// One block solve (solve_resolved_system, cpu_flow.cpp:379-461): flexible
// GMRES (flexible_gmres.hpp:33-108) on [A G; cs D] [u; p] = rhs from zero,
// the solution left in D.x and A applied to it in D.w; rhs (m entries) is
// uploaded here. Returns the momentum rows' relative residual; adds the
// device time to *seconds.
struct BlockSolve {
  uint32_t outer = 0;
  uint64_t inner = 0;
  double relative = 0.0;
  float momentum_relative = 0.0f;
};

BlockSolve block_solve(Context& ctx, ResDevice& D, const std::vector<double>& rhs_host,
                       float relative_tolerance, uint32_t max_outer, float inner_tolerance,
                       uint32_t max_inner, double* seconds) {
  const ResView& V = D.V;
  const uint64_t m = D.m;
  check(hipMemcpyAsync(D.rhs.data(), rhs_host.data(), m * sizeof(double), hipMemcpyHostToDevice,
                       ctx.stream),
        "upload resolved forcing");
  check(hipMemsetAsync(D.x.data(), 0, m * sizeof(double), ctx.stream), "clear solution");
  check(hipEventRecord(ctx.start, ctx.stream), "event");

  const double tolerance = static_cast<double>(relative_tolerance);
  BlockSolve out;
  double* slot = D.slots.data();
  device_dot(ctx, D, D.rhs.data(), D.rhs.data(), m, slot);
  const double rhs_norm = std::sqrt(read_scalar(ctx, slot));
  if (!std::isfinite(rhs_norm)) throw Error(MSIM_RUNTIME_ERROR, "FGMRES non-finite right-hand side");
  uint32_t iterations = 0;
  if (rhs_norm != 0.0) {
    for (;;) {
      apply_system(ctx, D, D.x.data(), D.w.data());
      k_res_residual<<<grid_of(m), kResBlock, 0, ctx.stream>>>(m, D.r.data(), D.rhs.data(),
                                                               D.w.data());
      device_dot(ctx, D, D.r.data(), D.r.data(), m, slot);
      const double beta = std::sqrt(read_scalar(ctx, slot));
      if (!std::isfinite(beta)) throw Error(MSIM_RUNTIME_ERROR, "FGMRES non-finite true residual");
      if (beta <= tolerance * rhs_norm) {
        out.relative = beta / rhs_norm;
        break;
      }
      if (iterations >= max_outer) {
        throw Error(MSIM_NOT_CONVERGED, "resolved-flow FGMRES did not converge: relative residual " +
                                            std::to_string(beta / rhs_norm));
      }
      const uint32_t count = std::min(kRestart, max_outer - iterations);
      std::vector<std::vector<double>> h(count + 1, std::vector<double>(count, 0.0));
      std::vector<double> cosine(count), sine(count), gv(count + 1, 0.0);
      gv[0] = beta;
      double* basis = D.basis.data();
      double* directions = D.directions.data();
      k_res_scaled<<<grid_of(m), kResBlock, 0, ctx.stream>>>(m, basis, 1 / beta, D.r.data());
      uint32_t used = 0;
      for (uint32_t j = 0; j < count; ++j) {
        double* direction = directions + static_cast<size_t>(j) * m;
        const double* vj = basis + static_cast<size_t>(j) * m;
        precondition(ctx, D, vj, direction, inner_tolerance, max_inner, out.inner);
        apply_system(ctx, D, direction, D.w.data());
        // Classical Gram-Schmidt twice, every scalar on the device: the
        // first pass's coefficients at slot[1 ..], the second's at
        // slot[1 + kRestart ..], the norm of what remains at slot[0].
        {
          const uint32_t blocks = chunks_of(m, V.chunk);
          const uint32_t count = j + 1;
          double* h1 = slot + 1;
          double* h2 = slot + 1 + kRestart;
          double* parts = D.partials.data();
          k_res_multidot<<<blocks, kResBlock, 0, ctx.stream>>>(m, V.chunk, basis, count,
                                                               D.w.data(), parts);
          k_res_multisum<<<count, kResBlock, 0, ctx.stream>>>(parts, blocks, h1);
          k_res_cgs<<<blocks, kResBlock, 0, ctx.stream>>>(m, V.chunk, basis, count, h1,
                                                          D.w.data(), 0, parts);
          k_res_multisum<<<count, kResBlock, 0, ctx.stream>>>(parts, blocks, h2);
          k_res_cgs<<<blocks, kResBlock, 0, ctx.stream>>>(m, V.chunk, basis, count, h2,
                                                          D.w.data(), 1, parts);
          k_res_sum<<<1, kResBlock, 0, ctx.stream>>>(parts, blocks, slot);
          check(hipGetLastError(), "launch Gram-Schmidt");
        }
        std::vector<double> col(1 + 2 * kRestart);
        check(hipMemcpyAsync(col.data(), slot, col.size() * sizeof(double),
                             hipMemcpyDeviceToHost, ctx.stream),
              "read Arnoldi column");
        check(hipStreamSynchronize(ctx.stream), "Arnoldi step");
        for (uint32_t i = 0; i <= j; ++i) {
          h[i][j] += col[1 + i];
          h[i][j] += col[1 + kRestart + i];
        }
        h[j + 1][j] = std::sqrt(col[0]);
        if (!std::isfinite(h[j + 1][j])) {
          throw Error(MSIM_RUNTIME_ERROR, "FGMRES non-finite Arnoldi vector");
        }
        const bool happy = h[j + 1][j] <= 1e-14;
        double* next = basis + static_cast<size_t>(j + 1) * m;
        k_res_scaled<<<grid_of(m), kResBlock, 0, ctx.stream>>>(
            m, next, happy ? 0.0 : 1 / h[j + 1][j], D.w.data());
        for (uint32_t i = 0; i < j; ++i) {
          const double upper = cosine[i] * h[i][j] + sine[i] * h[i + 1][j];
          h[i + 1][j] = -sine[i] * h[i][j] + cosine[i] * h[i + 1][j];
          h[i][j] = upper;
        }
        const double diagonal = std::hypot(h[j][j], h[j + 1][j]);
        if (diagonal == 0) throw Error(MSIM_RUNTIME_ERROR, "FGMRES Arnoldi breakdown");
        cosine[j] = h[j][j] / diagonal;
        sine[j] = h[j + 1][j] / diagonal;
        h[j][j] = diagonal;
        h[j + 1][j] = 0;
        gv[j + 1] = -sine[j] * gv[j];
        gv[j] *= cosine[j];
        ++iterations;
        used = j + 1;
        if (happy || std::abs(gv[j + 1]) <= tolerance * rhs_norm) break;
      }
      std::vector<double> weights(used);
      for (uint32_t i = used; i-- > 0;) {
        double value = gv[i];
        for (uint32_t jj = i + 1; jj < used; ++jj) value -= h[i][jj] * weights[jj];
        weights[i] = value / h[i][i];
      }
      UpdateWeights uw{};
      for (uint32_t i = 0; i < used; ++i) uw.w[i] = weights[i];
      k_res_update<<<grid_of(m), kResBlock, 0, ctx.stream>>>(m, D.x.data(), directions, uw, used);
      check(hipGetLastError(), "launch FGMRES update");
    }
  }
  out.outer = iterations;
  // The momentum rows' residual (cpu_flow.cpp:440-452) from A x in D.w.
  apply_system(ctx, D, D.x.data(), D.w.data());
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  std::vector<double> w(D.V.nf);
  check(hipMemcpyAsync(w.data(), D.w.data(), w.size() * sizeof(double), hipMemcpyDeviceToHost,
                       ctx.stream),
        "download resolved residual");
  check(hipStreamSynchronize(ctx.stream), "resolved flow");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  *seconds += ms * 1.0e-3;
  double momentum_square = 0.0, force_square = 0.0;
  for (uint64_t i = 0; i < D.V.nf; ++i) {
    momentum_square += (w[i] - rhs_host[i]) * (w[i] - rhs_host[i]);
    force_square += rhs_host[i] * rhs_host[i];
  }
  out.momentum_relative = static_cast<float>(std::sqrt(momentum_square / force_square));
  return out;
}

// The divergence RMS of a face field over the fluid sites.
double divergence_rms(const ResSystem& S, const std::vector<double>& u) {
  double square = 0.0;
  uint64_t fluid_count = 0;
  for (uint64_t site = 0; site < S.n; ++site) {
    if (S.fluid[site] == 0) continue;
    double d = 0.0;
    const uint32_t co[3] = {static_cast<uint32_t>(site / (static_cast<uint64_t>(S.dims[1]) * S.dims[2])),
                            static_cast<uint32_t>((site / S.dims[2]) % S.dims[1]),
                            static_cast<uint32_t>(site % S.dims[2])};
    for (int c = 0; c < 3; ++c) {
      uint32_t up[3] = {co[0], co[1], co[2]};
      ++up[c];
      d += (u[face_host(S, c, up[0], up[1], up[2])] - u[face_host(S, c, co[0], co[1], co[2])]) /
           static_cast<double>(S.h[c]);
    }
    square += d * d;
    ++fluid_count;
  }
  return fluid_count == 0 ? 0.0 : std::sqrt(square / static_cast<double>(fluid_count));
}

// The forcing of a port system's solve (flow_system.hpp:847-857): on each
// active face minus the gradient of the known pressure, the pressure held
// outside an exterior face standing on its missing side.
std::vector<double> port_force(const ResSystem& S, const std::vector<float>& held,
                               const std::vector<double>& known) {
  std::vector<double> rhs(S.nf + S.n, 0.0);
  for (uint64_t index = 0; index < S.nf; ++index) {
    if (S.active[index] == 0) continue;
    int c;
    uint32_t co[3];
    face_coords_host(S, index, &c, co);
    const auto side = [&](bool upper) {
      if (upper ? co[c] >= S.dims[c] : co[c] == 0) return static_cast<double>(held[index]);
      uint32_t t[3] = {co[0], co[1], co[2]};
      if (!upper) --t[c];
      return known.empty() ? 0.0 : known[site_host(S, t[0], t[1], t[2])];
    };
    rhs[index] = static_cast<double>(
        static_cast<float>(-(side(true) - side(false)) / static_cast<double>(S.h[c])));
  }
  return rhs;
}

}  // namespace

void resolved_flow(Context& ctx, const msim_grid& g, const std::vector<uint8_t>& mask,
                   const float* drag, const msim_resolved_params& params, float* faces_out,
                   msim_resolved_report& report) {
  const ResSystem S = build_system(g, mask, drag, params.axis);
  ResDevice D;
  setup_device(ctx, S, D);
  std::vector<double> rhs_host(D.m, 0.0);
  for (uint64_t i = 0; i < S.nf; ++i) rhs_host[i] = static_cast<double>(S.force[i]);
  double seconds = 0.0;
  const BlockSolve solved =
      block_solve(ctx, D, rhs_host, params.relative_tolerance, params.max_outer_iterations,
                  params.inner_relative_tolerance, params.max_inner_iterations, &seconds);
  std::vector<double> x(S.nf);
  check(hipMemcpyAsync(x.data(), D.x.data(), S.nf * sizeof(double), hipMemcpyDeviceToHost,
                       ctx.stream),
        "download resolved solution");
  check(hipStreamSynchronize(ctx.stream), "resolved flow");
  const double divergence = divergence_rms(S, x);
  // scale_velocity (flow_system.hpp:1368-1415) over the open inlet faces.
  std::vector<float> unscaled(S.nf);
  for (uint64_t i = 0; i < S.nf; ++i) unscaled[i] = static_cast<float>(x[i]);
  double inlet_sum = 0.0;
  uint64_t inlet_count = 0;
  float peak = 0.0f;
  for (uint64_t i = 0; i < S.nf; ++i) {
    if (!std::isfinite(unscaled[i])) {
      throw Error(MSIM_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
    }
    peak = std::max(peak, std::abs(unscaled[i]));
  }
  for (uint64_t i = S.off[S.axis]; i < S.off[S.axis] + S.cnt[S.axis]; ++i) {
    int c;
    uint32_t co[3];
    face_coords_host(S, i, &c, co);
    if (co[S.axis] == 0 && S.active[i] != 0) {
      inlet_sum += unscaled[i];
      ++inlet_count;
    }
  }
  if (inlet_count == 0) throw Error(MSIM_RUNTIME_ERROR, "flow solve has no open inlet faces");
  const float solved_mean = static_cast<float>(inlet_sum / static_cast<double>(inlet_count));
  if (peak == 0.0f || solved_mean <= 1.0e-9f * peak) {
    throw Error(MSIM_RUNTIME_ERROR, "the device carries no through-flow: the outlet is unreachable");
  }
  const float factor = (params.mean_inlet_speed - 0.0f) / solved_mean;
  float scaled_peak = 0.0f;
  for (uint64_t i = 0; i < S.nf; ++i) {
    const float v = unscaled[i] * factor + 0.0f;
    if (!std::isfinite(v)) throw Error(MSIM_RUNTIME_ERROR, "non-finite combined flow velocity");
    faces_out[i] = v;
    scaled_peak = std::max(scaled_peak, std::abs(v));
  }
  report.outer_iterations = solved.outer;
  report.inner_iterations = solved.inner;
  report.relative_residual = static_cast<float>(solved.relative);
  report.momentum_relative_residual = solved.momentum_relative;
  report.divergence_rms = static_cast<float>(divergence * std::abs(factor));
  report.mean_inlet_speed = params.mean_inlet_speed;
  report.max_speed = scaled_peak;
  report.min_gap_voxels = min_gap(S);
  report.seconds = seconds;
}

void resolved_port_flow(Context& ctx, const msim_grid& g, const std::vector<uint8_t>& mask,
                        const float* drag, const msim_flow_port* ports, uint32_t count,
                        const msim_resolved_port_params& params, const ShallowPressure& shallow,
                        float* faces_out, msim_resolved_port_report& report,
                        msim_port_state* states) {
  std::memset(&report, 0, sizeof(report));
  const uint32_t dims[3] = {g.nx, g.ny, g.nz};
  const float spacing[3] = {g.dx, g.dy, g.dz};
  const uint64_t nf = static_cast<uint64_t>(g.nx + 1) * g.ny * g.nz +
                      static_cast<uint64_t>(g.nx) * (g.ny + 1) * g.nz +
                      static_cast<uint64_t>(g.nx) * g.ny * (g.nz + 1);
  // Each port's faces (for_each_port_face, flow_system.hpp:221-236): all of
  // them open (port_flow_boundary); those onto fluid sites carry its flux
  // (make_port_faces, 1124-1130).
  std::vector<uint8_t> open(nf, 0);
  std::vector<std::vector<uint64_t>> all_faces(count), fluid_faces(count);
  std::vector<double> signed_area(count);
  for (uint32_t p = 0; p < count; ++p) {
    const PortGeometry geo = port_geometry(ports[p].face);
    for_each_port_face(dims, ports[p], [&](size_t face, size_t site) {
      open[face] = 1;
      all_faces[p].push_back(face);
      if (mask[site] == 0) fluid_faces[p].push_back(face);
    });
    const double area = static_cast<double>(spacing[geo.in_plane[0]]) * spacing[geo.in_plane[1]];
    signed_area[p] = geo.upper ? -area : area;
  }
  // The system's arrays do not depend on the pressures held: built once.
  const ResSystem S = build_system(g, mask, drag, 3, &open);
  ResDevice D;
  setup_device(ctx, S, D);
  double seconds = 0.0;
  std::vector<float> held_faces(nf, 0.0f);
  std::vector<double> x(S.nf);
  // One block solve holding each port at pressures[port]: its field, as
  // solved, rounded to float (superpose_port_flows' solve).
  auto solve = [&](const std::vector<float>& pressures) {
    for (uint32_t p = 0; p < count; ++p) {
      for (const uint64_t face : all_faces[p]) held_faces[face] = pressures[p];
    }
    std::vector<double> known;
    shallow(pressures, known);
    const std::vector<double> rhs = port_force(S, held_faces, known);
    const BlockSolve solved = block_solve(
        ctx, D, rhs, params.relative_tolerance, params.max_outer_iterations,
        params.inner_relative_tolerance, params.max_inner_iterations, &seconds);
    report.outer_iterations += solved.outer;
    report.inner_iterations += solved.inner;
    report.relative_residual =
        std::max(report.relative_residual, static_cast<float>(solved.relative));
    report.momentum_relative_residual =
        std::max(report.momentum_relative_residual, solved.momentum_relative);
    ++report.solves;
    check(hipMemcpyAsync(x.data(), D.x.data(), S.nf * sizeof(double), hipMemcpyDeviceToHost,
                         ctx.stream),
          "download resolved solution");
    check(hipStreamSynchronize(ctx.stream), "resolved port flow");
    std::vector<float> field(S.nf);
    for (uint64_t f = 0; f < S.nf; ++f) field[f] = static_cast<float>(x[f]);
    return field;
  };
  // superpose_port_flows (flow_system.hpp:1146-1252).
  auto flux = [&](uint32_t q, const auto& velocity) {
    double sum = 0.0;
    for (const uint64_t face : fluid_faces[q]) sum += static_cast<double>(velocity[face]);
    return sum * signed_area[q];
  };
  auto open_area = [&](uint32_t q) {
    return std::abs(signed_area[q]) * static_cast<double>(fluid_faces[q].size());
  };
  std::vector<uint32_t> driven;
  std::vector<float> held(count, 0.0f);
  bool holds_pressure = false;
  for (uint32_t p = 0; p < count; ++p) {
    if (ports[p].kind == MSIM_PORT_SPEED) {
      driven.push_back(p);
    } else {
      held[p] = ports[p].value;
      holds_pressure = holds_pressure || ports[p].value != 0.0f;
    }
  }
  std::vector<double> velocity;
  auto add = [&](const std::vector<float>& field, double weight) {
    if (velocity.empty()) velocity.assign(field.size(), 0.0);
    for (size_t f = 0; f < field.size(); ++f) velocity[f] += weight * static_cast<double>(field[f]);
  };
  if (holds_pressure) add(solve(held), 1.0);
  std::vector<float> pressures = held;
  if (!driven.empty()) {
    const size_t k = driven.size();
    std::vector<std::vector<float>> fields;
    std::vector<std::vector<double>> response(k, std::vector<double>(k + 1, 0.0));
    for (size_t j = 0; j < k; ++j) {
      std::vector<float> unit(count, 0.0f);
      unit[driven[j]] = 1.0f;
      fields.push_back(solve(unit));
      for (size_t m = 0; m < k; ++m) response[m][j] = flux(driven[m], fields.back());
    }
    for (size_t m = 0; m < k; ++m) {
      const double base = velocity.empty() ? 0.0 : flux(driven[m], velocity);
      response[m][k] = static_cast<double>(ports[driven[m]].value) * open_area(driven[m]) - base;
    }
    // Gaussian elimination with partial pivoting on the augmented matrix.
    for (size_t column = 0; column < k; ++column) {
      size_t pivot = column;
      for (size_t row = column + 1; row < k; ++row) {
        if (std::abs(response[row][column]) > std::abs(response[pivot][column])) pivot = row;
      }
      if (response[pivot][column] == 0.0) {
        throw Error(MSIM_RUNTIME_ERROR, "a speed port carries no flow: no pressure port is reachable");
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
      add(fields[j], pressure);
    }
  }
  float max_speed = 0.0f;
  std::vector<double> combined(S.nf);
  for (uint64_t f = 0; f < S.nf; ++f) {
    const float v = static_cast<float>(velocity[f]);
    if (!std::isfinite(v)) throw Error(MSIM_RUNTIME_ERROR, "flow solve produced a non-finite velocity");
    faces_out[f] = v;
    combined[f] = static_cast<double>(v);
    max_speed = std::max(max_speed, std::abs(v));
  }
  if (states != nullptr) {
    for (uint32_t p = 0; p < count; ++p) {
      const double f = flux(p, velocity);
      states[p].pressure = pressures[p];
      states[p].flux = static_cast<float>(f);
      states[p].mean_speed = static_cast<float>(f / open_area(p));
    }
  }
  // The divergence and gaps of the combined field do not depend on the
  // pressures held (flow_system.hpp:1523-1540).
  report.divergence_rms = static_cast<float>(divergence_rms(S, combined));
  report.max_speed = max_speed;
  report.min_gap_voxels = min_gap(S);
  report.seconds = seconds;
}

}  // namespace msim
