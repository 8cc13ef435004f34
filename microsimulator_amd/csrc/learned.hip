// Learned post-smoother of the multilevel preconditioner: a three-layer 3x3
// convolutional network on bfloat16 MFMA (no counterpart in the reference,
// whose only preconditioner is Jacobi, cpp/cpu/cpu_flow.cpp:65-69).
//
// On the finest level the cycle's post-smoothing
//     z = x + omega D^-1 r'            (r' = r - A x after the coarse correction)
// becomes one learned step
//     z = x + omega D^-1 r' + s N(r'),
// where N works tile by tile (16 x 32 sites, receptive region grown by 3):
//   u = D^-1 r',  s = max |u| over the region,  v = u / s,
//   in = [v, v cx, v cy, v cb]      (cx, cy, cb: the site's couplings times dinv)
//   a1 = clip(conv(in, W1)) * mask           4 -> 16 channels
//   a2 = clip(conv(a1, W2)) * mask           16 -> 16
//   N  = conv(a2, W3) * mask                 16 -> 1
// clip(h) = min(max(h, -1), 1); mask = on the grid and dinv > 0. The network
// is odd and has no biases, so the smoother is homogeneous of degree one in
// r'; with W3 = 0 it is the damped-Jacobi post-smoother. The clipped
// activation and its mask are one median-of-three per value:
// med3(h, -mask, mask).
//
// One workgroup of 256 threads handles one tile (three fit on a CU). The
// staged region is the tile grown by 4 (24 x 40 sites, row pitch 40); every
// layer is computed on whole 16-pixel runs of that linear layout, so tap
// (dx, dy) is a constant offset dx * 40 + dy. Pixels near the region's side
// columns come out as garbage that no kept pixel reads. A thread issues all
// of its global loads before it computes, so a tile pays one round trip. Each layer is a GEMM on
// __builtin_amdgcn_mfma_f32_16x16x16bf16_1k with the output channels in M,
// the pixels in N and (tap, input channel) in K. Lane l supplies
// A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j][l % 16] and receives
// D[4 (l / 16) + j][l % 16], so it reads and writes the four channels of one
// pixel as one 8-byte LDS access (activations kept as four planes of four
// channels, so a lane group's accesses are adjacent). The last layer puts the 9 taps in M,
// T[tap][p] = W3[tap] . a2[p], and gathers N[p] = sum_tap T[tap][p + tap].
//
// Training runs on the same tiles: forward, a loss kernel giving dL/dz, a
// backward kernel that writes per-tile weight gradients, a fixed-order
// reduction and Adam. The loss of a pair (r', e') with A e' = r' is
// (e' - z)^T A (e' - z) / e'^T A e', averaged over the batch.
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

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTX = 16;                  // tile rows (x)
constexpr int kTY = 32;                  // tile columns (y)
constexpr int kRows = kTX + 8;           // staged region: the tile grown by 4
constexpr int kPitch = kTY + 8;
constexpr int kPix = kRows * kPitch;     // 960
constexpr int kThreads = 256;            // forward kernels
constexpr int kBThreads = 512;           // backward kernel
constexpr int kC0 = 4, kC1 = 16;
constexpr int kW1 = 9 * kC0 * kC1;       // [tap][c][o]
constexpr int kW2 = 9 * kC1 * kC1;
constexpr int kW3 = 9 * kC1;
constexpr int kWeights = kW1 + kW2 + kW3;
static_assert(kWeights == MSIM_LEARNED_WEIGHTS, "weight count");
constexpr int kFrags = 3 + 9 + 1;        // packed A fragments: layer 1, layer 2, layer 3
constexpr float kOmega = 0.8f;
constexpr float kCorrection = 1.8f;

// Linear pixel ranges of each layer's outputs (whole rows of the region).
constexpr int kL1Begin = 2 * kPitch, kL1End = (kRows - 2) * kPitch;  // rows 2..21
constexpr int kL2Begin = 3 * kPitch, kL2End = (kRows - 3) * kPitch;  // rows 3..20
constexpr int kL1Tiles = (kL1End - kL1Begin) / 16;                  // 50
constexpr int kL2Tiles = (kL2End - kL2Begin) / 16;                  // 45
constexpr int kL2Pix = kL2End - kL2Begin;                           // 720
// a1 is kept on [kA1Base, kA1Base + kA1Pix): layer 2 reads up to 73 pixels
// beyond layer 1's range; the 16-pixel guards on each side are zero.
constexpr int kA1Base = kL1Begin - 16;
constexpr int kA1Pix = kL1End - kL1Begin + 32;
static_assert(kL2Begin - (kPitch + 1) >= kA1Base && kL2End + kPitch + 1 <= kA1Base + kA1Pix,
              "layer 2 reads inside the kept a1 range");
static_assert(kL1Begin - (kPitch + 1) >= 0 && kL1End + kPitch + 1 <= kPix,
              "layer 1 reads inside the region");
static_assert((kL1End - kL1Begin) % 16 == 0 && kL2Pix % 16 == 0, "whole pixel runs");
// Activations are stored as four planes of [pixel][4 channels] (bf16), one per
// group of four channels, so the 16 lanes of a lane group touch 16 adjacent
// 8-byte words: bank-conflict free.
constexpr int kA1Plane = 4 * kA1Pix;   // shorts per a1 plane
constexpr int kA2Plane = 4 * kL2Pix;   // shorts per a2 plane

__device__ constexpr int tap_offset(int t) { return (t / 3 - 1) * kPitch + (t % 3 - 1); }

__device__ inline unsigned short bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}
// The same rounding in one instruction for two values (v_cvt_pk_bf16_f32).
__device__ inline short to_bf16(float f) {
  return __builtin_bit_cast(short, static_cast<__bf16>(f));
}
__device__ inline float bf16_value(unsigned short b) {
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}
// bf16(clip(acc) * mask) of four channels: med3(h, -1, 1) on a kept pixel,
// med3(h, 0, 0) = 0 on a masked one.
__device__ inline bf16x4 activate(f32x4 acc, bool on) {
  const float keep = on ? 1.0f : 0.0f;
  bf16x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) o[j] = to_bf16(__builtin_amdgcn_fmed3f(acc[j], -keep, keep));
  return o;
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline bf16x4 lds4(const unsigned short* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}
__device__ inline void store4(unsigned short* p, bf16x4 v) {
  *reinterpret_cast<bf16x4*>(p) = v;
}

struct Geometry {
  const float* wx;
  const float* wy;
  const float* bc;
  const float* dinv;
  uint32_t nx, ny;
};

// Largest value over a block of NT threads; every thread gets it.
template <int NT>
__device__ inline float block_max_all(float v, float* lds) {
  for (int o = kWave / 2; o > 0; o /= 2) v = fmaxf(v, __shfl_xor(v, o));
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  float m = 0.0f;
  for (int w = 0; w < NT / kWave; ++w) m = fmaxf(m, lds[w]);
  __syncthreads();
  return m;
}

// Per-site terms of one region pixel whose residual r' is known: u and the
// three coupling features. Off the grid everything is zero.
__device__ inline float4 site_features(const Geometry& G, size_t i, uint32_t gx, uint32_t gy,
                                       float rprime, float d) {
  const float wxl = G.wx[i];
  const float wxu = gx + 1 < G.nx ? G.wx[i + G.ny] : 0.0f;
  const float wyl = G.wy[i];
  const float wyu = gy + 1 < G.ny ? G.wy[i + 1] : 0.0f;
  return make_float4(d * rprime, (wxl + wxu) * d, (wyl + wyu) * d, G.bc[i] * d);
}

__device__ inline bool region_site(int p, int fx0, int fy0, const Geometry& G, uint32_t* gx,
                                   uint32_t* gy) {
  const int x = fx0 - 4 + p / kPitch, y = fy0 - 4 + p % kPitch;
  *gx = static_cast<uint32_t>(x);
  *gy = static_cast<uint32_t>(y);
  return x >= 0 && y >= 0 && static_cast<uint32_t>(x) < G.nx && static_cast<uint32_t>(y) < G.ny;
}

__device__ inline bool feature_pixel(int p) {
  const int row = p / kPitch, col = p % kPitch;
  return row >= 1 && row < kRows - 1 && col >= 1 && col < kPitch - 1;
}

__device__ inline bool tile_pixel(int p) {
  const int row = p / kPitch, col = p % kPitch;
  return row >= 4 && row < 4 + kTX && col >= 4 && col < 4 + kTY;
}

// in = bf16(v, v cx, v cy, v cb) with v = u / s for every region pixel.
template <int NT>
__device__ inline void write_inputs(const float4* feat, unsigned short* in, float s) {
  const float inv = s > 0.0f ? 1.0f / s : 0.0f;
  for (int p = threadIdx.x; p < kPix; p += NT) {
    const float4 f = feat[p];
    const float v = f.x * inv;
    bf16x4 w;
    w[0] = to_bf16(v);
    w[1] = to_bf16(v * f.y);
    w[2] = to_bf16(v * f.z);
    w[3] = to_bf16(v * f.w);
    store4(in + 4 * p, w);
  }
}

// Layers 1 and 2: in -> a1 (kept on [kA1Base, ...)) -> a2 (kept on layer 2's
// range). `frag` holds the lane's A fragments. Each wave runs two pixel runs
// at a time (two independent accumulator chains) and loads a run's B
// fragments before its first MFMA.
template <int NT>
__device__ inline void layers12(const unsigned short* in, unsigned short* a1, unsigned short* a2,
                                const uint8_t* mask, const bf16x4* frag) {
  constexpr int kWaves = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int m = lane & 15, g = lane >> 4;
  // Zero guards of a1 (read only for pixels nothing keeps): 16 pixels at
  // each end of each plane.
  for (int k = threadIdx.x; k < 4 * 32 * 4; k += NT) {
    const int plane = k / 128, e = k % 128, px = e / 4 < 16 ? e / 4 : (kA1Pix - 32) + e / 4;
    a1[plane * kA1Plane + 4 * px + (e & 3)] = 0;
  }
  for (int mt = wave; mt < kL1Tiles; mt += 2 * kWaves) {
    const bool two = mt + kWaves < kL1Tiles;
    const int p0 = kL1Begin + 16 * mt + m;
    const int p1 = kL1Begin + 16 * (two ? mt + kWaves : mt) + m;
    bf16x4 b0[3], b1[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const int tap = 4 * s + g;
      b0[s] = bf16x4{0, 0, 0, 0};
      b1[s] = bf16x4{0, 0, 0, 0};
      if (tap < 9) {
        b0[s] = lds4(in + 4 * (p0 + tap_offset(tap)));
        b1[s] = lds4(in + 4 * (p1 + tap_offset(tap)));
      }
    }
    f32x4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      c0 = mfma(frag[s], b0[s], c0);
      c1 = mfma(frag[s], b1[s], c1);
    }
    store4(a1 + g * kA1Plane + 4 * (p0 - kA1Base), activate(c0, mask[p0] != 0));
    if (two) store4(a1 + g * kA1Plane + 4 * (p1 - kA1Base), activate(c1, mask[p1] != 0));
  }
  __syncthreads();
  for (int mt = wave; mt < kL2Tiles; mt += 2 * kWaves) {
    const bool two = mt + kWaves < kL2Tiles;
    const int p0 = kL2Begin + 16 * mt + m;
    const int p1 = kL2Begin + 16 * (two ? mt + kWaves : mt) + m;
    bf16x4 b0[9], b1[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      b0[t] = lds4(a1 + g * kA1Plane + 4 * (p0 + tap_offset(t) - kA1Base));
      b1[t] = lds4(a1 + g * kA1Plane + 4 * (p1 + tap_offset(t) - kA1Base));
    }
    f32x4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      c0 = mfma(frag[3 + t], b0[t], c0);
      c1 = mfma(frag[3 + t], b1[t], c1);
    }
    store4(a2 + g * kA2Plane + 4 * (p0 - kL2Begin), activate(c0, mask[p0] != 0));
    if (two) store4(a2 + g * kA2Plane + 4 * (p1 - kL2Begin), activate(c1, mask[p1] != 0));
  }
  __syncthreads();
}

// Layer 3 with the taps in M: T[tap][p] = W3[tap] . a2[p] over layer 2's range.
template <int NT>
__device__ inline void layer3(const unsigned short* a2, float* T, const bf16x4* frag) {
  constexpr int kWaves = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int m = lane & 15, g = lane >> 4;
  for (int mt = wave; mt < kL2Tiles; mt += kWaves) {
    const int q = 16 * mt + m;
    const bf16x4 b = lds4(a2 + g * kA2Plane + 4 * q);
    const f32x4 acc = mfma(frag[12], b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int tap = 4 * g + j;
      if (tap < 9) T[tap * kL2Pix + q] = acc[j];
    }
  }
  __syncthreads();
}

// N at region pixel p (a tile pixel) from T.
__device__ inline float gather(const float* T, int p) {
  float n = 0.0f;
#pragma unroll
  for (int t = 0; t < 9; ++t) n += T[t * kL2Pix + p + tap_offset(t) - kL2Begin];
  return n;
}

__device__ inline void load_frags(const bf16x4* packed, bf16x4* frag) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int f = 0; f < kFrags; ++f) frag[f] = packed[f * kWave + lane];
}

// Shared-memory plan of the forward kernels (bytes): 52.7 KB, three
// workgroups per CU.
constexpr int kOffA = 0;                                   // x (float) | in (bf16x4) | a2
constexpr int kSizeA = std::max({4 * kPix, 8 * kPix, 2 * 16 * kL2Pix});
constexpr int kOffB = kOffA + kSizeA;                      // feat (float4) | a1 | T
constexpr int kSizeB = std::max({16 * kPix, 2 * 16 * kA1Pix, 4 * 9 * kL2Pix});
constexpr int kOffZc = kOffB + kSizeB;                     // float[kTX * kTY]: classical z
constexpr int kOffMask = kOffZc + 4 * kTX * kTY;           // uint8[kPix]
constexpr int kOffRed = kOffMask + kPix;
constexpr int kForwardLds = kOffRed + 4 * 16;
static_assert(kOffB % 16 == 0 && kOffZc % 16 == 0 && kOffRed % 16 == 0, "aligned");
static_assert(3 * kForwardLds <= 160 * 1024, "three workgroups per CU");

struct UpOut {
  float* z;
  const float* q;   // previous A p; null: no dots
  double* zr;       // [2][blocks]
  double* zq;
};

// The finest level's up pass with the learned post-smoother:
// x = omega D^-1 r + c P e_c, r' = r - A x, z = x + omega D^-1 r' + s N(r');
// with dots z.r and z.q for flexible CG.
__global__ __launch_bounds__(kThreads, 3) void k_learn_up(Geometry G, const float* __restrict__ r,
                                                       const float* __restrict__ xp,
                                                       const float* __restrict__ ec, uint32_t nyc,
                                                       const CgState* st,
                                                       const bf16x4* __restrict__ packed,
                                                       UpOut out, uint32_t tiles_y,
                                                       uint32_t tiles) {
  __shared__ __align__(16) unsigned char lds[kForwardLds];
  __shared__ double s_dred[kThreads / kWave];
  float* s_x = reinterpret_cast<float*>(lds + kOffA);
  unsigned short* s_in = reinterpret_cast<unsigned short*>(lds + kOffA);
  unsigned short* s_a2 = s_in;
  unsigned short* s_a1 = reinterpret_cast<unsigned short*>(lds + kOffB);
  float* s_T = reinterpret_cast<float*>(lds + kOffB);
  float* s_zc = reinterpret_cast<float*>(lds + kOffZc);
  uint8_t* s_mask = lds + kOffMask;
  float* s_red = reinterpret_cast<float*>(lds + kOffRed);
  if (st->failure != 0 || st->halt != 0) return;
  bf16x4 frag[kFrags];
  load_frags(packed, frag);
  // A contiguous run of tiles per workgroup, and a contiguous band of tiles
  // per XCD, so tiles that share halo rows meet in one L2.
  const uint32_t per = (tiles + gridDim.x - 1) / gridDim.x;
  const uint32_t t0 = xcd_tile(blockIdx.x, gridDim.x) * per, t1 = min(tiles, t0 + per);
  double zr = 0.0, zq = 0.0;
  // Staging map: 240 threads cover the 24 x 40 region as 6 rows of 40
  // columns, four times over (rows r0, r0 + 6, r0 + 12, r0 + 18), so a thread
  // keeps one column and its four pixels differ by 6 rows.
  constexpr int kStageRows = 6, kStageThreads = kStageRows * kPitch, kPer = kRows / kStageRows;
  constexpr int kOutPer = kTX * kTY / kThreads;  // output pixels per thread
  static_assert(kOutPer * kThreads == kTX * kTY, "whole output pixels per thread");
  static_assert(kStageRows * kPer == kRows && kStageThreads <= kThreads, "staging map");
  const int scol = static_cast<int>(threadIdx.x) % kPitch;
  const int srow = static_cast<int>(threadIdx.x) / kPitch;
  const bool stager = threadIdx.x < kStageThreads;
  const bool feat_col = scol >= 1 && scol < kPitch - 1;
  // Every global value a thread's pixels of tile t need, in one round of
  // loads: r, dinv, e_c and, on feature pixels, the operator terms. The next
  // tile's round is issued before this tile's network runs.
  struct Stage {
    float r[kPer], d[kPer], x0[kPer], ec[kPer], wxl[kPer], wxu[kPer], wyl[kPer], wyu[kPer], b[kPer];
    float qo[kOutPer];  // q at the thread's output pixels
    uint32_t on, feat;
  };
  auto load_tile = [&](uint32_t t) {
    Stage v{};
    const int fx0 = static_cast<int>(t / tiles_y) * kTX;
    const int fy0 = static_cast<int>(t % tiles_y) * kTY;
    const int gy = fy0 - 4 + scol;
    const bool col_in = stager && gy >= 0 && gy < static_cast<int>(G.ny);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int row = srow + kStageRows * k;
      const int gx = fx0 - 4 + row;
      if (col_in && gx >= 0 && gx < static_cast<int>(G.nx)) {
        v.on |= 1u << k;
        const size_t i = static_cast<size_t>(gx) * G.ny + static_cast<uint32_t>(gy);
        v.r[k] = r[i];
        v.d[k] = G.dinv[i];
        v.x0[k] = xp != nullptr ? xp[i] : kOmega * v.d[k] * v.r[k];
        v.ec[k] = ec[static_cast<size_t>(gx / 2) * nyc + static_cast<uint32_t>(gy) / 2];
        if (feat_col && row >= 1 && row < kRows - 1) {
          v.feat |= 1u << k;
          v.wxl[k] = G.wx[i];
          v.wxu[k] = static_cast<uint32_t>(gx) + 1 < G.nx ? G.wx[i + G.ny] : 0.0f;
          v.wyl[k] = G.wy[i];
          v.wyu[k] = static_cast<uint32_t>(gy) + 1 < G.ny ? G.wy[i + 1] : 0.0f;
          v.b[k] = G.bc[i];
        }
      }
    }
#pragma unroll
    for (int m = 0; m < kOutPer; ++m) {
      const int k = static_cast<int>(threadIdx.x) + kThreads * m;
      const uint32_t gx = static_cast<uint32_t>(fx0 + k / kTY);
      const uint32_t gy = static_cast<uint32_t>(fy0 + k % kTY);
      if (out.q != nullptr && gx < G.nx && gy < G.ny) {
        const size_t i = static_cast<size_t>(gx) * G.ny + gy;
        v.qo[m] = out.q[i];
      }
    }
    return v;
  };
  Stage cur{};
  if (t0 < t1) cur = load_tile(t0);
  for (uint32_t t = t0; t < t1; ++t) {
    const int fx0 = static_cast<int>(t / tiles_y) * kTX;
    const int fy0 = static_cast<int>(t % tiles_y) * kTY;
    // x = x_pre + c P e_c over the region (x_pre: the pre-smoothed iterate,
    // omega D^-1 r after one sweep from zero).
    const uint32_t on = cur.on, feat = cur.feat;
    float xv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      xv[k] = ((on >> k) & 1u) ? cur.x0[k] + kCorrection * cur.ec[k] : 0.0f;
    }
    const float* rv = cur.r;
    const float* dv = cur.d;
    const float* wxl = cur.wxl;
    const float* wxu = cur.wxu;
    const float* wyl = cur.wyl;
    const float* wyu = cur.wyu;
    const float* bv = cur.b;
    if (stager) {
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int p = (srow + kStageRows * k) * kPitch + scol;
        s_x[p] = xv[k];
        s_mask[p] = ((on >> k) & 1u) && dv[k] > 0.0f ? 1 : 0;
      }
    }
    __syncthreads();
    // r' = r - A x, u = D^-1 r' and the features; the damped-Jacobi
    // post-smoother in the plain cycle's operation order.
    float uv[kPer], cxv[kPer], cyv[kPer], cbv[kPer];
    float umax = 0.0f;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int row = srow + kStageRows * k;
      const int p = row * kPitch + scol;
      uv[k] = cxv[k] = cyv[k] = cbv[k] = 0.0f;
      if ((feat >> k) & 1u) {
        const float xc = s_x[p];
        float ax = bv[k] * xc;
        ax += wxl[k] * (xc - s_x[p - kPitch]);
        ax += wxu[k] * (xc - s_x[p + kPitch]);
        ax += wyl[k] * (xc - s_x[p - 1]);
        ax += wyu[k] * (xc - s_x[p + 1]);
        const float d = dv[k];
        const float res = rv[k] - ax;
        uv[k] = d * res;
        cxv[k] = (wxl[k] + wxu[k]) * d;
        cyv[k] = (wyl[k] + wyu[k]) * d;
        cbv[k] = bv[k] * d;
        umax = fmaxf(umax, fabsf(uv[k]));
        if (row >= 4 && row < 4 + kTX && scol >= 4 && scol < 4 + kTY) {
          s_zc[(row - 4) * kTY + scol - 4] = xc + kOmega * d * res;
        }
      }
    }
    const Stage nxt = t + 1 < t1 ? load_tile(t + 1) : Stage{};
    const float s = block_max_all<kThreads>(umax, s_red);  // has barriers
    // in = bf16(v, v cx, v cy, v cb) with v = u / s; zero off the features.
    if (stager) {
      const float inv = s > 0.0f ? 1.0f / s : 0.0f;
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int p = (srow + kStageRows * k) * kPitch + scol;
        const float v = uv[k] * inv;
        bf16x4 w;
        w[0] = to_bf16(v);
        w[1] = to_bf16(v * cxv[k]);
        w[2] = to_bf16(v * cyv[k]);
        w[3] = to_bf16(v * cbv[k]);
        store4(s_in + 4 * p, w);
      }
    }
    __syncthreads();
    layers12<kThreads>(s_in, s_a1, s_a2, s_mask, frag);
    layer3<kThreads>(s_a2, s_T, frag);
#pragma unroll
    for (int m = 0; m < kOutPer; ++m) {
      const int k = static_cast<int>(threadIdx.x) + kThreads * m;
      const int lx = k / kTY, ly = k % kTY;
      const uint32_t gx = static_cast<uint32_t>(fx0 + lx), gy = static_cast<uint32_t>(fy0 + ly);
      if (gx >= G.nx || gy >= G.ny) continue;
      const int p = (lx + 4) * kPitch + ly + 4;
      const float n = s_mask[p] != 0 ? gather(s_T, p) : 0.0f;
      const float z = s_zc[k] + s * n;
      const size_t i = static_cast<size_t>(gx) * G.ny + gy;
      out.z[i] = z;
      if (out.q != nullptr) {
        zr += static_cast<double>(z) * static_cast<double>(r[i]);
        zq += static_cast<double>(z) * static_cast<double>(cur.qo[m]);
      }
    }
    __syncthreads();  // the next tile rewrites every shared array
    cur = nxt;
  }
  if (out.q != nullptr) {
    const uint32_t k = st->it_update + 1;
    const double a = block_sum<kThreads>(zr, s_dred);
    const double b = block_sum<kThreads>(zq, s_dred);
    if (threadIdx.x == 0) {
      out.zr[(k & 1) * gridDim.x + blockIdx.x] = a;
      out.zq[(k & 1) * gridDim.x + blockIdx.x] = b;
    }
  }
}

// ----------------------------------------------------------------------------
// Training set kernels. Pairs are rows of [pairs][n] arrays; a geometry is
// four rows of [geometries][4][n] (wx, wy, bc, dinv).
// ----------------------------------------------------------------------------

struct TrainView {
  const float* ops;
  const float* rp;
  const float* ep;
  const uint32_t* geometry;  // per pair
  const double* energy;      // e'^T A e' per pair
  uint32_t nx, ny;
};

__device__ inline Geometry geometry_of(const TrainView& V, uint32_t pair) {
  const size_t n = static_cast<size_t>(V.nx) * V.ny;
  const float* o = V.ops + static_cast<size_t>(V.geometry[pair]) * 4 * n;
  return Geometry{o, o + n, o + 2 * n, o + 3 * n, V.nx, V.ny};
}

// Stage u and the features of a pair's r' over the region; returns the
// thread's largest |u|.
template <int NT>
__device__ inline float stage_pair(const Geometry& G, const float* rp, int fx0, int fy0,
                                   float4* feat, uint8_t* mask) {
  float umax = 0.0f;
  for (int p = threadIdx.x; p < kPix; p += NT) {
    uint32_t gx, gy;
    float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint8_t on = 0;
    if (region_site(p, fx0, fy0, G, &gx, &gy)) {
      const size_t i = static_cast<size_t>(gx) * G.ny + gy;
      const float d = G.dinv[i];
      on = d > 0.0f ? 1 : 0;
      if (feature_pixel(p)) {
        f = site_features(G, i, gx, gy, rp[i], d);
        umax = fmaxf(umax, fabsf(f.x));
      }
    }
    feat[p] = f;
    mask[p] = on;
  }
  return umax;
}

// z = omega D^-1 r' + s N(r') for the pairs order[0 .. gridDim.z).
__global__ __launch_bounds__(kThreads) void k_learn_forward(TrainView V,
                                                            const uint32_t* __restrict__ order,
                                                            const bf16x4* __restrict__ packed,
                                                            float* __restrict__ zout) {
  __shared__ __align__(16) unsigned char lds[kForwardLds];
  unsigned short* s_in = reinterpret_cast<unsigned short*>(lds + kOffA);
  unsigned short* s_a2 = s_in;
  float4* s_feat = reinterpret_cast<float4*>(lds + kOffB);
  unsigned short* s_a1 = reinterpret_cast<unsigned short*>(lds + kOffB);
  float* s_T = reinterpret_cast<float*>(lds + kOffB);
  float* s_zc = reinterpret_cast<float*>(lds + kOffZc);
  uint8_t* s_mask = lds + kOffMask;
  float* s_red = reinterpret_cast<float*>(lds + kOffRed);
  bf16x4 frag[kFrags];
  load_frags(packed, frag);
  const uint32_t pair = order[blockIdx.z];
  const Geometry G = geometry_of(V, pair);
  const size_t n = static_cast<size_t>(V.nx) * V.ny;
  const float* rp = V.rp + static_cast<size_t>(pair) * n;
  float* z = zout + static_cast<size_t>(blockIdx.z) * n;
  const int fx0 = static_cast<int>(blockIdx.x) * kTX, fy0 = static_cast<int>(blockIdx.y) * kTY;
  const float umax = stage_pair<kThreads>(G, rp, fx0, fy0, s_feat, s_mask);
  const float s = block_max_all<kThreads>(umax, s_red);
  for (int p = threadIdx.x; p < kPix; p += kThreads) {
    if (tile_pixel(p)) s_zc[(p / kPitch - 4) * kTY + p % kPitch - 4] = kOmega * s_feat[p].x;
  }
  write_inputs<kThreads>(s_feat, s_in, s);
  __syncthreads();
  layers12<kThreads>(s_in, s_a1, s_a2, s_mask, frag);
  layer3<kThreads>(s_a2, s_T, frag);
  for (int k = threadIdx.x; k < kTX * kTY; k += kThreads) {
    const int lx = k / kTY, ly = k % kTY;
    const uint32_t gx = static_cast<uint32_t>(fx0 + lx), gy = static_cast<uint32_t>(fy0 + ly);
    if (gx >= G.nx || gy >= G.ny) continue;
    const int p = (lx + 4) * kPitch + ly + 4;
    const float nv = s_mask[p] != 0 ? gather(s_T, p) : 0.0f;
    z[static_cast<size_t>(gx) * G.ny + gy] = s_zc[k] + s * nv;
  }
}

// A applied to the field f at site (x, y) of geometry G (FP32, the order of
// the multilevel kernels).
__device__ inline float apply_at(const Geometry& G, const float* f, uint32_t x, uint32_t y) {
  const size_t i = static_cast<size_t>(x) * G.ny + y;
  const float c = f[i];
  float a = G.bc[i] * c;
  if (x > 0) a += G.wx[i] * (c - f[i - G.ny]);
  if (x + 1 < G.nx) a += G.wx[i + G.ny] * (c - f[i + G.ny]);
  if (y > 0) a += G.wy[i] * (c - f[i - 1]);
  if (y + 1 < G.ny) a += G.wy[i + 1] * (c - f[i + 1]);
  return a;
}

// d = e' - z, gz = dL/dz = -2 A d / (E * batch), and per-block sums of
// d^T A d / E. Grid: (blocks, batch).
__global__ __launch_bounds__(kStream) void k_learn_loss(TrainView V,
                                                        const uint32_t* __restrict__ order,
                                                        const float* __restrict__ d_all,
                                                        float* __restrict__ gz, double* parts) {
  __shared__ double s_red[kStream / kWave];
  const uint32_t pair = order[blockIdx.y];
  const Geometry G = geometry_of(V, pair);
  const size_t n = static_cast<size_t>(V.nx) * V.ny;
  const float* d = d_all + static_cast<size_t>(blockIdx.y) * n;
  const double E = V.energy[pair];
  const float scale = static_cast<float>(-2.0 / (E * gridDim.y));
  double acc = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kStream;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kStream + threadIdx.x; i < n; i += stride) {
    const uint32_t x = static_cast<uint32_t>(i / G.ny), y = static_cast<uint32_t>(i % G.ny);
    const float ad = apply_at(G, d, x, y);
    gz[static_cast<size_t>(blockIdx.y) * n + i] = scale * ad;
    acc += static_cast<double>(d[i]) * static_cast<double>(ad);
  }
  const double part = block_sum<kStream>(acc, s_red);
  if (threadIdx.x == 0) parts[static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x] = part / E;
}

// d = e' - z for the batch (the loss kernel reads d at neighbours).
__global__ __launch_bounds__(kStream) void k_learn_diff(TrainView V,
                                                        const uint32_t* __restrict__ order,
                                                        const float* __restrict__ zin,
                                                        float* __restrict__ d) {
  const uint32_t pair = order[blockIdx.y];
  const size_t n = static_cast<size_t>(V.nx) * V.ny;
  const float* e = V.ep + static_cast<size_t>(pair) * n;
  const float* z = zin + static_cast<size_t>(blockIdx.y) * n;
  float* o = d + static_cast<size_t>(blockIdx.y) * n;
  const size_t stride = static_cast<size_t>(gridDim.x) * kStream;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kStream + threadIdx.x; i < n; i += stride)
    o[i] = e[i] - z[i];
}

// Shared-memory plan of the backward kernel (bytes).
constexpr int kBOffMask = 0;                               // uint8[kPix]
constexpr int kBOffIn = kBOffMask + kPix;                  // bf16x4[kPix]
constexpr int kBOffA1 = kBOffIn + 8 * kPix;                // feat | a1 -> g_h1
constexpr int kBSizeA1 = std::max(16 * kPix, 2 * 16 * kA1Pix);
constexpr int kBOffA2 = kBOffA1 + kBSizeA1;                // a2
constexpr int kBOffG = kBOffA2 + 2 * 16 * kL2Pix;          // float[kPix]: s dL/dz on the tile
constexpr int kBOffGh2 = kBOffG + 4 * kPix;                // g_h2
constexpr int kBOffRed = kBOffGh2 + 2 * 16 * kL2Pix;
constexpr int kBackwardLds = kBOffRed + 4 * 16;
static_assert(kBOffIn % 16 == 0 && kBOffA1 % 16 == 0 && kBOffA2 % 16 == 0 &&
                  kBOffG % 16 == 0 && kBOffGh2 % 16 == 0,
              "aligned");
static_assert(kBackwardLds <= 160 * 1024, "fits the workgroup's LDS");

// Per-tile weight gradients of the batch loss for pairs order[blockIdx.z],
// written to grads[tile][kWeights].
__global__ __launch_bounds__(kBThreads) void k_learn_backward(TrainView V,
                                                             const uint32_t* __restrict__ order,
                                                             const float* __restrict__ w,
                                                             const bf16x4* __restrict__ packed,
                                                             const float* __restrict__ gz_all,
                                                             float* __restrict__ grads) {
  __shared__ __align__(16) unsigned char lds[kBackwardLds];
  uint8_t* s_mask = lds + kBOffMask;
  unsigned short* s_in = reinterpret_cast<unsigned short*>(lds + kBOffIn);
  float4* s_feat = reinterpret_cast<float4*>(lds + kBOffA1);
  unsigned short* s_a1 = reinterpret_cast<unsigned short*>(lds + kBOffA1);
  unsigned short* s_a2 = reinterpret_cast<unsigned short*>(lds + kBOffA2);
  float* s_g = reinterpret_cast<float*>(lds + kBOffG);
  unsigned short* s_gh2 = reinterpret_cast<unsigned short*>(lds + kBOffGh2);
  float* s_red = reinterpret_cast<float*>(lds + kBOffRed);
  constexpr int kWaves = kBThreads / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int m = lane & 15, g = lane >> 4;
  bf16x4 frag[kFrags];
  load_frags(packed, frag);
  const uint32_t pair = order[blockIdx.z];
  const Geometry G = geometry_of(V, pair);
  const size_t n = static_cast<size_t>(V.nx) * V.ny;
  const float* rp = V.rp + static_cast<size_t>(pair) * n;
  const float* gz = gz_all + static_cast<size_t>(blockIdx.z) * n;
  const uint32_t tiles = gridDim.x * gridDim.y;
  const uint32_t tile = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  float* out = grads + static_cast<size_t>(tile) * kWeights;
  (void)tiles;
  const int fx0 = static_cast<int>(blockIdx.x) * kTX, fy0 = static_cast<int>(blockIdx.y) * kTY;
  const float umax = stage_pair<kBThreads>(G, rp, fx0, fy0, s_feat, s_mask);
  const float s = block_max_all<kBThreads>(umax, s_red);
  write_inputs<kBThreads>(s_feat, s_in, s);
  __syncthreads();
  layers12<kBThreads>(s_in, s_a1, s_a2, s_mask, frag);
  // g = s dL/dz on the tile's masked sites, zero elsewhere.
  for (int p = threadIdx.x; p < kPix; p += kBThreads) {
    float v = 0.0f;
    uint32_t gx, gy;
    if (tile_pixel(p) && s_mask[p] != 0 && region_site(p, fx0, fy0, G, &gx, &gy))
      v = s * gz[static_cast<size_t>(gx) * G.ny + gy];
    s_g[p] = v;
  }
  __syncthreads();
  // dW3 (wave kWaves - 1) and g_h2 = (W3^T g shifted) * clip'(h2) (the other waves).
  if (wave == kWaves - 1) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < kL2Tiles; ++ks) {
      const int q0 = kL2Begin + 16 * ks;
      bf16x4 a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = q0 + 4 * g + j;
        a[j] = static_cast<short>(s_a2[(m >> 2) * kA2Plane + 4 * (q - kL2Begin) + (m & 3)]);
        b[j] = m < 9 ? to_bf16(s_g[q - tap_offset(m)]) : 0;
      }
      acc = mfma(a, b, acc);
    }
    if (m < 9) {
#pragma unroll
      for (int j = 0; j < 4; ++j) out[kW1 + kW2 + m * kC1 + 4 * g + j] = acc[j];
    }
  } else {
    bf16x4 wg3;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = 4 * g + j;
      wg3[j] = t < 9 ? to_bf16(w[kW1 + kW2 + t * kC1 + m]) : 0;
    }
    for (int mt = wave; mt < kL2Tiles; mt += kWaves - 1) {
      const int q = kL2Begin + 16 * mt + m;
      bf16x4 b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int t = 4 * g + j;
        b[j] = t < 9 ? to_bf16(s_g[q - tap_offset(t)]) : 0;
      }
      const f32x4 acc = mfma(wg3, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
      const bf16x4 av = lds4(s_a2 + g * kA2Plane + 4 * (q - kL2Begin));
      const bool on = s_mask[q] != 0;
      bf16x4 o;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // clip'(h) from a = clip(h): 1 inside (-1, 1), 0 where clipped.
        const bool pass = fabsf(bf16_value(static_cast<unsigned short>(av[j]))) < 1.0f;
        o[j] = on && pass ? to_bf16(acc[j]) : 0;
      }
      store4(s_gh2 + g * kA2Plane + 4 * (q - kL2Begin), o);
    }
  }
  __syncthreads();
  // dW2[t][c][o] = sum_p a1[p + t][c] g_h2[p][o], one tap per wave.
  for (int t = wave; t < 9; t += kWaves) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int off = tap_offset(t);
    for (int ks = 0; ks < kL2Tiles; ++ks) {
      const int p0 = kL2Begin + 16 * ks;
      bf16x4 a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = p0 + 4 * g + j;
        a[j] = static_cast<short>(s_a1[(m >> 2) * kA1Plane + 4 * (p + off - kA1Base) + (m & 3)]);
        b[j] = static_cast<short>(s_gh2[(m >> 2) * kA2Plane + 4 * (p - kL2Begin) + (m & 3)]);
      }
      acc = mfma(a, b, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) out[kW1 + t * kC1 * kC1 + (4 * g + j) * kC1 + m] = acc[j];
  }
  __syncthreads();
  // g_h1 = (sum_t W2[t] g_h2 shifted) * clip'(h1), in place over a1.
  {
    bf16x4 wb2[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        wb2[t][j] = to_bf16(w[kW1 + t * kC1 * kC1 + m * kC1 + 4 * g + j]);
    }
    for (int mt = wave; mt < kL1Tiles; mt += kWaves) {
      const int q = kL1Begin + 16 * mt + m;
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const int src = q - tap_offset(t);
        bf16x4 b = {0, 0, 0, 0};
        if (src >= kL2Begin && src < kL2End) b = lds4(s_gh2 + g * kA2Plane + 4 * (src - kL2Begin));
        acc = mfma(wb2[t], b, acc);
      }
      unsigned short* a1p = s_a1 + g * kA1Plane + 4 * (q - kA1Base);
      const bf16x4 av = lds4(a1p);
      const bool on = s_mask[q] != 0;
      bf16x4 o;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // clip'(h) from a = clip(h): 1 inside (-1, 1), 0 where clipped.
        const bool pass = fabsf(bf16_value(static_cast<unsigned short>(av[j]))) < 1.0f;
        o[j] = on && pass ? to_bf16(acc[j]) : 0;
      }
      store4(a1p, o);
    }
  }
  __syncthreads();
  // dW1[t][c][o] = sum_p in[p + t][c] g_h1[p][o]; N holds (tap, channel) pairs.
  for (int nt = wave; nt < 3; nt += kWaves) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int col = 16 * nt + m;
    const int t = col >> 2, c = col & 3;
    const int off = t < 9 ? tap_offset(t) : 0;
    for (int ks = 0; ks < kL1Tiles; ++ks) {
      const int p0 = kL1Begin + 16 * ks;
      bf16x4 a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = p0 + 4 * g + j;
        a[j] = static_cast<short>(s_a1[(m >> 2) * kA1Plane + 4 * (p - kA1Base) + (m & 3)]);
        b[j] = t < 9 ? static_cast<short>(s_in[4 * (p + off) + c]) : 0;
      }
      acc = mfma(a, b, acc);
    }
    if (t < 9) {
#pragma unroll
      for (int j = 0; j < 4; ++j) out[t * kC0 * kC1 + c * kC1 + 4 * g + j] = acc[j];
    }
  }
}

// grad[k] = sum over tiles of grads[tile][k] (fixed order, FP64); the last
// thread also folds the batch loss into losses[0].
__global__ __launch_bounds__(kStream) void k_learn_reduce(const float* __restrict__ grads,
                                                          uint32_t tiles,
                                                          const double* __restrict__ lparts,
                                                          uint32_t lcount, uint32_t batch,
                                                          double* __restrict__ grad,
                                                          float* __restrict__ loss) {
  const int k = static_cast<int>(blockIdx.x * kStream + threadIdx.x);
  if (k < kWeights) {
    double acc = 0.0;
    for (uint32_t t = 0; t < tiles; ++t) acc += static_cast<double>(grads[static_cast<size_t>(t) * kWeights + k]);
    grad[k] = acc;
  } else if (k == kWeights) {
    double acc = 0.0;
    for (uint32_t i = 0; i < lcount; ++i) acc += lparts[i];
    if (loss != nullptr) *loss = static_cast<float>(acc / batch);
  }
}

// Packs the A fragments of the forward from the FP32 weights (bf16, round to
// nearest even): layer 1 steps s = 0..2 (tap 4s + l/16, channel j, output
// l%16), layer 2 taps t (input 4 (l/16) + j, output l%16), layer 3 (tap l%16,
// input 4 (l/16) + j).
__device__ inline void pack_fragments(const float* w, unsigned short* packed) {
  for (int e = threadIdx.x; e < kFrags * kWave * 4; e += blockDim.x) {
    const int f = e / (kWave * 4), lane = (e / 4) % kWave, j = e % 4;
    const int m = lane & 15, g = lane >> 4;
    float v = 0.0f;
    if (f < 3) {
      const int tap = 4 * f + g;
      if (tap < 9) v = w[tap * kC0 * kC1 + j * kC1 + m];
    } else if (f < 12) {
      v = w[kW1 + (f - 3) * kC1 * kC1 + (4 * g + j) * kC1 + m];
    } else if (m < 9) {
      v = w[kW1 + kW2 + m * kC1 + 4 * g + j];
    }
    packed[e] = bf16_bits(v);
  }
}

__global__ __launch_bounds__(1024) void k_learn_pack(const float* __restrict__ w,
                                                     unsigned short* __restrict__ packed) {
  pack_fragments(w, packed);
}

// One Adam step on the FP32 weights, then the packed fragments.
__global__ __launch_bounds__(1024) void k_learn_adam(float* __restrict__ w, float* __restrict__ m1,
                                                     float* __restrict__ m2,
                                                     const double* __restrict__ grad, float lr,
                                                     float beta1, float beta2, float eps,
                                                     float correction1, float correction2,
                                                     unsigned short* __restrict__ packed) {
  for (int k = threadIdx.x; k < kWeights; k += blockDim.x) {
    const float gk = static_cast<float>(grad[k]);
    const float a = beta1 * m1[k] + (1.0f - beta1) * gk;
    const float b = beta2 * m2[k] + (1.0f - beta2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    w[k] = w[k] - lr * (a / correction1) / (sqrtf(b / correction2) + eps);
  }
  __syncthreads();
  pack_fragments(w, packed);
}

// e'^T A e' per pair, summed per block (grid: blocks x pairs), then folded.
__global__ __launch_bounds__(kStream) void k_learn_energy(TrainView V, double* parts) {
  __shared__ double s_red[kStream / kWave];
  const uint32_t pair = blockIdx.y;
  const Geometry G = geometry_of(V, pair);
  const size_t n = static_cast<size_t>(V.nx) * V.ny;
  const float* e = V.ep + static_cast<size_t>(pair) * n;
  double acc = 0.0;
  const size_t stride = static_cast<size_t>(gridDim.x) * kStream;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kStream + threadIdx.x; i < n; i += stride) {
    const uint32_t x = static_cast<uint32_t>(i / G.ny), y = static_cast<uint32_t>(i % G.ny);
    acc += static_cast<double>(e[i]) * static_cast<double>(apply_at(G, e, x, y));
  }
  const double part = block_sum<kStream>(acc, s_red);
  if (threadIdx.x == 0) parts[static_cast<size_t>(pair) * gridDim.x + blockIdx.x] = part;
}

__global__ void k_learn_energy_fold(const double* parts, uint32_t blocks, uint32_t pairs,
                                    double* energy) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pairs) return;
  double acc = 0.0;
  for (uint32_t b = 0; b < blocks; ++b) acc += parts[static_cast<size_t>(p) * blocks + b];
  energy[p] = acc;
}

// Harvest: r' = r - A x and e' = e - x from a snapshot residual r, the FP64
// solution e of A e = r and the cycle's part x before post-smoothing.
__global__ __launch_bounds__(kStream) void k_learn_pairs(Geometry G, const float* __restrict__ r,
                                                         const double* __restrict__ e,
                                                         const float* __restrict__ x,
                                                         float* __restrict__ rp,
                                                         float* __restrict__ ep) {
  const size_t n = static_cast<size_t>(G.nx) * G.ny;
  const size_t stride = static_cast<size_t>(gridDim.x) * kStream;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kStream + threadIdx.x; i < n; i += stride) {
    const uint32_t gx = static_cast<uint32_t>(i / G.ny), gy = static_cast<uint32_t>(i % G.ny);
    rp[i] = r[i] - apply_at(G, x, gx, gy);
    ep[i] = static_cast<float>(e[i] - static_cast<double>(x[i]));
  }
}

uint32_t stream_grid(size_t n) {
  return static_cast<uint32_t>(std::clamp<size_t>((n + kStream - 1) / kStream, 1, 1024));
}

TrainView train_view(const LearnedDevice& L) {
  return TrainView{L.ops.data(), L.rp.data(), L.ep.data(), L.geometry.data(), L.energy.data(),
                   L.nx, L.ny};
}

dim3 tile_grid(uint32_t nx, uint32_t ny, uint32_t batch) {
  return dim3((nx + kTX - 1) / kTX, (ny + kTY - 1) / kTY, batch);
}

void require_weights(const Context& ctx) {
  if (!ctx.learned.loaded) {
    throw Error(MSIM_INVALID_ARGUMENT, "the learned smoother has no weights: set them first");
  }
}

void require_set(const Context& ctx) {
  if (ctx.learned.pairs == 0) {
    throw Error(MSIM_INVALID_ARGUMENT, "no training set is uploaded");
  }
}

// Makes room for a batch of `batch` pairs.
void reserve_batch(Context& ctx, uint32_t batch) {
  auto& L = ctx.learned;
  const size_t n = static_cast<size_t>(L.nx) * L.ny;
  const dim3 tg = tile_grid(L.nx, L.ny, batch);
  L.z.reserve(batch * n);
  L.d.reserve(batch * n);
  L.gz.reserve(batch * n);
  L.grads.reserve(static_cast<size_t>(tg.x) * tg.y * batch * kWeights);
  L.lparts.reserve(static_cast<size_t>(stream_grid(n)) * batch);
  L.grad.reserve(kWeights);
}

// Forward, loss, backward and gradient reduction for the pairs order[0..batch).
void gradient_pass(Context& ctx, const uint32_t* order, uint32_t batch, float* loss) {
  auto& L = ctx.learned;
  const size_t n = static_cast<size_t>(L.nx) * L.ny;
  const TrainView V = train_view(L);
  const dim3 tg = tile_grid(L.nx, L.ny, batch);
  const auto* packed = reinterpret_cast<const bf16x4*>(L.packed.data());
  k_learn_forward<<<tg, kThreads, 0, ctx.stream>>>(V, order, packed, L.z.data());
  check(hipGetLastError(), "launch k_learn_forward");
  const uint32_t lb = stream_grid(n);
  k_learn_diff<<<dim3(lb, batch), kStream, 0, ctx.stream>>>(V, order, L.z.data(), L.d.data());
  check(hipGetLastError(), "launch k_learn_diff");
  k_learn_loss<<<dim3(lb, batch), kStream, 0, ctx.stream>>>(V, order, L.d.data(), L.gz.data(),
                                                            L.lparts.data());
  check(hipGetLastError(), "launch k_learn_loss");
  k_learn_backward<<<tg, kBThreads, 0, ctx.stream>>>(V, order, L.w.data(), packed, L.gz.data(),
                                                    L.grads.data());
  check(hipGetLastError(), "launch k_learn_backward");
  const uint32_t tiles = tg.x * tg.y * batch;
  k_learn_reduce<<<(kWeights + 1 + kStream - 1) / kStream, kStream, 0, ctx.stream>>>(
      L.grads.data(), tiles, L.lparts.data(), lb * batch, batch, L.grad.data(), loss);
  check(hipGetLastError(), "launch k_learn_reduce");
}

}  // namespace

void learned_set_weights(Context& ctx, const float* weights) {
  auto& L = ctx.learned;
  L.w.reserve(kWeights);
  L.m1.reserve(kWeights);
  L.m2.reserve(kWeights);
  L.packed.reserve(kFrags * kWave * 4);
  // New weights start a new Adam run.
  check(hipMemsetAsync(L.m1.data(), 0, kWeights * sizeof(float), ctx.stream), "clear moments");
  check(hipMemsetAsync(L.m2.data(), 0, kWeights * sizeof(float), ctx.stream), "clear moments");
  L.adam_steps = 0;
  check(hipMemcpyAsync(L.w.data(), weights, kWeights * sizeof(float), hipMemcpyHostToDevice,
                       ctx.stream),
        "upload weights");
  k_learn_pack<<<1, 1024, 0, ctx.stream>>>(L.w.data(), L.packed.data());
  check(hipGetLastError(), "launch k_learn_pack");
  check(hipStreamSynchronize(ctx.stream), "weights");
  L.loaded = true;
}

void learned_get_weights(Context& ctx, float* weights) {
  require_weights(ctx);
  check(hipMemcpyAsync(weights, ctx.learned.w.data(), kWeights * sizeof(float),
                       hipMemcpyDeviceToHost, ctx.stream),
        "download weights");
  check(hipStreamSynchronize(ctx.stream), "weights");
}

void learned_upload(Context& ctx, const msim_train_set& set) {
  auto& L = ctx.learned;
  const size_t n = static_cast<size_t>(set.nx) * set.ny;
  L.nx = set.nx;
  L.ny = set.ny;
  L.geometries = set.geometries;
  L.pairs = 0;
  L.ops.reserve(static_cast<size_t>(set.geometries) * 4 * n);
  L.rp.reserve(static_cast<size_t>(set.pairs) * n);
  L.ep.reserve(static_cast<size_t>(set.pairs) * n);
  L.geometry.reserve(set.pairs);
  L.energy.reserve(set.pairs);
  check(hipMemcpyAsync(L.ops.data(), set.operators, set.geometries * 4 * n * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload operators");
  check(hipMemcpyAsync(L.rp.data(), set.residuals, set.pairs * n * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload residuals");
  check(hipMemcpyAsync(L.ep.data(), set.errors, set.pairs * n * sizeof(float),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload errors");
  check(hipMemcpyAsync(L.geometry.data(), set.pair_geometry, set.pairs * sizeof(uint32_t),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload pair geometry");
  const uint32_t blocks = stream_grid(n);
  DeviceBuffer<double> parts;
  parts.reserve(static_cast<size_t>(blocks) * set.pairs);
  L.pairs = set.pairs;
  k_learn_energy<<<dim3(blocks, set.pairs), kStream, 0, ctx.stream>>>(train_view(L), parts.data());
  check(hipGetLastError(), "launch k_learn_energy");
  k_learn_energy_fold<<<(set.pairs + 255) / 256, 256, 0, ctx.stream>>>(parts.data(), blocks,
                                                                      set.pairs, L.energy.data());
  check(hipGetLastError(), "launch k_learn_energy_fold");
  std::vector<double> energy(set.pairs);
  check(hipMemcpyAsync(energy.data(), L.energy.data(), set.pairs * sizeof(double),
                       hipMemcpyDeviceToHost, ctx.stream),
        "download energies");
  check(hipStreamSynchronize(ctx.stream), "upload training set");
  for (uint32_t p = 0; p < set.pairs; ++p) {
    if (!(energy[p] > 0.0) || !std::isfinite(energy[p])) {
      L.pairs = 0;
      throw Error(MSIM_INVALID_ARGUMENT,
                  "training pair " + std::to_string(p) + " has a non-positive error energy");
    }
  }
}

void learned_train(Context& ctx, const msim_train_params& params, float* losses,
                   double* seconds) {
  require_weights(ctx);
  require_set(ctx);
  auto& L = ctx.learned;
  const uint32_t batch = params.batch;
  reserve_batch(ctx, batch);
  // Pair order: consecutive shuffles of all pairs (xorshift64*, seeded).
  std::vector<uint32_t> order(static_cast<size_t>(params.steps) * batch);
  std::vector<uint32_t> perm(L.pairs);
  uint64_t state = params.seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto next = [&]() {
    state ^= state >> 12;
    state ^= state << 25;
    state ^= state >> 27;
    return state * 0x2545F4914F6CDD1Dull;
  };
  size_t pos = perm.size();
  for (size_t k = 0; k < order.size(); ++k) {
    if (pos == perm.size()) {
      for (uint32_t i = 0; i < L.pairs; ++i) perm[i] = i;
      for (uint32_t i = L.pairs; i > 1; --i) std::swap(perm[i - 1], perm[next() % i]);
      pos = 0;
    }
    order[k] = perm[pos++];
  }
  L.order.reserve(order.size());
  L.losses.reserve(params.steps);
  check(hipMemcpyAsync(L.order.data(), order.data(), order.size() * sizeof(uint32_t),
                       hipMemcpyHostToDevice, ctx.stream),
        "upload order");
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  for (uint32_t step = 0; step < params.steps; ++step) {
    gradient_pass(ctx, L.order.data() + static_cast<size_t>(step) * batch, batch,
                  L.losses.data() + step);
    ++L.adam_steps;
    const double c1 = std::pow(static_cast<double>(params.beta1), L.adam_steps);
    const double c2 = std::pow(static_cast<double>(params.beta2), L.adam_steps);
    k_learn_adam<<<1, 1024, 0, ctx.stream>>>(
        L.w.data(), L.m1.data(), L.m2.data(), L.grad.data(), params.learning_rate, params.beta1,
        params.beta2, params.epsilon, static_cast<float>(1.0 - c1), static_cast<float>(1.0 - c2),
        L.packed.data());
    check(hipGetLastError(), "launch k_learn_adam");
  }
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  check(hipEventSynchronize(ctx.stop), "train");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  if (seconds != nullptr) *seconds = ms * 1.0e-3;
  if (losses != nullptr) {
    check(hipMemcpy(losses, L.losses.data(), params.steps * sizeof(float), hipMemcpyDeviceToHost),
          "download losses");
  }
}

void learned_gradients(Context& ctx, uint32_t pair, double* loss, double* gradient) {
  require_weights(ctx);
  require_set(ctx);
  auto& L = ctx.learned;
  if (pair >= L.pairs) throw Error(MSIM_INVALID_ARGUMENT, "pair index out of range");
  reserve_batch(ctx, 1);
  L.order.reserve(std::max<size_t>(L.order.size(), 1));
  L.losses.reserve(std::max<size_t>(L.losses.size(), 1));
  check(hipMemcpyAsync(L.order.data(), &pair, sizeof(uint32_t), hipMemcpyHostToDevice, ctx.stream),
        "upload order");
  gradient_pass(ctx, L.order.data(), 1, L.losses.data());
  float value = 0.0f;
  check(hipMemcpyAsync(&value, L.losses.data(), sizeof(float), hipMemcpyDeviceToHost, ctx.stream),
        "download loss");
  check(hipMemcpyAsync(gradient, L.grad.data(), kWeights * sizeof(double), hipMemcpyDeviceToHost,
                       ctx.stream),
        "download gradient");
  check(hipStreamSynchronize(ctx.stream), "gradients");
  *loss = value;
}

void learned_apply(Context& ctx, uint32_t pair, float* z) {
  require_weights(ctx);
  require_set(ctx);
  auto& L = ctx.learned;
  if (pair >= L.pairs) throw Error(MSIM_INVALID_ARGUMENT, "pair index out of range");
  reserve_batch(ctx, 1);
  L.order.reserve(std::max<size_t>(L.order.size(), 1));
  check(hipMemcpyAsync(L.order.data(), &pair, sizeof(uint32_t), hipMemcpyHostToDevice, ctx.stream),
        "upload order");
  const auto* packed = reinterpret_cast<const bf16x4*>(L.packed.data());
  k_learn_forward<<<tile_grid(L.nx, L.ny, 1), kThreads, 0, ctx.stream>>>(
      train_view(L), L.order.data(), packed, L.z.data());
  check(hipGetLastError(), "launch k_learn_forward");
  const size_t n = static_cast<size_t>(L.nx) * L.ny;
  check(hipMemcpyAsync(z, L.z.data(), n * sizeof(float), hipMemcpyDeviceToHost, ctx.stream),
        "download correction");
  check(hipStreamSynchronize(ctx.stream), "apply");
}

uint32_t learned_blocks(uint32_t nx, uint32_t ny) {
  const dim3 g = tile_grid(nx, ny, 1);
  return std::min<uint32_t>(g.x * g.y, 768);  // three per CU
}

void launch_up_learned(Context& ctx, const MgLevel& L0, const float* r, const float* xp,
                       const float* ec, uint32_t nyc, float* z, const CgState* st, const float* q,
                       double* zr, double* zq) {
  require_weights(ctx);
  const Geometry G{L0.wx.data(), L0.wy.data(), L0.bc.data(), L0.dinv.data(), L0.nx, L0.ny};
  const UpOut out{z, q, zr, zq};
  const dim3 g = tile_grid(L0.nx, L0.ny, 1);
  k_learn_up<<<learned_blocks(L0.nx, L0.ny), kThreads, 0, ctx.stream>>>(
      G, r, xp, ec, nyc, st, reinterpret_cast<const bf16x4*>(ctx.learned.packed.data()), out, g.y,
      g.x * g.y);
  check(hipGetLastError(), "launch k_learn_up");
}

void learned_make_pairs(Context& ctx, const MgLevel& L0, const float* r, const double* e,
                        const float* x, float* rp, float* ep) {
  const Geometry G{L0.wx.data(), L0.wy.data(), L0.bc.data(), L0.dinv.data(), L0.nx, L0.ny};
  k_learn_pairs<<<stream_grid(L0.cells()), kStream, 0, ctx.stream>>>(G, r, e, x, rp, ep);
  check(hipGetLastError(), "launch k_learn_pairs");
}

}  // namespace msim
