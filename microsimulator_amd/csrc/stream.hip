// A STREAM-style copy of device memory, timed with HIP events: the HBM
// bandwidth a streaming kernel actually reaches on this device, against which
// the stencil kernels' rooflines are also stated (beside the 8 TB/s figure).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"

namespace msim {
namespace {

constexpr int kCopyBlock = 256;
constexpr int kCopyUnroll = 4;

// dst = src, 16 bytes per access. Each thread copies kCopyUnroll vectors a
// block's width apart, issuing every load before its first store, so a wave
// keeps four loads in flight; one block per 16 KiB, no grid-stride loop.
__global__ __launch_bounds__(kCopyBlock) void k_stream_copy(const float4* __restrict__ src,
                                                            float4* __restrict__ dst,
                                                            size_t count) {
  const size_t base = static_cast<size_t>(blockIdx.x) * (kCopyBlock * kCopyUnroll) + threadIdx.x;
  float4 v[kCopyUnroll];
#pragma unroll
  for (int k = 0; k < kCopyUnroll; ++k) {
    const size_t i = base + static_cast<size_t>(k) * kCopyBlock;
    if (i < count) v[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < kCopyUnroll; ++k) {
    const size_t i = base + static_cast<size_t>(k) * kCopyBlock;
    if (i < count) dst[i] = v[k];
  }
}

}  // namespace

double stream_copy_bandwidth(Context& ctx, size_t bytes, uint32_t repeats) {
  const size_t count = std::max<size_t>(bytes / sizeof(float4), 1);
  DeviceBuffer<float4> src, dst;
  src.reserve(count);
  dst.reserve(count);
  check(hipMemsetAsync(src.data(), 0, count * sizeof(float4), ctx.stream), "clear copy source");
  const size_t per_block = static_cast<size_t>(kCopyBlock) * kCopyUnroll;
  const size_t blocks = (count + per_block - 1) / per_block;
  if (blocks > 0x7fffffffu) throw Error(MSIM_OVERFLOW, "copy probe: too many blocks");
  const uint32_t grid = static_cast<uint32_t>(blocks);
  // The kernel, then the runtime's own device-to-device copy; the faster of
  // the two is the ceiling.
  float best_ms = 0.0f;
  for (int form = 0; form < 2; ++form) {
    auto once = [&]() {
      if (form == 0) {
        k_stream_copy<<<grid, kCopyBlock, 0, ctx.stream>>>(src.data(), dst.data(), count);
        check(hipGetLastError(), "launch k_stream_copy");
      } else {
        check(hipMemcpyAsync(dst.data(), src.data(), count * sizeof(float4),
                             hipMemcpyDeviceToDevice, ctx.stream),
              "device copy");
      }
    };
    once();  // warm
    check(hipEventRecord(ctx.start, ctx.stream), "event");
    for (uint32_t k = 0; k < repeats; ++k) once();
    check(hipEventRecord(ctx.stop, ctx.stream), "event");
    check(hipEventSynchronize(ctx.stop), "event sync");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
    if (form == 0 || ms < best_ms) best_ms = ms;
  }
  // Bytes read plus bytes written per copy.
  return 2.0 * static_cast<double>(count * sizeof(float4)) * repeats / (best_ms * 1.0e-3) / 1.0e9;
}

}  // namespace msim
