// A STREAM-style copy of device memory, timed with HIP events: the HBM
// bandwidth a streaming kernel actually reaches on this device, against which
// the stencil kernels' rooflines are also stated (beside the 8 TB/s figure).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"

namespace msim {
namespace {

constexpr int kCopyBlock = 256;

// dst = src, 16 bytes per access, grid-stride.
__global__ __launch_bounds__(kCopyBlock) void k_stream_copy(const float4* __restrict__ src,
                                                            float4* __restrict__ dst,
                                                            size_t count) {
  const size_t stride = static_cast<size_t>(gridDim.x) * kCopyBlock;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kCopyBlock + threadIdx.x; i < count;
       i += stride) {
    dst[i] = src[i];
  }
}

}  // namespace

double stream_copy_bandwidth(Context& ctx, size_t bytes, uint32_t repeats) {
  const size_t count = std::max<size_t>(bytes / sizeof(float4), 1);
  DeviceBuffer<float4> src, dst;
  src.reserve(count);
  dst.reserve(count);
  check(hipMemsetAsync(src.data(), 0, count * sizeof(float4), ctx.stream), "clear copy source");
  // Four waves of blocks per CU on 256 CUs; every block walks its share.
  const uint32_t blocks = 256 * 16;
  k_stream_copy<<<blocks, kCopyBlock, 0, ctx.stream>>>(src.data(), dst.data(), count);  // warm
  check(hipGetLastError(), "launch k_stream_copy");
  check(hipEventRecord(ctx.start, ctx.stream), "event");
  for (uint32_t k = 0; k < repeats; ++k) {
    k_stream_copy<<<blocks, kCopyBlock, 0, ctx.stream>>>(src.data(), dst.data(), count);
  }
  check(hipGetLastError(), "launch k_stream_copy");
  check(hipEventRecord(ctx.stop, ctx.stream), "event");
  check(hipEventSynchronize(ctx.stop), "event sync");
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, ctx.start, ctx.stop), "event time");
  // Bytes read plus bytes written per copy.
  return 2.0 * static_cast<double>(count * sizeof(float4)) * repeats / (ms * 1.0e-3) / 1.0e9;
}

}  // namespace msim
