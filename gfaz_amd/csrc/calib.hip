// Measurement support: a plain device-to-device copy of a known byte count,
// used to scale the HBM traffic counters (FETCH_SIZE / WRITE_SIZE) that the
// profiler reports for the codec kernels in the same run.
#include "gfaz_amd.h"

#include "abi_types.hpp"

#include <string>

namespace {

__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src,
                                                uint4 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace


extern "C" int gfz_calibration_copy(gfz_ctx *ctx, void *dst, const void *src, uint64_t bytes,
                                    double *ms) {
  if (!ctx || !dst || !src || (bytes & 15))
    return GFZ_E_INVALID;
  try {
    gfz::Ctx &c = *ctx->ctx;
    gfz::DeviceGuard g(c.device());
    hipEvent_t a, b;
    GFZ_HIP(hipEventCreate(&a));
    GFZ_HIP(hipEventCreate(&b));
    const uint64_t n = bytes / 16;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(gfz::ceil_div(n, 256), 256ull * 64);
    GFZ_HIP(hipEventRecord(a, c.stream()));
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, c.stream(),
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
    GFZ_HIP(hipEventRecord(b, c.stream()));
    GFZ_HIP(hipEventSynchronize(b));
    float t = 0;
    GFZ_HIP(hipEventElapsedTime(&t, a, b));
    if (ms)
      *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return GFZ_OK;
  } catch (const gfz::Error &e) {
    return e.code;
  }
}
