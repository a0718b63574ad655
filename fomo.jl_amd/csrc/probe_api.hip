// probe_api.hip -- HBM calibration: a streaming copy whose bandwidth is the practical ceiling the
// stencil sweeps are compared with, and whose counter readings calibrate byte counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "fomo_common.h"
#include "plan.h"

namespace fomo {

// dst[i] = src[i] over n 16-byte elements: each lane keeps 4 loads in flight and the accesses
// carry the streaming (non-temporal) hint.
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kCopyUnroll = 4;

__global__ __launch_bounds__(256) void stream_copy_kernel(const f4v* __restrict__ src, f4v* __restrict__ dst,
                                                          size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x * kCopyUnroll;
  for (size_t i = (size_t)blockIdx.x * blockDim.x * kCopyUnroll + threadIdx.x; i < n; i += stride) {
    f4v v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const size_t k = i + (size_t)u * blockDim.x;
      if (k < n) v[u] = __builtin_nontemporal_load(src + k);
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const size_t k = i + (size_t)u * blockDim.x;
      if (k < n) __builtin_nontemporal_store(v[u], dst + k);
    }
  }
}

}  // namespace fomo

using namespace fomo;

extern "C" int fomo_stream_copy(int64_t bytes, int32_t iters, double* ms_per_copy) {
  if (bytes < 16 || iters < 1 || !ms_per_copy) return set_error(FOMO_EINVAL, "stream copy: bytes >= 16, iters >= 1");
  const size_t n = (size_t)bytes / 16;
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FOMO_OK;
  auto done = [&](int code) {
    if (s) (void)hipStreamSynchronize(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return code;
  };
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess)
    return done(set_error(FOMO_ENOMEM, "stream copy: hipMalloc failed"));
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    return done(set_error(FOMO_EHIP, "stream copy: stream/event creation failed"));
  if (hipMemsetAsync(a, 0, n * 16, s) != hipSuccess) return done(set_error(FOMO_EHIP, "stream copy: memset failed"));
  const int blocks = (int)std::min<size_t>(32768, (n + 1023) / 1024);
  hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, s, (const f4v*)a, (f4v*)b, n);  // warm
  (void)hipEventRecord(e0, s);
  for (int k = 0; k < iters; ++k)
    hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, s, (const f4v*)a, (f4v*)b, n);
  (void)hipEventRecord(e1, s);
  if (hipEventSynchronize(e1) != hipSuccess) return done(set_error(FOMO_EHIP, "stream copy failed"));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *ms_per_copy = ms / iters;
  (void)rc;
  return done(FOMO_OK);
}
