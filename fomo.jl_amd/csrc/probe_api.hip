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

// The access pattern of a sweep without its stencil: NR arrays read and NW written, one float per
// lane per access (as the sweeps load), four elements in flight per lane; output q is the sum of
// the inputs q, q + NW, q + 2 NW, ... so that every input is consumed.
constexpr int kMaxProbeR = 8, kMaxProbeW = 4;
struct ProbeArrays {
  const float* r[kMaxProbeR];
  float* w[kMaxProbeW];
};

template <int NR, int NW>
__global__ __launch_bounds__(256) void stream_probe_kernel(const ProbeArrays A, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x * kCopyUnroll;
  for (size_t i = (size_t)blockIdx.x * blockDim.x * kCopyUnroll + threadIdx.x; i < n; i += stride) {
    float v[kCopyUnroll][NR];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const size_t k = min(i + (size_t)u * blockDim.x, n - 1);  // clamped: every lane loads
#pragma unroll
      for (int q = 0; q < NR; ++q) v[u][q] = G(A.r[q])[k];
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const size_t k = i + (size_t)u * blockDim.x;
      if (k >= n) continue;
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        float acc = 0.0f;
#pragma unroll
        for (int e = q; e < NR; e += NW) acc += v[u][e];
        G(A.w[q])[k] = acc;
      }
    }
  }
}

template <int NR, int NW>
static void launch_probe(const ProbeArrays& a, size_t n, int blocks, hipStream_t s) {
  hipLaunchKernelGGL((stream_probe_kernel<NR, NW>), dim3(blocks), dim3(256), 0, s, a, n);
}

template <int NR>
static bool launch_probe_w(int nw, const ProbeArrays& a, size_t n, int blocks, hipStream_t s) {
  switch (nw) {
    case 1: launch_probe<NR, 1>(a, n, blocks, s); return true;
    case 2: launch_probe<NR, 2>(a, n, blocks, s); return true;
    case 3: launch_probe<NR, 3>(a, n, blocks, s); return true;
    case 4: launch_probe<NR, 4>(a, n, blocks, s); return true;
    default: return false;
  }
}

static bool launch_probe_rw(int nr, int nw, const ProbeArrays& a, size_t n, int blocks, hipStream_t s) {
  switch (nr) {
    case 1: return launch_probe_w<1>(nw, a, n, blocks, s);
    case 2: return launch_probe_w<2>(nw, a, n, blocks, s);
    case 3: return launch_probe_w<3>(nw, a, n, blocks, s);
    case 4: return launch_probe_w<4>(nw, a, n, blocks, s);
    case 5: return launch_probe_w<5>(nw, a, n, blocks, s);
    case 6: return launch_probe_w<6>(nw, a, n, blocks, s);
    case 7: return launch_probe_w<7>(nw, a, n, blocks, s);
    case 8: return launch_probe_w<8>(nw, a, n, blocks, s);
    default: return false;
  }
}

}  // namespace fomo

using namespace fomo;

extern "C" int fomo_stream_probe(int32_t n_read, int32_t n_write, int64_t bytes_per_array, int32_t iters,
                                 double* ms_per_pass) {
  if (n_read < 1 || n_read > kMaxProbeR || n_write < 1 || n_write > kMaxProbeW || n_write > n_read ||
      bytes_per_array < 4 || iters < 1 || !ms_per_pass)
    return set_error(FOMO_EINVAL, "stream probe: 1 <= n_write <= n_read <= 8, n_write <= 4, bytes >= 4, iters >= 1");
  const size_t n = (size_t)bytes_per_array / 4;
  const int na = n_read + n_write;
  void* buf[kMaxProbeR + kMaxProbeW] = {nullptr};
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto done = [&](int code) {
    if (s) (void)hipStreamSynchronize(s);
    for (void* b : buf)
      if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return code;
  };
  for (int q = 0; q < na; ++q)
    if (hipMalloc(&buf[q], n * 4) != hipSuccess) return done(set_error(FOMO_ENOMEM, "stream probe: hipMalloc failed"));
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    return done(set_error(FOMO_EHIP, "stream probe: stream/event creation failed"));
  ProbeArrays a{};
  for (int q = 0; q < n_read; ++q) {
    if (hipMemsetAsync(buf[q], 0, n * 4, s) != hipSuccess) return done(set_error(FOMO_EHIP, "stream probe: memset failed"));
    a.r[q] = (const float*)buf[q];
  }
  for (int q = 0; q < n_write; ++q) a.w[q] = (float*)buf[n_read + q];
  const int blocks = (int)std::min<size_t>(32768, (n + 1023) / 1024);
  if (!launch_probe_rw(n_read, n_write, a, n, blocks, s)) return done(set_error(FOMO_EINVAL, "stream probe: shape"));
  (void)hipEventRecord(e0, s);
  for (int k = 0; k < iters; ++k) launch_probe_rw(n_read, n_write, a, n, blocks, s);
  (void)hipEventRecord(e1, s);
  if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
    return done(set_error(FOMO_EHIP, "stream probe failed"));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *ms_per_pass = ms / iters;
  return done(FOMO_OK);
}

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
