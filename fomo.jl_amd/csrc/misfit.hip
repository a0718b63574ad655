// misfit.hip -- FP64 trace misfit and trace energy, reduced on the GPU over recorded traces.
//
// The reference measures a boundary's reflection as ||u - u_ref|| / ||u_ref|| over whole
// seismograms (test/verify_patch.jl:86-87) and checks solvers by per-trace relative L2; it forms
// both on the host.  Here the sums run on the device where the traces already are, in FP64 and in
// a fixed order: pass 1 splits the time axis into chunks (one block per 64 receivers x chunk, four
// wave rows striding over the chunk's samples, the rows added in row order), pass 2 adds the
// chunks of each receiver in chunk order.  The result is a pure function of the traces.
#include <hip/hip_runtime.h>

#include <string>

#include "plan.h"

namespace fomo {
namespace {

constexpr int kRecPerBlock = 64;  // receivers of one block (one wave row)
constexpr int kTRows = 4;         // wave rows striding over a chunk's time samples
constexpr int kMaxChunks = 64;
constexpr int kMinChunkLen = 256;

__global__ void __launch_bounds__(kRecPerBlock* kTRows)
    trace_partial_kernel(const float* __restrict__ a, const float* __restrict__ b, int n_rec, int t0, int t1,
                         int chunk, double* __restrict__ part) {
  __shared__ double sn[kTRows][kRecPerBlock];
  __shared__ double sd[kTRows][kRecPerBlock];
  const int lx = threadIdx.x, ly = threadIdx.y;
  const int r = blockIdx.x * kRecPerBlock + lx;
  const int rc = r < n_rec ? r : n_rec - 1;  // clamped: every lane loads; only lanes r < n_rec store
  const int c0 = t0 + blockIdx.y * chunk;
  const int c1 = min(c0 + chunk, t1);
  double num = 0.0, den = 0.0;
  for (int t = c0 + ly; t < c1; t += kTRows) {  // ly is uniform per wave: no divergent trip count
    const size_t o = (size_t)t * n_rec + rc;
    const double x = (double)a[o];
    const double y = b ? (double)b[o] : 0.0;  // b is a kernel argument: a uniform branch
    const double e = x - y;
    num += e * e;
    den += y * y;
  }
  sn[ly][lx] = num;
  sd[ly][lx] = den;
  __syncthreads();
  if (ly == 0 && r < n_rec) {
    double* q = part + ((size_t)blockIdx.y * n_rec + r) * 2;
    q[0] = ((sn[0][lx] + sn[1][lx]) + sn[2][lx]) + sn[3][lx];
    q[1] = ((sd[0][lx] + sd[1][lx]) + sd[2][lx]) + sd[3][lx];
  }
}

__global__ void trace_final_kernel(const double* __restrict__ part, int n_rec, int n_chunks, double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rec) return;
  double num = 0.0, den = 0.0;
  for (int c = 0; c < n_chunks; ++c) {
    const double* q = part + ((size_t)c * n_rec + r) * 2;
    num += q[0];
    den += q[1];
  }
  out[2 * r] = num;
  out[2 * r + 1] = den;
}

}  // namespace
}  // namespace fomo

using namespace fomo;

extern "C" int fomo_trace_misfit(fomo_plan* pa, int32_t shot_a, fomo_plan* pb, int32_t shot_b, int32_t slot,
                                 int32_t t0, int32_t t1, double* num, double* den) {
  if (!pa || !pa->impl) return set_error(FOMO_EINVAL, "plan is NULL");
  if (pb && !pb->impl) return set_error(FOMO_EINVAL, "second plan is NULL");
  if (!num || !den) return set_error(FOMO_EINVAL, "NULL output");
  PlanBase* A = pa->impl;
  PlanBase* B = pb ? pb->impl : nullptr;
  const int nt = A->d.nt;
  if (B && B->d.nt != nt) return set_error(FOMO_EINVAL, "plans record different numbers of time samples");
  if (t0 < 0 || t1 > nt || t0 > t1) return set_error(FOMO_EINVAL, "time window outside [0, nt]");
  const float* da = nullptr;
  const float* db = nullptr;
  int na = 0, nb = 0;
  int rc = A->seis_device(shot_a, slot, &da, &na);
  if (rc) return rc;
  if (B) {
    rc = B->seis_device(shot_b, slot, &db, &nb);
    if (rc) return rc;
    if (nb != na) return set_error(FOMO_EINVAL, "shots have different receiver counts");
  }
  if (na == 0) return FOMO_OK;
  if (t1 == t0) {
    for (int r = 0; r < na; ++r) num[r] = den[r] = 0.0;
    return FOMO_OK;
  }
  const int len = t1 - t0;
  int n_chunks = (len + kMinChunkLen - 1) / kMinChunkLen;
  if (n_chunks > kMaxChunks) n_chunks = kMaxChunks;
  const int chunk = (len + n_chunks - 1) / n_chunks;
  n_chunks = (len + chunk - 1) / chunk;
  if (B) FOMO_HIP_TRY(hipStreamSynchronize(B->stream));  // b's traces are complete before a's stream reads them
  double* work = nullptr;
  const size_t n_part = (size_t)n_chunks * na * 2;
  FOMO_HIP_TRY(hipMalloc(&work, (n_part + (size_t)na * 2) * sizeof(double)));
  double* out = work + n_part;
  auto fail = [&](hipError_t e, const char* what) {
    (void)hipStreamSynchronize(A->stream);
    (void)hipFree(work);
    return set_error(FOMO_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipLaunchKernelGGL(trace_partial_kernel, dim3((na + kRecPerBlock - 1) / kRecPerBlock, n_chunks),
                     dim3(kRecPerBlock, kTRows), 0, A->stream, da, db, na, t0, t1, chunk, work);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(e, "trace_partial_kernel");
  hipLaunchKernelGGL(trace_final_kernel, dim3((na + 255) / 256), dim3(256), 0, A->stream, work, na, n_chunks, out);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e, "trace_final_kernel");
  double* h = nullptr;
  if ((e = hipHostMalloc((void**)&h, (size_t)na * 2 * sizeof(double))) != hipSuccess) return fail(e, "hipHostMalloc");
  e = hipMemcpyAsync(h, out, (size_t)na * 2 * sizeof(double), hipMemcpyDeviceToHost, A->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(A->stream);
  if (e == hipSuccess)
    for (int r = 0; r < na; ++r) {
      num[r] = h[2 * r];
      den[r] = h[2 * r + 1];
    }
  (void)hipHostFree(h);
  if (e != hipSuccess) return fail(e, "trace misfit download");
  FOMO_HIP_TRY(hipFree(work));
  return FOMO_OK;
}
