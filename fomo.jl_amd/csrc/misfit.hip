// misfit.hip -- FP64 trace misfit and trace energy, reduced on the GPU over recorded traces.
//
// The reference measures a boundary's reflection as ||u - u_ref|| / ||u_ref|| over whole
// seismograms (test/verify_patch.jl:86-87) and checks solvers by per-trace relative L2; it forms
// both on the host.  Here the sums run on the device where the traces already are, in FP64 and in
// a fixed order: pass 1 splits the time axis into chunks (one block per 64 receivers x chunk, four
// wave rows striding over the chunk's samples, the rows added in row order), pass 2 adds the
// chunks of each receiver in chunk order.  The result is a pure function of the traces.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "plan.h"

namespace fomo {

// Scratch of fomo_trace_misfit_noalloc: the partial sums (2 doubles per receiver and time chunk)
// followed by the per-receiver results (2 doubles per receiver) of one call, at most
// kScratchDoubles doubles (16 MiB).  Each GPU the library runs on holds its own copy of this
// array until the process ends.  Host code never passes the array's name to a kernel: it uses
// the device address hipGetSymbolAddress returns for the current device (scratch_on_device).
constexpr size_t kScratchDoubles = (size_t)1 << 21;
static __device__ __align__(256) double scratch[kScratchDoubles];

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

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {nullptr};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_on_device(double** out) {
  int dev = 0;
  FOMO_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(FOMO_EUNSUP, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(FOMO_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return FOMO_OK;
}

struct MisfitJob {
  const float* da = nullptr;
  const float* db = nullptr;
  int na = 0, t0 = 0, t1 = 0, n_chunks = 0, chunk = 0;
  size_t n_part = 0;  // doubles of partial sums (the results follow them)
};

// Validates the arguments and resolves the traces; job.na == 0 or t1 == t0 needs no kernel.
int prepare_misfit(fomo_plan* pa, int32_t shot_a, fomo_plan* pb, int32_t shot_b, int32_t slot, int32_t t0,
                   int32_t t1, double* num, double* den, MisfitJob& job) {
  if (!pa || !pa->impl) return set_error(FOMO_EINVAL, "plan is NULL");
  if (pb && !pb->impl) return set_error(FOMO_EINVAL, "second plan is NULL");
  if (!num || !den) return set_error(FOMO_EINVAL, "NULL output");
  PlanBase* A = pa->impl;
  PlanBase* B = pb ? pb->impl : nullptr;
  const int nt = A->d.nt;
  if (B && B->d.nt != nt) return set_error(FOMO_EINVAL, "plans record different numbers of time samples");
  if (t0 < 0 || t1 > nt || t0 > t1) return set_error(FOMO_EINVAL, "time window outside [0, nt]");
  int na = 0, nb = 0;
  int rc = A->seis_device(shot_a, slot, &job.da, &na);
  if (rc) return rc;
  if (B) {
    rc = B->seis_device(shot_b, slot, &job.db, &nb);
    if (rc) return rc;
    if (nb != na) return set_error(FOMO_EINVAL, "shots have different receiver counts");
  }
  job.na = na;
  job.t0 = t0;
  job.t1 = t1;
  if (na == 0 || t1 == t0) return FOMO_OK;
  const int len = t1 - t0;
  int n_chunks = (len + kMinChunkLen - 1) / kMinChunkLen;
  if (n_chunks > kMaxChunks) n_chunks = kMaxChunks;
  job.chunk = (len + n_chunks - 1) / n_chunks;
  job.n_chunks = (len + job.chunk - 1) / job.chunk;
  job.n_part = (size_t)job.n_chunks * na * 2;
  return FOMO_OK;
}

// Both passes on A's stream into work (partials) and work + n_part (results).
hipError_t enqueue_misfit(const MisfitJob& j, double* work, hipStream_t s) {
  hipLaunchKernelGGL(trace_partial_kernel, dim3((j.na + kRecPerBlock - 1) / kRecPerBlock, j.n_chunks),
                     dim3(kRecPerBlock, kTRows), 0, s, j.da, j.db, j.na, j.t0, j.t1, j.chunk, work);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(trace_final_kernel, dim3((j.na + 255) / 256), dim3(256), 0, s, work, j.na, j.n_chunks,
                     work + j.n_part);
  return hipGetLastError();
}

}  // namespace
}  // namespace fomo

using namespace fomo;

extern "C" int fomo_trace_misfit(fomo_plan* pa, int32_t shot_a, fomo_plan* pb, int32_t shot_b, int32_t slot,
                                 int32_t t0, int32_t t1, double* num, double* den) {
  MisfitJob j;
  int rc = prepare_misfit(pa, shot_a, pb, shot_b, slot, t0, t1, num, den, j);
  if (rc) return rc;
  const int na = j.na;
  if (na == 0) return FOMO_OK;
  if (t1 == t0) {
    for (int r = 0; r < na; ++r) num[r] = den[r] = 0.0;
    return FOMO_OK;
  }
  PlanBase* A = pa->impl;
  if (pb) FOMO_HIP_TRY(hipStreamSynchronize(pb->impl->stream));  // b's traces are complete before a's stream reads them
  double* work = nullptr;
  FOMO_HIP_TRY(hipMalloc(&work, (j.n_part + (size_t)na * 2) * sizeof(double)));
  double* out = work + j.n_part;
  auto fail = [&](hipError_t e, const char* what) {
    (void)hipStreamSynchronize(A->stream);
    (void)hipFree(work);
    return set_error(FOMO_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e = enqueue_misfit(j, work, A->stream);
  if (e != hipSuccess) return fail(e, "trace misfit kernels");
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

// The same sums with the library's static device scratch instead of allocations: no hipMalloc,
// hipFree or pinned host buffer.  Calls that overlap on one GPU share its scratch and race; each
// call overwrites every scratch element it reads (pass 1 writes every partial of every chunk and
// receiver before pass 2 reads them), so nothing needs clearing between calls.
extern "C" int fomo_trace_misfit_noalloc(fomo_plan* pa, int32_t shot_a, fomo_plan* pb, int32_t shot_b, int32_t slot,
                                         int32_t t0, int32_t t1, double* num, double* den) {
  MisfitJob j;
  int rc = prepare_misfit(pa, shot_a, pb, shot_b, slot, t0, t1, num, den, j);
  if (rc) return rc;
  const int na = j.na;
  if (na == 0) return FOMO_OK;
  if (t1 == t0) {
    for (int r = 0; r < na; ++r) num[r] = den[r] = 0.0;
    return FOMO_OK;
  }
  const size_t need = j.n_part + (size_t)na * 2;
  if (need > kScratchDoubles)
    return set_error(FOMO_EINVAL, "trace misfit without allocation: (time chunks + 1) x 2 x receivers = " +
                                      std::to_string(need) + " doubles exceeds the static scratch of " +
                                      std::to_string(kScratchDoubles) + " (chunks = min(64, ceil(samples / 256)))");
  PlanBase* A = pa->impl;
  double* work = nullptr;
  if ((rc = scratch_on_device(&work))) return rc;
  if (pb) FOMO_HIP_TRY(hipStreamSynchronize(pb->impl->stream));
  FOMO_HIP_TRY(enqueue_misfit(j, work, A->stream));
  // the results are interleaved (num, den) per receiver: one copy each into the caller's arrays
  FOMO_HIP_TRY(hipMemcpy2DAsync(num, sizeof(double), work + j.n_part, 2 * sizeof(double), sizeof(double), na,
                                hipMemcpyDeviceToHost, A->stream));
  FOMO_HIP_TRY(hipMemcpy2DAsync(den, sizeof(double), work + j.n_part + 1, 2 * sizeof(double), sizeof(double), na,
                                hipMemcpyDeviceToHost, A->stream));
  FOMO_HIP_TRY(hipStreamSynchronize(A->stream));
  return FOMO_OK;
}
