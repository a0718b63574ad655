// fomo_api.hip -- host runtime and C-ABI of the gfx950 forward-modeling engine.
//
// Host-side setup restates the reference's setup code (padding, medium, HABC coefficients and
// weights, Ricker wavelet) in the same Float32 operation order; the library is compiled with
// -ffp-contract=off for host and device alike.  The time loop is a plan object holding all device
// state; every step is two sweeps and one gather, and pairs of steps are replayed from one
// captured hipGraph.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fomo_hip.h"
#include "fomo_common.h"
#include "engine3d.h"
#include "plan.h"
#include "plan2d.h"

using namespace fomo;

static thread_local std::string g_last_error;

namespace fomo {
int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
}  // namespace fomo

static int fail(int code, const std::string& msg) { return set_error(code, msg); }
#define HIP_TRY FOMO_HIP_TRY

extern "C" const char* fomo_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* fomo_version(void) { return "fomo-mi355x 0.1.0 (gfx950)"; }
extern "C" int fomo_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// ---------------------------------------------------------------------------------------------
// FD coefficients, src/utils/FD_utils.jl:5-11.  Julia's Float32[...] literal parses each decimal
// as Float64 and rounds to Float32; the double literals below are cast the same way.
static const double kFD[5][5] = {
    {1.0, 0, 0, 0, 0},
    {1.125, -0.041666667, 0, 0, 0},
    {1.171875, -0.065104167, 0.0046875, 0, 0},
    {1.1962890625, -0.079752604167, 0.0095703125, -0.000697544643, 0},
    {1.2115478515625, -0.089721679687, 0.0138427734375, -0.00176565987723, 0.0001186795166},
};

extern "C" int fomo_fd_coefficients(int32_t fd_order, float* out) {
  if (fd_order < 2 || fd_order > 10 || (fd_order & 1))
    return fail(FOMO_EINVAL, "Unsupported FD order: " + std::to_string(fd_order));
  const int N = fd_order / 2;
  for (int l = 0; l < N; ++l) out[l] = (float)kFD[N - 1][l];
  return FOMO_OK;
}

// Centred second-derivative weights, src/utils/FD_centered.jl:31-47: Julia parses each ratio in
// Float64 and rounds it to Float32; so do the double ratios below.
static const double kD2[5][5] = {
    {1.0, 0, 0, 0, 0},
    {4.0 / 3.0, -1.0 / 12.0, 0, 0, 0},
    {3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0, 0, 0},
    {8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0, 0},
    {5.0 / 3.0, -5.0 / 21.0, 5.0 / 126.0, -5.0 / 1008.0, 1.0 / 3150.0},
};
static const double kD2C[5] = {-2.0, -5.0 / 2.0, -49.0 / 18.0, -205.0 / 72.0, -5269.0 / 1800.0};

extern "C" int fomo_centered_d2(int32_t fd_order, float* c0, float* d2) {
  if (fd_order < 2 || fd_order > 10 || (fd_order & 1))
    return fail(FOMO_EINVAL, "Unsupported FD order: " + std::to_string(fd_order) + " (centered d2)");
  if (!c0 || !d2) return fail(FOMO_EINVAL, "NULL argument");
  const int N = fd_order / 2;
  *c0 = (float)kD2C[N - 1];
  for (int l = 0; l < N; ++l) d2[l] = (float)kD2[N - 1][l];
  return FOMO_OK;
}

// Centred first-derivative weights, src/utils/FD_centered.jl:14-20 (Float64 ratios rounded to
// Float32, as Julia's Float32[...] literal rounds them).
static const double kD1[5][5] = {
    {1.0 / 2.0, 0, 0, 0, 0},
    {2.0 / 3.0, -1.0 / 12.0, 0, 0, 0},
    {3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0, 0, 0},
    {4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0, 0},
    {5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0, -5.0 / 504.0, 1.0 / 1260.0},
};

extern "C" int fomo_centered_d1(int32_t fd_order, float* d1) {
  if (fd_order < 2 || fd_order > 10 || (fd_order & 1))
    return fail(FOMO_EINVAL, "Unsupported FD order: " + std::to_string(fd_order) + " (centered d1)");
  if (!d1) return fail(FOMO_EINVAL, "NULL argument");
  const int N = fd_order / 2;
  for (int l = 0; l < N; ++l) d1[l] = (float)kD1[N - 1][l];
  return FOMO_OK;
}

// ---------------------------------------------------------------------------------------------
// Ricker wavelet, src/utils/ricker_wavelet.jl:6-12:
//   t = (0:nt-1) .* dt; tau = t .- 1.5f0/f0; u = pi*f0 .* tau; w = (1 - 2u^2) exp(-u^2).
// Julia evaluates t, tau and u as lazy ranges: `(0:nt-1) .* dt` becomes range(0f0, step=dt),
// whose Float32 constructor looks for a small rational n/d equal to dt (Base.rat, denominators up
// to 2048) and then stores ref/step in Float64 (step = n/d); otherwise ref/step stay Float32.
// Shifting and scaling a range keep its storage type, and elements are T(ref + (k-1)*step)
// evaluated in the storage type.  The square and the product are Float32; exp is evaluated in
// double and rounded once.
static void julia_rat(float x, long* num, long* den) {
  float y = x;
  long a = 1, d = 1, b = 0, c = 0;
  const long m = 2048;  // maxintfloat(Float16, Int)
  while (std::fabs(y) <= (float)m) {
    const long f = (long)std::trunc(y);
    y -= (float)f;
    const long na = f * a + c, nb = f * b + d;
    c = a;
    a = na;
    d = b;
    b = nb;
    if (std::max(std::labs(a), std::labs(b)) > m) {
      *num = c;
      *den = d;
      return;
    }
    if ((float)a / (float)b == x) break;
    y = 1.0f / y;
  }
  *num = a;
  *den = b;
}

extern "C" int fomo_ricker_wavelet(float f0, float dt, int32_t nt, float* out) {
  if (nt < 0) return fail(FOMO_EINVAL, "nt must be >= 0");
  if (!(f0 > 0.0f)) return fail(FOMO_EINVAL, "f0 must be > 0");
  const float t0 = 1.0f / f0;
  const float shift = 1.5f * t0;
  const float pf = (float)M_PI * f0;  // Irrational * Float32 promotes pi to Float32
  long sn = 0, sd = 0;
  julia_rat(dt, &sn, &sd);
  bool wide = false;
  double step64 = 0.0;
  if (sd != 0 && (float)((double)sn / (double)sd) == dt) {
    const long den = sd;  // lcm(1, sd)
    const float dst = (float)den * dt;
    if (std::fabs(dst) <= 16777216.0f && std::fabs((float)den * 0.0f) <= 16777216.0f) {
      const long step_n = std::lrint(dst);  // round-half-even, as round(Int, x)
      step64 = (double)step_n / (double)den;
      wide = true;
    }
  }
  for (int k = 0; k < nt; ++k) {
    float u;
    if (wide) {
      const double ref = 0.0 - (double)shift;
      u = (float)((double)pf * ref + (double)k * ((double)pf * step64));
    } else {
      const float ref = 0.0f - shift;
      const float ur = pf * ref, us = pf * dt;
      u = ur + (float)k * us;
    }
    const float u2 = u * u;
    const float e = (float)std::exp(-(double)u2);
    out[k] = (1.0f - 2.0f * u2) * e;
  }
  return FOMO_OK;
}

// habc.jl:21-28
extern "C" int fomo_habc_coefficients(float v_ref, float dt, float dh, float* out3) {
  const float r = (v_ref * dt) / dh;
  const float b = 0.45f, beta = 1.0f, omb = 1.0f - b;
  out3[0] = (b * (beta + r) - r) / ((beta + r) * omb);
  out3[1] = (b * (beta + r) - beta) / ((beta + r) * omb);
  out3[2] = b / (b - 1.0f);
  return FOMO_OK;
}

// ---------------------------------------------------------------------------------------------
// Medium setup (2D), src/utils/pad_array.jl:8-67, acoustic2d/medium.jl:28-55,
// elastic2d/medium.jl:19-74.
namespace fomo {

// Edge padding of the window [x0, x0+nxi) x [0, nzi) of a model with row stride nx_model: clamp
// indexing into the window, as _pad_array does on the cropped model.
void pad2d(const float* in, int nxi, int nzi, int pad, int x0, int nx_model, std::vector<float>& out) {
  const int nx = nxi + 2 * pad, nz = nzi + 2 * pad;
  out.resize((size_t)nx * nz);
  for (int j = 0; j < nz; ++j) {
    const int jj = std::min(std::max(j - pad, 0), nzi - 1);
    for (int i = 0; i < nx; ++i) {
      const int ii = std::min(std::max(i - pad, 0), nxi - 1);
      out[(size_t)i + (size_t)j * nx] = in[(size_t)(x0 + ii) + (size_t)jj * nx_model];
    }
  }
}

static inline float harm_buoy(float r1, float r2) {
  if (r1 == 0.0f && r2 == 0.0f) return 0.0f;
  if (r1 == 0.0f) return 2.0f / r2;
  if (r2 == 0.0f) return 2.0f / r1;
  return 2.0f / (r1 + r2);
}

void buoyancy2d(const std::vector<float>& rho, int nx, int nz, std::vector<float>& bx, std::vector<float>& bz) {
  bx.assign((size_t)nx * nz, 0.0f);
  bz.assign((size_t)nx * nz, 0.0f);
  for (int j = 0; j < nz; ++j) {
    const size_t r = (size_t)j * nx;
    for (int i = 1; i < nx; ++i) bx[r + i] = harm_buoy(rho[r + i - 1], rho[r + i]);
    bx[r] = bx[r + 1];
  }
  for (int j = 0; j + 1 < nz; ++j)
    for (int i = 0; i < nx; ++i)
      bz[(size_t)i + (size_t)j * nx] = harm_buoy(rho[(size_t)i + (size_t)j * nx], rho[(size_t)i + (size_t)(j + 1) * nx]);
  for (int i = 0; i < nx; ++i) bz[(size_t)i + (size_t)(nz - 1) * nx] = bz[(size_t)i + (size_t)(nz - 2) * nx];
}

void elastic_params2d(const std::vector<float>& vp, const std::vector<float>& vs, const std::vector<float>& rho,
                      int nx, int nz, std::vector<float>& lam, std::vector<float>& lam2mu, std::vector<float>& muxz) {
  const size_t n = (size_t)nx * nz;
  std::vector<float> mu(n);
  lam.resize(n);
  lam2mu.resize(n);
  for (size_t c = 0; c < n; ++c) {
    mu[c] = rho[c] * (vs[c] * vs[c]);
    lam[c] = rho[c] * (vp[c] * vp[c]) - 2.0f * mu[c];
    lam2mu[c] = lam[c] + 2.0f * mu[c];
  }
  muxz.assign(n, 0.0f);
  for (int j = 0; j + 1 < nz; ++j)
    for (int i = 1; i < nx; ++i) {
      const float m1 = mu[(size_t)(i - 1) + (size_t)j * nx], m2 = mu[(size_t)i + (size_t)j * nx];
      const float m3 = mu[(size_t)(i - 1) + (size_t)(j + 1) * nx], m4 = mu[(size_t)i + (size_t)(j + 1) * nx];
      float v = 0.0f;
      if (!(m1 == 0.0f || m2 == 0.0f || m3 == 0.0f || m4 == 0.0f))
        v = 4.0f / (((1.0f / m1 + 1.0f / m2) + 1.0f / m3) + 1.0f / m4);
      muxz[(size_t)i + (size_t)j * nx] = v;
    }
  for (int j = 0; j < nz; ++j) muxz[(size_t)j * nx] = muxz[1 + (size_t)j * nx];
  for (int i = 0; i < nx; ++i) muxz[(size_t)i + (size_t)(nz - 1) * nx] = muxz[(size_t)i + (size_t)(nz - 2) * nx];
}

// habc.jl:30-34: w = Float32(clamp((dist - off)/(pad - 1), 0, 1)) evaluated in Float64.  The
// weight depends on the cell only through dist = distance to the nearest grid edge, so the
// device keeps one small table per staggering offset instead of full-grid arrays.
void weight_lut(int H, int pad, double off, float* out) {
  for (int d = 0; d <= H; ++d) {
    double w = ((double)d - off) / (double)(pad - 1);
    w = std::min(std::max(w, 0.0), 1.0);
    out[d] = (float)w;
  }
}

int upload_stored(PlanBase* p, int storage, void** dst, const float* h, size_t n, float enc) {
  if (storage == FOMO_STORE_F32) {
    float* d = nullptr;
    const int rc = p->upload(&d, h, n);
    *dst = d;
    return rc;
  }
  std::vector<unsigned short> tmp(n);
  for (size_t c = 0; c < n; ++c) {
    const float v = h[c] * enc;
    tmp[c] = storage == FOMO_STORE_F16 ? __half_as_ushort(__float2half_rn(v)) : __bfloat16_as_ushort(__float2bfloat16(v));
  }
  unsigned short* d = nullptr;
  const int rc = p->upload(&d, tmp.data(), n);
  *dst = d;
  return rc;
}

float max_abs(const float* v, size_t n) {
  float m = 0.0f;
  for (size_t c = 0; c < n; ++c) m = std::max(m, std::fabs(v[c]));
  return m;
}

// exp(-(factor*(nbc-d))^2) for d < nbc, 1 otherwise: factor*(nbc-d) and its square in Float32 as
// sponge.jl:44-46 evaluates them, the exponential in double rounded once to Float32.
void sponge_lut(int nbc, int pad, float factor, float* out) {
  for (int d = 0; d < pad; ++d) {
    if (d < nbc) {
      const float r = factor * (float)(nbc - d);
      out[d] = (float)std::exp((double)(-r * r));
    } else {
      out[d] = 1.0f;
    }
  }
}

float min_positive(const float* a, size_t n) {
  float m = INFINITY;
  for (size_t c = 0; c < n; ++c)
    if (a[c] > 0.0f && a[c] < m) m = a[c];
  return std::isfinite(m) ? m : 0.0f;
}

// ---------------------------------------------------------------------------------------------
PlanBase::~PlanBase() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  for (void* v : allocs) (void)hipFree(v);
  if (stream) (void)hipStreamDestroy(stream);
}

__global__ void tick_kernel(int* step, int m) {
  if (threadIdx.x == 0) *step += m;
}

// kGraphSteps steps (parities 0, 1, 0, ... offsets 0 .. kGraphSteps-1) and their counter tick
// are captured into one graph; a replay advances kGraphSteps steps from an even step.
int PlanBase::build_graph() {
  const long saved = steps_done;
  steps_done = 0;
  FOMO_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  hipError_t e1 = run_steps(kGraphSteps);
  steps_done = saved;
  hipGraph_t g = nullptr;
  hipError_t ec = hipStreamEndCapture(stream, &g);
  if (e1 != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return set_error(FOMO_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e1 != hipSuccess ? e1 : ec));
  }
  graph = g;
  FOMO_HIP_TRY(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0));
  return FOMO_OK;
}

int PlanBase::check_overflow() {
  FOMO_HIP_TRY(hipStreamSynchronize(stream));
  if (!ovf) return FOMO_OK;
  unsigned h = 0;
  FOMO_HIP_TRY(hipMemcpy(&h, ovf, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (h)
    return set_error(FOMO_EOVERFLOW,
                     "16-bit storage overflow: a wavefield value exceeded the storage format's range since the last "
                     "reset (stored saturated); the results are invalid.  Use Float32 storage or a smaller source "
                     "amplitude");
  return FOMO_OK;
}

hipError_t PlanBase::run_steps(int m, hipEvent_t* ev) {
  hipError_t e = hipSuccess;
  const bool per_step_gather = !gather_in_sweep();
  for (int k = 0; k < m && e == hipSuccess; ++k) {
    const int s = (int)((steps_done + k) % 2);
    for (int part = 0; part < (per_step_gather ? 3 : 2) && e == hipSuccess; ++part) {
      if (ev && k == 0 && part < 2) e = hipEventRecord(ev[part], stream);
      if (e == hipSuccess) e = enqueue_part(part, s, k);
    }
    if (e == hipSuccess && ev && k == 0) e = hipEventRecord(ev[2], stream);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(1), 0, stream, step, m);
    e = hipGetLastError();
  }
  if (e == hipSuccess && ev) e = hipEventRecord(ev[3], stream);
  if (e == hipSuccess) steps_done += m;
  return e;
}

int PlanBase::advance(int n) {
  if (n < 0) return set_error(FOMO_EINVAL, "step count must be >= 0");
  if (n == 0) return FOMO_OK;
  while (n > 0) {
    if (gexec && (steps_done % 2) == 0 && n >= kGraphSteps) {
      FOMO_HIP_TRY(hipGraphLaunch(gexec, stream));
      steps_done += kGraphSteps;
      n -= kGraphSteps;
    } else {
      // without a graph: all at once; with one: a single step to reach an even step, or the
      // tail shorter than a replay
      const int m = (gexec && (steps_done % 2)) ? 1 : (gexec ? std::min(n, kGraphSteps - 1) : n);
      FOMO_HIP_TRY(run_steps(m));
      n -= m;
    }
  }
  if (gather_in_sweep()) FOMO_HIP_TRY(enqueue_part(2, (int)(steps_done % 2), 0));
  return FOMO_OK;
}

}  // namespace fomo

// ---------------------------------------------------------------------------------------------
// C-ABI over plans.
static int validate_common(const fomo_desc* d) {
  if (!d) return fail(FOMO_EINVAL, "descriptor is NULL");
  if (d->fd_order < 2 || d->fd_order > 10 || (d->fd_order & 1))
    return fail(FOMO_EINVAL, "Unsupported FD order: " + std::to_string(d->fd_order));
  if (d->nx < 1 || d->nz < 1 || d->ny < 1) return fail(FOMO_EINVAL, "grid dimensions must be >= 1");
  if (d->nbc < 0) return fail(FOMO_EINVAL, "nbc must be >= 0");
  if (d->nbc + d->fd_order / 2 < 2) return fail(FOMO_EINVAL, "nbc + fd_order/2 must be >= 2");
  if (d->nt < 0) return fail(FOMO_EINVAL, "nt must be >= 0");
  if (d->n_src < 0 || d->n_rec < 0) return fail(FOMO_EINVAL, "n_src and n_rec must be >= 0");
  if (d->snap_interval < 0) return fail(FOMO_EINVAL, "snap_interval must be >= 0");
  if (!(d->dh > 0.0f) || !(d->dt > 0.0f)) return fail(FOMO_EINVAL, "dh and dt must be > 0");
  if (d->boundary != FOMO_BOUNDARY_HABC && d->boundary != FOMO_BOUNDARY_SPONGE)
    return fail(FOMO_EINVAL, "boundary must be FOMO_BOUNDARY_HABC or FOMO_BOUNDARY_SPONGE");
  if (d->boundary == FOMO_BOUNDARY_SPONGE) {
    if (!(d->sponge_factor > 0.0f)) return fail(FOMO_EINVAL, "sponge_factor must be > 0");
    if (d->eq != FOMO_ACOUSTIC2D && d->eq != FOMO_ELASTIC2D)
      return fail(FOMO_EUNSUP, "the sponge boundary is 2D only (as in the reference)");
  }
  // The sweeps address cells with 32-bit offsets: a 2D field, or one z plane of a 3D field (rows
  // at their pitch), may hold at most 2^30 padded cells (its Float32 byte offsets then stay below
  // 2^32).
  {
    const long long pad = (long long)d->nbc + d->fd_order / 2;
    long long px = d->nx + 2 * pad;
    const long long py = d->ny + 2 * pad, pz = d->nz + 2 * pad;
    const bool is3d = d->eq == FOMO_ACOUSTIC3D || d->eq == FOMO_ELASTIC3D;
    if (is3d && px < (1LL << 30))
      px = row_pitch3d((int)px, (int)pad, d->storage == FOMO_STORE_F32 ? 4 : 2).px;
    const long long cells = is3d ? px * py : px * pz;
    if (cells > (1LL << 30))
      return fail(FOMO_EUNSUP, std::string(is3d ? "3D plane" : "2D field") + " of " + std::to_string(cells) +
                                   " padded cells exceeds the 2^30 cells that 32-bit cell offsets address");
  }
  if (fomo_device_count() < 1) return fail(FOMO_EHIP, "no HIP device visible");
  return FOMO_OK;
}

extern "C" int fomo_batch_create(const fomo_desc* d, const float* vp, const float* vs, const float* rho,
                                 int32_t n_shots, const fomo_shot* shots, fomo_plan** out) {
  if (!out) return fail(FOMO_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = validate_common(d);
  if (rc) return rc;
  if (d->eq != FOMO_ACOUSTIC2D && d->eq != FOMO_ELASTIC2D && d->eq != FOMO_COUPLED2D)
    return fail(FOMO_EUNSUP, "batches are 2D (acoustic2d, elastic2d, coupled2d)");
  Plan2D* p = nullptr;
  if ((rc = create_plan2d(d, vp, vs, rho, n_shots, shots, &p))) return rc;
  *out = new fomo_plan{p};
  return FOMO_OK;
}

// A single-shot 2D plan of any 2D equation (acoustic2d, elastic2d, scalar2d, coupled2d).
static int create_single2d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, const fomo_shot* s,
                           fomo_plan** out) {
  if (d->eq == FOMO_SCALAR2D || d->eq == FOMO_COUPLED2D) {
    if (d->boundary != FOMO_BOUNDARY_HABC)
      return fail(FOMO_EUNSUP, std::string(d->eq == FOMO_SCALAR2D ? "scalar2d" : "coupled2d") +
                                   " uses the HABC (as in the reference)");
    Plan2D* p = nullptr;
    const int rc = create_plan2d(d, vp, vs, rho, 1, s, &p);
    if (rc) return rc;
    *out = new fomo_plan{p};
    return FOMO_OK;
  }
  return fomo_batch_create(d, vp, vs, rho, 1, s, out);
}

int fomo_plan_create3d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, const int32_t* sx,
                       const int32_t* sy, const int32_t* sz, const float* wavelet, const int32_t* rx,
                       const int32_t* ry, const int32_t* rz, fomo_plan** out);

extern "C" int fomo_plan_create(const fomo_desc* d, const float* vp, const float* vs, const float* rho,
                                const int32_t* sx, const int32_t* sy, const int32_t* sz, const float* wavelet,
                                const int32_t* rx, const int32_t* ry, const int32_t* rz, fomo_plan** out) {
  if (!out) return fail(FOMO_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = validate_common(d);
  if (rc) return rc;
  if (d->eq == FOMO_ACOUSTIC3D || d->eq == FOMO_ELASTIC3D)
    return fomo_plan_create3d(d, vp, vs, rho, sx, sy, sz, wavelet, rx, ry, rz, out);
  if (d->eq != FOMO_ACOUSTIC2D && d->eq != FOMO_ELASTIC2D && d->eq != FOMO_SCALAR2D && d->eq != FOMO_COUPLED2D)
    return fail(FOMO_EINVAL, "unknown equation " + std::to_string(d->eq));
  fomo_shot s{};
  s.x0 = 1;
  s.nx = d->nx;
  s.n_src = d->n_src;
  s.sx = sx;
  s.sz = sz;
  s.wavelet = wavelet;
  s.n_rec = d->n_rec;
  s.rx = rx;
  s.rz = rz;
  s.v_ref = d->v_ref;
  return create_single2d(d, vp, vs, rho, &s, out);
}

#define PLAN_OR_FAIL(p)                                         \
  do {                                                          \
    if (!(p) || !(p)->impl) return fail(FOMO_EINVAL, "plan is NULL"); \
  } while (0)

extern "C" int fomo_plan_reset(fomo_plan* p) {
  PLAN_OR_FAIL(p);
  return p->impl->reset();
}

extern "C" int fomo_plan_advance(fomo_plan* p, int32_t n) {
  PLAN_OR_FAIL(p);
  return p->impl->advance(n);
}

extern "C" int fomo_plan_timed_advance(fomo_plan* p, int32_t n, double* ms) {
  PLAN_OR_FAIL(p);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, p->impl->stream));
  int rc = p->impl->advance(n);
  HIP_TRY(hipEventRecord(e1, p->impl->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0.0f;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  return rc ? rc : p->impl->check_overflow();
}

extern "C" int fomo_plan_kernel_times(fomo_plan* p, int32_t n, double* out3) {
  PLAN_OR_FAIL(p);
  if (!out3) return fail(FOMO_EINVAL, "NULL argument");
  if (n < 1) return fail(FOMO_EINVAL, "step count must be >= 1");
  PlanBase* b = p->impl;
  hipEvent_t ev[2];
  for (int k = 0; k < 2; ++k) HIP_TRY(hipEventCreate(&ev[k]));
  // The two sweeps: n back-to-back launches of the same (idempotent) phase between two events.
  const int s = (int)(b->steps_done % 2);
  for (int part = 0; part < 2; ++part) {
    HIP_TRY(b->enqueue_part(part, s, 0));  // warm
    HIP_TRY(hipEventRecord(ev[0], b->stream));
    for (int t = 0; t < n; ++t) HIP_TRY(b->enqueue_part(part, s, 0));
    HIP_TRY(hipEventRecord(ev[1], b->stream));
    HIP_TRY(hipEventSynchronize(ev[1]));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
    out3[part] = ms / n;
  }
  // Everything else per step (receiver gather, counter ticks, launch gaps): n ordinary steps
  // through advance, less the two sweeps.
  HIP_TRY(hipEventRecord(ev[0], b->stream));
  int rc = b->advance(n);
  HIP_TRY(hipEventRecord(ev[1], b->stream));
  HIP_TRY(hipEventSynchronize(ev[1]));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (int k = 0; k < 2; ++k) (void)hipEventDestroy(ev[k]);
  if (rc) return rc;
  out3[2] = std::max(0.0, (double)ms / n - out3[0] - out3[1]);
  return FOMO_OK;
}

extern "C" int fomo_plan_read_snaps(fomo_plan* p, int32_t shot, int32_t* count, float* out) {
  PLAN_OR_FAIL(p);
  const float* src = nullptr;
  int n = 0;
  size_t cells = 0;
  int rc = p->impl->snap_buffer(shot, &src, &n, &cells);
  if (rc) return rc;
  if (count) *count = n;
  if (out && n) {
    HIP_TRY(hipStreamSynchronize(p->impl->stream));
    HIP_TRY(hipMemcpy(out, src, (size_t)n * cells * sizeof(float), hipMemcpyDeviceToHost));
  }
  return p->impl->check_overflow();
}

extern "C" int fomo_plan_launch_count(const fomo_plan* p, int32_t* n) {
  if (!p || !p->impl || !n) return fail(FOMO_EINVAL, "NULL argument");
  *n = p->impl->n_launches();
  return FOMO_OK;
}

extern "C" int fomo_plan_launch_info(const fomo_plan* p, int32_t k, const char** name, int64_t* cells,
                                     int32_t* bytes_per_cell) {
  if (!p || !p->impl) return fail(FOMO_EINVAL, "NULL plan");
  if (k < 0 || k >= p->impl->n_launches()) return fail(FOMO_EINVAL, "launch index out of range");
  const PlanBase::LaunchInfo li = p->impl->launch_info(k);
  if (name) *name = li.name;
  if (cells) *cells = li.cells;
  if (bytes_per_cell) *bytes_per_cell = li.bytes_per_cell;
  return FOMO_OK;
}

extern "C" int fomo_plan_launch_times(fomo_plan* p, int32_t n, double* ms) {
  PLAN_OR_FAIL(p);
  if (!ms) return fail(FOMO_EINVAL, "NULL argument");
  if (n < 1) return fail(FOMO_EINVAL, "repeat count must be >= 1");
  PlanBase* b = p->impl;
  hipEvent_t ev[2];
  for (int k = 0; k < 2; ++k) HIP_TRY(hipEventCreate(&ev[k]));
  const int s = (int)(b->steps_done % 2);
  int rc = FOMO_OK;
  for (int k = 0; k < b->n_launches() && rc == FOMO_OK; ++k) {
    hipError_t e = b->enqueue_launch(k, s);  // warm
    if (e == hipSuccess) e = hipEventRecord(ev[0], b->stream);
    for (int t = 0; t < n && e == hipSuccess; ++t) e = b->enqueue_launch(k, s);
    if (e == hipSuccess) e = hipEventRecord(ev[1], b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float t_ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t_ms, ev[0], ev[1]);
    if (e != hipSuccess) rc = fail(FOMO_EHIP, std::string("launch timing: ") + hipGetErrorString(e));
    ms[k] = t_ms / n;
  }
  for (int k = 0; k < 2; ++k) (void)hipEventDestroy(ev[k]);
  return rc;
}

extern "C" int fomo_plan_tuning(const fomo_plan* p, int32_t* out4) {
  if (!p || !p->impl || !out4) return fail(FOMO_EINVAL, "NULL argument");
  out4[0] = p->impl->tune_family;
  out4[1] = p->impl->tune_shape;
  out4[2] = p->impl->tune_extent;
  out4[3] = p->impl->tune_source;
  return FOMO_OK;
}

extern "C" int fomo_plan_sync(fomo_plan* p) {
  PLAN_OR_FAIL(p);
  return p->impl->check_overflow();
}

extern "C" int fomo_plan_steps_done(const fomo_plan* p, int32_t* out) {
  if (!p || !p->impl || !out) return fail(FOMO_EINVAL, "NULL argument");
  *out = (int32_t)p->impl->steps_done;
  return FOMO_OK;
}

extern "C" int fomo_plan_n_shots(const fomo_plan* p, int32_t* out) {
  if (!p || !p->impl || !out) return fail(FOMO_EINVAL, "NULL argument");
  *out = p->impl->n_shots();
  return FOMO_OK;
}

extern "C" int fomo_plan_read_shot_seis(fomo_plan* p, int32_t shot, float* s0, float* s1, float* s2, float* s3) {
  PLAN_OR_FAIL(p);
  float* dst[4] = {s0, s1, s2, s3};
  const int rc = p->impl->read_seis(shot, dst);
  return rc ? rc : p->impl->check_overflow();
}

extern "C" int fomo_plan_read_seis(fomo_plan* p, float* s0, float* s1, float* s2, float* s3) {
  return fomo_plan_read_shot_seis(p, 0, s0, s1, s2, s3);
}

extern "C" int fomo_plan_read_shot_field(fomo_plan* p, int32_t shot, int32_t which, float* out) {
  PLAN_OR_FAIL(p);
  if (!out) return fail(FOMO_EINVAL, "NULL argument");
  const int rc = p->impl->read_field(shot, which, out);
  return rc ? rc : p->impl->check_overflow();
}

extern "C" int fomo_plan_read_field(fomo_plan* p, int32_t which, float* out) {
  return fomo_plan_read_shot_field(p, 0, which, out);
}

extern "C" int fomo_plan_shot_dims(const fomo_plan* p, int32_t shot, int32_t* nx, int32_t* ny, int32_t* nz,
                                   int32_t* pad) {
  if (!p || !p->impl) return fail(FOMO_EINVAL, "plan is NULL");
  int a, b, c, e;
  int rc = p->impl->dims(shot, &a, &b, &c, &e);
  if (rc) return rc;
  if (nx) *nx = a;
  if (ny) *ny = b;
  if (nz) *nz = c;
  if (pad) *pad = e;
  return FOMO_OK;
}

extern "C" int fomo_plan_dims(const fomo_plan* p, int32_t* nx, int32_t* ny, int32_t* nz, int32_t* pad) {
  return fomo_plan_shot_dims(p, 0, nx, ny, nz, pad);
}

extern "C" int fomo_plan_destroy(fomo_plan* p) {
  if (p) {
    delete p->impl;
    delete p;
  }
  return FOMO_OK;
}

// ---------------------------------------------------------------------------------------------
// One-shot drop-in runs: set up, warm up one step (as the reference does), reset, time the loop,
// copy results back.  2D snapshots are written by the sweeps into a device buffer inside the step
// loop (graph replays included) and downloaded once after it.
static int timed_run(fomo_plan* p, const fomo_desc* d, float* snaps, double* kernel_time_s) {
  PlanBase* b = p->impl;
  int rc;
  if (d->nt > 0) {
    if ((rc = b->advance(1)) || (rc = fomo_plan_sync(p)) || (rc = b->reset())) return rc;
  }
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, b->stream));
  if ((rc = b->advance(d->nt))) return rc;
  HIP_TRY(hipEventRecord(e1, b->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (kernel_time_s) *kernel_time_s = ms * 1e-3;
  if (d->snap_interval > 0 && snaps) {
    const Shot2D& sh = static_cast<Plan2D*>(b)->shots[0];
    const size_t n = (size_t)sh.snap_n * d->nx * d->nz;
    if (!sh.snap) return fail(FOMO_EHIP, "snapshot buffer missing");
    if (n) HIP_TRY(hipMemcpy(snaps, sh.snap, n * sizeof(float), hipMemcpyDeviceToHost));
  }
  return FOMO_OK;
}

static int run_2d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, const int32_t* sx,
                  const int32_t* sz, const float* wavelet, const int32_t* rx, const int32_t* rz, float* s0, float* s1,
                  float* s2, float* snaps, double* kernel_time_s) {
  fomo_plan* p = nullptr;
  int rc = fomo_plan_create(d, vp, vs, rho, sx, nullptr, sz, wavelet, rx, nullptr, rz, &p);
  if (rc) return rc;
  rc = timed_run(p, d, snaps, kernel_time_s);
  if (!rc) rc = fomo_plan_read_seis(p, s0, s1, s2, nullptr);
  fomo_plan_destroy(p);
  return rc;
}

extern "C" int fomo_acoustic2d(const fomo_desc* d, const float* vp, const float* rho, const int32_t* sx,
                               const int32_t* sz, const float* wavelet, const int32_t* rx, const int32_t* rz,
                               float* seis_p, float* seis_vx, float* seis_vz, float* snaps, double* kernel_time_s) {
  if (!d || d->eq != FOMO_ACOUSTIC2D) return fail(FOMO_EINVAL, "descriptor eq must be FOMO_ACOUSTIC2D");
  return run_2d(d, vp, nullptr, rho, sx, sz, wavelet, rx, rz, seis_p, seis_vx, seis_vz, snaps, kernel_time_s);
}

extern "C" int fomo_elastic2d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, const int32_t* sx,
                              const int32_t* sz, const float* wavelet, const int32_t* rx, const int32_t* rz,
                              float* seis_vx, float* seis_vz, float* snaps, double* kernel_time_s) {
  if (!d || d->eq != FOMO_ELASTIC2D) return fail(FOMO_EINVAL, "descriptor eq must be FOMO_ELASTIC2D");
  return run_2d(d, vp, vs, rho, sx, sz, wavelet, rx, rz, seis_vx, seis_vz, nullptr, snaps, kernel_time_s);
}

extern "C" int fomo_scalar2d(const fomo_desc* d, const float* vp, const int32_t* sx, const int32_t* sz,
                             const float* wavelet, const int32_t* rx, const int32_t* rz, float* seis_u,
                             double* kernel_time_s) {
  if (!d || d->eq != FOMO_SCALAR2D) return fail(FOMO_EINVAL, "descriptor must have eq = FOMO_SCALAR2D");
  if (d->snap_interval != 0) return fail(FOMO_EINVAL, "scalar2d takes no snapshots (as in the reference)");
  return run_2d(d, vp, nullptr, nullptr, sx, sz, wavelet, rx, rz, seis_u, nullptr, nullptr, nullptr, kernel_time_s);
}

extern "C" int fomo_coupled2d(const fomo_desc* d, const float* vp, const float* vs, const int32_t* sx,
                              const int32_t* sz, const float* wavelet, const int32_t* rx, const int32_t* rz,
                              float* seis_P, float* seis_S, float* snaps_P, float* snaps_S, double* kernel_time_s) {
  if (!d || d->eq != FOMO_COUPLED2D) return fail(FOMO_EINVAL, "descriptor eq must be FOMO_COUPLED2D");
  fomo_plan* p = nullptr;
  int rc = fomo_plan_create(d, vp, vs, nullptr, sx, nullptr, sz, wavelet, rx, nullptr, rz, &p);
  if (rc) return rc;
  rc = timed_run(p, d, nullptr, kernel_time_s);
  if (!rc) rc = fomo_plan_read_seis(p, seis_P, seis_S, nullptr, nullptr);
  if (!rc && d->snap_interval > 0 && (snaps_P || snaps_S)) {
    // the plan's snapshot slots: every P snapshot, then every S snapshot
    const Shot2D& sh = static_cast<Plan2D*>(p->impl)->shots[0];
    const size_t per = (size_t)sh.snap_n * d->nx * d->nz;
    if (per && !sh.snap) rc = fail(FOMO_EHIP, "snapshot buffer missing");
    hipError_t e = hipSuccess;
    if (!rc && per && snaps_P) e = hipMemcpy(snaps_P, sh.snap, per * sizeof(float), hipMemcpyDeviceToHost);
    if (!rc && per && snaps_S && e == hipSuccess)
      e = hipMemcpy(snaps_S, sh.snap2, per * sizeof(float), hipMemcpyDeviceToHost);
    if (!rc && e != hipSuccess) rc = fail(FOMO_EHIP, std::string("snapshots: ") + hipGetErrorString(e));
  }
  fomo_plan_destroy(p);
  return rc;
}

// Reference batch form: every shot spans the whole model and shares the receivers and wavelet.
static int run_batch(const fomo_desc* d, const float* vp, const float* vs, const float* rho, int32_t n_shots,
                     const int32_t* sx, const int32_t* sz, const float* wavelet, const int32_t* rx, const int32_t* rz,
                     float* s0, float* s1, float* s2, double* kernel_time_s) {
  if (!d) return fail(FOMO_EINVAL, "descriptor is NULL");
  if (n_shots < 1) return fail(FOMO_EINVAL, "at least one shot is required");
  if (d->snap_interval != 0) return fail(FOMO_EINVAL, "batches take no snapshots");
  std::vector<fomo_shot> shots(n_shots);
  for (int s = 0; s < n_shots; ++s) {
    fomo_shot& x = shots[s];
    x = fomo_shot{};
    x.x0 = 1;
    x.nx = d->nx;
    x.n_src = d->n_src;
    x.sx = sx ? sx + (size_t)s * d->n_src : nullptr;
    x.sz = sz ? sz + (size_t)s * d->n_src : nullptr;
    x.wavelet = wavelet;
    x.n_rec = d->n_rec;
    x.rx = rx;
    x.rz = rz;
    x.v_ref = d->v_ref;
  }
  fomo_plan* p = nullptr;
  int rc = fomo_batch_create(d, vp, vs, rho, n_shots, shots.data(), &p);
  if (rc) return rc;
  rc = timed_run(p, d, nullptr, kernel_time_s);
  const size_t per = (size_t)d->n_rec * d->nt;
  for (int s = 0; s < n_shots && !rc; ++s)
    rc = fomo_plan_read_shot_seis(p, s, s0 ? s0 + s * per : nullptr, s1 ? s1 + s * per : nullptr,
                                  s2 ? s2 + s * per : nullptr, nullptr);
  fomo_plan_destroy(p);
  return rc;
}

extern "C" int fomo_acoustic2d_batch(const fomo_desc* d, const float* vp, const float* rho, int32_t n_shots,
                                     const int32_t* sx, const int32_t* sz, const float* wavelet, const int32_t* rx,
                                     const int32_t* rz, float* seis_p, float* seis_vx, float* seis_vz,
                                     double* kernel_time_s) {
  if (!d || d->eq != FOMO_ACOUSTIC2D) return fail(FOMO_EINVAL, "descriptor eq must be FOMO_ACOUSTIC2D");
  return run_batch(d, vp, nullptr, rho, n_shots, sx, sz, wavelet, rx, rz, seis_p, seis_vx, seis_vz, kernel_time_s);
}

extern "C" int fomo_elastic2d_batch(const fomo_desc* d, const float* vp, const float* vs, const float* rho,
                                    int32_t n_shots, const int32_t* sx, const int32_t* sz, const float* wavelet,
                                    const int32_t* rx, const int32_t* rz, float* seis_vx, float* seis_vz,
                                    double* kernel_time_s) {
  if (!d || d->eq != FOMO_ELASTIC2D) return fail(FOMO_EINVAL, "descriptor eq must be FOMO_ELASTIC2D");
  return run_batch(d, vp, vs, rho, n_shots, sx, sz, wavelet, rx, rz, seis_vx, seis_vz, nullptr, kernel_time_s);
}

