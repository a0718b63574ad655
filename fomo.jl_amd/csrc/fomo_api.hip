// fomo_api.hip -- host runtime and C-ABI of the gfx950 forward-modeling engine.
//
// Host-side setup restates the reference's setup code (padding, medium, HABC coefficients and
// weights, Ricker wavelet) in the same Float32 operation order; the library is compiled with
// -ffp-contract=off for host and device alike.  The time loop is a plan object holding all device
// state; every step is two sweeps and one gather, and pairs of steps are replayed from one
// captured hipGraph.
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

using namespace fomo;

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      return fail(_e == hipErrorOutOfMemory ? FOMO_ENOMEM : FOMO_EHIP,                       \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                        \
  } while (0)

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
// elastic2d/medium.jl:19-74.  Edge padding = clamp indexing of the interior.
static void pad2d(const float* in, int nxi, int nzi, int pad, std::vector<float>& out) {
  const int nx = nxi + 2 * pad, nz = nzi + 2 * pad;
  out.resize((size_t)nx * nz);
  for (int j = 0; j < nz; ++j) {
    const int jj = std::min(std::max(j - pad, 0), nzi - 1);
    for (int i = 0; i < nx; ++i) {
      const int ii = std::min(std::max(i - pad, 0), nxi - 1);
      out[(size_t)i + (size_t)j * nx] = in[(size_t)ii + (size_t)jj * nxi];
    }
  }
}

static inline float harm_buoy(float r1, float r2) {
  if (r1 == 0.0f && r2 == 0.0f) return 0.0f;
  if (r1 == 0.0f) return 2.0f / r2;
  if (r2 == 0.0f) return 2.0f / r1;
  return 2.0f / (r1 + r2);
}

static void buoyancy2d(const std::vector<float>& rho, int nx, int nz, std::vector<float>& bx,
                       std::vector<float>& bz) {
  bx.assign((size_t)nx * nz, 0.0f);
  bz.assign((size_t)nx * nz, 0.0f);
  for (int j = 0; j < nz; ++j) {
    const size_t r = (size_t)j * nx;
    for (int i = 1; i < nx; ++i) bx[r + i] = harm_buoy(rho[r + i - 1], rho[r + i]);
    bx[r] = bx[r + 1];
  }
  for (int j = 0; j + 1 < nz; ++j)
    for (int i = 0; i < nx; ++i) bz[(size_t)i + (size_t)j * nx] = harm_buoy(rho[(size_t)i + (size_t)j * nx], rho[(size_t)i + (size_t)(j + 1) * nx]);
  for (int i = 0; i < nx; ++i) bz[(size_t)i + (size_t)(nz - 1) * nx] = bz[(size_t)i + (size_t)(nz - 2) * nx];
}

static void elastic_params2d(const std::vector<float>& vp, const std::vector<float>& vs, const std::vector<float>& rho,
                             int nx, int nz, std::vector<float>& lam, std::vector<float>& lam2mu,
                             std::vector<float>& muxz) {
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
static void weight_lut(int H, int pad, double off, float* out) {
  for (int d = 0; d <= H; ++d) {
    double w = ((double)d - off) / (double)(pad - 1);
    w = std::min(std::max(w, 0.0), 1.0);
    out[d] = (float)w;
  }
}

// ---------------------------------------------------------------------------------------------
struct fomo_plan {
  fomo_desc d{};
  int eq = 0, N = 0, pad = 0, H = 0;
  int nx = 0, ny = 1, nz = 0;     // padded
  int nxi = 0, nyi = 1, nzi = 0;  // interior
  float a[kMaxN] = {0};
  float dtx = 0.0f, q = 0.0f, qt = 0.0f, qxt = 0.0f;
  int dev = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  int nfld = 0;
  float* fld[2][6] = {{nullptr}};
  float* mat[6] = {nullptr};
  float* wlut = nullptr;  // 4 tables of (H+1): vx, vz, vy, tau
  int* src_c = nullptr;
  float* wav = nullptr;
  int* rec_c = nullptr;
  int nseis = 0;
  float* seis[4] = {nullptr};
  int* step = nullptr;
  unsigned* done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long steps_done = 0;
};

template <class T>
static int dalloc(fomo_plan* p, T** ptr, size_t count) {
  void* v = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&v, bytes);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? FOMO_ENOMEM : FOMO_EHIP,
                "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
  p->allocs.push_back(v);
  *ptr = (T*)v;
  return FOMO_OK;
}

static int upload(fomo_plan* p, float** dst, const std::vector<float>& h) {
  int rc = dalloc(p, dst, h.size());
  if (rc) return rc;
  HIP_TRY(hipMemcpy(*dst, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  return FOMO_OK;
}

static void free_plan(fomo_plan* p) {
  if (!p) return;
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (void* v : p->allocs) (void)hipFree(v);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// Kernel arguments of one step from state parity s = steps_done % 2.
struct Step2D {
  Sweep2DArgs v, w;
  RecordArgs r;
  int ph_v, ph_s;
};

static void step_args2d(fomo_plan* p, int s, Step2D& st) {
  const int o = 1 - s;
  Sweep2DArgs& v = st.v;
  Sweep2DArgs& w = st.w;
  v = Sweep2DArgs{};
  w = Sweep2DArgs{};
  for (int l = 0; l < p->N; ++l) v.a[l] = w.a[l] = p->a[l];
  v.dtx = w.dtx = p->dtx;
  v.q = w.q = p->q;
  v.qt = w.qt = p->qt;
  v.qxt = w.qxt = p->qxt;
  v.nx = w.nx = p->nx;
  v.nz = w.nz = p->nz;
  v.H = w.H = p->H;
  v.n_src = w.n_src = p->d.n_src;
  v.src_c = w.src_c = p->src_c;
  v.wav = w.wav = p->wav;
  v.step = w.step = p->step;
  v.nt = w.nt = p->d.nt;
  const float* lut_vx = p->wlut;
  const float* lut_vz = p->wlut + (p->H + 1);
  const float* lut_tau = p->wlut + 3 * (p->H + 1);
  RecordArgs& r = st.r;
  r = RecordArgs{};
  r.rec_c = p->rec_c;
  r.n_rec = p->d.n_rec;
  r.nt = p->d.nt;
  r.step = p->step;
  r.done = p->done;
  int& ph_v = st.ph_v;
  int& ph_s = st.ph_s;
  if (p->eq == FOMO_ACOUSTIC2D) {
    // fields: 0 p, 1 vx, 2 vz
    ph_v = kAcVel;
    ph_s = kAcPrs;
    v.rd[0] = p->fld[s][0];
    v.old[0] = p->fld[s][1];
    v.old[1] = p->fld[s][2];
    v.out[0] = p->fld[o][1];
    v.out[1] = p->fld[o][2];
    v.mat[0] = p->mat[1];  // bx
    v.mat[1] = p->mat[2];  // bz
    v.wlut[0] = lut_vx;
    v.wlut[1] = lut_vz;
    v.inj_mask = 0;
    w.rd[0] = p->fld[o][1];
    w.rd[1] = p->fld[o][2];
    w.old[0] = p->fld[s][0];
    w.out[0] = p->fld[o][0];
    w.mat[0] = p->mat[0];  // kappa
    w.wlut[0] = lut_tau;
    w.inj_mask = p->d.n_src > 0 ? 1 : 0;
    r.nf = 3;
    for (int f = 0; f < 3; ++f) {
      r.f[f] = p->fld[o][f];
      r.seis[f] = p->seis[f];
    }
  } else {
    // fields: 0 vx, 1 vz, 2 txx, 3 tzz, 4 txz
    ph_v = kElVel;
    ph_s = kElStr;
    const bool fz = p->d.source_type == FOMO_SRC_FZ;
    v.rd[0] = p->fld[s][2];
    v.rd[1] = p->fld[s][3];
    v.rd[2] = p->fld[s][4];
    v.old[0] = p->fld[s][0];
    v.old[1] = p->fld[s][1];
    v.out[0] = p->fld[o][0];
    v.out[1] = p->fld[o][1];
    v.mat[0] = p->mat[3];  // bx
    v.mat[1] = p->mat[4];  // bz
    v.wlut[0] = lut_vx;
    v.wlut[1] = lut_vz;
    v.inj_mask = (fz && p->d.n_src > 0) ? 2 : 0;
    w.rd[0] = p->fld[o][0];
    w.rd[1] = p->fld[o][1];
    for (int f = 0; f < 3; ++f) {
      w.old[f] = p->fld[s][2 + f];
      w.out[f] = p->fld[o][2 + f];
      w.wlut[f] = lut_tau;
      w.mat[f] = p->mat[f];  // lam, lam2mu, muxz
    }
    w.inj_mask = (!fz && p->d.n_src > 0) ? 3 : 0;
    r.nf = 2;
    for (int f = 0; f < 2; ++f) {
      r.f[f] = p->fld[o][f];
      r.seis[f] = p->seis[f];
    }
  }
}

// Enqueue one step.  ev (optional, 4 events) brackets the three launches for per-kernel timing.
static hipError_t enqueue_step2d(fomo_plan* p, int s, hipEvent_t* ev = nullptr) {
  Step2D st;
  step_args2d(p, s, st);
  hipError_t e = hipSuccess;
  if (ev) e = hipEventRecord(ev[0], p->stream);
  if (e == hipSuccess) e = launch_sweep2d(st.ph_v, p->N, st.v, p->stream);
  if (e == hipSuccess && ev) e = hipEventRecord(ev[1], p->stream);
  if (e == hipSuccess) e = launch_sweep2d(st.ph_s, p->N, st.w, p->stream);
  if (e == hipSuccess && ev) e = hipEventRecord(ev[2], p->stream);
  if (e == hipSuccess) e = launch_record(st.r, p->stream);
  if (e == hipSuccess && ev) e = hipEventRecord(ev[3], p->stream);
  return e;
}

static int build_graph(fomo_plan* p) {
  HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e0 = enqueue_step2d(p, 0);
  hipError_t e1 = e0 == hipSuccess ? enqueue_step2d(p, 1) : e0;
  hipGraph_t g = nullptr;
  hipError_t ec = hipStreamEndCapture(p->stream, &g);
  if (e1 != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(FOMO_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e1 != hipSuccess ? e1 : ec));
  }
  p->graph = g;
  HIP_TRY(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
  return FOMO_OK;
}

static int validate(const fomo_desc* d) {
  if (!d) return fail(FOMO_EINVAL, "descriptor is NULL");
  if (d->eq != FOMO_ACOUSTIC2D && d->eq != FOMO_ELASTIC2D)
    return fail(FOMO_EUNSUP, "equation " + std::to_string(d->eq) + " is not supported by this build");
  if (d->fd_order < 2 || d->fd_order > 10 || (d->fd_order & 1))
    return fail(FOMO_EINVAL, "Unsupported FD order: " + std::to_string(d->fd_order));
  if (d->nx < 1 || d->nz < 1 || d->ny < 1) return fail(FOMO_EINVAL, "grid dimensions must be >= 1");
  if (d->nbc < 0) return fail(FOMO_EINVAL, "nbc must be >= 0");
  if (d->nbc + d->fd_order / 2 < 2) return fail(FOMO_EINVAL, "nbc + fd_order/2 must be >= 2");
  if (d->nt < 0) return fail(FOMO_EINVAL, "nt must be >= 0");
  if (d->n_src < 0 || d->n_rec < 0) return fail(FOMO_EINVAL, "n_src and n_rec must be >= 0");
  if (d->storage != FOMO_STORE_F32) return fail(FOMO_EUNSUP, "2D solvers store Float32 only");
  if (d->snap_interval < 0) return fail(FOMO_EINVAL, "snap_interval must be >= 0");
  return FOMO_OK;
}

static int check_points(const char* what, int n, const int32_t* x, const int32_t* z, int nx, int nz) {
  for (int s = 0; s < n; ++s)
    if (x[s] < 1 || x[s] > nx || z[s] < 1 || z[s] > nz)
      return fail(FOMO_EINVAL, std::string(what) + " index out of range: need 1 <= x <= " + std::to_string(nx) +
                                   ", 1 <= z <= " + std::to_string(nz));
  return FOMO_OK;
}

extern "C" int fomo_plan_create(const fomo_desc* d, const float* vp, const float* vs, const float* rho,
                                const int32_t* sx, const int32_t* sy, const int32_t* sz, const float* wavelet,
                                const int32_t* rx, const int32_t* ry, const int32_t* rz, fomo_plan** out) {
  (void)sy;
  (void)ry;
  if (!out) return fail(FOMO_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = validate(d);
  if (rc) return rc;
  if (!vp || !rho || (d->eq == FOMO_ELASTIC2D && !vs)) return fail(FOMO_EINVAL, "model arrays must not be NULL");
  if (d->n_src > 0 && (!sx || !sz || !wavelet)) return fail(FOMO_EINVAL, "source arrays must not be NULL");
  if (d->n_rec > 0 && (!rx || !rz)) return fail(FOMO_EINVAL, "receiver arrays must not be NULL");
  if ((rc = check_points("source", d->n_src, sx, sz, d->nx, d->nz))) return rc;
  if ((rc = check_points("receiver", d->n_rec, rx, rz, d->nx, d->nz))) return rc;
  if (fomo_device_count() < 1) return fail(FOMO_EHIP, "no HIP device visible");

  fomo_plan* p = new fomo_plan();
  p->d = *d;
  p->eq = d->eq;
  p->N = d->fd_order / 2;
  p->pad = d->nbc + p->N;
  p->H = p->pad - 1;
  p->nxi = d->nx;
  p->nzi = d->nz;
  p->nx = d->nx + 2 * p->pad;
  p->nz = d->nz + 2 * p->pad;
  fomo_fd_coefficients(d->fd_order, p->a);
  p->dtx = d->dt / d->dh;
  float hc[3];
  fomo_habc_coefficients(d->v_ref, d->dt, d->dh, hc);
  p->q = hc[0];
  p->qt = hc[1];
  p->qxt = hc[2];
  hipError_t e = hipGetDevice(&p->dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    free_plan(p);
    return fail(FOMO_EHIP, std::string("stream: ") + hipGetErrorString(e));
  }
  const size_t n = (size_t)p->nx * p->nz;
  auto bail = [&](int code) {
    free_plan(p);
    return code;
  };

  // medium
  std::vector<float> vpp, rhop, bx, bz;
  pad2d(vp, d->nx, d->nz, p->pad, vpp);
  pad2d(rho, d->nx, d->nz, p->pad, rhop);
  buoyancy2d(rhop, p->nx, p->nz, bx, bz);
  if (p->eq == FOMO_ACOUSTIC2D) {
    std::vector<float> kappa(n);
    for (size_t c = 0; c < n; ++c) kappa[c] = rhop[c] * (vpp[c] * vpp[c]);
    if ((rc = upload(p, &p->mat[0], kappa)) || (rc = upload(p, &p->mat[1], bx)) || (rc = upload(p, &p->mat[2], bz)))
      return bail(rc);
    p->nfld = 3;
    p->nseis = 3;
  } else {
    std::vector<float> vsp, lam, l2m, muxz;
    pad2d(vs, d->nx, d->nz, p->pad, vsp);
    elastic_params2d(vpp, vsp, rhop, p->nx, p->nz, lam, l2m, muxz);
    if ((rc = upload(p, &p->mat[0], lam)) || (rc = upload(p, &p->mat[1], l2m)) || (rc = upload(p, &p->mat[2], muxz)) ||
        (rc = upload(p, &p->mat[3], bx)) || (rc = upload(p, &p->mat[4], bz)))
      return bail(rc);
    p->nfld = 5;
    p->nseis = 2;
  }
  // HABC tables: vx (off 0), vz (off 0.5), vy (off 0.25), tau (off 0.75)
  {
    std::vector<float> lut(4 * (size_t)(p->H + 1));
    weight_lut(p->H, p->pad, 0.0, lut.data());
    weight_lut(p->H, p->pad, 0.5, lut.data() + (p->H + 1));
    weight_lut(p->H, p->pad, 0.25, lut.data() + 2 * (p->H + 1));
    weight_lut(p->H, p->pad, 0.75, lut.data() + 3 * (p->H + 1));
    if ((rc = upload(p, &p->wlut, lut))) return bail(rc);
  }
  // acquisition
  {
    std::vector<int> sc(std::max(d->n_src, 1), -1), rcell(std::max(d->n_rec, 1), 0);
    for (int s = 0; s < d->n_src; ++s) sc[s] = (sx[s] - 1 + p->pad) + (sz[s] - 1 + p->pad) * p->nx;
    for (int r = 0; r < d->n_rec; ++r) rcell[r] = (rx[r] - 1 + p->pad) + (rz[r] - 1 + p->pad) * p->nx;
    if ((rc = dalloc(p, &p->src_c, sc.size())) || (rc = dalloc(p, &p->rec_c, rcell.size()))) return bail(rc);
    e = hipMemcpy(p->src_c, sc.data(), sc.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->rec_c, rcell.data(), rcell.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(fail(FOMO_EHIP, std::string("upload acquisition: ") + hipGetErrorString(e)));
    const size_t nw = (size_t)d->n_src * d->nt;
    if ((rc = dalloc(p, &p->wav, nw))) return bail(rc);
    if (nw) {
      e = hipMemcpy(p->wav, wavelet, nw * sizeof(float), hipMemcpyHostToDevice);
      if (e != hipSuccess) return bail(fail(FOMO_EHIP, std::string("upload wavelet: ") + hipGetErrorString(e)));
    }
    for (int f = 0; f < p->nseis; ++f)
      if ((rc = dalloc(p, &p->seis[f], (size_t)d->n_rec * d->nt))) return bail(rc);
  }
  for (int b = 0; b < 2; ++b)
    for (int f = 0; f < p->nfld; ++f)
      if ((rc = dalloc(p, &p->fld[b][f], n))) return bail(rc);
  if ((rc = dalloc(p, &p->step, 1)) || (rc = dalloc(p, &p->done, 1))) return bail(rc);
  if ((rc = fomo_plan_reset(p))) return bail(rc);
  if (d->use_graph) {
    if ((rc = build_graph(p))) return bail(rc);
  }
  e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return bail(fail(FOMO_EHIP, std::string("setup: ") + hipGetErrorString(e)));
  *out = p;
  return FOMO_OK;
}

extern "C" int fomo_plan_reset(fomo_plan* p) {
  if (!p) return fail(FOMO_EINVAL, "plan is NULL");
  const size_t n = (size_t)p->nx * p->ny * p->nz;
  for (int b = 0; b < 2; ++b)
    for (int f = 0; f < p->nfld; ++f) HIP_TRY(hipMemsetAsync(p->fld[b][f], 0, n * sizeof(float), p->stream));
  for (int f = 0; f < p->nseis; ++f)
    HIP_TRY(hipMemsetAsync(p->seis[f], 0, std::max<size_t>((size_t)p->d.n_rec * p->d.nt, 1) * sizeof(float), p->stream));
  HIP_TRY(hipMemsetAsync(p->step, 0, sizeof(int), p->stream));
  HIP_TRY(hipMemsetAsync(p->done, 0, sizeof(unsigned), p->stream));
  p->steps_done = 0;
  return FOMO_OK;
}

extern "C" int fomo_plan_advance(fomo_plan* p, int32_t n) {
  if (!p) return fail(FOMO_EINVAL, "plan is NULL");
  if (n < 0) return fail(FOMO_EINVAL, "step count must be >= 0");
  while (n > 0) {
    if (p->gexec && (p->steps_done % 2) == 0 && n >= 2) {
      HIP_TRY(hipGraphLaunch(p->gexec, p->stream));
      p->steps_done += 2;
      n -= 2;
    } else {
      HIP_TRY(enqueue_step2d(p, (int)(p->steps_done % 2)));
      p->steps_done += 1;
      n -= 1;
    }
  }
  return FOMO_OK;
}

extern "C" int fomo_plan_timed_advance(fomo_plan* p, int32_t n, double* ms) {
  if (!p) return fail(FOMO_EINVAL, "plan is NULL");
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, p->stream));
  int rc = fomo_plan_advance(p, n);
  HIP_TRY(hipEventRecord(e1, p->stream));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0.0f;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  return rc;
}

extern "C" int fomo_plan_kernel_times(fomo_plan* p, int32_t n, double* out3) {
  if (!p || !out3) return fail(FOMO_EINVAL, "NULL argument");
  if (n < 1) return fail(FOMO_EINVAL, "step count must be >= 1");
  hipEvent_t ev[4];
  for (int k = 0; k < 4; ++k) HIP_TRY(hipEventCreate(&ev[k]));
  double acc[3] = {0.0, 0.0, 0.0};
  for (int t = 0; t < n; ++t) {
    HIP_TRY(enqueue_step2d(p, (int)(p->steps_done % 2), ev));
    p->steps_done += 1;
    HIP_TRY(hipEventSynchronize(ev[3]));
    for (int k = 0; k < 3; ++k) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      acc[k] += ms;
    }
  }
  for (int k = 0; k < 4; ++k) (void)hipEventDestroy(ev[k]);
  for (int k = 0; k < 3; ++k) out3[k] = acc[k] / n;
  return FOMO_OK;
}

extern "C" int fomo_plan_sync(fomo_plan* p) {
  if (!p) return fail(FOMO_EINVAL, "plan is NULL");
  HIP_TRY(hipStreamSynchronize(p->stream));
  return FOMO_OK;
}

extern "C" int fomo_plan_steps_done(const fomo_plan* p, int32_t* out) {
  if (!p || !out) return fail(FOMO_EINVAL, "NULL argument");
  *out = (int32_t)p->steps_done;
  return FOMO_OK;
}

extern "C" int fomo_plan_read_seis(fomo_plan* p, float* s0, float* s1, float* s2, float* s3) {
  if (!p) return fail(FOMO_EINVAL, "plan is NULL");
  HIP_TRY(hipStreamSynchronize(p->stream));
  float* dst[4] = {s0, s1, s2, s3};
  const size_t bytes = (size_t)p->d.n_rec * p->d.nt * sizeof(float);
  for (int f = 0; f < p->nseis; ++f)
    if (dst[f] && bytes) HIP_TRY(hipMemcpy(dst[f], p->seis[f], bytes, hipMemcpyDeviceToHost));
  return FOMO_OK;
}

extern "C" int fomo_plan_read_field(fomo_plan* p, int32_t which, float* out) {
  if (!p || !out) return fail(FOMO_EINVAL, "NULL argument");
  if (which < 0 || which >= p->nfld) return fail(FOMO_EINVAL, "field index out of range");
  HIP_TRY(hipStreamSynchronize(p->stream));
  const size_t n = (size_t)p->nx * p->ny * p->nz;
  HIP_TRY(hipMemcpy(out, p->fld[p->steps_done % 2][which], n * sizeof(float), hipMemcpyDeviceToHost));
  return FOMO_OK;
}

extern "C" int fomo_plan_dims(const fomo_plan* p, int32_t* nx, int32_t* ny, int32_t* nz, int32_t* pad) {
  if (!p) return fail(FOMO_EINVAL, "plan is NULL");
  if (nx) *nx = p->nx;
  if (ny) *ny = p->ny;
  if (nz) *nz = p->nz;
  if (pad) *pad = p->pad;
  return FOMO_OK;
}

extern "C" int fomo_plan_destroy(fomo_plan* p) {
  if (p) {
    (void)hipStreamSynchronize(p->stream);
    free_plan(p);
  }
  return FOMO_OK;
}

// ---------------------------------------------------------------------------------------------
// One-shot drop-in runs: set up, warm up one step (as the reference does), reset, time the loop,
// copy results back.  Snapshots are taken between groups of snap_interval steps by a 2D copy of
// the interior of the current buffer.
static int run_2d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, const int32_t* sx,
                  const int32_t* sz, const float* wavelet, const int32_t* rx, const int32_t* rz, float* s0, float* s1,
                  float* s2, float* snaps, double* kernel_time_s) {
  fomo_plan* p = nullptr;
  int rc = fomo_plan_create(d, vp, vs, rho, sx, nullptr, sz, wavelet, rx, nullptr, rz, &p);
  if (rc) return rc;
  auto bail = [&](int code) {
    fomo_plan_destroy(p);
    return code;
  };
  if (d->nt > 0) {
    if ((rc = fomo_plan_advance(p, 1)) || (rc = fomo_plan_sync(p)) || (rc = fomo_plan_reset(p))) return bail(rc);
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return bail(fail(FOMO_EHIP, "hipEventCreate failed"));
  (void)hipEventRecord(e0, p->stream);
  const int snap_field = d->eq == FOMO_ACOUSTIC2D ? 0 : 1;  // p or vz
  if (d->snap_interval > 0 && snaps) {
    const int ns = d->nt / d->snap_interval;
    for (int k = 0; k < ns; ++k) {
      if ((rc = fomo_plan_advance(p, d->snap_interval))) return bail(rc);
      const float* src = p->fld[p->steps_done % 2][snap_field] + (size_t)p->pad + (size_t)p->pad * p->nx;
      hipError_t e = hipMemcpy2DAsync(snaps + (size_t)k * d->nx * d->nz, (size_t)d->nx * sizeof(float), src,
                                      (size_t)p->nx * sizeof(float), (size_t)d->nx * sizeof(float), d->nz,
                                      hipMemcpyDeviceToHost, p->stream);
      if (e != hipSuccess) return bail(fail(FOMO_EHIP, std::string("snapshot copy: ") + hipGetErrorString(e)));
    }
    if ((rc = fomo_plan_advance(p, d->nt - ns * d->snap_interval))) return bail(rc);
  } else {
    if ((rc = fomo_plan_advance(p, d->nt))) return bail(rc);
  }
  (void)hipEventRecord(e1, p->stream);
  if (hipEventSynchronize(e1) != hipSuccess) return bail(fail(FOMO_EHIP, "step loop failed"));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (kernel_time_s) *kernel_time_s = ms * 1e-3;
  rc = fomo_plan_read_seis(p, s0, s1, s2, nullptr);
  return bail(rc);
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
