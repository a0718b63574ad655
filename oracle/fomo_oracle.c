/*
 * oracle/fomo_oracle.c -- CPU restatement of Fomo.jl's staggered-grid forward-modeling time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is imported by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing else.  The product
 * path (fomo.jl_amd/) never links or calls it.
 *
 * It restates the reference's *host-visible data flow* literally -- in-place updates, the strip
 * backup into *_old arrays, and the deterministic two-pass Higdon correction through a scratch
 * vector indexed by the reference's 1-D frame map -- so that it is an independent witness for the
 * GPU path, which uses a different data flow (ping-pong buffers, HABC fused into the sweep).
 *
 * Arithmetic: REAL is float (build liboracle_f32.so) or double (liboracle_f64.so).  The float
 * build is compiled with -ffp-contract=off and keeps every expression in the reference's Julia
 * evaluation order, which makes it bit-identical to the reference's uncontracted Float32 updates.
 *
 * Indexing: 1-based (i, j[, k]) on the padded grid, x fastest, exactly as the reference's
 * column-major Julia arrays: element (i,j) lives at (i-1) + (j-1)*nx.
 *
 * Reference files followed (paths relative to the Fomo.jl root):
 *   FD coefficients ........ src/utils/FD_utils.jl:5-11
 *   padding / buoyancy ..... src/utils/pad_array.jl:8-67, src/utils/pad_array_3d.jl:8-89
 *   acoustic medium ........ src/equations/acoustic2d/medium.jl:28-55
 *   elastic medium ......... src/equations/elastic2d/medium.jl:19-74
 *   HABC coefficients ...... src/boundary/habc/habc.jl:17-44, habc_3d.jl:23-55
 *   frame map .............. src/boundary/habc/kernels_fused.jl:28-52
 *   det two-pass HABC ...... src/boundary/habc/kernels_det.jl:22-172
 *   acoustic sweeps ........ src/equations/acoustic2d/fused_kernels.jl:27-114
 *   elastic sweeps ......... src/equations/elastic2d/fused_kernels.jl:21-124
 *   step order ............. src/equations/acoustic2d/loop_graph.jl:28-39,
 *                            src/equations/elastic2d/loop_graph.jl:30-41
 *   injection / gather ..... src/acquisition/device_indexed.jl:28-134
 *   3D acoustic ............ src/equations/acoustic3d/update_velocity.jl:15-52,
 *                            update_pressure.jl:13-48, acoustic3d.jl:134-185
 *   3D HABC ................ src/boundary/habc/kernels_3d.jl:117-189 (restated two-pass: every
 *                            read is the pre-correction value, the intent stated at
 *                            src/boundary/habc/kernels_det.jl:5-10)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#endif

typedef REAL real;

/* ------------------------------------------------------------------------------------------ */
/* FD_utils.jl:5-11 -- staggered coefficients, Float32 literals (decimal -> Float64 -> Float32,
 * as Julia's Float32[...] literal does).                                                       */
static const double FD_TABLE[5][5] = {
    {1.0, 0, 0, 0, 0},
    {1.125, -0.041666667, 0, 0, 0},
    {1.171875, -0.065104167, 0.0046875, 0, 0},
    {1.1962890625, -0.079752604167, 0.0095703125, -0.000697544643, 0},
    {1.2115478515625, -0.089721679687, 0.0138427734375, -0.00176565987723, 0.0001186795166},
};

int orc_fd_coeffs(int order, float *out) {
  if (order < 2 || order > 10 || (order & 1)) return -1;
  int N = order / 2;
  for (int l = 0; l < N; ++l) out[l] = (float)FD_TABLE[N - 1][l];
  return N;
}

static int load_coeffs(int order, real *a) {
  float tmp[5];
  int N = orc_fd_coeffs(order, tmp);
  for (int l = 0; l < N; ++l) a[l] = (real)tmp[l];
  return N;
}

/* ------------------------------------------------------------------------------------------ */
/* pad_array.jl:8-22 -- edge replication.  The reference first extends along x over all rows,
 * then copies whole rows along z, which equals clamp-indexing the interior.                    */
void orc_pad2d(const float *in, int nxi, int nzi, int pad, real *out) {
  int nx = nxi + 2 * pad, nz = nzi + 2 * pad;
  for (int j = 1; j <= nz; ++j) {
    int jj = j - pad;
    if (jj < 1) jj = 1;
    if (jj > nzi) jj = nzi;
    for (int i = 1; i <= nx; ++i) {
      int ii = i - pad;
      if (ii < 1) ii = 1;
      if (ii > nxi) ii = nxi;
      out[(size_t)(i - 1) + (size_t)(j - 1) * nx] = (real)in[(size_t)(ii - 1) + (size_t)(jj - 1) * nxi];
    }
  }
}

/* pad_array.jl:30-67 harmonic buoyancy with the vacuum special cases */
static inline real harm_buoy(real r1, real r2) {
  if (r1 == (real)0 && r2 == (real)0) return (real)0;
  if (r1 == (real)0) return (real)2 / r2;
  if (r2 == (real)0) return (real)2 / r1;
  return (real)2 / (r1 + r2);
}

void orc_buoyancy2d(const real *rho, int nx, int nz, real *bx, real *bz) {
#define R(i, j) rho[(size_t)((i)-1) + (size_t)((j)-1) * nx]
  memset(bx, 0, sizeof(real) * (size_t)nx * nz);
  memset(bz, 0, sizeof(real) * (size_t)nx * nz);
  for (int j = 1; j <= nz; ++j)
    for (int i = 2; i <= nx; ++i) bx[(size_t)(i - 1) + (size_t)(j - 1) * nx] = harm_buoy(R(i - 1, j), R(i, j));
  for (int j = 1; j <= nz; ++j) bx[(size_t)(j - 1) * nx] = bx[1 + (size_t)(j - 1) * nx];
  for (int j = 1; j <= nz - 1; ++j)
    for (int i = 1; i <= nx; ++i) bz[(size_t)(i - 1) + (size_t)(j - 1) * nx] = harm_buoy(R(i, j), R(i, j + 1));
  for (int i = 1; i <= nx; ++i) bz[(size_t)(i - 1) + (size_t)(nz - 1) * nx] = bz[(size_t)(i - 1) + (size_t)(nz - 2) * nx];
#undef R
}

/* elastic2d/medium.jl:56-74 */
void orc_elastic_params2d(const real *vp, const real *vs, const real *rho, int nx, int nz, real *lam,
                          real *mu, real *muxz, real *lam2mu) {
  size_t n = (size_t)nx * nz;
  for (size_t c = 0; c < n; ++c) {
    mu[c] = rho[c] * (vs[c] * vs[c]);
    lam[c] = rho[c] * (vp[c] * vp[c]) - (real)2 * mu[c];
    lam2mu[c] = lam[c] + (real)2 * mu[c];
  }
  memset(muxz, 0, sizeof(real) * n);
#define MU(i, j) mu[(size_t)((i)-1) + (size_t)((j)-1) * nx]
  for (int j = 1; j <= nz - 1; ++j)
    for (int i = 2; i <= nx; ++i) {
      real m1 = MU(i - 1, j), m2 = MU(i, j), m3 = MU(i - 1, j + 1), m4 = MU(i, j + 1);
      real v;
      if (m1 == (real)0 || m2 == (real)0 || m3 == (real)0 || m4 == (real)0)
        v = (real)0;
      else
        v = (real)4 / ((((real)1 / m1 + (real)1 / m2) + (real)1 / m3) + (real)1 / m4);
      muxz[(size_t)(i - 1) + (size_t)(j - 1) * nx] = v;
    }
#undef MU
  for (int j = 1; j <= nz; ++j) muxz[(size_t)(j - 1) * nx] = muxz[1 + (size_t)(j - 1) * nx];
  for (int i = 1; i <= nx; ++i) muxz[(size_t)(i - 1) + (size_t)(nz - 1) * nx] = muxz[(size_t)(i - 1) + (size_t)(nz - 2) * nx];
}

/* habc.jl:21-28: r = v_ref*dt/dh; q = (b(beta+r)-r)/((beta+r)(1-b)); qt = (b(beta+r)-beta)/(...);
 * qxt = b/(b-1); b = 0.45, beta = 1.  out = {q, qt, qxt}.                                     */
void orc_habc_coeffs(float v_ref, float dt, float dh, real *out) {
  real r = ((real)v_ref * (real)dt) / (real)dh;
  real b = (real)0.45f, beta = (real)1;
  real one_m_b = (real)1 - b;
  out[0] = (b * (beta + r) - r) / ((beta + r) * one_m_b);
  out[1] = (b * (beta + r) - beta) / ((beta + r) * one_m_b);
  out[2] = b / (b - (real)1);
}

/* habc.jl:30-34: weight = clamp((dist - off)/(pad-1), 0, 1) computed in Float64, stored as Float32
 * (the fp64 oracle keeps the Float64 value).  Returned here as a function of dist.           */
static inline real habc_weight(int dist, double off, int pad) {
  double w = ((double)dist - off) / (double)(pad - 1);
  if (w < 0.0) w = 0.0;
  if (w > 1.0) w = 1.0;
  return (real)w;
}

void orc_habc_weights2d(int nx, int nz, int pad, double off, real *w) {
  for (int j = 1; j <= nz; ++j)
    for (int i = 1; i <= nx; ++i) {
      int d = i - 1;
      if (nx - i < d) d = nx - i;
      if (j - 1 < d) d = j - 1;
      if (nz - j < d) d = nz - j;
      w[(size_t)(i - 1) + (size_t)(j - 1) * nx] = habc_weight(d, off, pad);
    }
}

/* kernels_fused.jl:28-52 -- frame map and size */
long orc_frame_total(int nx, int nz, int nbc) {
  long Wd = nx - 2, Hc = nz - 2L * nbc - 2;
  return 2 * Wd * nbc + 2L * nbc * Hc;
}

void orc_frame_ij(long idx, int nx, int nz, int nbc, int *pi, int *pj) {
  long Wd = nx - 2, Nh = Wd * nbc;
  if (idx <= Nh) {
    long q = (idx - 1) / Wd, r = (idx - 1) % Wd;
    *pi = (int)(2 + r);
    *pj = (int)(2 + q);
  } else if (idx <= 2 * Nh) {
    long t = idx - Nh - 1;
    long q = t / Wd, r = t % Wd;
    *pi = (int)(2 + r);
    *pj = (int)(nz - nbc + q);
  } else {
    long t = idx - 2 * Nh - 1;
    long q = t / (2L * nbc), r = t % (2L * nbc);
    *pj = (int)(nbc + 2 + q);
    *pi = (int)(r < nbc ? 2 + r : nx - nbc + (r - nbc));
  }
}

/* kernels_det.jl:22-57 -- Higdon value from pre-correction f and f_old */
static inline real habc_value(const real *f, const real *fo, real wt, int i, int j, real q, real qt, real qxt,
                              int nx, int nz, int nbc) {
#define F(a, b) f[(size_t)((a)-1) + (size_t)((b)-1) * nx]
#define FO(a, b) fo[(size_t)((a)-1) + (size_t)((b)-1) * nx]
  int is_left = i <= nbc + 1, is_right = i >= nx - nbc;
  int is_top = j <= nbc + 1, is_bottom = j >= nz - nbc;
  int in_x = is_left | is_right, in_z = is_top | is_bottom;
  real fc = F(i, j);
  real sx = 0, sz = 0;
  if (in_x)
    sx = is_left ? ((-q * F(i + 1, j) - qt * FO(i, j)) - qxt * FO(i + 1, j))
                 : ((-q * F(i - 1, j) - qt * FO(i, j)) - qxt * FO(i - 1, j));
  if (in_z)
    sz = is_top ? ((-q * F(i, j + 1) - qt * FO(i, j)) - qxt * FO(i, j + 1))
                : ((-q * F(i, j - 1) - qt * FO(i, j)) - qxt * FO(i, j - 1));
  if (in_x && in_z) return wt * fc + (((real)1 - wt) * (real)0.5f) * (sx + sz);
  if (in_x) return wt * fc + ((real)1 - wt) * sx;
  return wt * fc + ((real)1 - wt) * sz;
#undef F
#undef FO
}

/* kernels_det.jl:60-132: pass 1 into scratch, pass 2 back into f.  nf fields share the frame. */
static void habc_det(int nf, real **f, real **fo, const real **w, real *scr, real q, real qt, real qxt, int nx,
                     int nz, int nbc) {
  long total = orc_frame_total(nx, nz, nbc);
  for (int m = 0; m < nf; ++m) {
#pragma omp parallel for schedule(static)
    for (long idx = 1; idx <= total; ++idx) {
      int i, j;
      orc_frame_ij(idx, nx, nz, nbc, &i, &j);
      real wt = w[m][(size_t)(i - 1) + (size_t)(j - 1) * nx];
      scr[(size_t)m * total + idx - 1] = habc_value(f[m], fo[m], wt, i, j, q, qt, qxt, nx, nz, nbc);
    }
  }
  for (int m = 0; m < nf; ++m) {
#pragma omp parallel for schedule(static)
    for (long idx = 1; idx <= total; ++idx) {
      int i, j;
      orc_frame_ij(idx, nx, nz, nbc, &i, &j);
      f[m][(size_t)(i - 1) + (size_t)(j - 1) * nx] = scr[(size_t)m * total + idx - 1];
    }
  }
}

/* fused_kernels.jl:39-45: strip backup condition (nbc here is the HABC width pad-1) */
static inline int in_strip(int i, int j, int nx, int nz, int nbc) {
  return (i <= nbc + 2) | (i >= nx - nbc - 1) | (j <= nbc + 2) | (j >= nz - nbc - 1);
}

static void strip_backup(int nf, real **f, real **fo, int nx, int nz, int nbc) {
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nz; ++j)
    for (int i = 1; i <= nx; ++i)
      if (in_strip(i, j, nx, nz, nbc)) {
        size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
        for (int m = 0; m < nf; ++m) fo[m][c] = f[m][c];
      }
}

/* ------------------------------------------------------------------------------------------ */
/* Acoustic sweeps, fused_kernels.jl:27-63 and :86-114 (update region M < i <= nx-M, same in j) */
static void acoustic_velocity(real *vx, real *vz, const real *p, const real *bx, const real *bz, const real *a, int N,
                              real dtx, real dtz, int nx, int nz) {
  int M = N;
#pragma omp parallel for schedule(static)
  for (int j = M + 1; j <= nz - M; ++j)
    for (int i = M + 1; i <= nx - M; ++i) {
      size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
      real dpdx = 0, dpdz = 0;
      for (int l = 1; l <= N; ++l) {
        real cl = a[l - 1];
        dpdx += cl * (p[c + (l - 1)] - p[c - l]);
        dpdz += cl * (p[c + (size_t)l * nx] - p[c - (size_t)(l - 1) * nx]);
      }
      vx[c] += (bx[c] * dtx) * dpdx;
      vz[c] += (bz[c] * dtz) * dpdz;
    }
}

static void acoustic_pressure(real *p, const real *vx, const real *vz, const real *kappa, const real *a, int N,
                              real dtx, real dtz, int nx, int nz) {
  int M = N;
#pragma omp parallel for schedule(static)
  for (int j = M + 1; j <= nz - M; ++j)
    for (int i = M + 1; i <= nx - M; ++i) {
      size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
      real dvxdx = 0, dvzdz = 0;
      for (int l = 1; l <= N; ++l) {
        real cl = a[l - 1];
        dvxdx += cl * (vx[c + l] - vx[c - (l - 1)]);
        dvzdz += cl * (vz[c + (size_t)(l - 1) * nx] - vz[c - (size_t)l * nx]);
      }
      p[c] += kappa[c] * (dvxdx * dtx + dvzdz * dtz);
    }
}

/* Elastic sweeps, elastic2d/fused_kernels.jl:21-59 and :82-124 */
static void elastic_velocity(real *vx, real *vz, const real *txx, const real *tzz, const real *txz, const real *bx,
                             const real *bz, const real *a, int N, real dtx, real dtz, int nx, int nz) {
  int M = N;
#pragma omp parallel for schedule(static)
  for (int j = M + 1; j <= nz - M; ++j)
    for (int i = M + 1; i <= nx - M; ++i) {
      size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
      real dtxxdx = 0, dtxzdz = 0, dtxzdx = 0, dtzzdz = 0;
      for (int l = 1; l <= N; ++l) {
        real cl = a[l - 1];
        dtxxdx += cl * (txx[c + (l - 1)] - txx[c - l]);
        dtxzdz += cl * (txz[c + (size_t)(l - 1) * nx] - txz[c - (size_t)l * nx]);
        dtxzdx += cl * (txz[c + l] - txz[c - (l - 1)]);
        dtzzdz += cl * (tzz[c + (size_t)l * nx] - tzz[c - (size_t)(l - 1) * nx]);
      }
      vx[c] += bx[c] * (dtx * dtxxdx + dtz * dtxzdz);
      vz[c] += bz[c] * (dtx * dtxzdx + dtz * dtzzdz);
    }
}

static void elastic_stress(real *txx, real *tzz, real *txz, const real *vx, const real *vz, const real *lam,
                           const real *lam2mu, const real *muxz, const real *a, int N, real dtx, real dtz, int nx,
                           int nz) {
  int M = N;
#pragma omp parallel for schedule(static)
  for (int j = M + 1; j <= nz - M; ++j)
    for (int i = M + 1; i <= nx - M; ++i) {
      size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
      real dvxdx = 0, dvzdz = 0, dvxdz = 0, dvzdx = 0;
      for (int l = 1; l <= N; ++l) {
        real cl = a[l - 1];
        dvxdx += cl * (vx[c + l] - vx[c - (l - 1)]);
        dvzdz += cl * (vz[c + (size_t)(l - 1) * nx] - vz[c - (size_t)l * nx]);
        dvxdz += cl * (vx[c + (size_t)l * nx] - vx[c - (size_t)(l - 1) * nx]);
        dvzdx += cl * (vz[c + (l - 1)] - vz[c - l]);
      }
      real lv = lam[c], l2m = lam2mu[c];
      txx[c] += (l2m * dvxdx) * dtx + (lv * dvzdz) * dtz;
      tzz[c] += (lv * dvxdx) * dtx + (l2m * dvzdz) * dtz;
      txz[c] += muxz[c] * (dvxdz * dtz + dvzdx * dtx);
    }
}

/* ------------------------------------------------------------------------------------------ */
static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n, sz);
  return p;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Shared geometry for 2D runs. */
typedef struct {
  int nxi, nzi, M, pad, nx, nz, H, N;
  real a[5];
  real dtx;
  real q, qt, qxt;
} geom2d;

static int geom2d_init(geom2d *g, int nxi, int nzi, int nbc, int fd_order, float dt, float dh, float v_ref,
                       int habc_width) {
  g->N = load_coeffs(fd_order, g->a);
  if (g->N < 0) return -1;
  g->M = g->N;
  g->nxi = nxi;
  g->nzi = nzi;
  g->pad = nbc + g->M;
  g->nx = nxi + 2 * g->pad;
  g->nz = nzi + 2 * g->pad;
  /* HABC frame width: pad-1 (acoustic2d.jl:127-131, elastic path); the reference's earlier
   * acoustic path used nbc, which is what example/benchmark/output/report.txt was produced with. */
  g->H = habc_width > 0 ? habc_width : g->pad - 1;
  g->dtx = (real)dt / (real)dh;
  real c[3];
  orc_habc_coeffs(v_ref, dt, dh, c);
  g->q = c[0];
  g->qt = c[1];
  g->qxt = c[2];
  return 0;
}

/* Cerjan sponge (src/boundary/sponge.jl:30-50): damp = exp(-(factor*(nbc-d))^2) for d < nbc, d the
 * distance to the nearest padded edge (1-based min(i-1, nx-i, j-1, nz-j)), 1 elsewhere.  The product
 * factor*(nbc-d) and its square are Float32 as in the reference; exp is evaluated in double and
 * rounded once (Julia's Float32 exp agrees to the last ulp, which no test can pin). */
void orc_sponge_damp2d(int nx, int nz, int nbc, float factor, real *damp) {
  for (int j = 1; j <= nz; ++j)
    for (int i = 1; i <= nx; ++i) {
      int d = i - 1;
      if (nx - i < d) d = nx - i;
      if (j - 1 < d) d = j - 1;
      if (nz - j < d) d = nz - j;
      real v = (real)1;
      if (d < nbc) {
        float r = factor * (float)(nbc - d);
        v = (real)(float)exp((double)(-r * r));
      }
      damp[(size_t)(i - 1) + (size_t)(j - 1) * nx] = v;
    }
}

/* sponge.jl:56-78: field .*= damp over the whole padded grid */
static void sponge_apply(real *f, const real *damp, size_t n) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)n; ++c) f[c] = f[c] * damp[c];
}

/*
 * acoustic2d, boundary = :habc, det two-pass path (loop_graph.jl:28-39 step order).
 *   inputs  vp, rho: nxi*nzi (x fastest); sx,sz,rx,rz: 1-based interior indices;
 *           wavelet: n_src x nt column-major (already repeated / dt-scaled by the caller).
 *   outputs seis_p, seis_vx, seis_vz: n_rec x nt column-major; snaps: n_snap x (nxi*nzi) of p.
 *   final_state (optional, 3*nx*nz): p, vx, vz after the last step.
 *   habc_width <= 0 selects the reference's current frame width pad-1.
 *   sponge_factor > 0 selects boundary = :sponge instead (acoustic2d.jl:213-228: update, then
 *   multiply vx, vz / p by the damping profile; no HABC).
 */
int orc_acoustic2d(int nxi, int nzi, const float *vp, const float *rho, float dh, float dt, int nt, int nbc,
                   int fd_order, float v_ref, int n_src, const int *sx, const int *sz, const float *wavelet, int n_rec,
                   const int *rx, const int *rz, real *seis_p, real *seis_vx, real *seis_vz, int snap_interval,
                   real *snaps, real *final_state, int habc_width, float sponge_factor, int nthreads) {
  set_threads(nthreads);
  geom2d g;
  if (geom2d_init(&g, nxi, nzi, nbc, fd_order, dt, dh, v_ref, habc_width)) return -1;
  int nx = g.nx, nz = g.nz;
  size_t n = (size_t)nx * nz;
  real *vpp = xcalloc(n, sizeof(real)), *rhop = xcalloc(n, sizeof(real));
  real *kappa = xcalloc(n, sizeof(real)), *bx = xcalloc(n, sizeof(real)), *bz = xcalloc(n, sizeof(real));
  real *wvx = xcalloc(n, sizeof(real)), *wvz = xcalloc(n, sizeof(real)), *wtau = xcalloc(n, sizeof(real));
  real *vx = xcalloc(n, sizeof(real)), *vz = xcalloc(n, sizeof(real)), *p = xcalloc(n, sizeof(real));
  real *vxo = xcalloc(n, sizeof(real)), *vzo = xcalloc(n, sizeof(real)), *po = xcalloc(n, sizeof(real));
  long total = orc_frame_total(nx, nz, g.H);
  real *scr = xcalloc((size_t)(2 * total + 1), sizeof(real));
  orc_pad2d(vp, nxi, nzi, g.pad, vpp);
  orc_pad2d(rho, nxi, nzi, g.pad, rhop);
  for (size_t c = 0; c < n; ++c) kappa[c] = rhop[c] * (vpp[c] * vpp[c]);
  orc_buoyancy2d(rhop, nx, nz, bx, bz);
  orc_habc_weights2d(nx, nz, g.pad, 0.0, wvx);
  orc_habc_weights2d(nx, nz, g.pad, 0.5, wvz);
  orc_habc_weights2d(nx, nz, g.pad, 0.75, wtau);
  const int sponge = sponge_factor > 0.0f;
  real *damp = sponge ? xcalloc(n, sizeof(real)) : NULL;
  if (sponge) orc_sponge_damp2d(nx, nz, nbc, sponge_factor, damp);
  int snap_idx = 0;
  for (int it = 1; it <= nt; ++it) {
    /* A. velocity phase: backup + update, det HABC on (vx, w_vx), (vz, w_vz) -- or the sponge */
    real *fv[2] = {vx, vz}, *fvo[2] = {vxo, vzo};
    const real *wv[2] = {wvx, wvz};
    if (!sponge) strip_backup(2, fv, fvo, nx, nz, g.H);
    acoustic_velocity(vx, vz, p, bx, bz, g.a, g.N, g.dtx, g.dtx, nx, nz);
    if (sponge) {
      sponge_apply(vx, damp, n);
      sponge_apply(vz, damp, n);
    } else {
      habc_det(2, fv, fvo, wv, scr, g.q, g.qt, g.qxt, nx, nz, g.H);
    }
    /* B. pressure phase */
    real *fp[1] = {p}, *fpo[1] = {po};
    const real *wp[1] = {wtau};
    if (!sponge) strip_backup(1, fp, fpo, nx, nz, g.H);
    acoustic_pressure(p, vx, vz, kappa, g.a, g.N, g.dtx, g.dtx, nx, nz);
    if (sponge)
      sponge_apply(p, damp, n);
    else
      habc_det(1, fp, fpo, wp, scr, g.q, g.qt, g.qxt, nx, nz, g.H);
    /* C. raw-add injection after HABC */
    for (int s = 0; s < n_src; ++s) {
      size_t c = (size_t)(sx[s] + g.pad - 1) + (size_t)(sz[s] + g.pad - 1) * nx;
      p[c] += (real)wavelet[(size_t)s + (size_t)n_src * (it - 1)];
    }
    /* D. gather */
    for (int r = 0; r < n_rec; ++r) {
      size_t c = (size_t)(rx[r] + g.pad - 1) + (size_t)(rz[r] + g.pad - 1) * nx;
      size_t o = (size_t)r + (size_t)n_rec * (it - 1);
      if (seis_p) seis_p[o] = p[c];
      if (seis_vx) seis_vx[o] = vx[c];
      if (seis_vz) seis_vz[o] = vz[c];
    }
    /* E. snapshot of the interior of p */
    if (snap_interval > 0 && it % snap_interval == 0 && snaps) {
      real *dst = snaps + (size_t)snap_idx * nxi * nzi;
      for (int j = 1; j <= nzi; ++j)
        for (int i = 1; i <= nxi; ++i)
          dst[(size_t)(i - 1) + (size_t)(j - 1) * nxi] = p[(size_t)(i + g.pad - 1) + (size_t)(j + g.pad - 1) * nx];
      ++snap_idx;
    }
  }
  if (final_state) {
    memcpy(final_state, p, n * sizeof(real));
    memcpy(final_state + n, vx, n * sizeof(real));
    memcpy(final_state + 2 * n, vz, n * sizeof(real));
  }
  free(vpp); free(rhop); free(kappa); free(bx); free(bz); free(wvx); free(wvz); free(wtau);
  free(vx); free(vz); free(p); free(vxo); free(vzo); free(po); free(scr); free(damp);
  return 0;
}

/* Centred second-derivative weights, src/utils/FD_centered.jl:31-47: f'' ~ (c0 f[i] +
 * sum_l d2_l (f[i+l] + f[i-l])) / h^2; Float64 ratios rounded to Float32 as Julia's Float32[...]
 * literal rounds them. */
static const double D2_TABLE[5][5] = {
    {1.0, 0, 0, 0, 0},
    {4.0 / 3.0, -1.0 / 12.0, 0, 0, 0},
    {3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0, 0, 0},
    {8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0, 0},
    {5.0 / 3.0, -5.0 / 21.0, 5.0 / 126.0, -5.0 / 1008.0, 1.0 / 3150.0},
};
static const double D2_CENTER[5] = {-2.0, -5.0 / 2.0, -49.0 / 18.0, -205.0 / 72.0, -5269.0 / 1800.0};

int orc_centered_d2(int order, float *c0, float *d2) {
  if (order < 2 || order > 10 || (order & 1)) return -1;
  int N = order / 2;
  *c0 = (float)D2_CENTER[N - 1];
  for (int l = 0; l < N; ++l) d2[l] = (float)D2_TABLE[N - 1][l];
  return N;
}

/*
 * scalar2d (src/equations/scalar2d/scalar2d.jl:33-81, step order of _scalar2d_loop! at :58-81;
 * the reference's own CPU step is test/verify_scalar.jl:33-86):
 *   a) u_next = (2 uc - u_prev) + c2 lap on M < i <= nx-M, M < j <= nz-M, lap = (2 c0) uc +
 *      sum_l d2_l (((u[i+l] + u[i-l]) + u[j+l]) + u[j-l]); u_next = uc elsewhere;
 *   b) two-pass Higdon on u_next with f_old = u^n and the tau weights (kernels_det.jl:22-132);
 *   c) raw-add injection of wavelet[s, it] (already scaled by the caller as source_scale asks);
 *   d) gather u_next.
 * c2 = (vp_padded * Float32(dt/dh))^2.  final_state: u^nt (nx*nz).
 */
int orc_scalar2d(int nxi, int nzi, const float *vp, float dh, float dt, int nt, int nbc, int fd_order, float v_ref,
                 int n_src, const int *sx, const int *sz, const float *wavelet, int n_rec, const int *rx,
                 const int *rz, real *seis_u, real *final_state, int nthreads) {
  set_threads(nthreads);
  geom2d g;
  if (geom2d_init(&g, nxi, nzi, nbc, fd_order, dt, dh, v_ref, 0)) return -1;
  float c0f, d2f[5];
  const int N = orc_centered_d2(fd_order, &c0f, d2f);
  if (N < 0) return -1;
  const real c0 = (real)c0f;
  real d2[5];
  for (int l = 0; l < N; ++l) d2[l] = (real)d2f[l];
  const int M = N, nx = g.nx, nz = g.nz;
  const size_t n = (size_t)nx * nz;
  real *vpp = xcalloc(n, sizeof(real)), *c2 = xcalloc(n, sizeof(real)), *wtau = xcalloc(n, sizeof(real));
  real *ua = xcalloc(n, sizeof(real)), *ub = xcalloc(n, sizeof(real)), *uc_ = xcalloc(n, sizeof(real));
  long total = orc_frame_total(nx, nz, g.H);
  real *scr = xcalloc((size_t)(total + 1), sizeof(real));
  orc_pad2d(vp, nxi, nzi, g.pad, vpp);
  for (size_t c = 0; c < n; ++c) {
    real t = vpp[c] * g.dtx;
    c2[c] = t * t;
  }
  orc_habc_weights2d(nx, nz, g.pad, 0.75, wtau);
  real *u_prev = ua, *u_cur = ub, *u_next = uc_;
  for (int it = 1; it <= nt; ++it) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= nz; ++j)
      for (int i = 1; i <= nx; ++i) {
#define U(a, b) u_cur[(size_t)((a)-1) + (size_t)((b)-1) * nx]
        const size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
        real uc = u_cur[c], un = uc;
        if (i > M && i <= nx - M && j > M && j <= nz - M) {
          real lap = (real)2 * c0 * uc;
          for (int l = 1; l <= N; ++l) lap += d2[l - 1] * (((U(i + l, j) + U(i - l, j)) + U(i, j + l)) + U(i, j - l));
          un = ((real)2 * uc - u_prev[c]) + c2[c] * lap;
        }
        u_next[c] = un;
#undef U
      }
    real *fu[1] = {u_next}, *fuo[1] = {u_cur};
    const real *wu[1] = {wtau};
    habc_det(1, fu, fuo, wu, scr, g.q, g.qt, g.qxt, nx, nz, g.H);
    for (int s = 0; s < n_src; ++s) {
      size_t c = (size_t)(sx[s] + g.pad - 1) + (size_t)(sz[s] + g.pad - 1) * nx;
      u_next[c] += (real)wavelet[(size_t)s + (size_t)n_src * (it - 1)];
    }
    for (int r = 0; r < n_rec; ++r) {
      size_t c = (size_t)(rx[r] + g.pad - 1) + (size_t)(rz[r] + g.pad - 1) * nx;
      if (seis_u) seis_u[(size_t)r + (size_t)n_rec * (it - 1)] = u_next[c];
    }
    real *t = u_prev;
    u_prev = u_cur;
    u_cur = u_next;
    u_next = t;
  }
  if (final_state) memcpy(final_state, u_cur, n * sizeof(real));
  free(vpp); free(c2); free(wtau); free(ua); free(ub); free(uc_); free(scr);
  return 0;
}

/*
 * elastic2d, boundary = :habc, det two-pass path (elastic2d/loop_graph.jl:30-41 step order).
 *   source_type 0: explosive -- wavelet added to txx and tzz (elastic2d/fused_kernels.jl:147-159);
 *   source_type 1: vertical force -- wavelet added to vz at the end of the velocity phase,
 *                  after its HABC (an extension the reference does not have).
 *   outputs seis_vx, seis_vz (n_rec x nt); snaps of vz; final_state 5*nx*nz (vx,vz,txx,tzz,txz).
 *   sponge_factor > 0 selects boundary = :sponge (elastic2d.jl:180-195: vx, vz and txx, tzz, txz
 *   multiplied by the damping profile after their updates).
 */
int orc_elastic2d(int nxi, int nzi, const float *vp, const float *vs, const float *rho, float dh, float dt, int nt,
                  int nbc, int fd_order, float v_ref, int source_type, int n_src, const int *sx, const int *sz,
                  const float *wavelet, int n_rec, const int *rx, const int *rz, real *seis_vx, real *seis_vz,
                  int snap_interval, real *snaps, real *final_state, int habc_width, float sponge_factor,
                  int nthreads) {
  set_threads(nthreads);
  geom2d g;
  if (geom2d_init(&g, nxi, nzi, nbc, fd_order, dt, dh, v_ref, habc_width)) return -1;
  int nx = g.nx, nz = g.nz;
  size_t n = (size_t)nx * nz;
  real *vpp = xcalloc(n, sizeof(real)), *vsp = xcalloc(n, sizeof(real)), *rhop = xcalloc(n, sizeof(real));
  real *lam = xcalloc(n, sizeof(real)), *mu = xcalloc(n, sizeof(real)), *muxz = xcalloc(n, sizeof(real));
  real *l2m = xcalloc(n, sizeof(real)), *bx = xcalloc(n, sizeof(real)), *bz = xcalloc(n, sizeof(real));
  real *wvx = xcalloc(n, sizeof(real)), *wvz = xcalloc(n, sizeof(real)), *wtau = xcalloc(n, sizeof(real));
  real *vx = xcalloc(n, sizeof(real)), *vz = xcalloc(n, sizeof(real));
  real *txx = xcalloc(n, sizeof(real)), *tzz = xcalloc(n, sizeof(real)), *txz = xcalloc(n, sizeof(real));
  real *vxo = xcalloc(n, sizeof(real)), *vzo = xcalloc(n, sizeof(real));
  real *txxo = xcalloc(n, sizeof(real)), *tzzo = xcalloc(n, sizeof(real)), *txzo = xcalloc(n, sizeof(real));
  long total = orc_frame_total(nx, nz, g.H);
  real *scr = xcalloc((size_t)(3 * total + 1), sizeof(real));
  orc_pad2d(vp, nxi, nzi, g.pad, vpp);
  orc_pad2d(vs, nxi, nzi, g.pad, vsp);
  orc_pad2d(rho, nxi, nzi, g.pad, rhop);
  orc_buoyancy2d(rhop, nx, nz, bx, bz);
  orc_elastic_params2d(vpp, vsp, rhop, nx, nz, lam, mu, muxz, l2m);
  orc_habc_weights2d(nx, nz, g.pad, 0.0, wvx);
  orc_habc_weights2d(nx, nz, g.pad, 0.5, wvz);
  orc_habc_weights2d(nx, nz, g.pad, 0.75, wtau);
  const int sponge = sponge_factor > 0.0f;
  real *damp = sponge ? xcalloc(n, sizeof(real)) : NULL;
  if (sponge) orc_sponge_damp2d(nx, nz, nbc, sponge_factor, damp);
  int snap_idx = 0;
  for (int it = 1; it <= nt; ++it) {
    real *fv[2] = {vx, vz}, *fvo[2] = {vxo, vzo};
    const real *wv[2] = {wvx, wvz};
    if (!sponge) strip_backup(2, fv, fvo, nx, nz, g.H);
    elastic_velocity(vx, vz, txx, tzz, txz, bx, bz, g.a, g.N, g.dtx, g.dtx, nx, nz);
    if (sponge) {
      sponge_apply(vx, damp, n);
      sponge_apply(vz, damp, n);
    } else {
      habc_det(2, fv, fvo, wv, scr, g.q, g.qt, g.qxt, nx, nz, g.H);
    }
    /* vertical force (extension): raw add into vz at the end of the velocity phase */
    if (source_type == 1)
      for (int s = 0; s < n_src; ++s) {
        size_t c = (size_t)(sx[s] + g.pad - 1) + (size_t)(sz[s] + g.pad - 1) * nx;
        vz[c] += (real)wavelet[(size_t)s + (size_t)n_src * (it - 1)];
      }
    real *fs[3] = {txx, tzz, txz}, *fso[3] = {txxo, tzzo, txzo};
    const real *ws[3] = {wtau, wtau, wtau};
    if (!sponge) strip_backup(3, fs, fso, nx, nz, g.H);
    elastic_stress(txx, tzz, txz, vx, vz, lam, l2m, muxz, g.a, g.N, g.dtx, g.dtx, nx, nz);
    if (sponge) {
      sponge_apply(txx, damp, n);
      sponge_apply(tzz, damp, n);
      sponge_apply(txz, damp, n);
    } else {
      habc_det(3, fs, fso, ws, scr, g.q, g.qt, g.qxt, nx, nz, g.H);
    }
    for (int s = 0; s < n_src; ++s) {
      size_t c = (size_t)(sx[s] + g.pad - 1) + (size_t)(sz[s] + g.pad - 1) * nx;
      real wav = (real)wavelet[(size_t)s + (size_t)n_src * (it - 1)];
      if (source_type != 1) {
        txx[c] += wav;
        tzz[c] += wav;
      }
    }
    for (int r = 0; r < n_rec; ++r) {
      size_t c = (size_t)(rx[r] + g.pad - 1) + (size_t)(rz[r] + g.pad - 1) * nx;
      size_t o = (size_t)r + (size_t)n_rec * (it - 1);
      if (seis_vx) seis_vx[o] = vx[c];
      if (seis_vz) seis_vz[o] = vz[c];
    }
    if (snap_interval > 0 && it % snap_interval == 0 && snaps) {
      real *dst = snaps + (size_t)snap_idx * nxi * nzi;
      for (int j = 1; j <= nzi; ++j)
        for (int i = 1; i <= nxi; ++i)
          dst[(size_t)(i - 1) + (size_t)(j - 1) * nxi] = vz[(size_t)(i + g.pad - 1) + (size_t)(j + g.pad - 1) * nx];
      ++snap_idx;
    }
  }
  if (final_state) {
    memcpy(final_state, vx, n * sizeof(real));
    memcpy(final_state + n, vz, n * sizeof(real));
    memcpy(final_state + 2 * n, txx, n * sizeof(real));
    memcpy(final_state + 3 * n, tzz, n * sizeof(real));
    memcpy(final_state + 4 * n, txz, n * sizeof(real));
  }
  free(vpp); free(vsp); free(rhop); free(lam); free(mu); free(muxz); free(l2m); free(bx); free(bz);
  free(wvx); free(wvz); free(wtau); free(vx); free(vz); free(txx); free(tzz); free(txz);
  free(vxo); free(vzo); free(txxo); free(tzzo); free(txzo); free(scr); free(damp);
  return 0;
}

/* Single-sweep probes used by the known-answer tests: one velocity update of the acoustic
 * equations on a caller-given padded state (no HABC), and one HABC two-pass correction. */
void orc_probe_acoustic_velocity(int nx, int nz, int fd_order, real dtx, real *vx, real *vz, const real *p,
                                 const real *bx, const real *bz) {
  real a[5];
  int N = load_coeffs(fd_order, a);
  acoustic_velocity(vx, vz, p, bx, bz, a, N, dtx, dtx, nx, nz);
}

void orc_probe_acoustic_pressure(int nx, int nz, int fd_order, real dtx, real *p, const real *vx, const real *vz,
                                 const real *kappa) {
  real a[5];
  int N = load_coeffs(fd_order, a);
  acoustic_pressure(p, vx, vz, kappa, a, N, dtx, dtx, nx, nz);
}

/* One elastic velocity sweep and one elastic stress sweep (the functions the elastic2d time loop
 * calls) on a caller-given padded state, for the plane-wave and Hooke's-law known-answer tests. */
void orc_probe_elastic_velocity(int nx, int nz, int fd_order, real dtx, real *vx, real *vz, const real *txx,
                                const real *tzz, const real *txz, const real *bx, const real *bz) {
  real a[5];
  int N = load_coeffs(fd_order, a);
  elastic_velocity(vx, vz, txx, tzz, txz, bx, bz, a, N, dtx, dtx, nx, nz);
}

void orc_probe_elastic_stress(int nx, int nz, int fd_order, real dtx, real *txx, real *tzz, real *txz,
                              const real *vx, const real *vz, const real *lam, const real *lam2mu,
                              const real *muxz) {
  real a[5];
  int N = load_coeffs(fd_order, a);
  elastic_stress(txx, tzz, txz, vx, vz, lam, lam2mu, muxz, a, N, dtx, dtx, nx, nz);
}

void orc_probe_habc(int nx, int nz, int nbc, real q, real qt, real qxt, real *f, real *fo, const real *w) {
  long total = orc_frame_total(nx, nz, nbc);
  real *scr = xcalloc((size_t)total + 1, sizeof(real));
  real *ff[1] = {f}, *ffo[1] = {fo};
  const real *ww[1] = {w};
  habc_det(1, ff, ffo, ww, scr, q, qt, qxt, nx, nz, nbc);
  free(scr);
}

/* ------------------------------------------------------------------------------------------ */
/* coupled2d: the coupled P-S potential solver (src/equations/coupled2d/).                      */

/* Centred first-derivative weights, src/utils/FD_centered.jl:14-20: f' ~ sum_l d1_l (f[i+l] -
 * f[i-l]) / h; Float64 ratios rounded to Float32 as Julia's Float32[...] literal rounds them. */
static const double D1_TABLE[5][5] = {
    {1.0 / 2.0, 0, 0, 0, 0},
    {2.0 / 3.0, -1.0 / 12.0, 0, 0, 0},
    {3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0, 0, 0},
    {4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0, 0},
    {5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0, -5.0 / 504.0, 1.0 / 1260.0},
};

int orc_centered_d1(int order, float *d1) {
  if (order < 2 || order > 10 || (order & 1)) return -1;
  int N = order / 2;
  for (int l = 0; l < N; ++l) d1[l] = (float)D1_TABLE[N - 1][l];
  return N;
}

/* medium.jl:118-150 (_gaussian_smooth): separable Gaussian of radius ceil(3 sigma), taps
 * exp(-0.5 (k/sigma)^2) in Float64 rounded to the field type, normalised by their sum (summed
 * first-to-last, the order of Julia's sequential reduction for so few terms), x pass then z pass,
 * indices clamped to the grid.  tmp: nx*nz scratch. */
static void coupled_smooth(const real *f, real *out, real *tmp, int nx, int nz, double sigma) {
  int R = (int)ceil(3.0 * sigma);
  int nk = 2 * R + 1;
  real *ker = xcalloc((size_t)nk, sizeof(real));
  for (int k = -R; k <= R; ++k) {
    double t = (double)k / sigma;
    ker[k + R] = (real)exp(-0.5 * (t * t));
  }
  real s = ker[0];
  if (nk > 1) {
    s = ker[0] + ker[1];
    for (int k = 2; k < nk; ++k) s += ker[k];
  }
  for (int k = 0; k < nk; ++k) ker[k] = ker[k] / s;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nz; ++j)
    for (int i = 1; i <= nx; ++i) {
      real v = (real)0;
      for (int k = -R; k <= R; ++k) {
        int ii = i + k < 1 ? 1 : (i + k > nx ? nx : i + k);
        v += ker[k + R] * f[(size_t)(ii - 1) + (size_t)(j - 1) * nx];
      }
      tmp[(size_t)(i - 1) + (size_t)(j - 1) * nx] = v;
    }
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nz; ++j)
    for (int i = 1; i <= nx; ++i) {
      real v = (real)0;
      for (int k = -R; k <= R; ++k) {
        int jj = j + k < 1 ? 1 : (j + k > nz ? nz : j + k);
        v += ker[k + R] * tmp[(size_t)(i - 1) + (size_t)(jj - 1) * nx];
      }
      out[(size_t)(i - 1) + (size_t)(j - 1) * nx] = v;
    }
  free(ker);
}

/* medium.jl:156-193: x-gradient on M < i <= nx-M (every j), z-gradient on M < j <= nz-M (every
 * i), Laplacian on both ranges, each divided by h (gradients) or scaled by 1/(h h) (Laplacian);
 * zero elsewhere. */
static void coupled_derivs(const real *f, const real *d1, const real *d2, real c0, int N, real h, int nx, int nz,
                           real *gx, real *gz, real *lap) {
  const real ih2 = (real)1 / (h * h);
#define FF(a, b) f[(size_t)((a)-1) + (size_t)((b)-1) * nx]
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= nz; ++j)
    for (int i = 1; i <= nx; ++i) {
      const size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
      const int ix = i > N && i <= nx - N, iz = j > N && j <= nz - N;
      real vx = 0, vz = 0, vl = 0;
      if (ix) {
        real v = 0;
        for (int l = 1; l <= N; ++l) v += d1[l - 1] * (FF(i + l, j) - FF(i - l, j));
        vx = v / h;
      }
      if (iz) {
        real v = 0;
        for (int l = 1; l <= N; ++l) v += d1[l - 1] * (FF(i, j + l) - FF(i, j - l));
        vz = v / h;
      }
      if (ix && iz) {
        real ax = c0 * FF(i, j), az = c0 * FF(i, j);
        for (int l = 1; l <= N; ++l) ax += d2[l - 1] * (FF(i + l, j) + FF(i - l, j));
        for (int l = 1; l <= N; ++l) az += d2[l - 1] * (FF(i, j + l) + FF(i, j - l));
        vl = (ax + az) * ih2;
      }
      gx[c] = vx;
      gz[c] = vz;
      lap[c] = vl;
    }
#undef FF
}

/* init_coupled_medium, medium.jl:50-117, on the padded grid (nx, nz).  out: 8 arrays of nx*nz in
 * the order alpha = vp^2, beta = vs^2 (raw), then d alpha/dx, d alpha/dz, d beta/dx, d beta/dz,
 * lap alpha, lap beta of the smoothed alpha, beta (sigma <= 0: unsmoothed). */
static int coupled_medium(const real *vpp, const real *vsp, int nx, int nz, int fd_order, float dh, double sigma,
                          real *out) {
  float d1f[5], d2f[5], c0f;
  int N = orc_centered_d1(fd_order, d1f);
  if (N < 0 || orc_centered_d2(fd_order, &c0f, d2f) < 0) return -1;
  real d1[5], d2[5];
  for (int l = 0; l < N; ++l) {
    d1[l] = (real)d1f[l];
    d2[l] = (real)d2f[l];
  }
  const size_t n = (size_t)nx * nz;
  real *alpha = out, *beta = out + n;
  for (size_t c = 0; c < n; ++c) {
    alpha[c] = vpp[c] * vpp[c];
    beta[c] = vsp[c] * vsp[c];
  }
  real *as = xcalloc(n, sizeof(real)), *bs = xcalloc(n, sizeof(real)), *tmp = xcalloc(n, sizeof(real));
  if (sigma > 0) {
    coupled_smooth(alpha, as, tmp, nx, nz, sigma);
    coupled_smooth(beta, bs, tmp, nx, nz, sigma);
  } else {
    memcpy(as, alpha, n * sizeof(real));
    memcpy(bs, beta, n * sizeof(real));
  }
  const real h = (real)dh;
  coupled_derivs(as, d1, d2, (real)c0f, N, h, nx, nz, out + 2 * n, out + 3 * n, out + 6 * n);
  coupled_derivs(bs, d1, d2, (real)c0f, N, h, nx, nz, out + 4 * n, out + 5 * n, out + 7 * n);
  free(as); free(bs); free(tmp);
  return 0;
}

int orc_coupled_medium(int nxi, int nzi, const float *vp, const float *vs, float dh, int nbc, int fd_order,
                       double sigma, real *out) {
  if (fd_order < 2 || fd_order > 10 || (fd_order & 1)) return -1;
  const int pad = nbc + fd_order / 2, nx = nxi + 2 * pad, nz = nzi + 2 * pad;
  const size_t n = (size_t)nx * nz;
  real *vpp = xcalloc(n, sizeof(real)), *vsp = xcalloc(n, sizeof(real));
  orc_pad2d(vp, nxi, nzi, pad, vpp);
  orc_pad2d(vs, nxi, nzi, pad, vsp);
  int rc = coupled_medium(vpp, vsp, nx, nz, fd_order, dh, sigma, out);
  free(vpp); free(vsp);
  return rc;
}

/* update_fields.jl:27-112: the "velocity" half-step dPdt += dt rhs_P, dSdt += dt rhs_S on
 * M < i <= nx-M, M < j <= nz-M (in place: it reads only P, S and the cell's own rates).
 * med: the 8 arrays of coupled_medium. */
static void coupled_velocity(real *dPdt, real *dSdt, const real *P, const real *S, const real *med, const real *d1,
                             const real *d2, real c0, int N, real dt, real ih, real ih2, int nx, int nz) {
  const size_t n = (size_t)nx * nz;
  const real *al = med, *be = med + n, *dax = med + 2 * n, *daz = med + 3 * n, *dbx = med + 4 * n,
             *dbz = med + 5 * n, *la = med + 6 * n, *lb = med + 7 * n;
#define PP(a, b) P[(size_t)((a)-1) + (size_t)((b)-1) * nx]
#define SS(a, b) S[(size_t)((a)-1) + (size_t)((b)-1) * nx]
#pragma omp parallel for schedule(static)
  for (int j = N + 1; j <= nz - N; ++j)
    for (int i = N + 1; i <= nx - N; ++i) {
      const size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
      real dPdx = 0, dPdz = 0, lPx = c0 * PP(i, j), lPz = c0 * PP(i, j);
      real dSdx = 0, dSdz = 0, lSx = c0 * SS(i, j), lSz = c0 * SS(i, j);
      for (int l = 1; l <= N; ++l) {
        const real c1 = d1[l - 1], c2 = d2[l - 1];
        dPdx += c1 * (PP(i + l, j) - PP(i - l, j));
        dPdz += c1 * (PP(i, j + l) - PP(i, j - l));
        lPx += c2 * (PP(i + l, j) + PP(i - l, j));
        lPz += c2 * (PP(i, j + l) + PP(i, j - l));
        dSdx += c1 * (SS(i + l, j) - SS(i - l, j));
        dSdz += c1 * (SS(i, j + l) - SS(i, j - l));
        lSx += c2 * (SS(i + l, j) + SS(i - l, j));
        lSz += c2 * (SS(i, j + l) + SS(i, j - l));
      }
      dPdx = dPdx * ih;
      dPdz = dPdz * ih;
      dSdx = dSdx * ih;
      dSdz = dSdz * ih;
      const real lapP = (lPx + lPz) * ih2, lapS = (lSx + lSz) * ih2;
      const real Pv = PP(i, j);
      const real rP = (((Pv * la[c] + (real)2 * (dax[c] * dPdx + daz[c] * dPdz)) - ((real)2 * Pv) * lb[c]) +
                       (real)2 * (dbx[c] * dSdz - dbz[c] * dSdx)) +
                      al[c] * lapP;
      const real rS = ((real)2 * (dbx[c] * dSdx + dbz[c] * dSdz) + (real)2 * (dbz[c] * dPdx - dbx[c] * dPdz)) +
                      be[c] * lapS;
      dPdt[c] = dPdt[c] + dt * rP;
      dSdt[c] = dSdt[c] + dt * rS;
    }
#undef PP
#undef SS
}

/* update_fields.jl:124-144: the "position" half-step P += dt dPdt, S += dt dSdt on the same
 * region. */
static void coupled_position(real *P, real *S, const real *dPdt, const real *dSdt, real dt, int N, int nx, int nz) {
#pragma omp parallel for schedule(static)
  for (int j = N + 1; j <= nz - N; ++j)
    for (int i = N + 1; i <= nx - N; ++i) {
      const size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
      P[c] = P[c] + dt * dPdt[c];
      S[c] = S[c] + dt * dSdt[c];
    }
}

/* One velocity half-step and one position half-step on a caller-given padded state (no HABC,
 * no source), for the known-answer tests of the update equations. */
int orc_probe_coupled_velocity(int nx, int nz, int fd_order, float dt, float dh, real *dPdt, real *dSdt,
                               const real *P, const real *S, const real *med) {
  float d1f[5], d2f[5], c0f;
  int N = orc_centered_d1(fd_order, d1f);
  if (N < 0 || orc_centered_d2(fd_order, &c0f, d2f) < 0) return -1;
  real d1[5], d2[5];
  for (int l = 0; l < N; ++l) {
    d1[l] = (real)d1f[l];
    d2[l] = (real)d2f[l];
  }
  const real ih = (real)1 / (real)dh;
  coupled_velocity(dPdt, dSdt, P, S, med, d1, d2, (real)c0f, N, (real)dt, ih, ih * ih, nx, nz);
  return 0;
}

void orc_probe_coupled_position(int nx, int nz, int fd_order, float dt, real *P, real *S, const real *dPdt,
                                const real *dSdt) {
  coupled_position(P, S, dPdt, dSdt, (real)dt, fd_order / 2, nx, nz);
}

/*
 * coupled2d (src/equations/coupled2d/coupled2d.jl:53-260; step order of _coupled2d_loop! at
 * :197-260):
 *   A) backup dPdt, dSdt; velocity half-step; Higdon on dPdt (P coefficients, v_ref_p) and on
 *      dSdt (S coefficients, v_ref_s), both with the vx weights;
 *   B) backup P, S; position half-step; Higdon on P and S with the tau weights;
 *   C) raw add of wavelet[s, it] into P;  D) gather P and S;  E) snapshots of P and S.
 * The reference applies its in-place (racy) single-field Higdon kernel here
 * (src/boundary/habc/kernels.jl:122-190); this restatement uses the deterministic two-pass form
 * (every read is the pre-correction value), the form src/boundary/habc/kernels_det.jl:5-10
 * states as the intended semantics.
 *   outputs seis_P, seis_S (n_rec x nt); snaps_P, snaps_S (n_snap x nxi*nzi); final_state
 *   4*nx*nz (P, S, dPdt, dSdt).
 */
int orc_coupled2d(int nxi, int nzi, const float *vp, const float *vs, float dh, float dt, int nt, int nbc,
                  int fd_order, float v_ref_p, float v_ref_s, double sigma, int n_src, const int *sx, const int *sz,
                  const float *wavelet, int n_rec, const int *rx, const int *rz, int snap_interval, real *seis_P,
                  real *seis_S, real *snaps_P, real *snaps_S, real *final_state, int nthreads) {
  set_threads(nthreads);
  geom2d g;
  if (geom2d_init(&g, nxi, nzi, nbc, fd_order, dt, dh, v_ref_p, 0)) return -1;
  float d1f[5], d2f[5], c0f;
  const int N = orc_centered_d1(fd_order, d1f);
  if (N < 0 || orc_centered_d2(fd_order, &c0f, d2f) < 0) return -1;
  real d1[5], d2[5];
  for (int l = 0; l < N; ++l) {
    d1[l] = (real)d1f[l];
    d2[l] = (real)d2f[l];
  }
  real qs[3];
  orc_habc_coeffs(v_ref_s, dt, dh, qs);
  const int nx = g.nx, nz = g.nz;
  const size_t n = (size_t)nx * nz;
  real *med = xcalloc(8 * n, sizeof(real));
  if (orc_coupled_medium(nxi, nzi, vp, vs, dh, nbc, fd_order, sigma, med)) {
    free(med);
    return -1;
  }
  real *wvx = xcalloc(n, sizeof(real)), *wtau = xcalloc(n, sizeof(real));
  orc_habc_weights2d(nx, nz, g.pad, 0.0, wvx);
  orc_habc_weights2d(nx, nz, g.pad, 0.75, wtau);
  real *P = xcalloc(n, sizeof(real)), *S = xcalloc(n, sizeof(real));
  real *dP = xcalloc(n, sizeof(real)), *dS = xcalloc(n, sizeof(real));
  real *Po = xcalloc(n, sizeof(real)), *So = xcalloc(n, sizeof(real));
  real *dPo = xcalloc(n, sizeof(real)), *dSo = xcalloc(n, sizeof(real));
  long total = orc_frame_total(nx, nz, g.H);
  real *scr = xcalloc((size_t)(total + 1), sizeof(real));
  const real rdt = (real)dt, ih = (real)1 / (real)dh;
  int snap_idx = 0;
  for (int it = 1; it <= nt; ++it) {
    real *fv[2] = {dP, dS}, *fvo[2] = {dPo, dSo};
    strip_backup(2, fv, fvo, nx, nz, g.H);
    coupled_velocity(dP, dS, P, S, med, d1, d2, (real)c0f, N, rdt, ih, ih * ih, nx, nz);
    {
      real *f1[1] = {dP}, *f1o[1] = {dPo}, *f2[1] = {dS}, *f2o[1] = {dSo};
      const real *w1[1] = {wvx};
      habc_det(1, f1, f1o, w1, scr, g.q, g.qt, g.qxt, nx, nz, g.H);
      habc_det(1, f2, f2o, w1, scr, qs[0], qs[1], qs[2], nx, nz, g.H);
    }
    real *fp[2] = {P, S}, *fpo[2] = {Po, So};
    strip_backup(2, fp, fpo, nx, nz, g.H);
    coupled_position(P, S, dP, dS, rdt, N, nx, nz);
    {
      real *f1[1] = {P}, *f1o[1] = {Po}, *f2[1] = {S}, *f2o[1] = {So};
      const real *w1[1] = {wtau};
      habc_det(1, f1, f1o, w1, scr, g.q, g.qt, g.qxt, nx, nz, g.H);
      habc_det(1, f2, f2o, w1, scr, qs[0], qs[1], qs[2], nx, nz, g.H);
    }
    for (int s = 0; s < n_src; ++s) {
      size_t c = (size_t)(sx[s] + g.pad - 1) + (size_t)(sz[s] + g.pad - 1) * nx;
      P[c] += (real)wavelet[(size_t)s + (size_t)n_src * (it - 1)];
    }
    for (int r = 0; r < n_rec; ++r) {
      size_t c = (size_t)(rx[r] + g.pad - 1) + (size_t)(rz[r] + g.pad - 1) * nx;
      size_t o = (size_t)r + (size_t)n_rec * (it - 1);
      if (seis_P) seis_P[o] = P[c];
      if (seis_S) seis_S[o] = S[c];
    }
    if (snap_interval > 0 && it % snap_interval == 0) {
      for (int j = 1; j <= nzi; ++j)
        for (int i = 1; i <= nxi; ++i) {
          size_t c = (size_t)(i + g.pad - 1) + (size_t)(j + g.pad - 1) * nx;
          size_t o = (size_t)snap_idx * nxi * nzi + (size_t)(i - 1) + (size_t)(j - 1) * nxi;
          if (snaps_P) snaps_P[o] = P[c];
          if (snaps_S) snaps_S[o] = S[c];
        }
      ++snap_idx;
    }
  }
  if (final_state) {
    memcpy(final_state, P, n * sizeof(real));
    memcpy(final_state + n, S, n * sizeof(real));
    memcpy(final_state + 2 * n, dP, n * sizeof(real));
    memcpy(final_state + 3 * n, dS, n * sizeof(real));
  }
  free(med); free(wvx); free(wtau); free(P); free(S); free(dP); free(dS);
  free(Po); free(So); free(dPo); free(dSo); free(scr);
  return 0;
}
