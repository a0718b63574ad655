/*
 * oracle/fomo_oracle3d.c -- CPU restatement of Fomo.jl's acoustic3d time step.
 *
 * TEST INFRASTRUCTURE ONLY (see fomo_oracle.c).  Same conventions: 1-based (i,j,k) on the padded
 * cube, x fastest, REAL = float (-ffp-contract=off) or double.
 *
 * Reference files followed:
 *   padding / buoyancy ....... src/utils/pad_array_3d.jl:8-89
 *   medium ................... src/equations/acoustic3d/medium.jl:40-62
 *   HABC coefficients/weights  src/boundary/habc/habc_3d.jl:23-55
 *   strip backup ............. src/boundary/habc/kernels_3d.jl:15-27
 *   velocity / pressure ...... src/equations/acoustic3d/update_velocity.jl:15-52,
 *                              src/equations/acoustic3d/update_pressure.jl:13-48
 *   HABC ..................... src/boundary/habc/kernels_3d.jl:117-189.  The reference kernel
 *                              updates f in place while neighbours read it (a race); this
 *                              restatement reads only pre-correction values (two-pass), the
 *                              deterministic form the reference adopts in 2D
 *                              (src/boundary/habc/kernels_det.jl:5-10).
 *   step order ............... src/equations/acoustic3d/acoustic3d.jl:134-185
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

int orc_fd_coeffs(int order, float *out);
void orc_habc_coeffs(float v_ref, float dt, float dh, real *out);

#define ID3(i, j, k) ((size_t)((i)-1) + (size_t)((j)-1) * (size_t)nx + (size_t)((k)-1) * (size_t)nx * (size_t)ny)

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_pad3d(const float *in, int nxi, int nyi, int nzi, int pad, real *out) {
  int nx = nxi + 2 * pad, ny = nyi + 2 * pad, nz = nzi + 2 * pad;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        int ii = clampi(i - pad, 1, nxi), jj = clampi(j - pad, 1, nyi), kk = clampi(k - pad, 1, nzi);
        out[ID3(i, j, k)] = (real)in[(size_t)(ii - 1) + (size_t)(jj - 1) * nxi + (size_t)(kk - 1) * nxi * nyi];
      }
}

static inline real hb(real r1, real r2) {
  if (r1 == (real)0 && r2 == (real)0) return (real)0;
  if (r1 == (real)0) return (real)2 / r2;
  if (r2 == (real)0) return (real)2 / r1;
  return (real)2 / (r1 + r2);
}

/* pad_array_3d.jl:37-89 */
void orc_buoyancy3d(const real *rho, int nx, int ny, int nz, real *bx, real *by, real *bz) {
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        size_t c = ID3(i, j, k);
        bx[c] = i >= 2 ? hb(rho[ID3(i - 1, j, k)], rho[c]) : (real)0;
        by[c] = j >= 2 ? hb(rho[ID3(i, j - 1, k)], rho[c]) : (real)0;
        bz[c] = k <= nz - 1 ? hb(rho[c], rho[ID3(i, j, k + 1)]) : (real)0;
      }
  /* edge copies in the reference's order: buoy_vx[1,:,:] = buoy_vx[2,:,:], etc. */
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) bx[ID3(1, j, k)] = bx[ID3(2, j, k)];
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) by[ID3(i, 1, k)] = by[ID3(i, 2, k)];
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) bz[ID3(i, j, nz)] = bz[ID3(i, j, nz - 1)];
}

static inline real w3(int i, int j, int k, int nx, int ny, int nz, int pad, double off) {
  int d = i - 1;
  if (nx - i < d) d = nx - i;
  if (j - 1 < d) d = j - 1;
  if (ny - j < d) d = ny - j;
  if (k - 1 < d) d = k - 1;
  if (nz - k < d) d = nz - k;
  double w = ((double)d - off) / (double)(pad - 1);
  if (w < 0.0) w = 0.0;
  if (w > 1.0) w = 1.0;
  return (real)w;
}

static inline int strip3(int i, int j, int k, int nx, int ny, int nz, int nbc) {
  return i <= nbc + 2 || i >= nx - nbc - 1 || j <= nbc + 2 || j >= ny - nbc - 1 || k <= nbc + 2 || k >= nz - nbc - 1;
}

/* Two-pass 3D Higdon: fpre is the pre-correction copy, fo the strip backup, f receives results. */
static void habc3d(real *f, const real *fpre, const real *fo, int nx, int ny, int nz, int nbc, int pad, double off,
                   real q, real qt, real qxt) {
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        int il = i <= nbc + 1, ir = i >= nx - nbc, jf = j <= nbc + 1, jb = j >= ny - nbc;
        int kt = k <= nbc + 1, kb = k >= nz - nbc;
        int in_x = il | ir, in_y = jf | jb, in_z = kt | kb;
        if (!(in_x | in_y | in_z)) continue;
        size_t c = ID3(i, j, k);
        real wt = w3(i, j, k, nx, ny, nz, pad, off);
        real fc = fpre[c];
        real s = (real)0;
        int nd = 0;
        if (in_x) {
          int ii = il ? i + 1 : i - 1;
          s += (-q * fpre[ID3(ii, j, k)] - qt * fo[c]) - qxt * fo[ID3(ii, j, k)];
          ++nd;
        }
        if (in_y) {
          int jj = jf ? j + 1 : j - 1;
          s += (-q * fpre[ID3(i, jj, k)] - qt * fo[c]) - qxt * fo[ID3(i, jj, k)];
          ++nd;
        }
        if (in_z) {
          int kk = kt ? k + 1 : k - 1;
          s += (-q * fpre[ID3(i, j, kk)] - qt * fo[c]) - qxt * fo[ID3(i, j, kk)];
          ++nd;
        }
        f[c] = wt * fc + (((real)1 - wt) * s) / (real)nd;
      }
}

static void backup3(real *fo, const real *f, int nx, int ny, int nz, int nbc) {
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        if (strip3(i, j, k, nx, ny, nz, nbc)) fo[ID3(i, j, k)] = f[ID3(i, j, k)];
}

/* acoustic3d velocity sweep, update_velocity.jl:15-52: vx at (i-1/2,j,k), vy at (i,j-1/2,k), vz at
 * (i,j,k+1/2); interior cells M < i,j,k <= n - M. */
static void ac3_velocity(real *vx, real *vy, real *vz, const real *p, const real *bx, const real *by, const real *bz,
                         const real *a, int N, real dtdh, int nx, int ny, int nz) {
  int M = N;
#pragma omp parallel for schedule(static)
  for (int k = M + 1; k <= nz - M; ++k)
    for (int j = M + 1; j <= ny - M; ++j)
      for (int i = M + 1; i <= nx - M; ++i) {
        size_t c = ID3(i, j, k);
        size_t sy = (size_t)nx, sz = (size_t)nx * ny;
        real dpdx = 0, dpdy = 0, dpdz = 0;
        for (int l = 1; l <= N; ++l) {
          real cl = a[l - 1];
          dpdx += cl * (p[c + (l - 1)] - p[c - l]);
          dpdy += cl * (p[c + (size_t)(l - 1) * sy] - p[c - (size_t)l * sy]);
          dpdz += cl * (p[c + (size_t)l * sz] - p[c - (size_t)(l - 1) * sz]);
        }
        vx[c] += (bx[c] * dtdh) * dpdx;
        vy[c] += (by[c] * dtdh) * dpdy;
        vz[c] += (bz[c] * dtdh) * dpdz;
      }
}

/* acoustic3d pressure sweep, update_pressure.jl:13-48. */
static void ac3_pressure(real *p, const real *vx, const real *vy, const real *vz, const real *kappa, const real *a,
                         int N, real dtdh, int nx, int ny, int nz) {
  int M = N;
#pragma omp parallel for schedule(static)
  for (int k = M + 1; k <= nz - M; ++k)
    for (int j = M + 1; j <= ny - M; ++j)
      for (int i = M + 1; i <= nx - M; ++i) {
        size_t c = ID3(i, j, k);
        size_t sy = (size_t)nx, sz = (size_t)nx * ny;
        real dvxdx = 0, dvydy = 0, dvzdz = 0;
        for (int l = 1; l <= N; ++l) {
          real cl = a[l - 1];
          dvxdx += cl * (vx[c + l] - vx[c - (l - 1)]);
          dvydy += cl * (vy[c + (size_t)l * sy] - vy[c - (size_t)(l - 1) * sy]);
          dvzdz += cl * (vz[c + (size_t)(l - 1) * sz] - vz[c - (size_t)l * sz]);
        }
        p[c] += (kappa[c] * dtdh) * ((dvxdx + dvydy) + dvzdz);
      }
}

/*
 * acoustic3d (boundary HABC) -- inputs vp, rho: nxi*nyi*nzi (x fastest); 1-based interior
 * indices; wavelet n_src x nt column-major.  Outputs seis_{vx,vy,vz,p} n_rec x nt (any may be
 * NULL).  final_state (optional): p, vx, vy, vz.  max_steps < nt stops early (timing samples).
 */
int orc_acoustic3d(int nxi, int nyi, int nzi, const float *vp, const float *rho, float dh, float dt, int nt, int nbc,
                   int fd_order, float v_ref, int n_src, const int *sx, const int *sy, const int *sz,
                   const float *wavelet, int n_rec, const int *rx, const int *ry, const int *rz, real *seis_vx,
                   real *seis_vy, real *seis_vz, real *seis_p, real *final_state, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  float af[5];
  int N = orc_fd_coeffs(fd_order, af);
  if (N < 0) return -1;
  real a[5];
  for (int l = 0; l < N; ++l) a[l] = (real)af[l];
  int M = N, pad = nbc + M, H = pad - 1;
  int nx = nxi + 2 * pad, ny = nyi + 2 * pad, nz = nzi + 2 * pad;
  size_t n = (size_t)nx * ny * nz;
  real *rhop = calloc(n, sizeof(real)), *vpp = calloc(n, sizeof(real));
  real *kappa = calloc(n, sizeof(real)), *bx = calloc(n, sizeof(real)), *by = calloc(n, sizeof(real)),
       *bz = calloc(n, sizeof(real));
  real *p = calloc(n, sizeof(real)), *vx = calloc(n, sizeof(real)), *vy = calloc(n, sizeof(real)),
       *vz = calloc(n, sizeof(real));
  real *fo = calloc(n, sizeof(real)), *fpre = calloc(n, sizeof(real));
  if (!rhop || !vpp || !kappa || !bx || !by || !bz || !p || !vx || !vy || !vz || !fo || !fpre) return -2;
  orc_pad3d(vp, nxi, nyi, nzi, pad, vpp);
  orc_pad3d(rho, nxi, nyi, nzi, pad, rhop);
  for (size_t c = 0; c < n; ++c) kappa[c] = rhop[c] * (vpp[c] * vpp[c]);
  orc_buoyancy3d(rhop, nx, ny, nz, bx, by, bz);
  free(vpp);
  free(rhop);
  real hc[3];
  orc_habc_coeffs(v_ref, dt, dh, hc);
  real q = hc[0], qt = hc[1], qxt = hc[2];
  real dtdh = (real)dt / (real)dh;
  /* the velocity phase needs three backups; keep them in separate arrays */
  real *vxo = calloc(n, sizeof(real)), *vyo = calloc(n, sizeof(real)), *vzo = calloc(n, sizeof(real));
  if (!vxo || !vyo || !vzo) return -2;
  for (int it = 1; it <= nt; ++it) {
    backup3(vxo, vx, nx, ny, nz, H);
    backup3(vyo, vy, nx, ny, nz, H);
    backup3(vzo, vz, nx, ny, nz, H);
    ac3_velocity(vx, vy, vz, p, bx, by, bz, a, N, dtdh, nx, ny, nz);
    memcpy(fpre, vx, n * sizeof(real));
    habc3d(vx, fpre, vxo, nx, ny, nz, H, pad, 0.0, q, qt, qxt);
    memcpy(fpre, vy, n * sizeof(real));
    habc3d(vy, fpre, vyo, nx, ny, nz, H, pad, 0.25, q, qt, qxt);
    memcpy(fpre, vz, n * sizeof(real));
    habc3d(vz, fpre, vzo, nx, ny, nz, H, pad, 0.5, q, qt, qxt);

    backup3(fo, p, nx, ny, nz, H);
    ac3_pressure(p, vx, vy, vz, kappa, a, N, dtdh, nx, ny, nz);
    memcpy(fpre, p, n * sizeof(real));
    habc3d(p, fpre, fo, nx, ny, nz, H, pad, 0.75, q, qt, qxt);

    for (int s = 0; s < n_src; ++s)
      p[ID3(sx[s] + pad, sy[s] + pad, sz[s] + pad)] += (real)wavelet[(size_t)s + (size_t)n_src * (it - 1)];
    for (int r = 0; r < n_rec; ++r) {
      size_t c = ID3(rx[r] + pad, ry[r] + pad, rz[r] + pad);
      size_t o = (size_t)r + (size_t)n_rec * (it - 1);
      if (seis_vx) seis_vx[o] = vx[c];
      if (seis_vy) seis_vy[o] = vy[c];
      if (seis_vz) seis_vz[o] = vz[c];
      if (seis_p) seis_p[o] = p[c];
    }
  }
  if (final_state) {
    memcpy(final_state, p, n * sizeof(real));
    memcpy(final_state + n, vx, n * sizeof(real));
    memcpy(final_state + 2 * n, vy, n * sizeof(real));
    memcpy(final_state + 3 * n, vz, n * sizeof(real));
  }
  free(kappa); free(bx); free(by); free(bz); free(p); free(vx); free(vy); free(vz);
  free(fo); free(fpre); free(vxo); free(vyo); free(vzo);
  return 0;
}

/* ------------------------------------------------------------------------------------------------
 * elastic3d
 *   medium ................... src/equations/elastic3d/medium.jl:29-123
 *   velocity ................. src/equations/elastic3d/update_velocity.jl:17-84
 *   stress ................... src/equations/elastic3d/update_stress.jl:21-94
 *   HABC fields and weights .. src/boundary/habc/kernels_3d.jl:204-229 (vx/vy/vz own tables,
 *                              all six stresses the tau table)
 *   step order ............... src/equations/elastic3d/elastic3d.jl:122-162 (explosive source
 *                              added to txx, tyy, tzz after the stress HABC; vx, vy, vz recorded)
 */

static inline real hmean4(real m1, real m2, real m3, real m4) {
  if (m1 == (real)0 || m2 == (real)0 || m3 == (real)0 || m4 == (real)0) return (real)0;
  return (real)4 / ((((real)1 / m1 + (real)1 / m2) + (real)1 / m3) + (real)1 / m4);
}

/* lam, lam2mu, mu_txy, mu_txz, mu_tyz on the padded cube (medium.jl:76-123). */
void orc_elastic_params3d(const real *vp, const real *vs, const real *rho, int nx, int ny, int nz, real *lam,
                          real *lam2mu, real *mu_txy, real *mu_txz, real *mu_tyz) {
  size_t n = (size_t)nx * ny * nz;
  real *mu = malloc(n * sizeof(real));
  for (size_t c = 0; c < n; ++c) {
    mu[c] = rho[c] * (vs[c] * vs[c]);
    lam[c] = rho[c] * (vp[c] * vp[c]) - (real)2 * mu[c];
    lam2mu[c] = lam[c] + (real)2 * mu[c];
    mu_txy[c] = mu_txz[c] = mu_tyz[c] = (real)0;
  }
  for (int k = 1; k <= nz; ++k)
    for (int j = 2; j <= ny; ++j)
      for (int i = 2; i <= nx; ++i)
        mu_txy[ID3(i, j, k)] = hmean4(mu[ID3(i - 1, j - 1, k)], mu[ID3(i, j - 1, k)], mu[ID3(i - 1, j, k)],
                                      mu[ID3(i, j, k)]);
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) mu_txy[ID3(1, j, k)] = mu_txy[ID3(2, j, k)];
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) mu_txy[ID3(i, 1, k)] = mu_txy[ID3(i, 2, k)];
  for (int k = 1; k <= nz - 1; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 2; i <= nx; ++i)
        mu_txz[ID3(i, j, k)] = hmean4(mu[ID3(i - 1, j, k)], mu[ID3(i, j, k)], mu[ID3(i - 1, j, k + 1)],
                                      mu[ID3(i, j, k + 1)]);
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) mu_txz[ID3(1, j, k)] = mu_txz[ID3(2, j, k)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) mu_txz[ID3(i, j, nz)] = mu_txz[ID3(i, j, nz - 1)];
  for (int k = 1; k <= nz - 1; ++k)
    for (int j = 2; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        mu_tyz[ID3(i, j, k)] = hmean4(mu[ID3(i, j - 1, k)], mu[ID3(i, j, k)], mu[ID3(i, j - 1, k + 1)],
                                      mu[ID3(i, j, k + 1)]);
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) mu_tyz[ID3(i, 1, k)] = mu_tyz[ID3(i, 2, k)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) mu_tyz[ID3(i, j, nz)] = mu_tyz[ID3(i, j, nz - 1)];
  free(mu);
}

/* elastic3d velocity sweep, update_velocity.jl:17-84 (F = vx, vy, vz, txx, tyy, tzz, txy, txz, tyz). */
static void el3_velocity(real *const *F, const real *bx, const real *by, const real *bz, const real *a, int N,
                         real dtdh, int nx, int ny, int nz) {
  int M = N;
  const size_t sy_ = (size_t)nx, sz_ = (size_t)nx * ny;
  real *vx = F[0], *vy = F[1], *vz = F[2];
  const real *txx = F[3], *tyy = F[4], *tzz = F[5], *txy = F[6], *txz = F[7], *tyz = F[8];
#pragma omp parallel for schedule(static)
  for (int k = M + 1; k <= nz - M; ++k)
    for (int j = M + 1; j <= ny - M; ++j)
      for (int i = M + 1; i <= nx - M; ++i) {
        size_t c = ID3(i, j, k);
        real dtxxdx = 0, dtxydy = 0, dtxzdz = 0, dtxydx = 0, dtyydy = 0, dtyzdz = 0, dtxzdx = 0, dtyzdy = 0,
             dtzzdz = 0;
        for (int l = 1; l <= N; ++l) {
          real cl = a[l - 1];
          dtxxdx += cl * (txx[c + (l - 1)] - txx[c - l]);
          dtyydy += cl * (tyy[c + (l - 1) * sy_] - tyy[c - l * sy_]);
          dtzzdz += cl * (tzz[c + l * sz_] - tzz[c - (l - 1) * sz_]);
          dtxydy += cl * (txy[c + l * sy_] - txy[c - (l - 1) * sy_]);
          dtxydx += cl * (txy[c + l] - txy[c - (l - 1)]);
          dtxzdz += cl * (txz[c + (l - 1) * sz_] - txz[c - l * sz_]);
          dtxzdx += cl * (txz[c + l] - txz[c - (l - 1)]);
          dtyzdz += cl * (tyz[c + (l - 1) * sz_] - tyz[c - l * sz_]);
          dtyzdy += cl * (tyz[c + l * sy_] - tyz[c - (l - 1) * sy_]);
        }
        vx[c] += (bx[c] * dtdh) * ((dtxxdx + dtxydy) + dtxzdz);
        vy[c] += (by[c] * dtdh) * ((dtxydx + dtyydy) + dtyzdz);
        vz[c] += (bz[c] * dtdh) * ((dtxzdx + dtyzdy) + dtzzdz);
      }
}

/* elastic3d stress sweep, update_stress.jl:21-94. */
static void el3_stress(real *const *F, const real *lam, const real *l2m, const real *mxy, const real *mxz,
                       const real *myz, const real *a, int N, real dtdh, int nx, int ny, int nz) {
  int M = N;
  const size_t sy_ = (size_t)nx, sz_ = (size_t)nx * ny;
  const real *vx = F[0], *vy = F[1], *vz = F[2];
  real *txx = F[3], *tyy = F[4], *tzz = F[5], *txy = F[6], *txz = F[7], *tyz = F[8];
#pragma omp parallel for schedule(static)
  for (int k = M + 1; k <= nz - M; ++k)
    for (int j = M + 1; j <= ny - M; ++j)
      for (int i = M + 1; i <= nx - M; ++i) {
        size_t c = ID3(i, j, k);
        real dvxdx = 0, dvydy = 0, dvzdz = 0, dvxdy = 0, dvydx = 0, dvxdz = 0, dvzdx = 0, dvydz = 0, dvzdy = 0;
        for (int l = 1; l <= N; ++l) {
          real cl = a[l - 1];
          dvxdx += cl * (vx[c + l] - vx[c - (l - 1)]);
          dvydy += cl * (vy[c + l * sy_] - vy[c - (l - 1) * sy_]);
          dvzdz += cl * (vz[c + (l - 1) * sz_] - vz[c - l * sz_]);
          dvxdy += cl * (vx[c + (l - 1) * sy_] - vx[c - l * sy_]);
          dvydx += cl * (vy[c + (l - 1)] - vy[c - l]);
          dvxdz += cl * (vx[c + l * sz_] - vx[c - (l - 1) * sz_]);
          dvzdx += cl * (vz[c + (l - 1)] - vz[c - l]);
          dvydz += cl * (vy[c + l * sz_] - vy[c - (l - 1) * sz_]);
          dvzdy += cl * (vz[c + (l - 1) * sy_] - vz[c - l * sy_]);
        }
        real lv = lam[c], l2 = l2m[c];
        txx[c] += ((l2 * dvxdx + lv * dvydy) + lv * dvzdz) * dtdh;
        tyy[c] += ((lv * dvxdx + l2 * dvydy) + lv * dvzdz) * dtdh;
        tzz[c] += ((lv * dvxdx + lv * dvydy) + l2 * dvzdz) * dtdh;
        txy[c] += (mxy[c] * (dvxdy + dvydx)) * dtdh;
        txz[c] += (mxz[c] * (dvxdz + dvzdx)) * dtdh;
        tyz[c] += (myz[c] * (dvydz + dvzdy)) * dtdh;
      }
}

/*
 * elastic3d (boundary HABC) -- inputs vp, vs, rho: nxi*nyi*nzi; outputs seis_{vx,vy,vz} n_rec x nt
 * (any may be NULL); final_state (optional): vx, vy, vz, txx, tyy, tzz, txy, txz, tyz.
 */
int orc_elastic3d(int nxi, int nyi, int nzi, const float *vp, const float *vs, const float *rho, float dh, float dt,
                  int nt, int nbc, int fd_order, float v_ref, int n_src, const int *sx, const int *sy, const int *sz,
                  const float *wavelet, int n_rec, const int *rx, const int *ry, const int *rz, real *seis_vx,
                  real *seis_vy, real *seis_vz, real *final_state, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  float af[5];
  int N = orc_fd_coeffs(fd_order, af);
  if (N < 0) return -1;
  real a[5];
  for (int l = 0; l < N; ++l) a[l] = (real)af[l];
  int M = N, pad = nbc + M, H = pad - 1;
  int nx = nxi + 2 * pad, ny = nyi + 2 * pad, nz = nzi + 2 * pad;
  size_t n = (size_t)nx * ny * nz;
  real *arr[32];
  for (int q = 0; q < 29; ++q) {
    arr[q] = calloc(n, sizeof(real));
    if (!arr[q]) return -2;
  }
  real *vpp = arr[0], *vsp = arr[1], *rhop = arr[2];
  real *bx = arr[3], *by = arr[4], *bz = arr[5], *lam = arr[6], *l2m = arr[7], *mxy = arr[8], *mxz = arr[9],
       *myz = arr[10];
  real *F[9], *FO[9];
  for (int f = 0; f < 9; ++f) {
    F[f] = arr[11 + f];
    FO[f] = arr[20 + f];
  }
  real *vx = F[0], *vy = F[1], *vz = F[2];
  orc_pad3d(vp, nxi, nyi, nzi, pad, vpp);
  orc_pad3d(vs, nxi, nyi, nzi, pad, vsp);
  orc_pad3d(rho, nxi, nyi, nzi, pad, rhop);
  orc_buoyancy3d(rhop, nx, ny, nz, bx, by, bz);
  orc_elastic_params3d(vpp, vsp, rhop, nx, ny, nz, lam, l2m, mxy, mxz, myz);
  real *fpre = vpp; /* reused as the pre-correction copy */
  real hc[3];
  orc_habc_coeffs(v_ref, dt, dh, hc);
  real q = hc[0], qt = hc[1], qxt = hc[2];
  real dtdh = (real)dt / (real)dh;
  const double voff[3] = {0.0, 0.25, 0.5};
  for (int it = 1; it <= nt; ++it) {
    for (int f = 0; f < 3; ++f) backup3(FO[f], F[f], nx, ny, nz, H);
    el3_velocity(F, bx, by, bz, a, N, dtdh, nx, ny, nz);
    for (int f = 0; f < 3; ++f) {
      memcpy(fpre, F[f], n * sizeof(real));
      habc3d(F[f], fpre, FO[f], nx, ny, nz, H, pad, voff[f], q, qt, qxt);
    }

    for (int f = 3; f < 9; ++f) backup3(FO[f], F[f], nx, ny, nz, H);
    el3_stress(F, lam, l2m, mxy, mxz, myz, a, N, dtdh, nx, ny, nz);
    for (int f = 3; f < 9; ++f) {
      memcpy(fpre, F[f], n * sizeof(real));
      habc3d(F[f], fpre, FO[f], nx, ny, nz, H, pad, 0.75, q, qt, qxt);
    }
    for (int f = 3; f < 6; ++f)
      for (int s = 0; s < n_src; ++s)
        F[f][ID3(sx[s] + pad, sy[s] + pad, sz[s] + pad)] += (real)wavelet[(size_t)s + (size_t)n_src * (it - 1)];
    for (int r = 0; r < n_rec; ++r) {
      size_t c = ID3(rx[r] + pad, ry[r] + pad, rz[r] + pad);
      size_t o = (size_t)r + (size_t)n_rec * (it - 1);
      if (seis_vx) seis_vx[o] = vx[c];
      if (seis_vy) seis_vy[o] = vy[c];
      if (seis_vz) seis_vz[o] = vz[c];
    }
  }
  if (final_state)
    for (int f = 0; f < 9; ++f) memcpy(final_state + (size_t)f * n, F[f], n * sizeof(real));
  for (int q2 = 0; q2 < 29; ++q2) free(arr[q2]);
  return 0;
}

/* Single sweeps on a caller-given padded state (the functions the 3D time loops call), for the
 * plane-wave known-answer tests.  F: vx, vy, vz, txx, tyy, tzz, txy, txz, tyz (9 arrays). */
void orc_probe_acoustic3d_velocity(int nx, int ny, int nz, int fd_order, real dtdh, real *vx, real *vy, real *vz,
                                   const real *p, const real *bx, const real *by, const real *bz) {
  float af[5];
  real a[5];
  int N = orc_fd_coeffs(fd_order, af);
  for (int l = 0; l < N; ++l) a[l] = (real)af[l];
  ac3_velocity(vx, vy, vz, p, bx, by, bz, a, N, dtdh, nx, ny, nz);
}

void orc_probe_acoustic3d_pressure(int nx, int ny, int nz, int fd_order, real dtdh, real *p, const real *vx,
                                   const real *vy, const real *vz, const real *kappa) {
  float af[5];
  real a[5];
  int N = orc_fd_coeffs(fd_order, af);
  for (int l = 0; l < N; ++l) a[l] = (real)af[l];
  ac3_pressure(p, vx, vy, vz, kappa, a, N, dtdh, nx, ny, nz);
}

void orc_probe_elastic3d_velocity(int nx, int ny, int nz, int fd_order, real dtdh, real *const *F, const real *bx,
                                  const real *by, const real *bz) {
  float af[5];
  real a[5];
  int N = orc_fd_coeffs(fd_order, af);
  for (int l = 0; l < N; ++l) a[l] = (real)af[l];
  el3_velocity(F, bx, by, bz, a, N, dtdh, nx, ny, nz);
}

void orc_probe_elastic3d_stress(int nx, int ny, int nz, int fd_order, real dtdh, real *const *F, const real *lam,
                                const real *l2m, const real *mxy, const real *mxz, const real *myz) {
  float af[5];
  real a[5];
  int N = orc_fd_coeffs(fd_order, af);
  for (int l = 0; l < N; ++l) a[l] = (real)af[l];
  el3_stress(F, lam, l2m, mxy, mxz, myz, a, N, dtdh, nx, ny, nz);
}
