// engine2d.hip -- 2D staggered-grid sweeps for gfx950 with the deterministic Higdon boundary fused in.
//
// One launch per phase (velocity; pressure or stress) over a work list spanning every shot.  The
// user's interior of a shot (the padded grid minus the pad, which is exactly the absorbing frame
// plus the outer ring) is cut into strips 64 columns wide (one wave64 row) that march down z; the
// frame is cut into 64 x 8 tiles.  The fields a phase reads through FD stencils are staged in
// LDS; the phase's own fields (time n, ping-pong buffer) and the materials are read once,
// coalesced.
//
// Deterministic HABC without a second pass: a frame cell's Higdon value needs the inward
// neighbour's *pre-correction* value.  Frame tiles first compute pre-correction values on the
// tile plus a one-cell ring into LDS (an (N+1)-cell halo makes the ring's stencils available
// on-chip), then blend.  Every output is a pure function of read-only buffers, so results do not
// depend on launch geometry, strip length or the other shots of a batch.
//
// Arithmetic order follows the reference expressions term by term; the library is compiled
// with -ffp-contract=off, so FP32 results are bit-identical to the reference's uncontracted
// Float32 updates:
//   acoustic velocity / pressure ... src/equations/acoustic2d/fused_kernels.jl:47-61, :101-112
//   elastic velocity / stress ...... src/equations/elastic2d/fused_kernels.jl:39-57, :105-121
//   coupled P-S potentials ......... src/equations/coupled2d/update_fields.jl:27-144 (the
//                                    interior's position half-step rides in its velocity strips)
//   Higdon value ................... src/boundary/habc/kernels_det.jl:22-57
//   injection / gather ............. src/acquisition/device_indexed.jl:28-134 (step index read
//                                    from a device counter, as the reference's graph path does)
#include <algorithm>

#include <type_traits>

#include "fomo_common.h"
#include "storage.h"

namespace fomo {

// Resolved pointers of one phase of one shot.
struct PV {
  const void* sten[3];
  const void* old[3];
  void* out[3];
  const void* mat[8];
  float sdec[3], odec[3], oenc[3], mdec[8];  // scales of the stored arrays above
  gp<const float> lut[3];
  gp<const float> alpha[3];  // learned corrections of the distance-table weight (NULL: table only)
  gp<float> sens[3];         // f_pde - S at frame cells while training (NULL: not recorded)
};

template <int PH>
__device__ __forceinline__ void make_view(const Shot2D& S, int s, int it, PV& V) {
  const int o = 1 - s;
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    V.alpha[f] = nullptr;
    V.sens[f] = nullptr;
  }
  auto sten = [&](int k, int buf, int f) {
    V.sten[k] = S.fld[buf][f];
    V.sdec[k] = S.fdec[f];
  };
  auto old_out = [&](int k, int f) {
    V.old[k] = S.fld[s][f];
    V.odec[k] = S.fdec[f];
    V.out[k] = S.fld[o][f];
    V.oenc[k] = S.fenc[f];
  };
  auto mat = [&](int k, int m) {
    V.mat[k] = S.mat[m];
    V.mdec[k] = S.mdec[m];
  };
  const int L = S.H + 1;
  if (PH == kAcVel) {
    sten(0, s, 0);
    old_out(0, 1);
    old_out(1, 2);
    mat(0, 1);
    mat(1, 2);
    V.lut[0] = G(S.wlut);
    V.lut[1] = G(S.wlut + L);
    V.alpha[0] = V.alpha[1] = G((const float*)S.alpha[0]);
    V.sens[0] = G(S.sens[0]);
    V.sens[1] = G(S.sens[1]);
  } else if (PH == kAcPrs) {
    sten(0, o, 1);
    sten(1, o, 2);
    old_out(0, 0);
    mat(0, 0);
    V.lut[0] = G(S.wlut + 3 * L);
    V.alpha[0] = G((const float*)S.alpha[1]);
    V.sens[0] = G(S.sens[2]);
  } else if (PH == kElVel) {
    sten(0, s, 2);
    sten(1, s, 3);
    sten(2, s, 4);
    old_out(0, 0);
    old_out(1, 1);
    mat(0, 3);
    mat(1, 4);
    V.lut[0] = G(S.wlut);
    V.lut[1] = G(S.wlut + L);
  } else if (PH == kElStr) {
    sten(0, o, 0);
    sten(1, o, 1);
    for (int f = 0; f < 3; ++f) {
      old_out(f, 2 + f);
      mat(f, f);
      V.lut[f] = G(S.wlut + 3 * L);
    }
  } else if (PH == kCpVel) {
    // P, S through the stencils; dP/dt, dS/dt updated (vx weights); the eight medium arrays
    sten(0, s, 0);
    sten(1, s, 1);
    old_out(0, 2);
    old_out(1, 3);
    for (int m = 0; m < 8; ++m) mat(m, m);
    V.lut[0] = V.lut[1] = G(S.wlut);
  } else if (PH == kCpPos) {
    // the new dP/dt, dS/dt at the cell; P, S updated (tau weights)
    sten(0, o, 2);
    sten(1, o, 3);
    old_out(0, 0);
    old_out(1, 1);
    V.lut[0] = V.lut[1] = G(S.wlut + 3 * L);
  } else {
    // scalar leapfrog, step it: u^it is read (stencil and old value), u^{it-1} enters as the second
    // per-cell input, u^{it+1} is written; buffers rotate by it mod 3
    const int cur = it % 3, prv = (it + 2) % 3, nxt = (it + 1) % 3;
    V.sten[0] = S.fld[0][cur];
    V.sdec[0] = S.fdec[0];
    V.old[0] = S.fld[0][cur];
    V.odec[0] = S.fdec[0];
    V.out[0] = S.fld[0][nxt];
    V.oenc[0] = S.fenc[0];
    mat(0, 0);
    V.mat[1] = S.fld[0][prv];
    V.mdec[1] = S.fdec[0];
    V.lut[0] = G(S.wlut + 3 * L);
  }
}

// Stencil data staged in LDS: tile k covers global cells [oi, oi+TW) x [oj, oj+TH).
template <int TW, int TH>
struct TileRd {
  const float* t;
  int oi, oj;
  __device__ __forceinline__ float operator()(int k, int i, int j) const {
    return t[k * TW * TH + (j - oj) * TW + (i - oi)];
  }
};

// Phase physics.  NF output fields, NS stencil fields, NM materials.  `upd` is the FD update at
// padded cell (i, j) inside the update region N <= i < nx-N, N <= j < nz-N, from the stencil
// reader rd, the decoded old values o and materials m; outside that region the pre-correction
// value is the old value.  Loads are done by the callers, so a block can issue all of its loads
// before it computes anything.
template <int PH, int N, class T>
struct Phys;

template <int N, class T>
struct Phys<kAcVel, N, T> {
  static constexpr int NF = 2, NS = 1, NM = 2, HALO = N;
  template <class R>
  __device__ __forceinline__ static void upd(const R& rd, float dtx, const float* a, int i, int j,
                                             const float* o, const float* m, float* v) {
    float dx = 0.0f, dz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dx += cl * (rd(0, i + l - 1, j) - rd(0, i - l, j));
      dz += cl * (rd(0, i, j + l) - rd(0, i, j - l + 1));
    }
    v[0] = o[0] + (m[0] * dtx) * dx;
    v[1] = o[1] + (m[1] * dtx) * dz;
  }
};

template <int N, class T>
struct Phys<kAcPrs, N, T> {
  static constexpr int NF = 1, NS = 2, NM = 1, HALO = N;
  template <class R>
  __device__ __forceinline__ static void upd(const R& rd, float dtx, const float* a, int i, int j,
                                             const float* o, const float* m, float* v) {
    float dvxdx = 0.0f, dvzdz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dvxdx += cl * (rd(0, i + l, j) - rd(0, i - l + 1, j));
      dvzdz += cl * (rd(1, i, j + l - 1) - rd(1, i, j - l));
    }
    v[0] = o[0] + m[0] * (dvxdx * dtx + dvzdz * dtx);
  }
};

template <int N, class T>
struct Phys<kElVel, N, T> {
  static constexpr int NF = 2, NS = 3, NM = 2, HALO = N;
  template <class R>
  __device__ __forceinline__ static void upd(const R& rd, float dtx, const float* a, int i, int j,
                                             const float* o, const float* m, float* v) {
    float dtxxdx = 0.0f, dtxzdz = 0.0f, dtxzdx = 0.0f, dtzzdz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dtxxdx += cl * (rd(0, i + l - 1, j) - rd(0, i - l, j));
      dtxzdz += cl * (rd(2, i, j + l - 1) - rd(2, i, j - l));
      dtxzdx += cl * (rd(2, i + l, j) - rd(2, i - l + 1, j));
      dtzzdz += cl * (rd(1, i, j + l) - rd(1, i, j - l + 1));
    }
    const float dt = dtx;
    v[0] = o[0] + m[0] * (dt * dtxxdx + dt * dtxzdz);
    v[1] = o[1] + m[1] * (dt * dtxzdx + dt * dtzzdz);
  }
};

template <int N, class T>
struct Phys<kElStr, N, T> {
  static constexpr int NF = 3, NS = 2, NM = 3, HALO = N;
  template <class R>
  __device__ __forceinline__ static void upd(const R& rd, float dtx, const float* a, int i, int j,
                                             const float* o, const float* m, float* v) {
    float dvxdx = 0.0f, dvzdz = 0.0f, dvxdz = 0.0f, dvzdx = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dvxdx += cl * (rd(0, i + l, j) - rd(0, i - l + 1, j));
      dvzdz += cl * (rd(1, i, j + l - 1) - rd(1, i, j - l));
      dvxdz += cl * (rd(0, i, j + l) - rd(0, i, j - l + 1));
      dvzdx += cl * (rd(1, i + l - 1, j) - rd(1, i - l, j));
    }
    const float dt = dtx;
    const float lv = m[0], l2m = m[1];
    v[0] = o[0] + ((l2m * dvxdx) * dt + (lv * dvzdz) * dt);
    v[1] = o[1] + ((lv * dvxdx) * dt + (l2m * dvzdz) * dt);
    v[2] = o[2] + m[2] * (dvxdz * dt + dvzdx * dt);
  }
};

// Second-order scalar (constant-density) equation, src/equations/scalar2d/scalar2d.jl:33-55:
// u^{n+1} = (2 u - u^{n-1}) + c2 * lap, lap = (2 c0) u + sum_l d2_l (((u[i+l] + u[i-l]) +
// u[j+l]) + u[j-l]); o[0] = u^n at the cell, m = {c2, u^{n-1}}; c0 is passed in dtx's place.
template <int N, class T>
struct Phys<kScal, N, T> {
  static constexpr int NF = 1, NS = 1, NM = 2, HALO = N;
  template <class R>
  __device__ __forceinline__ static void upd(const R& rd, float c0, const float* a, int i, int j,
                                             const float* o, const float* m, float* v) {
    const float uc = o[0];
    float lap = (2.0f * c0) * uc;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      lap += cl * (((rd(0, i + l, j) + rd(0, i - l, j)) + rd(0, i, j + l)) + rd(0, i, j - l));
    }
    v[0] = (2.0f * uc - m[1]) + m[0] * lap;
  }
};

// Per-shot step constants read once per block: Float32(dt/dh), and the coupled equations' dt, 1/dh
// and (1/dh)^2.
struct UpdC {
  float dtx, dt, ih, ih2;
  __device__ __forceinline__ explicit UpdC(const Shot2D& S) : dtx(S.dtx), dt(S.dt), ih(S.ih), ih2(S.ih2) {}
};

// Coupled P-S potentials, "velocity" half-step (src/equations/coupled2d/update_fields.jl:27-112):
// centred first derivatives (weights A.a) and Laplacians (A.a2, centre A.c0) of P and S, then
//   rhs_P = P lap_alpha + 2 (grad alpha . grad P) - 2 P lap_beta + 2 (dbeta/dx dS/dz - dbeta/dz dS/dx)
//           + alpha lap P,
//   rhs_S = 2 (grad beta . grad S) + 2 (dbeta/dz dP/dx - dbeta/dx dP/dz) + beta lap S,
// dP/dt += dt rhs_P, dS/dt += dt rhs_S, each sum in the reference's left-to-right order.
// m = {alpha, beta, dalpha/dx, dalpha/dz, dbeta/dx, dbeta/dz, lap alpha, lap beta}.
template <int N, class T>
struct Phys<kCpVel, N, T> {
  static constexpr int NF = 2, NS = 2, NM = 8, HALO = N;
  template <class R>
  __device__ __forceinline__ static void upd(const R& rd, const SweepArgs& A, const UpdC& K, int i, int j,
                                             const float* o, const float* m, float* v) {
    const float c0 = A.c0;
    const float Pc = rd(0, i, j), Sc = rd(1, i, j);
    float dPdx = 0.0f, dPdz = 0.0f, lPx = c0 * Pc, lPz = c0 * Pc;
    float dSdx = 0.0f, dSdz = 0.0f, lSx = c0 * Sc, lSz = c0 * Sc;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float c1 = A.a[l - 1], c2 = A.a2[l - 1];
      const float pr = rd(0, i + l, j), pl = rd(0, i - l, j), pd = rd(0, i, j + l), pu = rd(0, i, j - l);
      const float sr = rd(1, i + l, j), sl = rd(1, i - l, j), sd = rd(1, i, j + l), su = rd(1, i, j - l);
      dPdx += c1 * (pr - pl);
      dPdz += c1 * (pd - pu);
      lPx += c2 * (pr + pl);
      lPz += c2 * (pd + pu);
      dSdx += c1 * (sr - sl);
      dSdz += c1 * (sd - su);
      lSx += c2 * (sr + sl);
      lSz += c2 * (sd + su);
    }
    const float ih = K.ih, ih2 = K.ih2, dt = K.dt;
    dPdx = dPdx * ih;
    dPdz = dPdz * ih;
    dSdx = dSdx * ih;
    dSdz = dSdz * ih;
    const float lapP = (lPx + lPz) * ih2, lapS = (lSx + lSz) * ih2;
    const float rP = (((Pc * m[6] + 2.0f * (m[2] * dPdx + m[3] * dPdz)) - (2.0f * Pc) * m[7]) +
                      2.0f * (m[4] * dSdz - m[5] * dSdx)) +
                     m[0] * lapP;
    const float rS = (2.0f * (m[4] * dSdx + m[5] * dSdz) + 2.0f * (m[5] * dPdx - m[4] * dPdz)) + m[1] * lapS;
    v[0] = o[0] + dt * rP;
    v[1] = o[1] + dt * rS;
  }
};

// Coupled "position" half-step (update_fields.jl:124-144): P += dt dP/dt, S += dt dS/dt with the
// new rates at the cell itself (no stencil: the staged tile is the one-cell ring).
template <int N, class T>
struct Phys<kCpPos, N, T> {
  static constexpr int NF = 2, NS = 2, NM = 0, HALO = 0;
  template <class R>
  __device__ __forceinline__ static void upd(const R& rd, const SweepArgs& A, const UpdC& K, int i, int j,
                                             const float* o, const float* m, float* v) {
    (void)A;
    (void)m;
    v[0] = o[0] + K.dt * rd(0, i, j);
    v[1] = o[1] + K.dt * rd(1, i, j);
  }
};

// FD update of phase P at (i, j): the coupled phases take their weights and steps from the
// launch and the shot, the others from (dtx or c0, a).
template <int PH, class P, class R>
__device__ __forceinline__ void upd_any(const R& rd, const SweepArgs& A, const UpdC& K, int i, int j,
                                        const float* o, const float* m, float* v) {
  if constexpr (PH == kCpVel || PH == kCpPos)
    P::upd(rd, A, K, i, j, o, m, v);
  else
    P::upd(rd, PH == kScal ? A.c0 : K.dtx, A.a, i, j, o, m, v);
}

// Old values and materials of cell c (decoded).  Fields are stored as T, materials as TM.
template <class P, class T, class TM>
__device__ __forceinline__ void load_om(const PV& V, unsigned c, float* o, float* m) {
#pragma unroll
  for (int f = 0; f < P::NF; ++f) o[f] = Cvt<T>::ld(V.old[f], c, V.odec[f]);
#pragma unroll
  for (int k = 0; k < P::NM; ++k) m[k] = Cvt<TM>::ld(V.mat[k], c, V.mdec[k]);
}

// Which output fields receive the raw-add source after the HABC.
template <int PH>
__device__ __forceinline__ int inj_mask(const Shot2D& S) {
  if (PH == kAcPrs) return 1;
  if (PH == kElVel) return S.src_fz ? 2 : 0;
  if (PH == kElStr) return S.src_fz ? 0 : 3;
  if (PH == kScal) return 1;
  if (PH == kCpVel) return 1;  // into the new P, which the interior strips of this phase compute
  return 0;
}

template <int NF>
__device__ __forceinline__ void inject(const Shot2D& S, int mask, int c, int it, int nt, float* v) {
  if (mask == 0 || it >= nt) return;
  for (int s = 0; s < S.n_src; ++s) {
    if (G(S.src_c)[s] == c) {
      const float w = G(S.wav)[(size_t)s + (size_t)S.n_src * (size_t)it];
#pragma unroll
      for (int f = 0; f < NF; ++f)
        if (mask & (1 << f)) v[f] += w;
    }
  }
}

// Reader of the strip window for the update of row j.  The window is circular over RS rows (row
// r in slot (r - jb) mod RS) and keeps a copy of its first 2N slots after the last one, so rows
// j - N .. j + N lie in consecutive slots from the slot of row j - N: every stencil read is a
// constant offset from one base per row (an LDS immediate), not a slot computed per read.
template <int RW, int RSX, int N>
struct StripRd {
  const float* t;  // slot of row j - N, shifted so that column i is t[i]
  int j;
  __device__ __forceinline__ float operator()(int k, int i, int jj) const {
    return t[k * RSX * RW + (jj - j + N) * RW + i];
  }
};

// Frame tile: at most 64 x kFrameRows cells of the absorbing frame.  Every global load of the tile
// is issued at once: the stencil fields with an (N+1) halo, old values and materials of the tile
// plus a one-cell ring, and the old values each frame cell's Higdon term needs.  Pre-correction
// values of the tile and ring then go to LDS, and each frame cell blends with its inward
// neighbour's pre-correction value.  BZ rows of a tile, TY thread rows of the block (64 x TY
// threads, BZ / TY rows per thread).
template <int PH, int N, class T, class TM, int BZ = kFrameRows, int TY = 4>
__device__ __forceinline__ void frame_tile(const SweepArgs& A, const Work2D& W, const Shot2D& S, const PV& V,
                                           float* lds, int it) {
  using P = Phys<PH, N, T>;
  constexpr int NF = P::NF, NS = P::NS, NM = P::NM, NMA = NM > 0 ? NM : 1, HN = P::HALO;
  constexpr int BX = kTileW, ZPT = BZ / TY, NTH = BX * TY;
  static_assert(BZ % TY == 0, "whole rows per thread");
  constexpr int TW = BX + 2 * HN + 2, TH = BZ + 2 * HN + 2;
  constexpr int EW = BX + 2, EH = BZ + 2;
  constexpr int KS = (TW * TH + NTH - 1) / NTH;  // staged cells per thread
  constexpr int KR = (EW * EH + NTH - 1) / NTH;  // tile + ring cells per thread
  float* tile = lds;
  float* pre_s = lds + NS * TW * TH;
  const int nx = S.nx, nz = S.nz, H = S.H;
  const UpdC K(S);
  // Higdon coefficients: the coupled S fields have their own reference velocity
  constexpr bool two_sets = PH == kCpVel || PH == kCpPos;
  const float qa = S.q, qta = S.qt, qxta = S.qxt;
  const float qb = two_sets ? S.q2 : qa, qtb = two_sets ? S.qt2 : qta, qxtb = two_sets ? S.qxt2 : qxta;
  const float* damp = S.damp;
  const int i0 = W.i0, j0 = W.j0;
  const int tid = threadIdx.x + threadIdx.y * BX;
  const int oi = i0 - HN - 1, oj = j0 - HN - 1;
  // 1. issue every load
  float st[KS][NS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int idx = tid + k * NTH;
    const int r = idx / TW, cc = idx - r * TW;
    const int gi = oi + cc, gj = oj + r;
    const bool in = idx < TW * TH && gi >= 0 && gi < nx && gj >= 0 && gj < nz;
    const unsigned g = in ? (unsigned)(gi + gj * nx) : 0u;
    // Unconditional load of a valid cell, zeroed by a multiply outside the grid: a load whose
    // value is selected would be moved under a branch and waited for there.  (Cells outside the
    // grid never enter an update; zeroing them only keeps the staged tile clean.)
    const float zm = in ? 1.0f : 0.0f;
#pragma unroll
    for (int f = 0; f < NS; ++f) st[k][f] = Cvt<T>::ld(V.sten[f], g, V.sdec[f]) * zm;
  }
  float ro[KR][NF], rm[KR][NMA];
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int idx = tid + k * NTH;
    const int r = idx / EW, cc = idx - r * EW;
    const int gi = i0 - 1 + cc, gj = j0 - 1 + r;
    const bool ok = idx < EW * EH && gi >= 0 && gi < nx && gj >= 0 && gj < nz && cc <= W.w + 1 && r <= W.h + 1;
    const unsigned g = ok ? (unsigned)(gi + gj * nx) : 0u;
    load_om<P, T, TM>(V, g, ro[k], rm[k]);
  }
  const int i = i0 + (int)threadIdx.x;
  const bool lane_ok = (int)threadIdx.x < W.w;
  // Old values at the cell and at its inward x / z neighbours, the distance-table weights, the
  // learned corrections and the sponge factors.  Every load has a valid index and no lane-dependent
  // condition (such a load would be waited for in its own branch, draining all loads in flight).
  float fo[ZPT][NF], fx[ZPT][NF], fz[ZPT][NF], lw[ZPT][NF], aw[ZPT][NF], dm[ZPT];
#pragma unroll
  for (int k = 0; k < ZPT; ++k) {
    const int jl = (int)threadIdx.y * ZPT + k;
    const int j = j0 + jl;
    const bool ok = lane_ok && jl < W.h;
    const int dist = ok ? min(min(i, nx - 1 - i), min(j, nz - 1 - j)) : 0;  // < pad in a frame tile
    const bool Lf = i <= H, Rf = i >= nx - 1 - H, Tf = j <= H, Bf = j >= nz - 1 - H;
    const bool blend = ok && !damp && dist >= 1 && dist <= H;
    const bool inx = blend && (Lf || Rf), inz = blend && (Tf || Bf);
    const unsigned c = ok ? (unsigned)(i + j * nx) : 0u;
    const unsigned cx = inx ? (unsigned)(i + (Lf ? 1 : -1) + j * nx) : c;
    const unsigned cz = inz ? (unsigned)(i + (j + (Tf ? 1 : -1)) * nx) : c;
    const int dl = min(dist, H);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      fo[k][f] = Cvt<T>::ld(V.old[f], c, V.odec[f]);
      fx[k][f] = Cvt<T>::ld(V.old[f], cx, V.odec[f]);
      fz[k][f] = Cvt<T>::ld(V.old[f], cz, V.odec[f]);
      lw[k][f] = V.lut[f][dl];
      aw[k][f] = V.alpha[f] ? V.alpha[f][c] : 0.0f;
    }
    dm[k] = damp ? G(damp)[dist] : 1.0f;
  }
  // 2. stage the stencil fields
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int idx = tid + k * NTH;
    if (idx < TW * TH) {
#pragma unroll
      for (int f = 0; f < NS; ++f) tile[f * TW * TH + idx] = st[k][f];
    }
  }
  __syncthreads();
  // 3. pre-correction values of the tile and its ring
  const TileRd<TW, TH> rd{tile, oi, oj};
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int idx = tid + k * NTH;
    const int r = idx / EW, cc = idx - r * EW;
    const int gi = i0 - 1 + cc, gj = j0 - 1 + r;
    if (idx >= EW * EH || gi < 0 || gi >= nx || gj < 0 || gj >= nz || cc > W.w + 1 || r > W.h + 1) continue;
    float v[NF];
    if (gi < N || gi >= nx - N || gj < N || gj >= nz - N) {
#pragma unroll
      for (int f = 0; f < NF; ++f) v[f] = ro[k][f];
    } else {
      upd_any<PH, P>(rd, A, K, gi, gj, ro[k], rm[k], v);
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) pre_s[f * EW * EH + idx] = v[f];
  }
  __syncthreads();
  // 4. blend and store (no source lies in the frame: sources are validated against the user's
  // interior, which interior strips cover)
  if (!lane_ok) return;
#pragma unroll
  for (int k = 0; k < ZPT; ++k) {
    const int jl = (int)threadIdx.y * ZPT + k;
    if (jl >= W.h) break;
    const int j = j0 + jl;
    const int li = (int)threadIdx.x + 1, lj = jl + 1;
    const unsigned c = (unsigned)(i + j * nx);
    float v[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) v[f] = pre_s[f * EW * EH + lj * EW + li];
    const int dist = min(min(i, nx - 1 - i), min(j, nz - 1 - j));
    if (damp) {
      // Cerjan sponge, src/boundary/sponge.jl:56-78: field *= damp after the update
#pragma unroll
      for (int f = 0; f < NF; ++f) v[f] = v[f] * dm[k];
    } else if (dist >= 1 && dist <= H) {
      // Region flags (0-based): left i <= H, right i >= nx-1-H, top j <= H, bottom j >= nz-1-H.
      const bool Lf = i <= H, Rf = i >= nx - 1 - H, Tf = j <= H, Bf = j >= nz - 1 - H;
      const bool inx = Lf || Rf, inz = Tf || Bf;
      const int di = Lf ? 1 : -1, dj = Tf ? 1 : -1;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const bool sb = two_sets && f == 1;
        const float q = sb ? qb : qa, qt = sb ? qtb : qta, qxt = sb ? qxtb : qxta;
        const float foc = fo[k][f];
        float sx = 0.0f, sz = 0.0f;
        if (inx) {
          const float fn = pre_s[f * EW * EH + lj * EW + li + di];
          sx = (-q * fn - qt * foc) - qxt * fx[k][f];
        }
        if (inz) {
          const float fn = pre_s[f * EW * EH + (lj + dj) * EW + li];
          sz = (-q * fn - qt * foc) - qxt * fz[k][f];
        }
        const float w = V.alpha[f] ? fminf(fmaxf(lw[k][f] + aw[k][f], 0.0f), 1.0f) : lw[k][f];
        if (V.sens[f]) V.sens[f][c] = v[f] - (inx && inz ? 0.5f * (sx + sz) : (inx ? sx : sz));
        if (inx && inz)
          v[f] = w * v[f] + ((1.0f - w) * 0.5f) * (sx + sz);
        else if (inx)
          v[f] = w * v[f] + (1.0f - w) * sx;
        else
          v[f] = w * v[f] + (1.0f - w) * sz;
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) st_field<T>(V.out[f], c, v[f], V.oenc[f], S.ovf);
  }
}

// Interior strip: at most 64 columns by W.h rows of the user's interior (no frame cell, every
// cell in the update region).  The block walks down z in groups of eight rows (two rows per
// wave, four rows apart) through a circular LDS window of RS rows holding the stencil inputs.
// While a group is computed, the stencil rows, old values and materials of the next group are in
// flight (sixteen to twenty loads per lane), so each block waits on one round of loads per eight
// rows.  The window spans exactly the rows one group reads (GR + 2N) and the rows the next group
// adds (GR), so the rows being written never overlap the rows being read and one barrier per
// group suffices; RSX adds the copy of the first 2N slots (see StripRd).
constexpr int kStripGroup = 8;
template <int N>
struct StripRing {
  static constexpr int RS = 2 * kStripGroup + 2 * N;
  static constexpr int RSX = RS + 2 * N;
};

template <int PH, int N, class T, class TM>
__device__ __forceinline__ void interior_strip(const SweepArgs& A, const Work2D& W, const Shot2D& S, const PV& V,
                                               float* lds, int it) {
  using P = Phys<PH, N, T>;
  constexpr int NF = P::NF, NS = P::NS, NM = P::NM, NMA = NM > 0 ? NM : 1;
  constexpr int BX = kTileW, RW = BX + 2 * N, RS = StripRing<N>::RS, RSX = StripRing<N>::RSX;
  constexpr int GR = kStripGroup, RPT = GR / 4;
  constexpr int W0 = GR + 2 * N, R0 = (W0 + 3) / 4;  // initial window rows, per thread row
  const int nx = S.nx, nz = S.nz;
  const UpdC K(S);
  const int i0 = W.i0, j0 = W.j0, jend = W.j0 + W.h;
  const int tx = threadIdx.x, ty = threadIdx.y;
  // Rows of this strip where this lane's column holds a source, scanned before any other load so
  // that the row loop never waits on the source table.
  const int mask = inj_mask<PH>(S);
  int src_lo = 1 << 30, src_hi = -1;
  if (mask && it < A.nt) {
    const int n_src = S.n_src;
    const int* src_c = S.src_c;
    const int ic = i0 + tx;
    for (int q = 0; q < n_src; ++q) {
      const int cs = G(src_c)[q];
      const int js = cs / nx, is = cs - js * nx;
      if (is == ic && js >= j0 && js < jend) {
        src_lo = min(src_lo, js);
        src_hi = max(src_hi, js);
      }
    }
  }
  const int oi = i0 - N, jb = j0 - N;
  // columns loaded by this lane: oi + tx and, for tx < 2N, oi + 64 + tx
  const int ca = oi + tx, cb = oi + BX + tx;
  const bool has_b = tx < 2 * N;
  const bool ina = ca >= 0 && ca < nx, inb = has_b && cb >= 0 && cb < nx;
  // Unconditional loads of valid cells; cells outside the grid are zeroed by a multiply when the
  // row is written to LDS (see frame_tile), so nothing between the loads and that write waits.
  auto load_row = [&](int j, float (&ra)[NS], float (&rb)[NS]) {
    const bool inj = j >= 0 && j < nz;
    const unsigned rowo = (unsigned)((inj ? j : 0) * nx);
    const unsigned ga = rowo + (unsigned)(ina ? ca : 0), gb = rowo + (unsigned)(inb ? cb : 0);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      ra[k] = Cvt<T>::ld(V.sten[k], ga, V.sdec[k]);
      rb[k] = Cvt<T>::ld(V.sten[k], gb, V.sdec[k]);
    }
  };
  auto store_row = [&](int j, const float (&ra)[NS], const float (&rb)[NS]) {
    const bool inj = j >= 0 && j < nz;
    const float za = (inj && ina) ? 1.0f : 0.0f, zb = (inj && inb) ? 1.0f : 0.0f;
    const int slot = (int)((unsigned)(j - jb) % (unsigned)RS);  // j >= jb
    const bool mirror = slot < 2 * N;  // uniform per wave (one thread row per wave)
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const float va = ra[k] * za, vb = rb[k] * zb;
      float* row = lds + k * RSX * RW + slot * RW;
      row[tx] = va;
      if (has_b) row[BX + tx] = vb;
      if (mirror) {
        row[RS * RW + tx] = va;
        if (has_b) row[RS * RW + BX + tx] = vb;
      }
    }
  };
  const int i = i0 + tx;
  const bool col_ok = tx < W.w;
  const int n_grp = (W.h + GR - 1) / GR;
  // Snapshot after this step (acoustic: p from the pressure phase; elastic: vz from the velocity
  // phase): the interior cells are exactly the strips' cells.
  constexpr int SF = PH == kAcPrs ? 0 : (PH == kElVel ? 1 : -1);
  const int pad = S.H + 1, nxi = nx - 2 * pad;
  float* snap_k = nullptr;  // snapshot k of this step (coupled: of P; of S at the same offset from snap2)
  if ((SF >= 0 || PH == kCpVel) && S.snap && (it + 1) % S.snap_every == 0 && (it + 1) / S.snap_every <= S.snap_n)
    snap_k = S.snap + (size_t)((it + 1) / S.snap_every - 1) * (size_t)nxi * (size_t)(nz - 2 * pad);
  // Loads of group k: the window rows it adds (j0 + GR*k + N + ty + 4r) and the old values and
  // materials of its rows j0 + GR*k + ty + 4r.
  float ra[RPT][NS], rb[RPT][NS], no[RPT][NF], nm[RPT][NMA];
  auto issue = [&](int k) {
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      load_row(j0 + GR * k + N + ty + 4 * r, ra[r], rb[r]);
      const int j = j0 + GR * k + ty + 4 * r;
      load_om<P, T, TM>(V, (col_ok && j < jend) ? (unsigned)(i + j * nx) : 0u, no[r], nm[r]);
    }
  };
  float co[RPT][NF], cm[RPT][NMA];
  {
    // initial window (rows j0-N .. j0+GR-1+N) and the old values / materials of group 0
    float ia[R0][NS], ib[R0][NS];
#pragma unroll
    for (int k = 0; k < R0; ++k) load_row(ty + 4 * k < W0 ? jb + ty + 4 * k : -1, ia[k], ib[k]);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int j = j0 + ty + 4 * r;
      load_om<P, T, TM>(V, (col_ok && j < jend) ? (unsigned)(i + j * nx) : 0u, co[r], cm[r]);
    }
#pragma unroll
    for (int k = 0; k < R0; ++k)
      if (ty + 4 * k < W0) store_row(jb + ty + 4 * k, ia[k], ib[k]);
  }
  __syncthreads();
  for (int k = 0; k < n_grp; ++k) {
    const bool more = k + 1 < n_grp;
    if (more) issue(k + 1);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int j = j0 + GR * k + ty + 4 * r;
      if (col_ok && j < jend) {
        float v[NF];
        const StripRd<RW, RSX, N> rd{lds + (int)((unsigned)(j - N - jb) % (unsigned)RS) * RW - oi, j};  // j - N >= jb
        upd_any<PH, P>(rd, A, K, i, j, co[r], cm[r], v);
        const unsigned c = (unsigned)(i + j * nx);
        if constexpr (PH == kCpVel) {
          // The position half-step of this interior cell needs only its own new rates (interior
          // cells take no Higdon blend): P += dt dP/dt, S += dt dS/dt, then the source into P.
          const int s = A.parity, o = 1 - s;
          float pn = rd(0, i, j) + K.dt * v[0];
          const float sn = rd(1, i, j) + K.dt * v[1];
          if (j >= src_lo && j <= src_hi) inject<1>(S, 1, (int)c, it, A.nt, &pn);
          st_field<T>(S.fld[o][0], c, pn, S.fenc[0], S.ovf);
          st_field<T>(S.fld[o][1], c, sn, S.fenc[1], S.ovf);
          if (snap_k) {
            const size_t so = (size_t)(i - pad) + (size_t)(j - pad) * nxi;
            G(snap_k)[so] = pn;
            G(snap_k + (S.snap2 - S.snap))[so] = sn;
          }
        } else {
          if (j >= src_lo && j <= src_hi) inject<NF>(S, mask, (int)c, it, A.nt, v);
        }
#pragma unroll
        for (int f = 0; f < NF; ++f) st_field<T>(V.out[f], c, v[f], V.oenc[f], S.ovf);
        if constexpr (SF >= 0) {
          if (snap_k) G(snap_k)[(size_t)(i - pad) + (size_t)(j - pad) * nxi] = round_trip<T>(v[SF], V.oenc[SF], V.odec[SF]);
        }
      }
    }
    if (!more) break;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      store_row(j0 + GR * (k + 1) + N + ty + 4 * r, ra[r], rb[r]);
#pragma unroll
      for (int f = 0; f < NF; ++f) co[r][f] = no[r][f];
#pragma unroll
      for (int q = 0; q < NM; ++q) cm[r][q] = nm[r][q];
    }
    __syncthreads();
  }
}

// Fused acoustic interior strip: the whole step (new vx, vz, then new p) of a strip of
// 257 - 2N columns by W.h rows in one pass down its rows.  The separate sweeps move 48 B per cell
// (read p, vx, vz, bx, bz, write vx, vz; read vx, vz, p, kappa, write p); here each of p, vx, vz,
// bx, bz, kappa is read and each of vx, vz, p written once (36 B), plus halos shared through L2.
//
// Lane k of the 256 owns column x0 + k, x0 = i0 - (2N - 1), and keeps that column's p rows
// tau - LP .. tau + N and new vz rows tau - LP - N .. tau in registers, so the z-derivatives of
// the velocity and pressure updates need no shared memory.  At row tau the block
//   A. computes new vz(tau) (lanes of the strip's columns use it), stages p(tau) with its x-halo
//      into one of two LDS rows, and computes new p at row t = tau - LP on the strip's columns
//      from the vx ring, its vz queue and its p queue;
//   B. after one barrier, computes new vx(tau) on the 256 columns i0 - N + 1 .. i0 + W + N - 1
//      the pressure's x-stencil reads, into a ring of LP + 1 LDS rows.
// The pressure row lags the velocity row by LP = max(N - 1, 2): its z-stencil reaches new vz(t +
// N - 1), and new vx(t), written in phase B of row t, must be behind a barrier (the one of row
// t + 1) before phase A of row t + 2 reads it.
// Strips of odd row bands march up instead (UP: rows in decreasing order, the pressure row t =
// tau + LP with LP = max(N, 2), since its z-stencil now trails at new vz(t - N); p queue rows
// tau - N + 1 .. tau + LP, vz queue rows tau .. tau + LP + N - 1).  Vertically adjacent strips
// then reach the rows they both read at the same time, at the start or at the end of their
// marches, so the second read finds them in the caches (the Infinity Cache, going by the L2 miss
// counts, which do not fall) instead of HBM.  New velocities of cells in the user's interior
// are recomputed from the old state (buffers [s], read-only during the launch); those of frame
// cells are read from the frame velocity launch, which runs first and writes only frame cells
// (buffers [1-s]; this launch writes only interior cells).  Rows are the same for every lane, so
// row bounds, row offsets and the choice of source buffer are scalar.  Every load of a row is
// unconditional (rows past the grid read clamped cells, zeroed where the separate sweeps' windows
// hold zeros), so the number of loads in flight is the same on every path and the wait before a
// use counts only the older loads.  Values are the separate sweeps' expressions in their order
// (and, with 16-bit storage, the stored values the pressure sweep would read back), so the result
// is bit-identical to them.
template <int N, bool UP = false>
struct MarchGeom {
  static constexpr int NT = 256;                      // lanes
  static constexpr int HX = 2 * N - 1;                // lanes left of the strip's first column
  static constexpr int W = NT + 1 - 2 * N;            // columns of a strip
  static constexpr int SW = NT + 2 * N - 1;           // staged p row: columns x0 .. x0 + SW - 1
  // pressure row = velocity row - LP (down) or + LP (up)
  static constexpr int LP = UP ? (N > 2 ? N : 2) : (N - 1 > 2 ? N - 1 : 2);
  static constexpr int NQ = UP ? LP + N : LP + N + 1;  // p queue: rows tau - LP .. tau + N (up: tau + LP .. tau - N + 1)
  static constexpr int NZ = UP ? LP + N : LP + N + 1;  // vz queue: rows tau - LP - N .. tau (up: tau + LP + N - 1 .. tau)
  static constexpr int NR = LP + 1;                   // rows of the vx ring
  // rows of loads in flight: at 8192^2 FD8 the interior launch takes 0.561 / 0.557 / 0.561 ms
  // with 2 / 3 / 4 rows in Float32 and 0.521 / 0.545 / 0.603 ms in Float16
  static constexpr int D = 2;
  static constexpr int LDS = 2 * SW + NR * NT;
};

int fused2d_width(int eq, int N) { return eq == FOMO_ACOUSTIC2D ? 257 - 2 * N : 256 - 2 * N; }

template <int N, class T, class TM, bool UP>
__device__ __forceinline__ void march_strip(const SweepArgs& A, const Work2D& W, const Shot2D& S, float* lds,
                                            int it) {
  using Gm = MarchGeom<N, UP>;
  constexpr int NT = Gm::NT, HX = Gm::HX, SW = Gm::SW, LP = Gm::LP, NQ = Gm::NQ, NZ = Gm::NZ, NR = Gm::NR,
                D = Gm::D;
  const int nx = S.nx, nz = S.nz, pad = S.H + 1;
  const float dtx = S.dtx;
  const int s = A.parity, o = 1 - s;
  const int i0 = W.i0, j0 = W.j0, h = W.h, w = W.w;
  const int k = threadIdx.x;
  const int x0 = i0 - HX;
  const int cp = x0 + k;                      // this lane's p / vz / pressure column
  const int cv = x0 + N + k;                  // this lane's vx column
  const int ce = x0 + NT + min(k, HX - 1);    // extra staged p column (lanes k < HX)
  auto clampx = [=](int i) { return min(max(i, 0), nx - 1); };
  auto rowc = [=](int j) { return min(max(j, 0), nz - 1); };
  auto in_z = [=](int j) { return j >= 0 && j < nz; };
  auto core_z = [=](int j) { return j >= pad && j < nz - pad; };
  const bool inp = cp >= 0 && cp < nx, inv = cv >= 0 && cv < nx, ine = ce >= 0 && ce < nx;
  const int cpc = clampx(cp), cvc = clampx(cv), cec = clampx(ce);
  const bool own = k >= HX && k - HX < w;     // a column of the strip
  const bool core_v = cv >= pad && cv < nx - pad;
  const bool store_v = cv >= i0 && cv < i0 + w;
  const void* const p_s = S.fld[s][0];
  const void* const vz_s = S.fld[s][2];
  const void* const vz_o = S.fld[o][2];
  const void* const vx_src = core_v ? S.fld[s][1] : S.fld[o][1];
  void* const p_w = S.fld[o][0];
  void* const vx_w = S.fld[o][1];
  void* const vz_w = S.fld[o][2];
  auto ld = [=](const void* base, int f, int j, int i) -> float {
    return Cvt<T>::ld(base, (unsigned)(j * nx) + (unsigned)i, S.fdec[f]);
  };
  auto ldm = [=](int m, int j, int i) -> float {
    return Cvt<TM>::ld(S.mat[m], (unsigned)(j * nx) + (unsigned)i, S.mdec[m]);
  };
  float a[N];
#pragma unroll
  for (int l = 0; l < N; ++l) a[l] = A.a[l];

  // Sources in this lane's column within the strip, scanned once before the loop: the rows and
  // amplitudes of this step of at most two (the host keeps the separate sweeps for plans with more
  // sources in one column), in source order, so that sources sharing a cell add in that order.
  int srow0 = -1, srow1 = -1;
  float sw0 = 0.0f, sw1 = 0.0f;
  if (it < A.nt && own) {
    int q2 = 0;
    for (int q = 0; q < S.n_src; ++q) {
      const int cs = G(S.src_c)[q];
      const int js = cs / nx, is = cs - js * nx;
      if (is == cp && js >= j0 && js < j0 + h) {
        const float wq = G(S.wav)[(size_t)q + (size_t)S.n_src * (size_t)it];
        if (q2 == 0) {
          srow0 = js;
          sw0 = wq;
        } else {
          srow1 = js;
          sw1 = wq;
        }
        ++q2;
      }
    }
  }
  // snapshot of this step (p), written by the pressure rows
  const int nxi = nx - 2 * pad;
  float* snap_k = nullptr;
  if (S.snap && (it + 1) % S.snap_every == 0 && (it + 1) / S.snap_every <= S.snap_n)
    snap_k = S.snap + (size_t)((it + 1) / S.snap_every - 1) * (size_t)nxi * (size_t)(nz - 2 * pad);

  float* const stg = lds;            // two staged p rows (row tau in stg + (tau & 1) * SW)
  float* const vxr = lds + 2 * SW;   // new vx of row r in vxr + (r mod NR) * NT

  // Row offsets from the velocity row tau: DIR the march step, TOFF the pressure row, LEAD the p row
  // that enters the queue.  Queue slots in march order (oldest first): qi / zi of a row offset.
  constexpr int DIR = UP ? -1 : 1, TOFF = UP ? LP : -LP, LEAD = UP ? 1 - N : N;
  auto qi = [](int off) constexpr { return UP ? LP - off : off + LP; };
  auto zi = [](int off) constexpr { return UP ? LP + N - 1 - off : off + LP + N; };
  // loads of row rho: p(rho + LEAD), extra p(rho), old / frame-final vx(rho) and bx(rho) at the vx
  // column, old / frame-final vz(rho) and bz(rho), kappa(rho + TOFF)
  struct Pre {
    float p, pe, vxo, bx, vzo, bz, kap;
  };
  auto prefetch = [=](int rho, Pre& P) __attribute__((always_inline)) {
    const int rp = rowc(rho + LEAD), r0 = rowc(rho), rk = rowc(rho + TOFF);
    P.p = ld(p_s, 0, rp, cpc);
    P.pe = ld(p_s, 0, r0, cec);
    P.vxo = ld(vx_src, 1, r0, cvc);
    P.bx = ldm(1, r0, cvc);
    P.vzo = ld(core_z(rho) ? vz_s : vz_o, 2, r0, cpc);
    P.bz = ldm(2, r0, cpc);
    P.kap = ldm(0, rk, cpc);
  };

  // rows j0 - N .. j0 + h - 1 + LP down, or j0 + h + N - 2 .. j0 - LP up (the new vz rows the
  // pressure rows j0 .. j0 + h - 1 read, and the velocity rows of those pressure rows)
  const int tau0 = UP ? j0 + h + N - 2 : j0 - N;
  const int nit = (h + N + LP - (UP ? 1 : 0) + D - 1) / D;
  Pre pre[D];
#pragma unroll
  for (int d = 0; d < D; ++d) prefetch(tau0 + DIR * d, pre[d]);
  // p queue before the first row: slot i >= 1 holds p at row offset (qi^-1)(i) from tau0 - DIR
  float Q[NQ], Z[NZ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) Q[i] = 0.0f;
#pragma unroll
  for (int i = 0; i < NZ; ++i) Z[i] = 0.0f;
  {
    float q0[NQ - 1];
#pragma unroll
    for (int i = 1; i < NQ; ++i) q0[i - 1] = ld(p_s, 0, rowc(tau0 - DIR + (UP ? LP - i : i - LP)), cpc);
#pragma unroll
    for (int i = 1; i < NQ; ++i) Q[i] = (inp && in_z(tau0 - DIR + (UP ? LP - i : i - LP))) ? q0[i - 1] : 0.0f;
  }

  for (int g = 0; g < nit; ++g) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int tau = tau0 + DIR * (g * D + d);
      Pre& P = pre[d];
      // A. p queue, new vz(tau), staged p(tau), new p(tau + TOFF)
#pragma unroll
      for (int i = 0; i < NQ - 1; ++i) Q[i] = Q[i + 1];
      Q[NQ - 1] = (inp && in_z(tau + LEAD)) ? P.p : 0.0f;
      float vz = in_z(tau) ? P.vzo : 0.0f;
      if (core_z(tau)) {
        float dz = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) dz += a[l - 1] * (Q[qi(l)] - Q[qi(1 - l)]);
        vz = P.vzo + (P.bz * dtx) * dz;
        if (own && tau >= j0 && tau < j0 + h)
          st_field<T>(vz_w, (unsigned)(tau * nx) + (unsigned)cp, vz, S.fenc[2], S.ovf);
        vz = round_trip<T>(vz, S.fenc[2], S.fdec[2]);
      }
#pragma unroll
      for (int i = 0; i < NZ - 1; ++i) Z[i] = Z[i + 1];
      Z[NZ - 1] = vz;
      float* const sb = stg + (tau & 1) * SW;
      sb[k] = Q[qi(0)];
      if (k < HX) sb[NT + k] = (ine && in_z(tau)) ? P.pe : 0.0f;
      const int t = tau + TOFF;
      if (t >= j0 && t < j0 + h && own) {
        const float* VX = vxr + (t % NR) * NT + k - N;
        float dvxdx = 0.0f, dvzdz = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) {
          const float cl = a[l - 1];
          dvxdx += cl * (VX[l] - VX[1 - l]);
          dvzdz += cl * (Z[zi(TOFF + l - 1)] - Z[zi(TOFF - l)]);
        }
        float v = Q[qi(TOFF)] + P.kap * (dvxdx * dtx + dvzdz * dtx);
        if (t == srow0) v += sw0;
        if (t == srow1) v += sw1;
        st_field<T>(p_w, (unsigned)(t * nx) + (unsigned)cp, v, S.fenc[0], S.ovf);
        if (snap_k) G(snap_k)[(size_t)(cp - pad) + (size_t)(t - pad) * nxi] = round_trip<T>(v, S.fenc[0], S.fdec[0]);
      }
      __syncthreads();
      // B. new vx(tau) on the 256 vx columns
      if (tau >= j0 && tau < j0 + h) {
        const float* PS = sb + k + N;
        float v = inv ? P.vxo : 0.0f;
        if (core_v) {
          float dx = 0.0f;
#pragma unroll
          for (int l = 1; l <= N; ++l) dx += a[l - 1] * (PS[l - 1] - PS[-l]);
          v = P.vxo + (P.bx * dtx) * dx;
          if (store_v) st_field<T>(vx_w, (unsigned)(tau * nx) + (unsigned)cv, v, S.fenc[1], S.ovf);
          v = round_trip<T>(v, S.fenc[1], S.fdec[1]);
        }
        vxr[(tau % NR) * NT + k] = v;
      }
      prefetch(tau + DIR * D, P);
    }
  }
}

// Fused elastic interior strip: the whole step (new vx, vz, then new txx, tzz, txz) of a strip of
// 256 - 2N columns by W.h rows in one pass down its rows.  The separate sweeps move 80 B per cell
// (velocity: read txx, tzz, txz, vx, vz, bx, bz, write vx, vz; stress: read vx, vz, txx, tzz, txz,
// lam, lam2mu, muxz, write txx, tzz, txz); here each of the ten arrays is read and each of the
// five fields written once (60 B), plus halos shared through L2.
//
// Lane k of the 256 owns column c = i0 - N + k; the strip's columns are lanes N .. N+w-1.  Each lane
// reads each stress of its column once, into register queues that reach from the stress row
// t = tau - LP (its old value) to the velocities' z-stencils (txz rows up to tau + N - 1, tzz up
// to tau + N), and keeps its column's new vx and vz of the last 2N + 1 rows (the stresses'
// z-stencils).  At row tau the block
//   A. stages txx(tau) and txz(tau) from the queues (plus 2N halo columns) into one of two LDS
//      rows, and computes the new stresses of row t on the strip's columns from the velocity rings
//      (x) and queues (z);
//   B. after one barrier, computes new vx(tau), vz(tau) on all 256 columns (the stresses' x-stencils
//      reach N columns past the strip) into rings of LP + 1 LDS rows.
// The stress row lags by LP = N + 1: its z-stencil reaches new vx(t + N), computed in phase B of
// row t + N and in the lane's queue from row t + N + 1 on; and the rings' row t, written in phase
// B of row t, is behind the barrier of row t + 1 before phase A of row t + 2 reads it.  Strips of
// odd row bands march up (UP: the stress row t = tau + LP, whose z-stencil trails at new vz(t - N);
// txz queue rows tau + LP .. tau - N, tzz tau + LP .. tau - N + 1, new velocities tau + 2N + 1 ..
// tau + 1), so that vertically adjacent strips read their shared rows at the same time and the
// second read finds them in the caches.  As in the
// acoustic march, velocities of cells in the user's interior are recomputed from the old state
// (buffers [s]) and those of frame cells read from the frame velocity launch, which runs first
// (buffers [1-s]); every value is the separate sweeps' expression in their order
// (src/equations/elastic2d/fused_kernels.jl:21-124), and 16-bit storage rounds the velocities as
// the stress sweep would read them back, so the result is bit-identical to the separate sweeps.
template <int N, bool UP = false>
struct MarchGeomEl {
  static constexpr int NT = 256;
  static constexpr int W = NT - 2 * N;       // columns of a strip
  static constexpr int SW = NT + 2 * N;      // staged row: columns c(0) - N .. c(255) + N
  static constexpr int LP = N + 1;           // stress row = velocity row - LP (up: + LP)
  static constexpr int NXX = LP + 1;         // txx rows t .. tau
  static constexpr int NXZ = UP ? 2 * N + 2 : 2 * N + 1;  // txz rows t .. tau + N - 1 (up: t .. tau - N)
  static constexpr int NZZ = UP ? 2 * N + 1 : 2 * N + 2;  // tzz rows t .. tau + N (up: t .. tau - N + 1)
  static constexpr int NV = 2 * N + 1;       // new vx / vz queues: rows tau - NV .. tau - 1 at phase A (up: tau + NV .. tau + 1)
  static constexpr int NR = LP + 1;          // rows of the velocity rings
  static constexpr int D = 2;                // rows of loads in flight
  static constexpr int LDS = 2 * 2 * SW + 2 * NR * NT;
};

template <int N, class T, class TM, bool UP>
__device__ __forceinline__ void march_strip_el(const SweepArgs& A, const Work2D& W, const Shot2D& S, float* lds,
                                               int it) {
  using Gm = MarchGeomEl<N, UP>;
  constexpr int NT = Gm::NT, SW = Gm::SW, LP = Gm::LP, NXX = Gm::NXX, NXZ = Gm::NXZ, NZZ = Gm::NZZ, NV = Gm::NV,
                NR = Gm::NR, D = Gm::D;
  const int nx = S.nx, nz = S.nz, pad = S.H + 1;
  const float dtx = S.dtx;
  const int s = A.parity, o = 1 - s;
  const int i0 = W.i0, j0 = W.j0, h = W.h, w = W.w;
  const int k = threadIdx.x;
  const int c0 = i0 - N;
  const int c = c0 + k;                       // this lane's column
  // halo column staged by lanes k < 2N (slot k: columns left of lane 0; slot NT + k: right of lane
  // 255); the other lanes repeat lane 2N - 1's cell (one cache line) and stage nothing
  const int ke = min(k, 2 * N - 1);
  const int ce = ke < N ? c0 - N + ke : c0 + NT + (ke - N);
  const int se = ke < N ? ke : NT + ke;
  auto clampx = [=](int i) { return min(max(i, 0), nx - 1); };
  auto rowc = [=](int j) { return min(max(j, 0), nz - 1); };
  auto in_z = [=](int j) { return j >= 0 && j < nz; };
  auto core_z = [=](int j) { return j >= pad && j < nz - pad; };
  const bool inc = c >= 0 && c < nx, ine = ce >= 0 && ce < nx;
  const int cc = clampx(c), cce = clampx(ce);
  const bool own = k >= N && k - N < w;
  const bool core_x = c >= pad && c < nx - pad;
  // array bases and scales, read from the shot table once: old state [s], this launch's outputs and
  // the frame velocity launch's results [o], moduli and buoyancies
  const void* const vx_s = S.fld[s][0];
  const void* const vz_s = S.fld[s][1];
  const void* const xx_s = S.fld[s][2];
  const void* const zz_s = S.fld[s][3];
  const void* const xz_s = S.fld[s][4];
  void* const vx_o = S.fld[o][0];
  void* const vz_o = S.fld[o][1];
  void* const xx_o = S.fld[o][2];
  void* const zz_o = S.fld[o][3];
  void* const xz_o = S.fld[o][4];
  const void* const m_lam = S.mat[0];
  const void* const m_l2m = S.mat[1];
  const void* const m_mu = S.mat[2];
  const void* const m_bx = S.mat[3];
  const void* const m_bz = S.mat[4];
  const float fdec[5] = {S.fdec[0], S.fdec[1], S.fdec[2], S.fdec[3], S.fdec[4]};
  const float fenc[5] = {S.fenc[0], S.fenc[1], S.fenc[2], S.fenc[3], S.fenc[4]};
  const float mdec[5] = {S.mdec[0], S.mdec[1], S.mdec[2], S.mdec[3], S.mdec[4]};
  unsigned* const ovf = S.ovf;
  auto ld = [=](const void* base, int f, int j, int i) -> float {
    return Cvt<T>::ld(base, (unsigned)(j * nx) + (unsigned)i, fdec[f]);
  };
  auto ldm = [=](const void* base, int m, int j, int i) -> float {
    return Cvt<TM>::ld(base, (unsigned)(j * nx) + (unsigned)i, mdec[m]);
  };
  float a[N];
#pragma unroll
  for (int l = 0; l < N; ++l) a[l] = A.a[l];

  // Sources in this lane's column (at most two per column, in source order, so that sources
  // sharing a cell add in that order): a vertical force goes into vz on every lane that computes
  // that column's velocity, an explosive source into txx and tzz of the strip's own cells.
  int srow0 = -1, srow1 = -1;
  float sw0 = 0.0f, sw1 = 0.0f;
  if (it < A.nt && (S.src_fz || own)) {
    int q2 = 0;
    for (int q = 0; q < S.n_src; ++q) {
      const int csrc = G(S.src_c)[q];
      const int js = csrc / nx, is = csrc - js * nx;
      if (is == c) {
        const float wq = G(S.wav)[(size_t)q + (size_t)S.n_src * (size_t)it];
        if (q2 == 0) {
          srow0 = js;
          sw0 = wq;
        } else if (q2 == 1) {
          srow1 = js;
          sw1 = wq;
        }
        ++q2;
      }
    }
  }
  const bool force = S.src_fz != 0;
  // snapshot of this step (vz), written by the velocity rows
  const int nxi = nx - 2 * pad;
  float* snap_k = nullptr;
  if (S.snap && (it + 1) % S.snap_every == 0 && (it + 1) / S.snap_every <= S.snap_n)
    snap_k = S.snap + (size_t)((it + 1) / S.snap_every - 1) * (size_t)nxi * (size_t)(nz - 2 * pad);

  float* const stg = lds;                    // two staged rows of txx and txz: stg + (tau & 1) * 2 * SW
  float* const vxr = lds + 4 * SW;           // new vx of row r at vxr + (r mod NR) * NT
  float* const vzr = vxr + NR * NT;          // new vz likewise

  // Row offsets from the velocity row tau: DIR the march step, TOFF the stress row, LXZ / LZZ the
  // txz / tzz rows that enter their queues.  Queue slots in march order (oldest first): sq / sv of
  // a row offset (stress queues, new-velocity queues).
  constexpr int DIR = UP ? -1 : 1, TOFF = UP ? LP : -LP, LXZ = UP ? -N : N - 1, LZZ = UP ? 1 - N : N;
  auto sq = [](int off) constexpr { return UP ? LP - off : off + LP; };
  auto sv = [](int off) constexpr { return UP ? NV - off : off + NV; };
  // loads of row rho: txx(c, rho) and the halo column's txx, txz at rho; txz(c, rho + LXZ) and
  // tzz(c, rho + LZZ) for the queues; old (interior) or frame-final vx, vz and bx, bz at (c, rho);
  // the moduli at (c, rho + TOFF)
  struct Pre {
    float xx, xxe, xze, qxz, qzz, vxo, vzo, bx, bz, lam, l2m, mu;
  };
  auto prefetch = [=](int rho, Pre& P) __attribute__((always_inline)) {
    const int r0 = rowc(rho), rq = rowc(rho + LXZ), rz = rowc(rho + LZZ), rt = rowc(rho + TOFF);
    P.xx = ld(xx_s, 2, r0, cc);
    P.xxe = ld(xx_s, 2, r0, cce);
    P.xze = ld(xz_s, 4, r0, cce);
    P.qxz = ld(xz_s, 4, rq, cc);
    P.qzz = ld(zz_s, 3, rz, cc);
    const bool cz = core_x && core_z(rho);
    P.vxo = ld(cz ? vx_s : (const void*)vx_o, 0, r0, cc);
    P.vzo = ld(cz ? vz_s : (const void*)vz_o, 1, r0, cc);
    P.bx = ldm(m_bx, 3, r0, cc);
    P.bz = ldm(m_bz, 4, r0, cc);
    P.lam = ldm(m_lam, 0, rt, cc);
    P.l2m = ldm(m_l2m, 1, rt, cc);
    P.mu = ldm(m_mu, 2, rt, cc);
  };

  // rows j0 - N .. j0 + h + N down, or j0 + h + N - 1 .. j0 - N - 1 up: the velocity rows the stress
  // rows j0 .. j0 + h - 1 read, and the velocity rows of those stress rows
  const int tau0 = UP ? j0 + h + N - 1 : j0 - N;
  const int nit = (h + N + LP + D - 1) / D;
  Pre pre[D];
#pragma unroll
  for (int d = 0; d < D; ++d) prefetch(tau0 + DIR * d, pre[d]);
  // Queues at phase A of row tau (after that row's shift): Qxx[sq(o)] = txx(tau + o), Qxz[sq(o)] =
  // txz(tau + o), Qzz[sq(o)] = tzz(tau + o); QVX[sv(o)] / QVZ[sv(o)] hold new vx / vz of row
  // tau + o.  Before the first row they hold the rows one earlier; txx rows before tau0 only ever
  // serve stress rows outside the strip and start as zeros.
  float Qxx[NXX], Qxz[NXZ], Qzz[NZZ], QVX[NV], QVZ[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) QVX[i] = QVZ[i] = 0.0f;
#pragma unroll
  for (int i = 0; i < NXX; ++i) Qxx[i] = 0.0f;
  Qxz[0] = Qzz[0] = 0.0f;
  {
    float q0[NXZ - 1], q1[NZZ - 1];
    // slot i >= 1 holds the row at offset (sq^-1)(i) from tau0 - DIR
    auto row_of = [&](int i) { return tau0 - DIR + (UP ? LP - i : i - LP); };
#pragma unroll
    for (int i = 1; i < NXZ; ++i) q0[i - 1] = ld(xz_s, 4, rowc(row_of(i)), cc);
#pragma unroll
    for (int i = 1; i < NZZ; ++i) q1[i - 1] = ld(zz_s, 3, rowc(row_of(i)), cc);
#pragma unroll
    for (int i = 1; i < NXZ; ++i) Qxz[i] = (inc && in_z(row_of(i))) ? q0[i - 1] : 0.0f;
#pragma unroll
    for (int i = 1; i < NZZ; ++i) Qzz[i] = (inc && in_z(row_of(i))) ? q1[i - 1] : 0.0f;
  }

  for (int g = 0; g < nit; ++g) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int tau = tau0 + DIR * (g * D + d);
      Pre& P = pre[d];
      // A. queues, staged txx / txz of row tau, new stresses of row t = tau + TOFF
      const bool iz = in_z(tau);
#pragma unroll
      for (int i = 0; i < NXX - 1; ++i) Qxx[i] = Qxx[i + 1];
#pragma unroll
      for (int i = 0; i < NXZ - 1; ++i) Qxz[i] = Qxz[i + 1];
#pragma unroll
      for (int i = 0; i < NZZ - 1; ++i) Qzz[i] = Qzz[i + 1];
      Qxx[NXX - 1] = (inc && iz) ? P.xx : 0.0f;
      Qxz[NXZ - 1] = (inc && in_z(tau + LXZ)) ? P.qxz : 0.0f;
      Qzz[NZZ - 1] = (inc && in_z(tau + LZZ)) ? P.qzz : 0.0f;
      float* const sb = stg + (tau & 1) * 2 * SW;
      sb[k + N] = Qxx[sq(0)];        // txx(c, tau)
      sb[SW + k + N] = Qxz[sq(0)];   // txz(c, tau)
      if (k < 2 * N) {
        sb[se] = (ine && iz) ? P.xxe : 0.0f;
        sb[SW + se] = (ine && iz) ? P.xze : 0.0f;
      }
      const int t = tau + TOFF;
      if (t >= j0 && t < j0 + h && own) {
        const float* VX = vxr + (t % NR) * NT + k;
        const float* VZ = vzr + (t % NR) * NT + k;
        float dvxdx = 0.0f, dvzdz = 0.0f, dvxdz = 0.0f, dvzdx = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) {
          const float cl = a[l - 1];
          dvxdx += cl * (VX[l] - VX[1 - l]);
          dvzdz += cl * (QVZ[sv(TOFF + l - 1)] - QVZ[sv(TOFF - l)]);
          dvxdz += cl * (QVX[sv(TOFF + l)] - QVX[sv(TOFF + 1 - l)]);
          dvzdx += cl * (VZ[l - 1] - VZ[-l]);
        }
        const float dt = dtx;
        float txx = Qxx[sq(TOFF)] + ((P.l2m * dvxdx) * dt + (P.lam * dvzdz) * dt);
        float tzz = Qzz[sq(TOFF)] + ((P.lam * dvxdx) * dt + (P.l2m * dvzdz) * dt);
        const float txz = Qxz[sq(TOFF)] + P.mu * (dvxdz * dt + dvzdx * dt);
        if (!force) {
          if (t == srow0) {
            txx += sw0;
            tzz += sw0;
          }
          if (t == srow1) {
            txx += sw1;
            tzz += sw1;
          }
        }
        const unsigned co = (unsigned)(t * nx) + (unsigned)c;
        st_field<T>(xx_o, co, txx, fenc[2], ovf);
        st_field<T>(zz_o, co, tzz, fenc[3], ovf);
        st_field<T>(xz_o, co, txz, fenc[4], ovf);
      }
      __syncthreads();
      // B. new vx(tau), vz(tau) on the 256 columns, into the rings and the lane's queues
      {
        float vx = inc && iz ? P.vxo : 0.0f, vz = inc && iz ? P.vzo : 0.0f;
        if (core_x && core_z(tau)) {
          const float* XX = sb + k + N;
          const float* XZ = sb + SW + k + N;
          float dtxxdx = 0.0f, dtxzdz = 0.0f, dtxzdx = 0.0f, dtzzdz = 0.0f;
#pragma unroll
          for (int l = 1; l <= N; ++l) {
            const float cl = a[l - 1];
            dtxxdx += cl * (XX[l - 1] - XX[-l]);
            dtxzdz += cl * (Qxz[sq(l - 1)] - Qxz[sq(-l)]);
            dtxzdx += cl * (XZ[l] - XZ[1 - l]);
            dtzzdz += cl * (Qzz[sq(l)] - Qzz[sq(1 - l)]);
          }
          const float dt = dtx;
          vx = P.vxo + P.bx * (dt * dtxxdx + dt * dtxzdz);
          vz = P.vzo + P.bz * (dt * dtxzdx + dt * dtzzdz);
          if (force) {
            if (tau == srow0) vz += sw0;
            if (tau == srow1) vz += sw1;
          }
          if (own && tau >= j0 && tau < j0 + h) {
            const unsigned co = (unsigned)(tau * nx) + (unsigned)c;
            st_field<T>(vx_o, co, vx, fenc[0], ovf);
            st_field<T>(vz_o, co, vz, fenc[1], ovf);
          }
          vx = round_trip<T>(vx, fenc[0], fdec[0]);
          vz = round_trip<T>(vz, fenc[1], fdec[1]);
          if (snap_k && own && tau >= j0 && tau < j0 + h)
            G(snap_k)[(size_t)(c - pad) + (size_t)(tau - pad) * nxi] = vz;
        }
        // an up strip's last row can be -1 (nbc = 0, j0 = N): its ring slot is NR - 1
        const int rs = (tau + NR) % NR;
        vxr[rs * NT + k] = vx;
        vzr[rs * NT + k] = vz;
#pragma unroll
        for (int i = 0; i < NV - 1; ++i) {
          QVX[i] = QVX[i + 1];
          QVZ[i] = QVZ[i + 1];
        }
        QVX[NV - 1] = vx;
        QVZ[NV - 1] = vz;
      }
      prefetch(tau + DIR * D, P);
    }
  }
}

// Scalar interior strip (scalar2d leapfrog, u^{it+1} = (2 u^it - u^{it-1}) + c2 lap u^it with the
// separate sweep's expression order, src/equations/scalar2d/scalar2d.jl): one pass down (or, kind
// 2, up) the rows of a strip of 256 - 2N columns.  Lane k owns column c = i0 - N + k, the strip's
// columns are lanes N .. N+w-1, so every x-neighbour of a strip cell is another lane's column and
// nothing past the 256 columns is staged.  Each lane keeps u^it of its column over rows tau - N ..
// tau + N in a register queue (the z-stencil), stages u^it(tau) in one of two LDS rows (the
// x-stencil) and, after one barrier, writes u^{it+1}(tau): the update reads only the state before
// the step, so there is no lag and one barrier per row.  Sources (at most two per column, in
// source order) add to u^{it+1} as in the sweep, 16-bit storage as there too; like the sweep (and
// the reference's scalar2d) it records no snapshots.
template <int N, class T, bool UP>
__device__ __forceinline__ void march_strip_sc(const SweepArgs& A, const Work2D& W, const Shot2D& S, float* lds,
                                               int it) {
  constexpr int NT = 256, NQ = 2 * N + 1, D = 2;
  constexpr int DIR = UP ? -1 : 1, LEAD = UP ? -N : N;
  auto qi = [](int off) constexpr { return UP ? N - off : N + off; };  // queue slot of row tau + off
  const int nx = S.nx, nz = S.nz;
  const int i0 = W.i0, j0 = W.j0, h = W.h, w = W.w;
  const int k = threadIdx.x;
  const int c = i0 - N + k;
  const bool inc = c >= 0 && c < nx;
  const int cc = min(max(c, 0), nx - 1);
  const bool own = k >= N && k - N < w;
  auto rowc = [=](int j) { return min(max(j, 0), nz - 1); };
  auto in_z = [=](int j) { return j >= 0 && j < nz; };
  const int cur = it % 3, prv = (it + 2) % 3, nxt = (it + 1) % 3;
  const void* const u_s = S.fld[0][cur];
  const void* const up_s = S.fld[0][prv];
  void* const u_o = S.fld[0][nxt];
  const void* const c2_s = S.mat[0];
  const float dec = S.fdec[0], enc = S.fenc[0], mdec = S.mdec[0];
  unsigned* const ovf = S.ovf;
  const float c0 = A.c0;
  auto ld = [=](const void* base, int j, int i) -> float {
    return Cvt<T>::ld(base, (unsigned)(j * nx) + (unsigned)i, dec);
  };
  float a[N];
#pragma unroll
  for (int l = 0; l < N; ++l) a[l] = A.a[l];
  int srow0 = -1, srow1 = -1;
  float sw0 = 0.0f, sw1 = 0.0f;
  if (it < A.nt && own) {
    int q2 = 0;
    for (int q = 0; q < S.n_src; ++q) {
      const int cs = G(S.src_c)[q];
      const int js = cs / nx, is = cs - js * nx;
      if (is == c && js >= j0 && js < j0 + h) {
        const float wq = G(S.wav)[(size_t)q + (size_t)S.n_src * (size_t)it];
        if (q2 == 0) {
          srow0 = js;
          sw0 = wq;
        } else {
          srow1 = js;
          sw1 = wq;
        }
        ++q2;
      }
    }
  }
  // loads of row rho: u^it(rho + LEAD) for the queue, u^{it-1}(rho) and c2(rho)
  struct Pre {
    float u, up, c2;
  };
  auto prefetch = [=](int rho, Pre& P) __attribute__((always_inline)) {
    P.u = ld(u_s, rowc(rho + LEAD), cc);
    P.up = ld(up_s, rowc(rho), cc);
    P.c2 = Cvt<T>::ld(c2_s, (unsigned)(rowc(rho) * nx) + (unsigned)cc, mdec);
  };
  const int tau0 = UP ? j0 + h - 1 : j0;
  const int nit = (h + D - 1) / D;
  Pre pre[D];
#pragma unroll
  for (int d = 0; d < D; ++d) prefetch(tau0 + DIR * d, pre[d]);
  // queue before the first row: slot i >= 1 holds row tau0 - DIR + (UP ? N - i : i - N)
  float Q[NQ];
  Q[0] = 0.0f;
  {
    float q0[NQ - 1];
#pragma unroll
    for (int i = 1; i < NQ; ++i) q0[i - 1] = ld(u_s, rowc(tau0 - DIR + (UP ? N - i : i - N)), cc);
#pragma unroll
    for (int i = 1; i < NQ; ++i) Q[i] = (inc && in_z(tau0 - DIR + (UP ? N - i : i - N))) ? q0[i - 1] : 0.0f;
  }
  for (int g = 0; g < nit; ++g) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int tau = tau0 + DIR * (g * D + d);
      Pre& P = pre[d];
      if (g * D + d < h) {
#pragma unroll
        for (int i = 0; i < NQ - 1; ++i) Q[i] = Q[i + 1];
        Q[NQ - 1] = (inc && in_z(tau + LEAD)) ? P.u : 0.0f;
        float* const sb = lds + (tau & 1) * NT;
        sb[k] = Q[qi(0)];
        __syncthreads();
        if (own) {
          const float uc = Q[qi(0)];
          float lap = (2.0f * c0) * uc;
#pragma unroll
          for (int l = 1; l <= N; ++l) {
            const float cl = a[l - 1];
            lap += cl * (((sb[k + l] + sb[k - l]) + Q[qi(l)]) + Q[qi(-l)]);
          }
          float v = (2.0f * uc - P.up) + P.c2 * lap;
          if (tau == srow0) v += sw0;
          if (tau == srow1) v += sw1;
          st_field<T>(u_o, (unsigned)(tau * nx) + (unsigned)c, v, enc, ovf);
        }
      }
      prefetch(tau + DIR * D, P);
    }
  }
}

template <int N, class T>
__global__ __launch_bounds__(256) void fused2d_sc_kernel(const SweepArgs A) {
  __shared__ float lds[2 * 256];
  if ((int)blockIdx.x >= A.n_work) return;
  const Work2D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  const Shot2D& S = A.shots[W.shot];
  if (W.kind == 2)
    march_strip_sc<N, T, true>(A, W, S, lds, *A.step + A.koff);
  else
    march_strip_sc<N, T, false>(A, W, S, lds, *A.step + A.koff);
}

// Work items of kind 2 march up (see march_strip_el), those of kind 1 down.
template <int N, class T, class TM>
__global__ __launch_bounds__(256) void fused2d_el_kernel(const SweepArgs A) {
  __shared__ float lds[std::max(MarchGeomEl<N, false>::LDS, MarchGeomEl<N, true>::LDS)];
  if ((int)blockIdx.x >= A.n_work) return;
  const Work2D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  const Shot2D& S = A.shots[W.shot];
  if (W.kind == 2)
    march_strip_el<N, T, TM, true>(A, W, S, lds, *A.step + A.koff);
  else
    march_strip_el<N, T, TM, false>(A, W, S, lds, *A.step + A.koff);
}

// Work items of kind 2 march up (see march_strip), those of kind 1 down.
template <int N, class T, class TM>
__global__ __launch_bounds__(256) void fused2d_kernel(const SweepArgs A) {
  __shared__ float lds[std::max(MarchGeom<N, false>::LDS, MarchGeom<N, true>::LDS)];
  if ((int)blockIdx.x >= A.n_work) return;
  const Work2D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  const Shot2D& S = A.shots[W.shot];
  if (W.kind == 2)
    march_strip<N, T, TM, true>(A, W, S, lds, *A.step + A.koff);
  else
    march_strip<N, T, TM, false>(A, W, S, lds, *A.step + A.koff);
}

template <int N, class T, class TM>
static hipError_t launch_fused2d_n(const SweepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((fused2d_kernel<N, T, TM>), dim3(a.n_work), dim3(MarchGeom<N>::NT), 0, s, a);
  return hipGetLastError();
}

template <int N, class T, class TM>
static hipError_t launch_fused2d_el_n(const SweepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((fused2d_el_kernel<N, T, TM>), dim3(a.n_work), dim3(MarchGeomEl<N>::NT), 0, s, a);
  return hipGetLastError();
}

template <class T, class TM>
static hipError_t launch_fused2d_t(bool el, int N, const SweepArgs& a, hipStream_t s) {
  switch (N) {
    case 1: return el ? launch_fused2d_el_n<1, T, TM>(a, s) : launch_fused2d_n<1, T, TM>(a, s);
    case 2: return el ? launch_fused2d_el_n<2, T, TM>(a, s) : launch_fused2d_n<2, T, TM>(a, s);
    case 3: return el ? launch_fused2d_el_n<3, T, TM>(a, s) : launch_fused2d_n<3, T, TM>(a, s);
    case 4: return el ? launch_fused2d_el_n<4, T, TM>(a, s) : launch_fused2d_n<4, T, TM>(a, s);
    case 5: return el ? launch_fused2d_el_n<5, T, TM>(a, s) : launch_fused2d_n<5, T, TM>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template <class T>
static hipError_t launch_fused2d_sc(int N, const SweepArgs& a, hipStream_t s) {
  const dim3 grid(a.n_work), blk(256);
  switch (N) {
    case 1: hipLaunchKernelGGL((fused2d_sc_kernel<1, T>), grid, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((fused2d_sc_kernel<2, T>), grid, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((fused2d_sc_kernel<3, T>), grid, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((fused2d_sc_kernel<4, T>), grid, blk, 0, s, a); break;
    case 5: hipLaunchKernelGGL((fused2d_sc_kernel<5, T>), grid, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fused2d(int storage, int eq, int N, const SweepArgs& a, hipStream_t s) {
  if (a.n_work <= 0) return hipSuccess;
  if (eq == FOMO_SCALAR2D) {
    switch (storage) {
      case FOMO_STORE_F32: return launch_fused2d_sc<float>(N, a, s);
      case FOMO_STORE_F16: return launch_fused2d_sc<__half>(N, a, s);
      case FOMO_STORE_BF16: return launch_fused2d_sc<__hip_bfloat16>(N, a, s);
      default: return hipErrorInvalidValue;  // the scalar keeps a wavefield among its materials
    }
  }
  if (eq != FOMO_ACOUSTIC2D && eq != FOMO_ELASTIC2D) return hipErrorInvalidValue;
  const bool el = eq == FOMO_ELASTIC2D;
  switch (storage) {
    case FOMO_STORE_F32: return launch_fused2d_t<float, float>(el, N, a, s);
    case FOMO_STORE_F16: return launch_fused2d_t<__half, __half>(el, N, a, s);
    case FOMO_STORE_BF16: return launch_fused2d_t<__hip_bfloat16, __hip_bfloat16>(el, N, a, s);
    case FOMO_STORE_F32_MAT16: return launch_fused2d_t<float, __half>(el, N, a, s);
    default: return hipErrorInvalidValue;
  }
}

// Receiver entry e: the value of every slot's field in buffer [parity] into column col
// (src/acquisition/device_indexed.jl:100-134).
template <class T>
__device__ __forceinline__ void gather_entry(const GatherArgs& g, int e, int col, int parity) {
  if (e >= g.n_entries || col < 0 || col >= g.nt) return;
  const Shot2D& S = g.shots[g.rec_shot[e]];
  const unsigned c = (unsigned)g.rec_c[e];
  const long long o = g.rec_out[e] + (long long)g.rec_stride[e] * col;
  if (S.nbuf == 3) {  // scalar: u after step col is u^{col+1}
    G(g.seis[0])[o] = Cvt<T>::ld(S.fld[0][(col + 1) % 3], c, S.fdec[0]);
    return;
  }
  for (int k = 0; k < g.nslots; ++k) G(g.seis[k])[o] = Cvt<T>::ld(S.fld[parity][k], c, S.fdec[k]);
}

// The host lays the work list out per XCD: hardware hands block b to XCD b % 8, and the list
// gives each XCD a contiguous, cost-balanced run of the grid so that halo rows stay in its L2.
// Padding entries have kind -1.  Blocks past the work list record the previous step's receivers.
template <int PH, int N, class T, class TM>
__device__ __forceinline__ void sweep2d_body(const SweepArgs& A) {
  using P = Phys<PH, N, T>;
  constexpr int NF = P::NF, NS = P::NS, HN = P::HALO;
  constexpr int FT = NS * (kTileW + 2 * HN + 2) * (kFrameRows + 2 * HN + 2) + NF * (kTileW + 2) * (kFrameRows + 2);
  constexpr int IS = NS * StripRing<N>::RSX * (kTileW + 2 * N);
  __shared__ float lds[FT > IS ? FT : IS];
  const int it = *A.step + A.koff;
  if ((int)blockIdx.x >= A.n_work) {
    const int e = ((int)blockIdx.x - A.n_work) * kGatherBlock + (int)(threadIdx.x + threadIdx.y * kTileW);
    gather_entry<T>(A.g, e, it - 1, A.parity);
    return;
  }
  const Work2D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  const Shot2D& S = A.shots[W.shot];
  PV V;
  make_view<PH>(S, A.parity, it, V);
  if (W.kind == 0)
    frame_tile<PH, N, T, TM>(A, W, S, V, lds, it);
  else
    interior_strip<PH, N, T, TM>(A, W, S, V, lds, it);
}

template <int PH, int N, class T>
__global__ __launch_bounds__(256) void sweep2d_kernel(const SweepArgs A) {
  sweep2d_body<PH, N, T, T>(A);
}

// FP32 fields with FP16 materials (velocity-stress phases only: the scalar phase keeps a field
// among its "materials").
template <int PH, int N>
__global__ __launch_bounds__(256) void sweep2d_mat16_kernel(const SweepArgs A) {
  sweep2d_body<PH, N, float, __half>(A);
}

// The frame launches of the fused families: frame tiles of kFrameRowsF rows in blocks of 64 x 8
// threads (the same two rows per thread as the 8-row tiles of sweep2d_kernel, whose interior strips
// set its block shape), so that a round of loads covers twice the cells and the staged halo rows
// weigh half as much; blocks past the work list record the previous step's receivers.
template <int PH, int N, class T, class TM>
__global__ __launch_bounds__(kTileW * 8) void frame2d_kernel(const SweepArgs A) {
  using P = Phys<PH, N, T>;
  constexpr int NF = P::NF, NS = P::NS, HN = P::HALO;
  __shared__ float lds[NS * (kTileW + 2 * HN + 2) * (kFrameRowsF + 2 * HN + 2) + NF * (kTileW + 2) * (kFrameRowsF + 2)];
  const int it = *A.step + A.koff;
  if ((int)blockIdx.x >= A.n_work) {
    const int t = (int)(threadIdx.x + threadIdx.y * kTileW);
    if (t < kGatherBlock) gather_entry<T>(A.g, ((int)blockIdx.x - A.n_work) * kGatherBlock + t, it - 1, A.parity);
    return;
  }
  const Work2D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  const Shot2D& S = A.shots[W.shot];
  PV V;
  make_view<PH>(S, A.parity, it, V);
  frame_tile<PH, N, T, TM, kFrameRowsF, 8>(A, W, S, V, lds, it);
}

template <class T>
__global__ __launch_bounds__(256) void gather2d_kernel(const GatherArgs g, const int* step, int koff, int parity) {
  gather_entry<T>(g, (int)(blockIdx.x * blockDim.x + threadIdx.x), *step + koff - 1, parity);
}

template <class T>
__global__ void decode2d_kernel(const void* src, float* dst, size_t n, float dec) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
    dst[c] = Cvt<T>::ld(src, c, dec);
}

int strip_rows(int tile_cfg) { return tile_cfg == 0 ? 16 : (tile_cfg == 1 ? 32 : 64); }

// Launch tag of the FP32-field, FP16-material instantiation.
struct Mat16 {};

template <int PH, int N, class T>
static hipError_t launch_cfg(const SweepArgs& a, hipStream_t s) {
  static_assert(kTileW * 4 == kGatherBlock, "one receiver entry per thread of a gather block");
  if constexpr (std::is_same<T, Mat16>::value)
    hipLaunchKernelGGL((sweep2d_mat16_kernel<PH, N>), dim3(a.n_work + a.n_gather), dim3(kTileW, 4), 0, s, a);
  else
    hipLaunchKernelGGL((sweep2d_kernel<PH, N, T>), dim3(a.n_work + a.n_gather), dim3(kTileW, 4), 0, s, a);
  return hipGetLastError();
}

template <int PH, class T>
static hipError_t launch_phase(int N, const SweepArgs& a, hipStream_t s) {
  switch (N) {
    case 1: return launch_cfg<PH, 1, T>(a, s);
    case 2: return launch_cfg<PH, 2, T>(a, s);
    case 3: return launch_cfg<PH, 3, T>(a, s);
    case 4: return launch_cfg<PH, 4, T>(a, s);
    case 5: return launch_cfg<PH, 5, T>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template <class T>
static hipError_t launch_t2(int phase, int N, const SweepArgs& a, hipStream_t s) {
  switch (phase) {
    case kAcVel: return launch_phase<kAcVel, T>(N, a, s);
    case kAcPrs: return launch_phase<kAcPrs, T>(N, a, s);
    case kElVel: return launch_phase<kElVel, T>(N, a, s);
    case kElStr: return launch_phase<kElStr, T>(N, a, s);
    case kScal:
      if constexpr (std::is_same<T, Mat16>::value)
        return hipErrorInvalidValue;
      else
        return launch_phase<kScal, T>(N, a, s);
    case kCpVel:  // the coupled equations run with Float32 storage only
      if constexpr (std::is_same<T, float>::value)
        return launch_phase<kCpVel, T>(N, a, s);
      else
        return hipErrorInvalidValue;
    case kCpPos:
      if constexpr (std::is_same<T, float>::value)
        return launch_phase<kCpPos, T>(N, a, s);
      else
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

// tile_cfg only shapes the work list (the strip length lives in each item).
hipError_t launch_sweep2d(int storage, int phase, int N, int tile_cfg, const SweepArgs& a, hipStream_t s) {
  (void)tile_cfg;
  if (a.n_work + a.n_gather <= 0) return hipSuccess;
  switch (storage) {
    case FOMO_STORE_F32: return launch_t2<float>(phase, N, a, s);
    case FOMO_STORE_F16: return launch_t2<__half>(phase, N, a, s);
    case FOMO_STORE_BF16: return launch_t2<__hip_bfloat16>(phase, N, a, s);
    case FOMO_STORE_F32_MAT16: return launch_t2<Mat16>(phase, N, a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int PH, class T, class TM>
static hipError_t launch_frame_n(int N, const SweepArgs& a, hipStream_t s) {
  const dim3 grid(a.n_work + a.n_gather), blk(kTileW, 8);
  switch (N) {
    case 1: hipLaunchKernelGGL((frame2d_kernel<PH, 1, T, TM>), grid, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((frame2d_kernel<PH, 2, T, TM>), grid, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((frame2d_kernel<PH, 3, T, TM>), grid, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((frame2d_kernel<PH, 4, T, TM>), grid, blk, 0, s, a); break;
    case 5: hipLaunchKernelGGL((frame2d_kernel<PH, 5, T, TM>), grid, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class T, class TM>
static hipError_t launch_frame_t(int phase, int N, const SweepArgs& a, hipStream_t s) {
  switch (phase) {
    case kAcVel: return launch_frame_n<kAcVel, T, TM>(N, a, s);
    case kAcPrs: return launch_frame_n<kAcPrs, T, TM>(N, a, s);
    case kElVel: return launch_frame_n<kElVel, T, TM>(N, a, s);
    case kScal:
      if constexpr (std::is_same<T, TM>::value)
        return launch_frame_n<kScal, T, TM>(N, a, s);
      else
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_frame2d(int storage, int phase, int N, const SweepArgs& a, hipStream_t s) {
  if (a.n_work + a.n_gather <= 0) return hipSuccess;
  switch (storage) {
    case FOMO_STORE_F32: return launch_frame_t<float, float>(phase, N, a, s);
    case FOMO_STORE_F16: return launch_frame_t<__half, __half>(phase, N, a, s);
    case FOMO_STORE_BF16: return launch_frame_t<__hip_bfloat16, __hip_bfloat16>(phase, N, a, s);
    case FOMO_STORE_F32_MAT16: return launch_frame_t<float, __half>(phase, N, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_gather2d(int storage, const GatherArgs& g, const int* step, int koff, int parity, hipStream_t s) {
  if (g.n_entries <= 0) return hipSuccess;
  const int blocks = (g.n_entries + kGatherBlock - 1) / kGatherBlock;
  switch (storage) {
    case FOMO_STORE_F32:
      hipLaunchKernelGGL(gather2d_kernel<float>, dim3(blocks), dim3(kGatherBlock), 0, s, g, step, koff, parity);
      break;
    case FOMO_STORE_F16:
      hipLaunchKernelGGL(gather2d_kernel<__half>, dim3(blocks), dim3(kGatherBlock), 0, s, g, step, koff, parity);
      break;
    case FOMO_STORE_BF16:
      hipLaunchKernelGGL(gather2d_kernel<__hip_bfloat16>, dim3(blocks), dim3(kGatherBlock), 0, s, g, step, koff,
                         parity);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_decode(int storage, const void* src, float* dst, size_t n, float dec, hipStream_t s) {
  const int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  if (blocks == 0) return hipSuccess;
  switch (storage) {
    case FOMO_STORE_F32: hipLaunchKernelGGL(decode2d_kernel<float>, dim3(blocks), dim3(256), 0, s, src, dst, n, dec); break;
    case FOMO_STORE_F16: hipLaunchKernelGGL(decode2d_kernel<__half>, dim3(blocks), dim3(256), 0, s, src, dst, n, dec); break;
    case FOMO_STORE_BF16:
      hipLaunchKernelGGL(decode2d_kernel<__hip_bfloat16>, dim3(blocks), dim3(256), 0, s, src, dst, n, dec);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace fomo
