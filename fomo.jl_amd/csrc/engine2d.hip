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
//   Higdon value ................... src/boundary/habc/kernels_det.jl:22-57
//   injection / gather ............. src/acquisition/device_indexed.jl:28-134 (step index read
//                                    from a device counter, as the reference's graph path does)
#include <algorithm>

#include "fomo_common.h"
#include "storage.h"

namespace fomo {

// Resolved pointers of one phase of one shot.
struct PV {
  const void* sten[3];
  const void* old[3];
  void* out[3];
  const void* mat[3];
  float sdec[3], odec[3], oenc[3], mdec[3];  // scales of the stored arrays above
  gp<const float> lut[3];
  gp<const float> alpha[3];  // learned corrections of the distance-table weight (NULL: table only)
  gp<float> sens[3];         // f_pde - S at frame cells while training (NULL: not recorded)
};

template <int PH>
__device__ __forceinline__ void make_view(const Shot2D& S, int s, PV& V) {
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
  } else {
    sten(0, o, 0);
    sten(1, o, 1);
    for (int f = 0; f < 3; ++f) {
      old_out(f, 2 + f);
      mat(f, f);
      V.lut[f] = G(S.wlut + 3 * L);
    }
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

// Pre-correction values of all output fields of a phase at padded cell (i, j): the FD update
// inside the update region N <= i < nx-N, N <= j < nz-N, the old value elsewhere.
template <int PH, int N, class T>
struct Phys;

template <int N, class T>
struct Phys<kAcVel, N, T> {
  static constexpr int NF = 2, NS = 1;
  template <class R>
  __device__ __forceinline__ static void pre_all(const R& rd, const PV& V, const Shot2D& S, const float* a, int i,
                                                 int j, float* v) {
    const unsigned c = (unsigned)(i + j * S.nx);
    v[0] = Cvt<T>::ld(V.old[0], c, V.odec[0]);
    v[1] = Cvt<T>::ld(V.old[1], c, V.odec[1]);
    if (i < N || i >= S.nx - N || j < N || j >= S.nz - N) return;
    float dx = 0.0f, dz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dx += cl * (rd(0, i + l - 1, j) - rd(0, i - l, j));
      dz += cl * (rd(0, i, j + l) - rd(0, i, j - l + 1));
    }
    v[0] = v[0] + (Cvt<T>::ld(V.mat[0], c, V.mdec[0]) * S.dtx) * dx;
    v[1] = v[1] + (Cvt<T>::ld(V.mat[1], c, V.mdec[1]) * S.dtx) * dz;
  }
};

template <int N, class T>
struct Phys<kAcPrs, N, T> {
  static constexpr int NF = 1, NS = 2;
  template <class R>
  __device__ __forceinline__ static void pre_all(const R& rd, const PV& V, const Shot2D& S, const float* a, int i,
                                                 int j, float* v) {
    const unsigned c = (unsigned)(i + j * S.nx);
    v[0] = Cvt<T>::ld(V.old[0], c, V.odec[0]);
    if (i < N || i >= S.nx - N || j < N || j >= S.nz - N) return;
    float dvxdx = 0.0f, dvzdz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dvxdx += cl * (rd(0, i + l, j) - rd(0, i - l + 1, j));
      dvzdz += cl * (rd(1, i, j + l - 1) - rd(1, i, j - l));
    }
    v[0] = v[0] + Cvt<T>::ld(V.mat[0], c, V.mdec[0]) * (dvxdx * S.dtx + dvzdz * S.dtx);
  }
};

template <int N, class T>
struct Phys<kElVel, N, T> {
  static constexpr int NF = 2, NS = 3;
  template <class R>
  __device__ __forceinline__ static void pre_all(const R& rd, const PV& V, const Shot2D& S, const float* a, int i,
                                                 int j, float* v) {
    const unsigned c = (unsigned)(i + j * S.nx);
    v[0] = Cvt<T>::ld(V.old[0], c, V.odec[0]);
    v[1] = Cvt<T>::ld(V.old[1], c, V.odec[1]);
    if (i < N || i >= S.nx - N || j < N || j >= S.nz - N) return;
    float dtxxdx = 0.0f, dtxzdz = 0.0f, dtxzdx = 0.0f, dtzzdz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dtxxdx += cl * (rd(0, i + l - 1, j) - rd(0, i - l, j));
      dtxzdz += cl * (rd(2, i, j + l - 1) - rd(2, i, j - l));
      dtxzdx += cl * (rd(2, i + l, j) - rd(2, i - l + 1, j));
      dtzzdz += cl * (rd(1, i, j + l) - rd(1, i, j - l + 1));
    }
    const float dt = S.dtx;
    v[0] = v[0] + Cvt<T>::ld(V.mat[0], c, V.mdec[0]) * (dt * dtxxdx + dt * dtxzdz);
    v[1] = v[1] + Cvt<T>::ld(V.mat[1], c, V.mdec[1]) * (dt * dtxzdx + dt * dtzzdz);
  }
};

template <int N, class T>
struct Phys<kElStr, N, T> {
  static constexpr int NF = 3, NS = 2;
  template <class R>
  __device__ __forceinline__ static void pre_all(const R& rd, const PV& V, const Shot2D& S, const float* a, int i,
                                                 int j, float* v) {
    const unsigned c = (unsigned)(i + j * S.nx);
    v[0] = Cvt<T>::ld(V.old[0], c, V.odec[0]);
    v[1] = Cvt<T>::ld(V.old[1], c, V.odec[1]);
    v[2] = Cvt<T>::ld(V.old[2], c, V.odec[2]);
    if (i < N || i >= S.nx - N || j < N || j >= S.nz - N) return;
    float dvxdx = 0.0f, dvzdz = 0.0f, dvxdz = 0.0f, dvzdx = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dvxdx += cl * (rd(0, i + l, j) - rd(0, i - l + 1, j));
      dvzdz += cl * (rd(1, i, j + l - 1) - rd(1, i, j - l));
      dvxdz += cl * (rd(0, i, j + l) - rd(0, i, j - l + 1));
      dvzdx += cl * (rd(1, i + l - 1, j) - rd(1, i - l, j));
    }
    const float dt = S.dtx;
    const float lv = Cvt<T>::ld(V.mat[0], c, V.mdec[0]), l2m = Cvt<T>::ld(V.mat[1], c, V.mdec[1]);
    v[0] = v[0] + ((l2m * dvxdx) * dt + (lv * dvzdz) * dt);
    v[1] = v[1] + ((lv * dvxdx) * dt + (l2m * dvzdz) * dt);
    v[2] = v[2] + Cvt<T>::ld(V.mat[2], c, V.mdec[2]) * (dvxdz * dt + dvzdx * dt);
  }
};

// Which output fields receive the raw-add source after the HABC.
template <int PH>
__device__ __forceinline__ int inj_mask(const Shot2D& S) {
  if (PH == kAcPrs) return 1;
  if (PH == kElVel) return S.src_fz ? 2 : 0;
  if (PH == kElStr) return S.src_fz ? 0 : 3;
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

// Stencil data in a circular window of 16 rows (N <= 6): row j lives in slot (j - jb) & 15.
template <int RW>
struct WinRd {
  const float* t;
  int oi, jb;
  __device__ __forceinline__ float operator()(int k, int i, int j) const {
    return t[k * 16 * RW + ((j - jb) & 15) * RW + (i - oi)];
  }
};

// Frame tile: at most 64 x kFrameRows cells of the absorbing frame.  Stencil inputs are staged
// with an (N+1) halo; pre-correction values of the tile plus a one-cell ring go to LDS; each
// frame cell then blends with its inward neighbour's pre-correction value.
template <int PH, int N, class T>
__device__ __forceinline__ void frame_tile(const SweepArgs& A, const Work2D& W, const Shot2D& S, const PV& V,
                                           float* lds) {
  using P = Phys<PH, N, T>;
  constexpr int NF = P::NF, NS = P::NS;
  constexpr int BX = kTileW, BZ = kFrameRows, ZPT = BZ / 4;
  constexpr int TW = BX + 2 * N + 2, TH = BZ + 2 * N + 2;
  constexpr int EW = BX + 2, EH = BZ + 2;
  float* tile = lds;
  float* pre_s = lds + NS * TW * TH;
  const int nx = S.nx, nz = S.nz, H = S.H;
  const int i0 = W.i0, j0 = W.j0;
  const int tid = threadIdx.x + threadIdx.y * BX;
  const int oi = i0 - N - 1, oj = j0 - N - 1;
  for (int idx = tid; idx < TW * TH; idx += BX * 4) {
    const int r = idx / TW, cc = idx - r * TW;
    const int gi = oi + cc, gj = oj + r;
    const bool in = gi >= 0 && gi < nx && gj >= 0 && gj < nz;
    const unsigned g = in ? (unsigned)(gi + gj * nx) : 0u;
#pragma unroll
    for (int k = 0; k < NS; ++k) tile[k * TW * TH + idx] = in ? Cvt<T>::ld(V.sten[k], g, V.sdec[k]) : 0.0f;
  }
  __syncthreads();
  const TileRd<TW, TH> rd{tile, oi, oj};
  for (int idx = tid; idx < EW * EH; idx += BX * 4) {
    const int r = idx / EW, cc = idx - r * EW;
    const int gi = i0 - 1 + cc, gj = j0 - 1 + r;
    if (gi < 0 || gi >= nx || gj < 0 || gj >= nz || cc > W.w + 1 || r > W.h + 1) continue;
    float v[NF];
    P::pre_all(rd, V, S, A.a, gi, gj, v);
#pragma unroll
    for (int f = 0; f < NF; ++f) pre_s[f * EW * EH + idx] = v[f];
  }
  __syncthreads();
  if ((int)threadIdx.x >= W.w) return;
  const int i = i0 + (int)threadIdx.x;
  const int it = *A.step;
  const int mask = inj_mask<PH>(S);
  const float q = S.q, qt = S.qt, qxt = S.qxt;
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
    if (S.damp) {
      // Cerjan sponge, src/boundary/sponge.jl:56-78: field *= damp after the update
      const float dm = G(S.damp)[dist];
#pragma unroll
      for (int f = 0; f < NF; ++f) v[f] = v[f] * dm;
    } else if (dist >= 1 && dist <= H) {
      // Region flags (0-based): left i <= H, right i >= nx-1-H, top j <= H, bottom j >= nz-1-H.
      const bool Lf = i <= H, Rf = i >= nx - 1 - H, Tf = j <= H, Bf = j >= nz - 1 - H;
      const bool inx = Lf || Rf, inz = Tf || Bf;
      const int di = Lf ? 1 : -1, dj = Tf ? 1 : -1;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const void* fo = V.old[f];
        const float od = V.odec[f];
        const float foc = Cvt<T>::ld(fo, c, od);
        float sx = 0.0f, sz = 0.0f;
        if (inx) {
          const float fn = pre_s[f * EW * EH + lj * EW + li + di];
          sx = (-q * fn - qt * foc) - qxt * Cvt<T>::ld(fo, (unsigned)(i + di + j * nx), od);
        }
        if (inz) {
          const float fn = pre_s[f * EW * EH + (lj + dj) * EW + li];
          sz = (-q * fn - qt * foc) - qxt * Cvt<T>::ld(fo, (unsigned)(i + (j + dj) * nx), od);
        }
        const float wt = V.alpha[f] ? fminf(fmaxf(V.lut[f][dist] + V.alpha[f][c], 0.0f), 1.0f) : V.lut[f][dist];
        if (V.sens[f]) V.sens[f][c] = v[f] - (inx && inz ? 0.5f * (sx + sz) : (inx ? sx : sz));
        if (inx && inz)
          v[f] = wt * v[f] + ((1.0f - wt) * 0.5f) * (sx + sz);
        else if (inx)
          v[f] = wt * v[f] + (1.0f - wt) * sx;
        else
          v[f] = wt * v[f] + (1.0f - wt) * sz;
      }
    }
    inject<NF>(S, mask, (int)c, it, A.nt, v);
#pragma unroll
    for (int f = 0; f < NF; ++f) Cvt<T>::st(V.out[f], c, v[f], V.oenc[f]);
  }
}

// Interior strip: at most 64 columns by W.h rows of the user's interior (no frame cell, every
// cell in the update region).  The block walks down z four rows at a time (one row per wave)
// through a 16-row circular LDS window of the stencil inputs; the rows of the next step are
// loaded into registers while the current rows are computed.
template <int PH, int N, class T>
__device__ __forceinline__ void interior_strip(const SweepArgs& A, const Work2D& W, const Shot2D& S, const PV& V,
                                               float* lds) {
  using P = Phys<PH, N, T>;
  constexpr int NF = P::NF, NS = P::NS;
  constexpr int BX = kTileW, RW = BX + 2 * N;
  static_assert(4 + 2 * N <= 16, "window of 16 rows");
  const int nx = S.nx, nz = S.nz;
  const int i0 = W.i0, j0 = W.j0, jend = W.j0 + W.h;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int oi = i0 - N, jb = j0 - N;
  // columns loaded by this lane: oi + tx and, for tx < 2N, oi + 64 + tx
  const int ca = oi + tx, cb = oi + BX + tx;
  const bool has_b = tx < 2 * N;
  const bool ina = ca >= 0 && ca < nx, inb = has_b && cb >= 0 && cb < nx;
  auto load_row = [&](int j, float (&ra)[NS], float (&rb)[NS]) {
    const bool inj = j >= 0 && j < nz;
    const unsigned rowo = (unsigned)((inj ? j : 0) * nx);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      ra[k] = (inj && ina) ? Cvt<T>::ld(V.sten[k], rowo + (unsigned)ca, V.sdec[k]) : 0.0f;
      rb[k] = (inj && inb) ? Cvt<T>::ld(V.sten[k], rowo + (unsigned)cb, V.sdec[k]) : 0.0f;
    }
  };
  auto store_row = [&](int j, const float (&ra)[NS], const float (&rb)[NS]) {
    const int slot = (j - jb) & 15;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      lds[k * 16 * RW + slot * RW + tx] = ra[k];
      if (has_b) lds[k * 16 * RW + slot * RW + BX + tx] = rb[k];
    }
  };
  // initial window: rows j0-N .. j0+3+N
  for (int r = ty; r < 4 + 2 * N; r += 4) {
    float ra[NS], rb[NS];
    load_row(jb + r, ra, rb);
    store_row(jb + r, ra, rb);
  }
  __syncthreads();
  const WinRd<RW> rd{lds, oi, jb};
  const int it = *A.step;
  const int mask = inj_mask<PH>(S);
  const int i = i0 + tx;
  const bool col_ok = tx < W.w;
  for (int jr = j0; jr < jend; jr += 4) {
    const bool more = jr + 4 < jend;
    const int jn = jr + 4 + N + ty;
    float na[NS], nb[NS];
    if (more) load_row(jn, na, nb);
    const int j = jr + ty;
    if (col_ok && j < jend) {
      float v[NF];
      P::pre_all(rd, V, S, A.a, i, j, v);
      const unsigned c = (unsigned)(i + j * nx);
      inject<NF>(S, mask, (int)c, it, A.nt, v);
#pragma unroll
      for (int f = 0; f < NF; ++f) Cvt<T>::st(V.out[f], c, v[f], V.oenc[f]);
    }
    __syncthreads();
    if (more) store_row(jn, na, nb);
    __syncthreads();
  }
}

// The host lays the work list out per XCD: hardware hands block b to XCD b % 8, and the list
// gives each XCD a contiguous, cost-balanced run of the grid so that halo rows stay in its L2.
// Padding entries have kind -1.
template <int PH, int N, class T>
__global__ __launch_bounds__(256) void sweep2d_kernel(const SweepArgs A) {
  using P = Phys<PH, N, T>;
  constexpr int NF = P::NF, NS = P::NS;
  constexpr int FT = NS * (kTileW + 2 * N + 2) * (kFrameRows + 2 * N + 2) + NF * (kTileW + 2) * (kFrameRows + 2);
  constexpr int IS = NS * 16 * (kTileW + 2 * N);
  __shared__ float lds[FT > IS ? FT : IS];
  if ((int)blockIdx.x >= A.n_work) return;
  const Work2D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  const Shot2D& S = A.shots[W.shot];
  PV V;
  make_view<PH>(S, A.parity, V);
  if (W.kind == 0)
    frame_tile<PH, N, T>(A, W, S, V, lds);
  else
    interior_strip<PH, N, T>(A, W, S, V, lds);
}

template <class T>
__global__ __launch_bounds__(256) void gather2d_kernel(const GatherArgs A) {
  const int it = *A.step;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < A.n_entries && it < A.nt) {
    const Shot2D& S = A.shots[A.rec_shot[e]];
    const unsigned c = (unsigned)A.rec_c[e];
    const long long o = A.rec_out[e] + (long long)A.rec_stride[e] * it;
    for (int k = 0; k < A.nslots; ++k) {
      const int f = A.field_of_slot[k];
      G(A.seis[k])[o] = Cvt<T>::ld(S.fld[A.parity_out][f], c, S.fdec[f]);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned t = atomicAdd(A.done, 1u);
    if (t == gridDim.x - 1) {
      atomicExch(A.done, 0u);
      atomicExch(A.step, it + 1);
    }
  }
}

template <class T>
__global__ void decode2d_kernel(const void* src, float* dst, size_t n, float dec) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
    dst[c] = Cvt<T>::ld(src, c, dec);
}

int strip_rows(int tile_cfg) { return tile_cfg == 0 ? 16 : (tile_cfg == 1 ? 32 : 64); }

template <int PH, int N, class T>
static hipError_t launch_cfg(const SweepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((sweep2d_kernel<PH, N, T>), dim3(a.n_work), dim3(kTileW, 4), 0, s, a);
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
    default: return hipErrorInvalidValue;
  }
}

// tile_cfg only shapes the work list (the strip length lives in each item).
hipError_t launch_sweep2d(int storage, int phase, int N, int tile_cfg, const SweepArgs& a, hipStream_t s) {
  (void)tile_cfg;
  if (a.n_work <= 0) return hipSuccess;
  switch (storage) {
    case FOMO_STORE_F32: return launch_t2<float>(phase, N, a, s);
    case FOMO_STORE_F16: return launch_t2<__half>(phase, N, a, s);
    case FOMO_STORE_BF16: return launch_t2<__hip_bfloat16>(phase, N, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_gather2d(int storage, const GatherArgs& a, hipStream_t s) {
  const int blocks = a.n_entries > 0 ? (a.n_entries + 255) / 256 : 1;
  switch (storage) {
    case FOMO_STORE_F32: hipLaunchKernelGGL(gather2d_kernel<float>, dim3(blocks), dim3(256), 0, s, a); break;
    case FOMO_STORE_F16: hipLaunchKernelGGL(gather2d_kernel<__half>, dim3(blocks), dim3(256), 0, s, a); break;
    case FOMO_STORE_BF16: hipLaunchKernelGGL(gather2d_kernel<__hip_bfloat16>, dim3(blocks), dim3(256), 0, s, a); break;
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
