// engine3d.hip -- 3D staggered-grid sweeps (acoustic3d, elastic3d) for gfx950 with the
// deterministic Higdon boundary fused in.
//
// Padded cube, x fastest, then y, then z; cell (i,j,k) at i + j*px + k*px*ny (px: the row pitch,
// a multiple of 128 bytes with the interior's first column aligned).  One launch per
// phase over a work list of column tiles (64 x rows cells in x-y) that march down z.  A tile
// keeps, per thread, the fields read through z-stencils in register queues, stages the planes
// read through x/y stencils in LDS with their halos, and reads every other array once.
//
// Higdon boundary without a second pass.  A frame cell's blend needs the pre-correction values
// of its inward neighbours in x, y and z.  The march runs one plane behind: at plane t it
// computes the pre-correction values of plane t, then finalises plane t-1, whose z-neighbours
// t-2, t (registers) and x/y-neighbours (an LDS copy of plane t-1) are all on chip.  Tiles whose
// footprint touches the x/y frame also compute a one-cell ring of pre-correction values around
// their output cells (ring tiles); the other tiles need no ring.  Every output is a pure function
// of the read-only buffers (ping-pong), so the result does not depend on the tiling.
//
// Storage is Float32, Float16 or BFloat16 (per-array scale); arithmetic is Float32.  With Float32
// storage and -ffp-contract=off the result is bit-identical to the reference's Float32 updates,
// evaluated term by term in the reference's order:
//   acoustic velocity / pressure ... src/equations/acoustic3d/update_velocity.jl:31-48,
//                                    src/equations/acoustic3d/update_pressure.jl:29-45
//   elastic velocity / stress ...... src/equations/elastic3d/update_velocity.jl:30-82,
//                                    src/equations/elastic3d/update_stress.jl:34-92
//   Higdon ......................... src/boundary/habc/kernels_3d.jl:131-186 on pre-correction
//                                    values (the race-free form of src/boundary/habc/kernels_det.jl:5-10)
//   injection ...................... src/acquisition/inject_source_3d.jl:5-17 (after the HABC)
#include <algorithm>
#include <type_traits>

#include "engine3d.h"
#include "storage.h"

namespace fomo {

enum Phase3 { kA3Vel = 0, kA3Prs = 1, kE3Vel = 2, kE3Str = 3 };

// Per phase: NF output fields, NS staged planes, NQ z-queues, NM materials, TR / TRC thread rows of
// the frame / core blocks (64 x TR threads), YPT rows per thread (a tile is 64 x TR*YPT cells),
// whether the stencil
// inputs are this step's new velocities (buffer 1-s), and the field ids involved.
template <int PH>
struct P3;
template <>
struct P3<kA3Vel> {
  static constexpr int NF = 3, NS = 1, NQ = 1, YPT = 2, NM = 3, TR = 8, TRC = 16;
  __device__ static constexpr int mat_id(int k) { return 1 + k; }  // bx by bz
  static constexpr bool kNew = false;
  __device__ static constexpr int out_f(int k) { return 1 + k; }                       // vx vy vz
  __device__ static constexpr int lut(int k) { return k == 0 ? 0 : (k == 1 ? 2 : 1); }  // vx vy vz tables
  __device__ static constexpr bool inj(int) { return false; }
  __device__ static constexpr int st_f(int) { return 0; }  // p
  __device__ static constexpr bool st_xh(int) { return true; }
  __device__ static constexpr bool st_yh(int) { return true; }
  __device__ static constexpr int st_q(int) { return 0; }  // own column from queue 0
  __device__ static constexpr int q_f(int) { return 0; }   // p
  __device__ static constexpr bool q_back(int) { return false; }
};
template <>
struct P3<kA3Prs> {
  static constexpr int NF = 1, NS = 2, NQ = 1, YPT = 2, NM = 1, TR = 8, TRC = 16;
  __device__ static constexpr int mat_id(int) { return 0; }  // kappa
  static constexpr bool kNew = true;
  __device__ static constexpr int out_f(int) { return 0; }
  __device__ static constexpr int lut(int) { return 3; }
  __device__ static constexpr bool inj(int) { return true; }
  __device__ static constexpr int st_f(int s) { return s == 0 ? 1 : 2; }  // vx (x), vy (y)
  __device__ static constexpr bool st_xh(int s) { return s == 0; }
  __device__ static constexpr bool st_yh(int s) { return s == 1; }
  __device__ static constexpr int st_q(int) { return -1; }
  __device__ static constexpr int q_f(int) { return 3; }  // vz
  __device__ static constexpr bool q_back(int) { return true; }
};
template <>
struct P3<kE3Vel> {
  static constexpr int NF = 3, NS = 5, NQ = 3, YPT = 1, NM = 3, TR = 8, TRC = 8;
  __device__ static constexpr int mat_id(int k) { return k; }  // bx by bz
  static constexpr bool kNew = false;
  __device__ static constexpr int out_f(int k) { return k; }  // vx vy vz
  __device__ static constexpr int lut(int k) { return k == 0 ? 0 : (k == 1 ? 2 : 1); }
  __device__ static constexpr bool inj(int) { return false; }
  // staged: txx (x), txy (x, y), tyy (y), txz (x), tyz (y)
  __device__ static constexpr int st_f(int s) { return s == 0 ? 3 : (s == 1 ? 6 : (s == 2 ? 4 : (s == 3 ? 7 : 8))); }
  __device__ static constexpr bool st_xh(int s) { return s == 0 || s == 1 || s == 3; }
  __device__ static constexpr bool st_yh(int s) { return s == 1 || s == 2 || s == 4; }
  __device__ static constexpr int st_q(int s) { return s == 3 ? 1 : (s == 4 ? 2 : -1); }
  // queues: tzz (forward), txz (backward), tyz (backward)
  __device__ static constexpr int q_f(int q) { return q == 0 ? 5 : (q == 1 ? 7 : 8); }
  __device__ static constexpr bool q_back(int q) { return q != 0; }
};
template <>
struct P3<kE3Str> {
  static constexpr int NF = 6, NS = 3, NQ = 3, YPT = 1, NM = 5, TR = 8, TRC = 8;
  __device__ static constexpr int mat_id(int k) { return 3 + k; }  // lam lam2mu mu_txy mu_txz mu_tyz
  static constexpr bool kNew = true;
  __device__ static constexpr int out_f(int k) { return 3 + k; }  // txx tyy tzz txy txz tyz
  __device__ static constexpr int lut(int) { return 3; }
  __device__ static constexpr bool inj(int k) { return k < 3; }  // explosive: txx, tyy, tzz
  __device__ static constexpr int st_f(int s) { return s; }      // vx, vy, vz (x and y halos)
  __device__ static constexpr bool st_xh(int) { return true; }
  __device__ static constexpr bool st_yh(int) { return true; }
  __device__ static constexpr int st_q(int s) { return s; }
  // queues: vx (forward), vy (forward), vz (backward)
  __device__ static constexpr int q_f(int q) { return q; }
  __device__ static constexpr bool q_back(int q) { return q == 2; }
};

template <int PH, bool LAG>
constexpr int kTR = LAG ? P3<PH>::TR : P3<PH>::TRC;  // thread rows of a block
template <int PH, bool LAG>
constexpr int kRows = kTR<PH, LAG> * P3<PH>::YPT;  // tile rows
static_assert(kRows<kA3Vel, true> == kRowsAc3 && kRows<kA3Prs, true> == kRowsAc3, "acoustic frame tile rows");
static_assert(kRows<kA3Vel, false> == kRowsAc3Core && kRows<kA3Prs, false> == kRowsAc3Core, "acoustic core tile rows");
static_assert(kRows<kE3Vel, true> == kRowsEl3 && kRows<kE3Str, true> == kRowsEl3, "elastic frame tile rows");
static_assert(kRows<kE3Vel, false> == kRowsEl3Core && kRows<kE3Str, false> == kRowsEl3Core, "elastic core tile rows");

// LDS of a tile: the staged planes (with halos) of two consecutive levels and, for the frame
// march, two-plane rings of the pre-correction values and of the old values of the outputs
// (planes t-1 and t).  Level t uses staging buffer t % 2; with both buffers and both rings
// alternating, one block barrier per plane orders every LDS write after the reads it replaces.
template <int PH, int N, bool LAG>
struct Lds3 {
  static constexpr int ROWS = kRows<PH, LAG>, PW = kTileW + 2 * N, PR = ROWS + 2 * N;
  static constexpr int STG = P3<PH>::NS * PR * PW, SLOT = P3<PH>::NF * ROWS * kTileW;
  static constexpr int CORE = 2 * STG, SIZE = 2 * STG + 4 * SLOT;
};

// LAG = true: frame march (one plane behind, ring tiles, Higdon blend, bounds-checked loads).
// LAG = false: core march over the user's interior (no frame cell anywhere in the item): every
// cell is updated and stored at once, and every stencil read lies inside the grid.
template <int PH, int N, class T, bool LAG>
__device__ __forceinline__ void march3d(const Sweep3DArgs& A, const Work3D& W, float* lds) {
  using P = P3<PH>;
  using L = Lds3<PH, N, LAG>;
  constexpr int YPT = P::YPT, ROWS = L::ROWS, NF = P::NF, NS = P::NS, NQ = P::NQ, NM = P::NM;
  constexpr int TR = kTR<PH, LAG>, PW = L::PW, PR = L::PR, HYR = (2 * N + TR - 1) / TR;
  float* ring = lds + 2 * L::STG;  // pre-correction values of plane t in slot t % 2, old values in slot 2 + t % 2
  const Shot3D& S = A.shot;
  const int nx = S.nx, ny = S.ny, nz = S.nz, H = S.H, px = S.px;
  const size_t sz = (size_t)px * ny;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int s = A.parity, o = 1 - s;
  const int sb = P::kNew ? o : s;
  const int rx = W.ring & 1, ry = (W.ring >> 1) & 1;
  const int pi0 = W.i0 - rx, pj0 = W.j0 - ry;
  const int i = pi0 + tx;
  const float* __restrict__ a = A.a;
  const float dtdh = S.dtdh;
  const int it = *A.step + A.koff;
  const bool i_upd = i >= N && i < nx - N;

  // Loads of field f at (ii, jj) of plane kk: scalar plane base + 32-bit in-plane offset.  Frame
  // tiles reach outside the grid (halos, the ring) and, where the band is narrower than the tile,
  // past what their cells need; there the cell is clamped to the columns and rows the tile's
  // outputs, ring and stencils use (inside the grid) instead of the load being skipped: a load
  // under a per-lane condition is waited for in its own branch.  A clamped load repeats a cell the
  // tile loads anyway, so it costs no HBM traffic, and its value enters no stored result (those
  // cells are not stored, and every stencil of an output or ring cell stays inside the range).
  const int cx_lo = LAG ? max(pi0 - N, 0) : 0, cx_hi = LAG ? min(W.i0 + W.w + rx - 1 + N, nx - 1) : nx - 1;
  const int cy_lo = LAG ? max(pj0 - N, 0) : 0, cy_hi = LAG ? min(W.j0 + W.hy + ry - 1 + N, ny - 1) : ny - 1;
  auto ld = [&](int buf, int f, int ii, int jj, int kk) -> float {
    if constexpr (LAG) {
      ii = min(max(ii, cx_lo), cx_hi);
      jj = min(max(jj, cy_lo), cy_hi);
    }
    kk = min(max(kk, 0), nz - 1);  // uniform: scalar instructions (the loads of a level past an item's last)
    const unsigned off = (unsigned)(ii + jj * px);
    return Cvt<T>::ld(plane_of<T>(S.fld[buf][f], kk, sz), off, S.fdec[f]);
  };

  // z-queues: window of 2N planes starting at t-N (backward stencils) or t-N+1 (forward)
  float q[NQ][YPT][2 * N];
  const int t0 = LAG ? W.k0 - 1 : W.k0;
#pragma unroll
  for (int qq = 0; qq < NQ; ++qq) {
    const int lo = P::q_back(qq) ? -N : -N + 1;
#pragma unroll
    for (int r = 0; r < YPT; ++r)
#pragma unroll
      for (int m = 0; m < 2 * N; ++m) q[qq][r][m] = ld(sb, P::q_f(qq), i, pj0 + ty + TR * r, t0 + lo + m);
  }

  // Registers prefetched one plane ahead: own-column values of staged fields that have no queue,
  // x-halo values (lanes tx < N), y-halo rows, old values of the outputs and the materials.
  // NSET register sets: level t uses set (t - t0) % NSET, refilled with level t + NSET as soon as
  // it is consumed.  The 16-bit acoustic frame march keeps two planes of loads in flight: it runs
  // at 4 waves per SIMD under its 128-VGPR cap either way, and the second set took the 512^3 FP16
  // step from 3.13 to 2.65 ms.  In Float32 the second set needs 142 VGPRs (one 512-thread block
  // per CU) or a spilling cap, and both were slower (3.16 -> 3.25 ms).
  constexpr int NSET = LAG && (PH == kA3Vel || PH == kA3Prs) && !std::is_same<T, float>::value ? 2 : 1;
  // Every level issues its next loads (past the item's last level too), so that every path issues
  // the same loads and the compiler's wait before a use counts only the older ones instead of
  // draining all of them (512^3: elastic3d FP16 6.94 -> 6.13 ms with the two-set acoustic frames
  // above, acoustic3d FP16 3.13 -> 2.61 ms).  Two phases keep their loads under `t < t_end`:
  // issued unconditionally they held a few more registers and lost a wave per SIMD (the elastic
  // stress frame 124 -> 129 VGPRs in FP32, spills in FP16; the FP32 acoustic pressure frame
  // 80 -> 81 VGPRs, 6 -> 5 waves, 0.49 -> 0.54 ms).
  constexpr bool kAll = PH != kE3Str && !(PH == kA3Prs && LAG && std::is_same<T, float>::value);
  float sv[NSET][NS][YPT], hx[NSET][NS][YPT][2], hy[NSET][NS][HYR], oldF[NSET][YPT][NF], matF[NSET][YPT][NM],
      qn[NSET][NQ][YPT];
  auto prefetch_stage = [&](int tp, int st) {
#pragma unroll
    for (int sp = 0; sp < NS; ++sp) {
      const int f = P::st_f(sp);
#pragma unroll
      for (int r = 0; r < YPT; ++r) {
        const int j = pj0 + ty + TR * r;
        if (P::st_q(sp) < 0) sv[st][sp][r] = ld(sb, f, i, j, tp);
        if (P::st_xh(sp)) {  // used by lanes tx < N; the others load a clamped cell
          hx[st][sp][r][0] = ld(sb, f, max(pi0 - N + tx, 0), j, tp);
          hx[st][sp][r][1] = ld(sb, f, min(pi0 + kTileW + tx, nx - 1), j, tp);
        }
      }
      if (P::st_yh(sp)) {
#pragma unroll
        for (int m = 0; m < HYR; ++m) {
          const int hr = ty + TR * m;
          const int jl = hr < N ? hr - N : ROWS + (hr - N);  // used by rows hr < 2N
          hy[st][sp][m] = ld(sb, f, min(i, nx - 1), min(max(pj0 + jl, 0), ny - 1), tp);
        }
      }
    }
  };
  auto prefetch_cells = [&](int tp, int st) {
#pragma unroll
    for (int r = 0; r < YPT; ++r) {
      const int j = pj0 + ty + TR * r;
      // cells outside the grid load a clamped cell (see ld); materials of cells outside the update
      // region are loaded too and not used
      const int ic = LAG ? min(max(i, cx_lo), cx_hi) : i, jc = LAG ? min(max(j, cy_lo), cy_hi) : j;
      const int tc = min(max(tp, 0), nz - 1);
      const unsigned co = (unsigned)(ic + jc * px);
#pragma unroll
      for (int f = 0; f < NF; ++f)
        oldF[st][r][f] = Cvt<T>::ld(plane_of<T>(S.fld[s][P::out_f(f)], tc, sz), co, S.fdec[P::out_f(f)]);
#pragma unroll
      for (int k = 0; k < NM; ++k)
        matF[st][r][k] = Cvt<T>::ld(plane_of<T>(S.mat[P::mat_id(k)], tc, sz), co, S.mdec[P::mat_id(k)]);
    }
  };
  // the queue entry that completes the window of level t + 1 (consumed at the end of level t)
  auto prefetch_q = [&](int t, int st) {
#pragma unroll
    for (int qq = 0; qq < NQ; ++qq) {
      const int lo = P::q_back(qq) ? -N : -N + 1;
#pragma unroll
      for (int r = 0; r < YPT; ++r) qn[st][qq][r] = ld(sb, P::q_f(qq), i, pj0 + ty + TR * r, t + 1 + lo + 2 * N - 1);
    }
  };

  float preN[YPT][NF], oldT[YPT][NF], preP[YPT][NF], oldP[YPT][NF];

  // Core items: does a source lie in this thread's column and rows within the item's planes?
  // Scanned once, so the plane loop does not wait on the source table.
  bool src_here = false;
  if constexpr (!LAG) {
    bool inj_any = false;
#pragma unroll
    for (int f = 0; f < NF; ++f) inj_any = inj_any || P::inj(f);
    if (inj_any && it < A.nt) {
      for (int q = 0; q < S.n_src; ++q) {
        const long long cs = (long long)G(S.src_c)[q];
        const long long ks = cs / (long long)sz, rem = cs - ks * (long long)sz;
        const int js = (int)(rem / px), is = (int)(rem - (long long)js * px);
        if (is != i || ks < W.k0 || ks >= W.k0 + W.hz) continue;
#pragma unroll
        for (int r = 0; r < YPT; ++r) src_here = src_here || js == pj0 + ty + TR * r;
      }
    }
  }
  const int t_end = LAG ? W.k0 + W.hz : W.k0 + W.hz - 1;
  // Loads of levels past t_end are issued too (planes clamped into the grid).
#pragma unroll
  for (int st = 0; st < NSET; ++st) {
    prefetch_stage(t0 + st, st);
    prefetch_cells(t0 + st, st);
    if constexpr (NSET == 2) prefetch_q(t0 + st, st);
  }
  auto level = [&](const int t, const int st) __attribute__((always_inline)) {
    float* stg = lds + (t & 1) * L::STG;
    // 1. stage the x/y-stencil planes of level t from the prefetched registers
#pragma unroll
    for (int sp = 0; sp < NS; ++sp) {
      float* pl = stg + sp * PR * PW;
#pragma unroll
      for (int r = 0; r < YPT; ++r) {
        const int jl = ty + TR * r;
        const int qs = P::st_q(sp) < 0 ? 0 : P::st_q(sp);
        const float v = P::st_q(sp) >= 0 ? q[qs][r][P::q_back(qs) ? N : N - 1] : sv[st][sp][r];
        pl[(jl + N) * PW + tx + N] = v;
        if (P::st_xh(sp) && tx < N) {
          pl[(jl + N) * PW + tx] = hx[st][sp][r][0];
          pl[(jl + N) * PW + kTileW + N + tx] = hx[st][sp][r][1];
        }
      }
      if (P::st_yh(sp)) {
#pragma unroll
        for (int m = 0; m < HYR; ++m) {
          const int hr = ty + TR * m;
          if (hr < 2 * N) {
            const int jl = hr < N ? hr - N : ROWS + (hr - N);
            pl[(jl + N) * PW + tx + N] = hy[st][sp][m];
          }
        }
      }
    }
    // 2. issue the staging loads of level t + NSET into the set just consumed (one set: also the
    // queue entry this level consumes at its end; two sets: that entry was loaded two levels ago)
    if (kAll || t < t_end) {
      prefetch_stage(t + NSET, st);
      if constexpr (NSET == 1) prefetch_q(t, st);
    }
    __syncthreads();

    // 3. pre-correction values of level t for this thread's cells
    const bool k_upd = t >= N && t < nz - N;
#pragma unroll
    for (int r = 0; r < YPT; ++r) {
      const int jl = ty + TR * r, j = pj0 + jl;
      const bool upd = !LAG || (i_upd && k_upd && j >= N && j < ny - N);
      const int row = (jl + N) * PW + tx + N;
      if constexpr (PH == kA3Vel) {
        const float* P0 = stg + row;
        float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) {
          const float cl = a[l - 1];
          dx += cl * (P0[l - 1] - P0[-l]);
          dy += cl * (P0[(l - 1) * PW] - P0[-l * PW]);
          dz += cl * (q[0][r][N - 1 + l] - q[0][r][N - l]);
        }
        preN[r][0] = oldF[st][r][0] + (matF[st][r][0] * dtdh) * dx;
        preN[r][1] = oldF[st][r][1] + (matF[st][r][1] * dtdh) * dy;
        preN[r][2] = oldF[st][r][2] + (matF[st][r][2] * dtdh) * dz;
      } else if constexpr (PH == kA3Prs) {
        const float* VX = stg + row;
        const float* VY = stg + PR * PW + row;
        float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) {
          const float cl = a[l - 1];
          dx += cl * (VX[l] - VX[-(l - 1)]);
          dy += cl * (VY[l * PW] - VY[-(l - 1) * PW]);
          dz += cl * (q[0][r][N + l - 1] - q[0][r][N - l]);
        }
        preN[r][0] = oldF[st][r][0] + (matF[st][r][0] * dtdh) * ((dx + dy) + dz);
      } else if constexpr (PH == kE3Vel) {
        const float* TXX = stg + row;
        const float* TXY = stg + PR * PW + row;
        const float* TYY = stg + 2 * PR * PW + row;
        const float* TXZ = stg + 3 * PR * PW + row;
        const float* TYZ = stg + 4 * PR * PW + row;
        float dtxxdx = 0.0f, dtxydy = 0.0f, dtxzdz = 0.0f, dtxydx = 0.0f, dtyydy = 0.0f, dtyzdz = 0.0f;
        float dtxzdx = 0.0f, dtyzdy = 0.0f, dtzzdz = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) {
          const float cl = a[l - 1];
          dtxxdx += cl * (TXX[l - 1] - TXX[-l]);
          dtyydy += cl * (TYY[(l - 1) * PW] - TYY[-l * PW]);
          dtzzdz += cl * (q[0][r][N - 1 + l] - q[0][r][N - l]);
          dtxydy += cl * (TXY[l * PW] - TXY[-(l - 1) * PW]);
          dtxydx += cl * (TXY[l] - TXY[-(l - 1)]);
          dtxzdz += cl * (q[1][r][N + l - 1] - q[1][r][N - l]);
          dtxzdx += cl * (TXZ[l] - TXZ[-(l - 1)]);
          dtyzdz += cl * (q[2][r][N + l - 1] - q[2][r][N - l]);
          dtyzdy += cl * (TYZ[l * PW] - TYZ[-(l - 1) * PW]);
        }
        preN[r][0] = oldF[st][r][0] + (matF[st][r][0] * dtdh) * ((dtxxdx + dtxydy) + dtxzdz);
        preN[r][1] = oldF[st][r][1] + (matF[st][r][1] * dtdh) * ((dtxydx + dtyydy) + dtyzdz);
        preN[r][2] = oldF[st][r][2] + (matF[st][r][2] * dtdh) * ((dtxzdx + dtyzdy) + dtzzdz);
      } else {
        const float* VX = stg + row;
        const float* VY = stg + PR * PW + row;
        const float* VZ = stg + 2 * PR * PW + row;
        float dvxdx = 0.0f, dvydy = 0.0f, dvzdz = 0.0f, dvxdy = 0.0f, dvydx = 0.0f, dvxdz = 0.0f;
        float dvzdx = 0.0f, dvydz = 0.0f, dvzdy = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) {
          const float cl = a[l - 1];
          dvxdx += cl * (VX[l] - VX[-(l - 1)]);
          dvydy += cl * (VY[l * PW] - VY[-(l - 1) * PW]);
          dvzdz += cl * (q[2][r][N + l - 1] - q[2][r][N - l]);
          dvxdy += cl * (VX[(l - 1) * PW] - VX[-l * PW]);
          dvydx += cl * (VY[l - 1] - VY[-l]);
          dvxdz += cl * (q[0][r][N - 1 + l] - q[0][r][N - l]);
          dvzdx += cl * (VZ[l - 1] - VZ[-l]);
          dvydz += cl * (q[1][r][N - 1 + l] - q[1][r][N - l]);
          dvzdy += cl * (VZ[(l - 1) * PW] - VZ[-l * PW]);
        }
        const float lv = matF[st][r][0], l2 = matF[st][r][1];
        preN[r][0] = oldF[st][r][0] + ((l2 * dvxdx + lv * dvydy) + lv * dvzdz) * dtdh;
        preN[r][1] = oldF[st][r][1] + ((lv * dvxdx + l2 * dvydy) + lv * dvzdz) * dtdh;
        preN[r][2] = oldF[st][r][2] + ((lv * dvxdx + lv * dvydy) + l2 * dvzdz) * dtdh;
        preN[r][3] = oldF[st][r][3] + (matF[st][r][2] * (dvxdy + dvydx)) * dtdh;
        preN[r][4] = oldF[st][r][4] + (matF[st][r][3] * (dvxdz + dvzdx)) * dtdh;
        preN[r][5] = oldF[st][r][5] + (matF[st][r][4] * (dvydz + dvzdy)) * dtdh;
      }
      if (!upd) {
#pragma unroll
        for (int f = 0; f < NF; ++f) preN[r][f] = oldF[st][r][f];
      }
    }
    if constexpr (!LAG) {
      // core march: no frame cell, store at once (after the next plane's loads are issued)
      if (kAll || t < t_end) prefetch_cells(t + NSET, st);
      const bool col_out = tx < W.w;
#pragma unroll
      for (int r = 0; r < YPT; ++r) {
        const int jl = ty + TR * r;
        if (!col_out || jl >= W.hy) continue;
        const unsigned co = (unsigned)(i + (pj0 + jl) * px);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          float v = preN[r][f];
          if (P::inj(f) && src_here) inject3d(S, (size_t)co + (size_t)t * sz, it, A.nt, v);
          st_field<T>(plane_of<T>(S.fld[o][P::out_f(f)], t, sz), co, v, S.fenc[P::out_f(f)], S.ovf);
        }
      }
    } else {
    // publish plane t's pre-correction and old values; then load the old values and materials of t+1
    {
      float* sp_ = ring + (t & 1) * L::SLOT;
      float* so_ = ring + (2 + (t & 1)) * L::SLOT;
#pragma unroll
      for (int r = 0; r < YPT; ++r)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const int e = (f * ROWS + ty + TR * r) * kTileW + tx;
          sp_[e] = preN[r][f];
          so_[e] = oldF[st][r][f];
          oldT[r][f] = oldF[st][r][f];
        }
    }
    if (kAll || t < t_end) prefetch_cells(t + NSET, st);

    // 4. finalise plane t-1 (its rings were written before this level's barrier): Higdon blend of frame cells, injection, store
    if (t > W.k0) {
      const int kk = t - 1;
      const bool kt = kk <= H, kb = kk >= nz - 1 - H;
      const bool il = i <= H, ir = i >= nx - 1 - H;
      const bool col_out = tx >= rx && tx < rx + W.w;
      const float* curp = ring + ((t - 1) & 1) * L::SLOT;        // pre-correction values of plane t-1
      const float* curo = ring + (2 + ((t - 1) & 1)) * L::SLOT;  // old values of plane t-1
#pragma unroll
      for (int r = 0; r < YPT; ++r) {
        const int jl = ty + TR * r, j = pj0 + jl;
        if (!col_out || jl < ry || jl >= ry + W.hy) continue;
        const unsigned co = (unsigned)(i + j * px);
        const bool jf = j <= H, jb = j >= ny - 1 - H;
        const int dist = min(min(min(i, nx - 1 - i), min(j, ny - 1 - j)), min(kk, nz - 1 - kk));
        const bool frame = dist <= H;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const int F = P::out_f(f);
          const int e = (f * ROWS + jl) * kTileW + tx;
          float v = curp[e];
          // loaded for every cell (clamped distance), not under the per-lane frame condition
          const float wt = G(S.wlut)[P::lut(f) * (H + 1) + min(dist, H)];
          if (frame) {
            const float foc = curo[e];
            float sum = 0.0f;
            int nd = 0;
            if (il || ir) {
              const int d = il ? 1 : -1;
              sum += (-S.q * curp[e + d] - S.qt * foc) - S.qxt * curo[e + d];
              ++nd;
            }
            if (jf || jb) {
              const int d = jf ? kTileW : -kTileW;
              sum += (-S.q * curp[e + d] - S.qt * foc) - S.qxt * curo[e + d];
              ++nd;
            }
            if (kt || kb) {
              const float pz = kt ? preN[r][f] : preP[r][f];
              const float oz = kt ? oldT[r][f] : oldP[r][f];
              sum += (-S.q * pz - S.qt * foc) - S.qxt * oz;
              ++nd;
            }
            v = wt * v + ((1.0f - wt) * sum) / (float)nd;
          }
          // no injection here: sources are validated against the user's interior (core items)
          st_field<T>(plane_of<T>(S.fld[o][F], kk, sz), co, v, S.fenc[F], S.ovf);
        }
      }
    }
    // own values of plane t-1 become the z-neighbours below the next output plane
    {
      const float* curp = ring + ((t - 1) & 1) * L::SLOT;
      const float* curo = ring + (2 + ((t - 1) & 1)) * L::SLOT;
#pragma unroll
      for (int r = 0; r < YPT; ++r)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const int e = (f * ROWS + ty + TR * r) * kTileW + tx;
          preP[r][f] = curp[e];
          oldP[r][f] = curo[e];
        }
    }
    }  // LAG

    // 5. advance the queues to the window of t+1
#pragma unroll
    for (int qq = 0; qq < NQ; ++qq)
#pragma unroll
      for (int r = 0; r < YPT; ++r) {
#pragma unroll
        for (int m = 0; m < 2 * N - 1; ++m) q[qq][r][m] = q[qq][r][m + 1];
        q[qq][r][2 * N - 1] = qn[st][qq][r];
      }
    if constexpr (NSET == 2) prefetch_q(t + NSET, st);
  };
  for (int t = t0; t <= t_end; t += NSET) {
    level(t, 0);
    if constexpr (NSET == 2) {
      if (t + 1 <= t_end) level(t + 1, 1);
    }
  }
}

template <int PH, int N, class T, bool LAG>
__device__ __forceinline__ void sweep3d_body(const Sweep3DArgs& A) {
  __shared__ float lds[LAG ? Lds3<PH, N, LAG>::SIZE : Lds3<PH, N, LAG>::CORE];
  if ((int)blockIdx.x >= A.n_work) return;
  const Work3D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  march3d<PH, N, T, LAG>(A, W, lds);
}

// Float32 storage: the compiler's register budget.  16-bit storage: at most 128 VGPRs, 4 waves
// per SIMD (the elastic stress core otherwise needs a few more and runs 3 waves per SIMD: its
// 512^3 step took 7.47 ms against 6.60 ms with the cap; Float32 loses 1 % to the spill the cap
// costs it, so it has none).
template <int PH, int N, class T, bool LAG>
__global__ __launch_bounds__((kTileW * kTR<PH, LAG>)) void sweep3d_kernel(const Sweep3DArgs A) {
  sweep3d_body<PH, N, T, LAG>(A);
}
template <int PH, int N, class T, bool LAG>
__global__ __launch_bounds__((kTileW * kTR<PH, LAG>)) __attribute__((amdgpu_waves_per_eu(4))) void sweep3d_kernel_w4(
    const Sweep3DArgs A) {
  sweep3d_body<PH, N, T, LAG>(A);
}

// Interior cells of a slice: a_n x b_n with a fastest (xz: x, z; xy: x, y; yz: y, z).
__device__ __forceinline__ void slice_dims(int plane, int ni, int nj, int nk, int& a_n, int& b_n) {
  a_n = plane == 2 ? nj : ni;
  b_n = plane == 1 ? nj : nk;
}

template <class T>
__global__ __launch_bounds__(256) void gather3d_kernel(const Gather3DArgs A) {
  const int it = *A.step + A.koff;
  const Shot3D& S = A.shot;
  if (it >= A.nt) return;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < A.n_rec; r += gridDim.x * blockDim.x) {
    const size_t c = (size_t)G(A.rec_c)[r];
    const size_t o = (size_t)r + (size_t)A.n_rec * (size_t)it;
    for (int k = 0; k < A.nslots; ++k) {
      const int f = A.slot_field[k];
      G(A.seis[k])[o] = Cvt<T>::ld(S.fld[A.parity_out][f], c, S.fdec[f]);
    }
  }
  if (!A.snap) return;  // snap_every is set (>= 1) only with a snapshot buffer
  const int ks = (it + 1) / A.snap_every;
  if ((it + 1) % A.snap_every == 0 && ks <= A.snap_n) {
    const int pad = S.pad, ni = S.nx - 2 * pad, nj = S.ny - 2 * pad, nk = S.nz - 2 * pad;
    int a_n, b_n;
    slice_dims(A.snap_plane, ni, nj, nk, a_n, b_n);
    const size_t n = (size_t)a_n * b_n;
    float* dst = A.snap + (size_t)(ks - 1) * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
      const int aa = (int)(e % a_n), bb = (int)(e / a_n);
      const int i = A.snap_plane == 2 ? A.snap_idx : pad + aa;
      const int j = A.snap_plane == 0 ? A.snap_idx : (A.snap_plane == 1 ? pad + bb : pad + aa);
      const int k = A.snap_plane == 1 ? A.snap_idx : pad + bb;
      const size_t c = (size_t)i + (size_t)j * S.px + (size_t)k * S.px * S.ny;
      G(dst)[e] = Cvt<T>::ld(S.fld[A.parity_out][A.snap_f], c, S.fdec[A.snap_f]);
    }
  }
}

// Decode a stored array to Float32 without its pitch: dst[i + nx*r] = value at i + px*r of the
// array base, for rows r < rows.
template <class T>
__global__ void decode_kernel(const void* src, float* dst, int nx, size_t rows, int px, float dec) {
  const size_t n = (size_t)nx * rows;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
    const size_t r = c / (size_t)nx, i = c - r * (size_t)nx;
    dst[c] = Cvt<T>::ld(src, i + r * (size_t)px, dec);
  }
}

hipError_t launch_decode3d(int storage, const void* src, float* dst, int nx, size_t rows, int px, float dec,
                           hipStream_t s) {
  const size_t n = (size_t)nx * rows;
  const int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  switch (storage) {
    case FOMO_STORE_F32:
      hipLaunchKernelGGL(decode_kernel<float>, dim3(blocks), dim3(256), 0, s, src, dst, nx, rows, px, dec);
      break;
    case FOMO_STORE_F16:
      hipLaunchKernelGGL(decode_kernel<__half>, dim3(blocks), dim3(256), 0, s, src, dst, nx, rows, px, dec);
      break;
    case FOMO_STORE_BF16:
      hipLaunchKernelGGL(decode_kernel<__hip_bfloat16>, dim3(blocks), dim3(256), 0, s, src, dst, nx, rows, px, dec);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int PH, int N, class T, bool LAG>
static void launch_one(const Sweep3DArgs& a, hipStream_t s) {
  const dim3 blk(kTileW, kTR<PH, LAG>);
  if constexpr (std::is_same<T, float>::value)
    hipLaunchKernelGGL((sweep3d_kernel<PH, N, T, LAG>), dim3(a.n_work), blk, 0, s, a);
  else
    hipLaunchKernelGGL((sweep3d_kernel_w4<PH, N, T, LAG>), dim3(a.n_work), blk, 0, s, a);
}

template <int PH, class T, bool LAG>
static hipError_t launch_n(int N, const Sweep3DArgs& a, hipStream_t s) {
  if (a.n_work <= 0) return hipSuccess;
  switch (N) {
    case 1: launch_one<PH, 1, T, LAG>(a, s); break;
    case 2: launch_one<PH, 2, T, LAG>(a, s); break;
    case 3: launch_one<PH, 3, T, LAG>(a, s); break;
    case 4: launch_one<PH, 4, T, LAG>(a, s); break;
    case 5: launch_one<PH, 5, T, LAG>(a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int PH, class T>
static hipError_t launch_ph(int N, const Sweep3DArgs& core, const Sweep3DArgs& frame, hipStream_t s) {
  hipError_t e = launch_n<PH, T, false>(N, core, s);
  if (e == hipSuccess) e = launch_n<PH, T, true>(N, frame, s);
  return e;
}

template <class T>
static hipError_t launch_t(int eq, int phase, int N, const Sweep3DArgs& c, const Sweep3DArgs& f, hipStream_t s) {
  if (eq == FOMO_ACOUSTIC3D) return phase == 0 ? launch_ph<kA3Vel, T>(N, c, f, s) : launch_ph<kA3Prs, T>(N, c, f, s);
  return phase == 0 ? launch_ph<kE3Vel, T>(N, c, f, s) : launch_ph<kE3Str, T>(N, c, f, s);
}

hipError_t launch_sweep3d(int eq, int storage, int phase, int N, const Sweep3DArgs& core, const Sweep3DArgs& frame,
                          hipStream_t s) {
  if (eq != FOMO_ACOUSTIC3D && eq != FOMO_ELASTIC3D) return hipErrorInvalidValue;
  switch (storage) {
    case FOMO_STORE_F32: return launch_t<float>(eq, phase, N, core, frame, s);
    case FOMO_STORE_F16: return launch_t<__half>(eq, phase, N, core, frame, s);
    case FOMO_STORE_BF16: return launch_t<__hip_bfloat16>(eq, phase, N, core, frame, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_gather3d(int storage, const Gather3DArgs& a, hipStream_t s) {
  int blocks = a.n_rec > 0 ? std::min(1024, (a.n_rec + 255) / 256) : 1;
  if (a.snap) {
    const int pad = a.shot.pad;
    const long long cells = a.snap_plane == 0 ? (long long)(a.shot.nx - 2 * pad) * (a.shot.nz - 2 * pad)
                            : a.snap_plane == 1 ? (long long)(a.shot.nx - 2 * pad) * (a.shot.ny - 2 * pad)
                                                : (long long)(a.shot.ny - 2 * pad) * (a.shot.nz - 2 * pad);
    blocks = std::max(blocks, (int)std::min<long long>(512, (cells + 1023) / 1024));
  }
  switch (storage) {
    case FOMO_STORE_F32: hipLaunchKernelGGL(gather3d_kernel<float>, dim3(blocks), dim3(256), 0, s, a); break;
    case FOMO_STORE_F16: hipLaunchKernelGGL(gather3d_kernel<__half>, dim3(blocks), dim3(256), 0, s, a); break;
    case FOMO_STORE_BF16:
      hipLaunchKernelGGL(gather3d_kernel<__hip_bfloat16>, dim3(blocks), dim3(256), 0, s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace fomo
