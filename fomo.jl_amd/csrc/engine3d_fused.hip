// engine3d_fused.hip -- acoustic3d: one launch takes the user's interior (the cells outside the
// Higdon frame) through a whole time step, velocity and then pressure.
//
// The separate sweeps read p, vx, vy, vz, bx, by, bz and write vx, vy, vz, then read vx, vy, vz, p,
// kappa and write p (64 B per cell in FP32).  Here each column tile (64 x kFTY cells in x-y) marches
// down z and, at every plane, first computes the new velocities it needs from p (staged in LDS with
// its x/y halos) and then the new pressure of the plane N-1 levels behind, whose z-stencil of new
// vz is complete by then.  The new velocities never leave the chip on their way to the pressure
// update: HBM sees each of p, vx, vy, vz, bx, by, bz, kappa read once and vx, vy, vz, p written
// once (48 B per cell), plus the halos neighbouring tiles share through L2.
//
// The pressure update of a cell needs new velocities up to N cells away.  Where such a cell is in
// the user's interior the tile recomputes it from the old state (buffers of parity s, read-only
// during the launch); where it lies in the frame it reads the value the frame velocity launch,
// which runs first, wrote (buffers of parity 1-s; the fused launch writes only interior cells, so
// nothing it reads changes under it).  The frame pressure launch runs afterwards.  Every value is
// the same Float32 expression, in the same order, as in the separate sweeps
// (src/equations/acoustic3d/update_velocity.jl:31-48, update_pressure.jl:29-45), so the result is
// bit-identical to them, and to the reference's updates.
#include <type_traits>

#include "engine3d.h"
#include "storage.h"

namespace fomo {

constexpr int kFTR = 16;                  // thread rows of a block (64 x kFTR threads)
constexpr int kFThreads = kTileW * kFTR;  // 1024

template <int N>
struct FusedGeom {
  static constexpr int TY = kFTY;
  static constexpr int VXL = N - 1, VXW = kTileW + 2 * N - 1;  // new vx: columns [-(N-1), 64+N)
  static constexpr int VYT = N - 1, VYH = TY + 2 * N - 1;      // new vy: rows [-(N-1), TY+N)
  static constexpr int PH = 2 * N - 1;                         // p halo (x and y) of the stage
  static constexpr int PSW = kTileW + 2 * PH, PSH = TY + 2 * PH;  // stage bounding box
  static constexpr int STG = PSW * PSH;                        // one p stage
  static constexpr int VXS = TY * VXW, VYS = VYH * kTileW;     // one ring slot
  static constexpr int LDS = 2 * STG + N * (VXS + VYS);
  // per-level work items and passes of the block's threads
  static constexpr int NVX = VXS, NVY = VYS, NOWN = TY * kTileW;
  static constexpr int NPH = 2 * TY * PH + 2 * PH * kTileW;    // p halo items of the stage
  static constexpr int PVX = (NVX + kFThreads - 1) / kFThreads, PVY = (NVY + kFThreads - 1) / kFThreads;
  static constexpr int POWN = NOWN / kFThreads, PPH = (NPH + kFThreads - 1) / kFThreads;
};
static_assert(kFTY * kTileW % kFThreads == 0, "own cells split evenly over the block");

template <int N, class T>
__device__ __forceinline__ void fused3d_body(const Sweep3DArgs& A, const Work3D& W, float* lds) {
  using Gm = FusedGeom<N>;
  constexpr int TY = Gm::TY, PH = Gm::PH, PSW = Gm::PSW, VXW = Gm::VXW;
  constexpr int POWN = Gm::POWN, PVX = Gm::PVX, PVY = Gm::PVY, PPH = Gm::PPH;
  const Shot3D& S = A.shot;
  const int nx = S.nx, ny = S.ny, nz = S.nz, pad = S.pad, px = S.px;
  const size_t sz = (size_t)px * ny;
  const int tid = threadIdx.x + kTileW * threadIdx.y;
  const int s = A.parity, o = 1 - s;
  const int i0 = W.i0, j0 = W.j0, k0 = W.k0, w = W.w, hy = W.hy, hz = W.hz;
  const float* __restrict__ a = A.a;
  const float dtdh = S.dtdh;
  const int it = *A.step + A.koff;
  float* stg0 = lds;
  float* vxr = lds + 2 * Gm::STG;
  float* vyr = vxr + N * Gm::VXS;

  auto clampi = [](int v, int lo, int hi) { return min(max(v, lo), hi); };
  // Halo item e of the p stage -> tile row r and column c: first the left / right bands of the
  // tile's rows (2 PH cells per row), then the PH rows above and below over the tile's columns.
  auto halo_item = [](int e, int& r, int& c) {
    if (e < 2 * TY * PH) {
      r = e / (2 * PH);
      const int m = e % (2 * PH);
      c = m < PH ? m - PH : kTileW + (m - PH);
    } else {
      const int e2 = e - 2 * TY * PH;
      const int br = e2 / kTileW;
      c = e2 % kTileW;
      r = br < PH ? br - PH : TY + (br - PH);
    }
  };
  auto core_x = [&](int i) { return i >= pad && i < nx - pad; };
  auto core_y = [&](int j) { return j >= pad && j < ny - pad; };
  auto core_z = [&](int k) { return k >= pad && k < nz - pad; };
  // Array bases (uniform): old state (parity s), the launch's outputs (parity 1 - s), materials.
  // A per-lane choice between old and new buffers selects between two of these pointers.
  const void* const old_f[4] = {S.fld[s][0], S.fld[s][1], S.fld[s][2], S.fld[s][3]};
  void* const new_f[4] = {S.fld[o][0], S.fld[o][1], S.fld[o][2], S.fld[o][3]};
  const void* const mat[4] = {S.mat[0], S.mat[1], S.mat[2], S.mat[3]};
  const float fdec[4] = {S.fdec[0], S.fdec[1], S.fdec[2], S.fdec[3]};
  const float fenc[4] = {S.fenc[0], S.fenc[1], S.fenc[2], S.fenc[3]};
  const float mdec[4] = {S.mdec[0], S.mdec[1], S.mdec[2], S.mdec[3]};
  auto ldp = [&](const void* base, float dec, int i, int j, int k) -> float {
    return Cvt<T>::ld(plane_of<T>(base, k, sz), (unsigned)(i + j * px), dec);
  };
  auto ldf = [&](bool old, int f, int i, int j, int k) -> float {
    return ldp(old ? old_f[f] : (const void*)new_f[f], fdec[f], i, j, k);
  };
  auto ldm = [&](int m, int i, int j, int k) -> float { return ldp(mat[m], mdec[m], i, j, k); };

  // Own cells: thread tid owns tile cells e = tid + kFThreads * r (row e / 64, column e % 64).
  // Cells of a partial tile beyond its width or height load a clamped cell and store nothing.
  int oi[POWN], oj[POWN];
  bool own_ok[POWN];
#pragma unroll
  for (int r = 0; r < POWN; ++r) {
    const int e = tid + kFThreads * r;
    oi[r] = min(i0 + (e % kTileW), nx - 1);
    oj[r] = min(j0 + e / kTileW, ny - 1);
    own_ok[r] = (e % kTileW) < w && e / kTileW < hy;
  }
  // Does a source lie in one of this thread's columns within the item's planes?
  bool src_here = false;
  if (it < A.nt) {
    for (int q = 0; q < S.n_src; ++q) {
      const long long cs = (long long)G(S.src_c)[q];
      const long long ks = cs / (long long)sz, rem = cs - ks * (long long)sz;
      const int js = (int)(rem / px), is = (int)(rem - (long long)js * px);
      if (ks < k0 || ks >= k0 + hz) continue;
#pragma unroll
      for (int r = 0; r < POWN; ++r) src_here = src_here || (own_ok[r] && is == oi[r] && js == oj[r]);
    }
  }

  // Levels tau = k0 - N .. k0 + hz + N - 2.  At level tau: new vz of plane tau (own cells); new vx,
  // vy of plane tau (with halos) when tau is an output plane; new p of plane t = tau - N + 1 when t
  // is an output plane.  Register queues per own cell: old p of planes tau-N+1 .. tau+N (the
  // z-stencil of vz), new vz of planes tau-2N+1 .. tau (the z-stencil of p).
  const int tau0 = k0 - N, tau_end = k0 + hz + N - 2;
  float pq[POWN][2 * N], vzq[POWN][2 * N];
#pragma unroll
  for (int r = 0; r < POWN; ++r)
#pragma unroll
    for (int m = 0; m < 2 * N; ++m) {
      pq[r][m] = ldf(true, 0, oi[r], oj[r], clampi(tau0 - N + 1 + m, 0, nz - 1));
      vzq[r][m] = 0.0f;
    }

  // Values prefetched for a level: p halo items of the stage, old (or, in the frame, new) vx / vy
  // / vz with their buoyancies, kappa of the pressure plane, and the next p-queue plane.  Two sets
  // alternate: a level consumes its set and then refills it with the loads of the level two
  // ahead, so the loads have a whole level in flight.  Every lane of a wave that has items in a
  // pass loads (cells clamped into the grid); a wave whose lanes are all past the items skips the
  // pass, so no load waits under a per-lane condition.
  const int wave0 = tid & ~(kTileW - 1);
  struct Pre {
    float ph[PPH], vx[PVX], bx[PVX], vy[PVY], by[PVY], vz[POWN], bz[POWN], kap[POWN], pq[POWN];
  };
  auto prefetch = [&](int tau, Pre& P) {
    const bool outp = tau >= k0 && tau < k0 + hz;  // an output plane: x/y work at this level
    const int kc = clampi(tau, 0, nz - 1);
    if (outp) {
#pragma unroll
      for (int q = 0; q < PPH; ++q) {
        if (kFThreads * q + wave0 >= Gm::NPH) continue;
        const int e = tid + kFThreads * q;
        int r, c;
        halo_item(e, r, c);
        P.ph[q] = ldf(true, 0, clampi(i0 + c, 0, nx - 1), clampi(j0 + r, 0, ny - 1), kc);
      }
#pragma unroll
      for (int q = 0; q < PVX; ++q) {
        if (kFThreads * q + wave0 >= Gm::NVX) continue;
        const int e = tid + kFThreads * q;
        const int r = e / VXW, c = e % VXW - Gm::VXL;
        const int i = clampi(i0 + c, 0, nx - 1), j = clampi(j0 + r, 0, ny - 1);
        P.vx[q] = ldf(core_x(i0 + c), 1, i, j, kc);
        P.bx[q] = ldm(1, i, j, kc);
      }
#pragma unroll
      for (int q = 0; q < PVY; ++q) {
        if (kFThreads * q + wave0 >= Gm::NVY) continue;
        const int e = tid + kFThreads * q;
        const int r = e / kTileW - Gm::VYT, c = e % kTileW;
        const int i = clampi(i0 + c, 0, nx - 1), j = clampi(j0 + r, 0, ny - 1);
        P.vy[q] = ldf(core_y(j0 + r), 2, i, j, kc);
        P.by[q] = ldm(2, i, j, kc);
      }
    }
    const bool cz = core_z(tau);
    const int kp = clampi(tau - N + 1, 0, nz - 1);  // pressure plane
    const int kq = clampi(tau + N + 1, 0, nz - 1);  // next p-queue plane
#pragma unroll
    for (int r = 0; r < POWN; ++r) {
      P.vz[r] = ldf(cz, 3, oi[r], oj[r], kc);
      P.bz[r] = ldm(3, oi[r], oj[r], kc);
      P.kap[r] = ldm(0, oi[r], oj[r], kp);
      P.pq[r] = ldf(true, 0, oi[r], oj[r], kq);
    }
  };

  auto level = [&](int tau, Pre& P) {
    const bool outp = tau >= k0 && tau < k0 + hz;
    const int t = tau - N + 1;
    const bool pout = t >= k0;  // t < k0 + hz always holds
    float* stg = stg0 + (tau & 1) * Gm::STG;
    // 1. stage p of plane tau (own cells from the queue, halos from the prefetch)
    if (outp) {
#pragma unroll
      for (int r = 0; r < POWN; ++r) {
        const int e = tid + kFThreads * r;
        stg[(e / kTileW + PH) * PSW + e % kTileW + PH] = pq[r][N - 1];
      }
#pragma unroll
      for (int q = 0; q < PPH; ++q) {
        const int e = tid + kFThreads * q;
        if (e >= Gm::NPH) continue;
        int rr, c;
        halo_item(e, rr, c);
        stg[(rr + PH) * PSW + c + PH] = P.ph[q];
      }
    }
    __syncthreads();

    // 2. new vz of plane tau, own cells (computed in the interior, read from the frame launch's output elsewhere)
    const bool cz = core_z(tau);
#pragma unroll
    for (int r = 0; r < POWN; ++r) {
      float v = P.vz[r];
      if (cz) {
        float dz = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) dz += a[l - 1] * (pq[r][N - 1 + l] - pq[r][N - l]);
        v = P.vz[r] + (P.bz[r] * dtdh) * dz;
      }
      if (outp && own_ok[r]) st_field<T>(plane_of<T>(new_f[3], tau, sz), (unsigned)(oi[r] + oj[r] * px), v, fenc[3], S.ovf);
      // the pressure update reads what the separate sweeps would read back: the stored value
      if (cz) v = round_trip<T>(v, fenc[3], fdec[3]);
#pragma unroll
      for (int m = 0; m < 2 * N - 1; ++m) vzq[r][m] = vzq[r][m + 1];
      vzq[r][2 * N - 1] = v;
    }
    // 3. new vx, vy of plane tau with their halos, into the ring; own cells to memory
    if (outp) {
      float* rx_ = vxr + (tau % N) * Gm::VXS;
      float* ry_ = vyr + (tau % N) * Gm::VYS;
#pragma unroll
      for (int q = 0; q < PVX; ++q) {
        const int e = tid + kFThreads * q;
        if (e >= Gm::NVX) continue;
        const int r = e / VXW, cc = e % VXW, c = cc - Gm::VXL;
        float v = P.vx[q];
        if (core_x(i0 + c)) {
          const float* P0 = stg + (r + PH) * PSW + c + PH;
          float dx = 0.0f;
#pragma unroll
          for (int l = 1; l <= N; ++l) dx += a[l - 1] * (P0[l - 1] - P0[-l]);
          v = P.vx[q] + (P.bx[q] * dtdh) * dx;
          if (c >= 0 && c < w && r < hy)
            st_field<T>(plane_of<T>(new_f[1], tau, sz), (unsigned)(i0 + c + (j0 + r) * px), v, fenc[1], S.ovf);
          v = round_trip<T>(v, fenc[1], fdec[1]);
        }
        rx_[r * VXW + cc] = v;
      }
#pragma unroll
      for (int q = 0; q < PVY; ++q) {
        const int e = tid + kFThreads * q;
        if (e >= Gm::NVY) continue;
        const int rr = e / kTileW, r = rr - Gm::VYT, c = e % kTileW;
        float v = P.vy[q];
        if (core_y(j0 + r)) {
          const float* P0 = stg + (r + PH) * PSW + c + PH;
          float dy = 0.0f;
#pragma unroll
          for (int l = 1; l <= N; ++l) dy += a[l - 1] * (P0[(l - 1) * PSW] - P0[-l * PSW]);
          v = P.vy[q] + (P.by[q] * dtdh) * dy;
          if (r >= 0 && r < hy && c < w)
            st_field<T>(plane_of<T>(new_f[2], tau, sz), (unsigned)(i0 + c + (j0 + r) * px), v, fenc[2], S.ovf);
          v = round_trip<T>(v, fenc[2], fdec[2]);
        }
        ry_[rr * kTileW + c] = v;
      }
    }
    if constexpr (N == 1) __syncthreads();  // with N = 1 the slot is read in the level that writes it
    // 4. new p of plane t = tau - N + 1 from the ring slot written N - 1 levels ago and the vz queue
    if (pout) {
      const float* rx_ = vxr + (t % N) * Gm::VXS;
      const float* ry_ = vyr + (t % N) * Gm::VYS;
#pragma unroll
      for (int r = 0; r < POWN; ++r) {
        const int e = tid + kFThreads * r;
        const int row = e / kTileW, col = e % kTileW;
        const float* VX = rx_ + row * VXW + col + Gm::VXL;
        const float* VY = ry_ + (row + Gm::VYT) * kTileW + col;
        float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
        for (int l = 1; l <= N; ++l) {
          const float cl = a[l - 1];
          dx += cl * (VX[l] - VX[-(l - 1)]);
          dy += cl * (VY[l * kTileW] - VY[-(l - 1) * kTileW]);
          dz += cl * (vzq[r][N + l - 1] - vzq[r][N - l]);
        }
        float v = pq[r][0] + (P.kap[r] * dtdh) * ((dx + dy) + dz);
        const unsigned co = (unsigned)(oi[r] + oj[r] * px);
        if (src_here) inject3d(S, (size_t)co + (size_t)t * sz, it, A.nt, v);
        if (own_ok[r]) st_field<T>(plane_of<T>(new_f[0], t, sz), co, v, fenc[0], S.ovf);
      }
    }
    // 5. advance the p queue; refill this set with the loads of level tau + 2
#pragma unroll
    for (int r = 0; r < POWN; ++r) {
#pragma unroll
      for (int m = 0; m < 2 * N - 1; ++m) pq[r][m] = pq[r][m + 1];
      pq[r][2 * N - 1] = P.pq[r];
    }
    if (tau + 2 <= tau_end) prefetch(tau + 2, P);
  };

  Pre P0, P1;
  prefetch(tau0, P0);
  prefetch(tau0 + 1, P1);  // tau_end >= tau0 + 2N - 1 >= tau0 + 1
  for (int tau = tau0; tau <= tau_end; tau += 2) {
    level(tau, P0);
    if (tau + 1 <= tau_end) level(tau + 1, P1);
  }
}

template <int N, class T>
__global__ __launch_bounds__(kFThreads) void fused3d_kernel(const Sweep3DArgs A) {
  __shared__ float lds[FusedGeom<N>::LDS];
  if ((int)blockIdx.x >= A.n_work) return;
  const Work3D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  fused3d_body<N, T>(A, W, lds);
}

template <int N, class T>
static void launch_fused_n(const Sweep3DArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((fused3d_kernel<N, T>), dim3(a.n_work), dim3(kTileW, kFTR), 0, s, a);
}

template <class T>
static hipError_t launch_fused_t(int N, const Sweep3DArgs& a, hipStream_t s) {
  switch (N) {
    case 1: launch_fused_n<1, T>(a, s); break;
    case 2: launch_fused_n<2, T>(a, s); break;
    case 3: launch_fused_n<3, T>(a, s); break;
    case 4: launch_fused_n<4, T>(a, s); break;
    case 5: launch_fused_n<5, T>(a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fused3d(int storage, int N, const Sweep3DArgs& a, hipStream_t s) {
  if (a.n_work <= 0) return hipSuccess;
  switch (storage) {
    case FOMO_STORE_F32: return launch_fused_t<float>(N, a, s);
    case FOMO_STORE_F16: return launch_fused_t<__half>(N, a, s);
    case FOMO_STORE_BF16: return launch_fused_t<__hip_bfloat16>(N, a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fomo
