// sweep2d.hip -- 2D staggered-grid sweeps for gfx950 with the deterministic Higdon boundary fused in.
//
// One launch per phase (velocity; pressure or stress).  Each output cell is a pure function of
// read-only buffers:
//   * the other phase's fields (read through the FD stencils),
//   * this phase's fields at time n (ping-pong buffer `old`),
//   * the materials and a per-distance weight table.
// A frame cell's Higdon value needs the *pre-correction* value of its inward neighbour; the thread
// recomputes that neighbour's update itself instead of waiting for another thread, so there is no
// scratch buffer, no second pass, no strip backup and no ordering between threads.  Results are
// therefore independent of launch geometry and bit-reproducible.
//
// Arithmetic order follows the reference expressions term by term and the library is compiled
// with -ffp-contract=off, so the FP32 result is bit-identical to the reference's uncontracted
// Float32 updates:
//   acoustic velocity / pressure ... src/equations/acoustic2d/fused_kernels.jl:47-61, :101-112
//   elastic velocity / stress ...... src/equations/elastic2d/fused_kernels.jl:39-57, :105-121
//   Higdon value ................... src/boundary/habc/kernels_det.jl:22-57
//   frame / region predicates ...... src/boundary/habc/kernels_fused.jl:28-52 (1-based there)
#include "fomo_common.h"

namespace fomo {

// Pre-correction value of output field F of a phase at padded cell (i, j): the FD update inside
// the update region M <= i < nx-M, M <= j < nz-M, the old value elsewhere.
template <int PH, int N>
struct Phys;

template <int N>
struct Phys<kAcVel, N> {
  static constexpr int NF = 2;
  __device__ static float pre(const Sweep2DArgs& A, int f, int i, int j) {
    const size_t nx = A.nx, c = (size_t)i + (size_t)j * nx;
    const float o = A.old[f][c];
    if (i < N || i >= A.nx - N || j < N || j >= A.nz - N) return o;
    const float* __restrict__ p = A.rd[0];
    float d = 0.0f;
    if (f == 0) {
#pragma unroll
      for (int l = 1; l <= N; ++l) d += A.a[l - 1] * (p[c + (l - 1)] - p[c - l]);
      return o + (A.mat[0][c] * A.dtx) * d;
    }
#pragma unroll
    for (int l = 1; l <= N; ++l) d += A.a[l - 1] * (p[c + (size_t)l * nx] - p[c - (size_t)(l - 1) * nx]);
    return o + (A.mat[1][c] * A.dtx) * d;
  }
  __device__ static void pre_all(const Sweep2DArgs& A, int i, int j, float* v) {
    const size_t nx = A.nx, c = (size_t)i + (size_t)j * nx;
    v[0] = A.old[0][c];
    v[1] = A.old[1][c];
    if (i < N || i >= A.nx - N || j < N || j >= A.nz - N) return;
    const float* __restrict__ p = A.rd[0];
    float dx = 0.0f, dz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = A.a[l - 1];
      dx += cl * (p[c + (l - 1)] - p[c - l]);
      dz += cl * (p[c + (size_t)l * nx] - p[c - (size_t)(l - 1) * nx]);
    }
    v[0] = v[0] + (A.mat[0][c] * A.dtx) * dx;
    v[1] = v[1] + (A.mat[1][c] * A.dtx) * dz;
  }
};

template <int N>
struct Phys<kAcPrs, N> {
  static constexpr int NF = 1;
  __device__ static float pre(const Sweep2DArgs& A, int f, int i, int j) {
    (void)f;
    const size_t nx = A.nx, c = (size_t)i + (size_t)j * nx;
    const float o = A.old[0][c];
    if (i < N || i >= A.nx - N || j < N || j >= A.nz - N) return o;
    const float* __restrict__ vx = A.rd[0];
    const float* __restrict__ vz = A.rd[1];
    float dvxdx = 0.0f, dvzdz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = A.a[l - 1];
      dvxdx += cl * (vx[c + l] - vx[c - (l - 1)]);
      dvzdz += cl * (vz[c + (size_t)(l - 1) * nx] - vz[c - (size_t)l * nx]);
    }
    return o + A.mat[0][c] * (dvxdx * A.dtx + dvzdz * A.dtx);
  }
  __device__ static void pre_all(const Sweep2DArgs& A, int i, int j, float* v) { v[0] = pre(A, 0, i, j); }
};

template <int N>
struct Phys<kElVel, N> {
  static constexpr int NF = 2;
  __device__ static float pre(const Sweep2DArgs& A, int f, int i, int j) {
    const size_t nx = A.nx, c = (size_t)i + (size_t)j * nx;
    const float o = A.old[f][c];
    if (i < N || i >= A.nx - N || j < N || j >= A.nz - N) return o;
    const float* __restrict__ txx = A.rd[0];
    const float* __restrict__ tzz = A.rd[1];
    const float* __restrict__ txz = A.rd[2];
    float d1 = 0.0f, d2 = 0.0f;
    if (f == 0) {
#pragma unroll
      for (int l = 1; l <= N; ++l) {
        const float cl = A.a[l - 1];
        d1 += cl * (txx[c + (l - 1)] - txx[c - l]);
        d2 += cl * (txz[c + (size_t)(l - 1) * nx] - txz[c - (size_t)l * nx]);
      }
      return o + A.mat[0][c] * (A.dtx * d1 + A.dtx * d2);
    }
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = A.a[l - 1];
      d1 += cl * (txz[c + l] - txz[c - (l - 1)]);
      d2 += cl * (tzz[c + (size_t)l * nx] - tzz[c - (size_t)(l - 1) * nx]);
    }
    return o + A.mat[1][c] * (A.dtx * d1 + A.dtx * d2);
  }
  __device__ static void pre_all(const Sweep2DArgs& A, int i, int j, float* v) {
    const size_t nx = A.nx, c = (size_t)i + (size_t)j * nx;
    v[0] = A.old[0][c];
    v[1] = A.old[1][c];
    if (i < N || i >= A.nx - N || j < N || j >= A.nz - N) return;
    const float* __restrict__ txx = A.rd[0];
    const float* __restrict__ tzz = A.rd[1];
    const float* __restrict__ txz = A.rd[2];
    float dtxxdx = 0.0f, dtxzdz = 0.0f, dtxzdx = 0.0f, dtzzdz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = A.a[l - 1];
      dtxxdx += cl * (txx[c + (l - 1)] - txx[c - l]);
      dtxzdz += cl * (txz[c + (size_t)(l - 1) * nx] - txz[c - (size_t)l * nx]);
      dtxzdx += cl * (txz[c + l] - txz[c - (l - 1)]);
      dtzzdz += cl * (tzz[c + (size_t)l * nx] - tzz[c - (size_t)(l - 1) * nx]);
    }
    v[0] = v[0] + A.mat[0][c] * (A.dtx * dtxxdx + A.dtx * dtxzdz);
    v[1] = v[1] + A.mat[1][c] * (A.dtx * dtxzdx + A.dtx * dtzzdz);
  }
};

template <int N>
struct Phys<kElStr, N> {
  static constexpr int NF = 3;
  __device__ static void derivs(const Sweep2DArgs& A, size_t c, float* d) {
    const size_t nx = A.nx;
    const float* __restrict__ vx = A.rd[0];
    const float* __restrict__ vz = A.rd[1];
    float dvxdx = 0.0f, dvzdz = 0.0f, dvxdz = 0.0f, dvzdx = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = A.a[l - 1];
      dvxdx += cl * (vx[c + l] - vx[c - (l - 1)]);
      dvzdz += cl * (vz[c + (size_t)(l - 1) * nx] - vz[c - (size_t)l * nx]);
      dvxdz += cl * (vx[c + (size_t)l * nx] - vx[c - (size_t)(l - 1) * nx]);
      dvzdx += cl * (vz[c + (l - 1)] - vz[c - l]);
    }
    d[0] = dvxdx;
    d[1] = dvzdz;
    d[2] = dvxdz;
    d[3] = dvzdx;
  }
  __device__ static float combine(const Sweep2DArgs& A, int f, size_t c, float o, const float* d) {
    const float dt = A.dtx;
    if (f == 0) {
      const float lv = A.mat[0][c], l2m = A.mat[1][c];
      return o + ((l2m * d[0]) * dt + (lv * d[1]) * dt);
    }
    if (f == 1) {
      const float lv = A.mat[0][c], l2m = A.mat[1][c];
      return o + ((lv * d[0]) * dt + (l2m * d[1]) * dt);
    }
    return o + A.mat[2][c] * (d[2] * dt + d[3] * dt);
  }
  __device__ static float pre(const Sweep2DArgs& A, int f, int i, int j) {
    const size_t c = (size_t)i + (size_t)j * (size_t)A.nx;
    const float o = A.old[f][c];
    if (i < N || i >= A.nx - N || j < N || j >= A.nz - N) return o;
    float d[4];
    derivs(A, c, d);
    return combine(A, f, c, o, d);
  }
  __device__ static void pre_all(const Sweep2DArgs& A, int i, int j, float* v) {
    const size_t c = (size_t)i + (size_t)j * (size_t)A.nx;
    v[0] = A.old[0][c];
    v[1] = A.old[1][c];
    v[2] = A.old[2][c];
    if (i < N || i >= A.nx - N || j < N || j >= A.nz - N) return;
    float d[4];
    derivs(A, c, d);
    v[0] = combine(A, 0, c, v[0], d);
    v[1] = combine(A, 1, c, v[1], d);
    v[2] = combine(A, 2, c, v[2], d);
  }
};

// Deterministic Higdon blend of field f at frame cell (i, j) with pre-correction value fc.
// Region flags in 0-based form: left i <= H, right i >= nx-1-H, top j <= H, bottom j >= nz-1-H.
template <int PH, int N>
__device__ __forceinline__ float higdon(const Sweep2DArgs& A, int f, int i, int j, int dist, float fc) {
  const int nx = A.nx, nz = A.nz, H = A.H;
  const bool L = i <= H, R = i >= nx - 1 - H, T = j <= H, B = j >= nz - 1 - H;
  const bool inx = L || R, inz = T || B;
  const float* __restrict__ fo = A.old[f];
  const size_t c = (size_t)i + (size_t)j * (size_t)nx;
  const float foc = fo[c];
  float sx = 0.0f, sz = 0.0f;
  if (inx) {
    const int ni = L ? i + 1 : i - 1;
    sx = (-A.q * Phys<PH, N>::pre(A, f, ni, j) - A.qt * foc) - A.qxt * fo[(size_t)ni + (size_t)j * nx];
  }
  if (inz) {
    const int nj = T ? j + 1 : j - 1;
    sz = (-A.q * Phys<PH, N>::pre(A, f, i, nj) - A.qt * foc) - A.qxt * fo[(size_t)i + (size_t)nj * nx];
  }
  const float wt = A.wlut[f][dist];
  if (inx && inz) return wt * fc + ((1.0f - wt) * 0.5f) * (sx + sz);
  if (inx) return wt * fc + (1.0f - wt) * sx;
  return wt * fc + (1.0f - wt) * sz;
}

constexpr int kBX = 64;  // one wave64 along x: coalesced 256-byte rows
constexpr int kBZ = 4;

template <int PH, int N>
__global__ __launch_bounds__(kBX* kBZ) void sweep2d_kernel(const Sweep2DArgs A) {
  constexpr int NF = Phys<PH, N>::NF;
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBZ + threadIdx.y;
  if (i >= A.nx || j >= A.nz) return;
  float v[NF];
  Phys<PH, N>::pre_all(A, i, j, v);
  const int dist = min(min(i, A.nx - 1 - i), min(j, A.nz - 1 - j));
  if (dist >= 1 && dist <= A.H) {
#pragma unroll
    for (int f = 0; f < NF; ++f) v[f] = higdon<PH, N>(A, f, i, j, dist, v[f]);
  }
  const size_t c = (size_t)i + (size_t)j * (size_t)A.nx;
  if (A.inj_mask) {
    const int it = *A.step;
    if (it < A.nt) {
      for (int s = 0; s < A.n_src; ++s) {
        if (A.src_c[s] == (int)c) {
          const float w = A.wav[(size_t)s + (size_t)A.n_src * (size_t)it];
#pragma unroll
          for (int f = 0; f < NF; ++f)
            if (A.inj_mask & (1 << f)) v[f] += w;
        }
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) A.out[f][c] = v[f];
}

__global__ __launch_bounds__(256) void record_kernel(const RecordArgs A) {
  const int it = *A.step;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < A.n_rec && it < A.nt) {
    const size_t c = (size_t)A.rec_c[r];
    const size_t o = (size_t)r + (size_t)A.n_rec * (size_t)it;
    for (int f = 0; f < A.nf; ++f) A.seis[f][o] = A.f[f][c];
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

template <int PH>
static hipError_t launch_phase(int N, const Sweep2DArgs& a, hipStream_t s) {
  dim3 blk(kBX, kBZ);
  dim3 grd((a.nx + kBX - 1) / kBX, (a.nz + kBZ - 1) / kBZ);
  switch (N) {
    case 1: hipLaunchKernelGGL((sweep2d_kernel<PH, 1>), grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((sweep2d_kernel<PH, 2>), grd, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((sweep2d_kernel<PH, 3>), grd, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((sweep2d_kernel<PH, 4>), grd, blk, 0, s, a); break;
    case 5: hipLaunchKernelGGL((sweep2d_kernel<PH, 5>), grd, blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sweep2d(int phase, int N, const Sweep2DArgs& a, hipStream_t s) {
  switch (phase) {
    case kAcVel: return launch_phase<kAcVel>(N, a, s);
    case kAcPrs: return launch_phase<kAcPrs>(N, a, s);
    case kElVel: return launch_phase<kElVel>(N, a, s);
    case kElStr: return launch_phase<kElStr>(N, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_record(const RecordArgs& a, hipStream_t s) {
  const int blocks = a.n_rec > 0 ? (a.n_rec + 255) / 256 : 1;
  hipLaunchKernelGGL(record_kernel, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace fomo
