// engine3d.hip -- 3D acoustic sweeps for gfx950 with the deterministic Higdon boundary fused in.
//
// Padded cube, x fastest, then y, then z; cell (i,j,k) at i + j*nx + k*nx*ny.  One launch per
// phase over a work list with two kinds of items:
//   * interior column tiles (64 x 16 cells in x-y, the user's interior only) that march down z.
//     The field read through the z-stencil lives in a per-thread register queue, the planes read
//     through x/y stencils are staged in LDS with their halos; every other array is read once.
//   * frame boxes (64 x 4 cells of one plane) covering the six pad slabs.  Each thread computes
//     its cell's pre-correction values and, for the Higdon blend, those of its inward neighbours
//     (recomputed from cached global reads), so the result needs no second pass.
// Storage is Float32, Float16 or BFloat16 (per-array scale); arithmetic is Float32.  With Float32
// storage and -ffp-contract=off the result is bit-identical to the reference's Float32 updates:
//   velocity ... src/equations/acoustic3d/update_velocity.jl:31-48
//   pressure ... src/equations/acoustic3d/update_pressure.jl:29-45
//   Higdon ..... src/boundary/habc/kernels_3d.jl:131-186, evaluated on pre-correction values
//                (the race-free form, as src/boundary/habc/kernels_det.jl:5-10 describes for 2D)
#include <algorithm>

#include "engine3d.h"

namespace fomo {

template <class T>
struct Cvt;
template <>
struct Cvt<float> {
  __device__ __forceinline__ static float ld(const void* p, size_t c, float) { return ((const float*)p)[c]; }
  __device__ __forceinline__ static void st(void* p, size_t c, float v, float) { ((float*)p)[c] = v; }
};
template <>
struct Cvt<__half> {
  __device__ __forceinline__ static float ld(const void* p, size_t c, float dec) {
    return __half2float(((const __half*)p)[c]) * dec;
  }
  __device__ __forceinline__ static void st(void* p, size_t c, float v, float enc) {
    ((__half*)p)[c] = __float2half_rn(v * enc);
  }
};
template <>
struct Cvt<__hip_bfloat16> {
  __device__ __forceinline__ static float ld(const void* p, size_t c, float dec) {
    return __bfloat162float(((const __hip_bfloat16*)p)[c]) * dec;
  }
  __device__ __forceinline__ static void st(void* p, size_t c, float v, float enc) {
    ((__hip_bfloat16*)p)[c] = __float2bfloat16(v * enc);
  }
};

struct G3 {
  int nx, ny, nz;
  size_t sy, sz;
  __device__ __forceinline__ size_t at(int i, int j, int k) const {
    return (size_t)i + (size_t)j * sy + (size_t)k * sz;
  }
  __device__ __forceinline__ bool in(int i, int j, int k) const {
    return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz;
  }
};

// --- cell-by-cell physics on global memory (frame boxes) ------------------------------------
template <int N, class T>
struct Cell3D {
  const Shot3D& S;
  const float* a;
  int s;  // read parity
  G3 g;
  __device__ __forceinline__ bool upd(int i, int j, int k) const {
    return i >= N && i < g.nx - N && j >= N && j < g.ny - N && k >= N && k < g.nz - N;
  }
  __device__ __forceinline__ float fld(int buf, int f, size_t c) const {
    return Cvt<T>::ld(S.fld[buf][f], c, S.fdec[f]);
  }
  __device__ __forceinline__ float mat(int m, size_t c) const { return Cvt<T>::ld(S.mat[m], c, S.mdec[m]); }
  // pre-correction velocity component f (1 vx, 2 vy, 3 vz) at (i,j,k)
  __device__ float pre_v(int f, int i, int j, int k) const {
    const size_t c = g.at(i, j, k);
    const float o = fld(s, f, c);
    if (!upd(i, j, k)) return o;
    float d = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      float hi, lo;
      if (f == 1) {
        hi = fld(s, 0, c + (l - 1));
        lo = fld(s, 0, c - l);
      } else if (f == 2) {
        hi = fld(s, 0, c + (size_t)(l - 1) * g.sy);
        lo = fld(s, 0, c - (size_t)l * g.sy);
      } else {
        hi = fld(s, 0, c + (size_t)l * g.sz);
        lo = fld(s, 0, c - (size_t)(l - 1) * g.sz);
      }
      d += a[l - 1] * (hi - lo);
    }
    return o + (mat(f, c) * S.dtdh) * d;
  }
  // pre-correction pressure at (i,j,k): reads the new velocities (buffer 1-s)
  __device__ float pre_p(int i, int j, int k) const {
    const int o_ = 1 - s;
    const size_t c = g.at(i, j, k);
    const float o = fld(s, 0, c);
    if (!upd(i, j, k)) return o;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
    for (int l = 1; l <= N; ++l) {
      const float cl = a[l - 1];
      dx += cl * (fld(o_, 1, c + l) - fld(o_, 1, c - (l - 1)));
      dy += cl * (fld(o_, 2, c + (size_t)l * g.sy) - fld(o_, 2, c - (size_t)(l - 1) * g.sy));
      dz += cl * (fld(o_, 3, c + (size_t)(l - 1) * g.sz) - fld(o_, 3, c - (size_t)l * g.sz));
    }
    return o + (mat(0, c) * S.dtdh) * ((dx + dy) + dz);
  }
};

// Higdon blend (kernels_3d.jl:131-186) of field f at frame cell (i,j,k): directional terms in
// x, y, z order summed from 0, then wt*f + ((1-wt)*sum)/n_dirs.  pre(i,j,k) gives the
// pre-correction value; old values come from the read buffer.
template <class PreFn, class T>
__device__ __forceinline__ float higdon3d(const Shot3D& S, const G3& g, int s, int f, int lut, int i, int j, int k,
                                         float fc, const PreFn& pre) {
  const int H = S.H;
  const bool il = i <= H, ir = i >= g.nx - 1 - H, jf = j <= H, jb = j >= g.ny - 1 - H;
  const bool kt = k <= H, kb = k >= g.nz - 1 - H;
  const size_t c = g.at(i, j, k);
  const float foc = Cvt<T>::ld(S.fld[s][f], c, S.fdec[f]);
  float sum = 0.0f;
  int nd = 0;
  if (il || ir) {
    const int ii = il ? i + 1 : i - 1;
    sum += (-S.q * pre(ii, j, k) - S.qt * foc) - S.qxt * Cvt<T>::ld(S.fld[s][f], g.at(ii, j, k), S.fdec[f]);
    ++nd;
  }
  if (jf || jb) {
    const int jj = jf ? j + 1 : j - 1;
    sum += (-S.q * pre(i, jj, k) - S.qt * foc) - S.qxt * Cvt<T>::ld(S.fld[s][f], g.at(i, jj, k), S.fdec[f]);
    ++nd;
  }
  if (kt || kb) {
    const int kk = kt ? k + 1 : k - 1;
    sum += (-S.q * pre(i, j, kk) - S.qt * foc) - S.qxt * Cvt<T>::ld(S.fld[s][f], g.at(i, j, kk), S.fdec[f]);
    ++nd;
  }
  const int dist = min(min(min(i, g.nx - 1 - i), min(j, g.ny - 1 - j)), min(k, g.nz - 1 - k));
  const float wt = S.wlut[lut * (H + 1) + dist];
  return wt * fc + ((1.0f - wt) * sum) / (float)nd;
}

__device__ __forceinline__ void inject3d(const Shot3D& S, size_t c, int it, int nt, float& v) {
  if (it >= nt) return;
  for (int q = 0; q < S.n_src; ++q)
    if ((size_t)S.src_c[q] == c) v += S.wav[(size_t)q + (size_t)S.n_src * (size_t)it];
}

template <int N, class T>
__device__ void frame_box(const Sweep3DArgs& A, const Work3D& W, int phase) {
  const Shot3D& S = *A.shot;
  const G3 g{S.nx, S.ny, S.nz, (size_t)S.nx, (size_t)S.nx * S.ny};
  const int tx = threadIdx.x, ty = threadIdx.y;
  if (tx >= W.w || ty >= W.hy) return;
  const int i = W.i0 + tx, j = W.j0 + ty, k = W.k0;
  const int s = A.parity, o = 1 - s;
  const Cell3D<N, T> C{S, A.a, s, g};
  const size_t c = g.at(i, j, k);
  const int dist = min(min(min(i, g.nx - 1 - i), min(j, g.ny - 1 - j)), min(k, g.nz - 1 - k));
  const bool frame = dist >= 1 && dist <= S.H;
  if (phase == 0) {
    // vx, vy, vz with weight tables vx (0), vy (2), vz (1)
    const int luts[4] = {0, 0, 2, 1};
#pragma unroll
    for (int f = 1; f <= 3; ++f) {
      float v = C.pre_v(f, i, j, k);
      if (frame) {
        auto pre = [&](int a_, int b_, int c_) { return C.pre_v(f, a_, b_, c_); };
        v = higdon3d<decltype(pre), T>(S, g, s, f, luts[f], i, j, k, v, pre);
      }
      Cvt<T>::st(S.fld[o][f], c, v, S.fenc[f]);
    }
  } else {
    float v = C.pre_p(i, j, k);
    if (frame) {
      auto pre = [&](int a_, int b_, int c_) { return C.pre_p(a_, b_, c_); };
      v = higdon3d<decltype(pre), T>(S, g, s, 0, 3, i, j, k, v, pre);
    }
    inject3d(S, c, *A.step, A.nt, v);
    Cvt<T>::st(S.fld[o][0], c, v, S.fenc[0]);
  }
}

// --- interior column tiles -------------------------------------------------------------------
// Velocity: p through x, y (LDS plane with halo N) and z (register queue of 2N planes).
template <int N, class T>
__device__ void interior_vel(const Sweep3DArgs& A, const Work3D& W, float* lds) {
  constexpr int TY = kTileY3, YPT = TY / 4, PW = kTileW + 2 * N;
  const Shot3D& S = *A.shot;
  const G3 g{S.nx, S.ny, S.nz, (size_t)S.nx, (size_t)S.nx * S.ny};
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int s = A.parity, o = 1 - s;
  const int i = W.i0 + tx;
  const float* __restrict__ a = A.a;
  float q[YPT][2 * N];  // planes k-N+1 .. k+N
  auto ldp = [&](int ii, int jj, int kk) -> float {
    return g.in(ii, jj, kk) ? Cvt<T>::ld(S.fld[s][0], g.at(ii, jj, kk), S.fdec[0]) : 0.0f;
  };
#pragma unroll
  for (int r = 0; r < YPT; ++r)
#pragma unroll
    for (int m = 0; m < 2 * N; ++m) q[r][m] = ldp(i, W.j0 + ty + 4 * r, W.k0 - N + 1 + m);
  const float dtdh = S.dtdh;
  for (int k = W.k0; k < W.k0 + W.hz; ++k) {
    // plane k into LDS: own columns from the queue, x-halos and y-halo rows from global
#pragma unroll
    for (int r = 0; r < YPT; ++r) {
      const int jl = ty + 4 * r;
      lds[(jl + N) * PW + tx + N] = q[r][N - 1];
      if (tx < N) {
        lds[(jl + N) * PW + tx] = ldp(W.i0 - N + tx, W.j0 + jl, k);
        lds[(jl + N) * PW + kTileW + N + tx] = ldp(W.i0 + kTileW + tx, W.j0 + jl, k);
      }
    }
    for (int hr = ty; hr < 2 * N; hr += 4) {
      const int jl = hr < N ? hr - N : TY + (hr - N);
      lds[(jl + N) * PW + tx + N] = ldp(i, W.j0 + jl, k);
    }
    __syncthreads();
    const bool col_ok = tx < W.w;
#pragma unroll
    for (int r = 0; r < YPT; ++r) {
      const int jl = ty + 4 * r;
      if (!col_ok || jl >= W.hy) continue;
      const int j = W.j0 + jl;
      const size_t c = g.at(i, j, k);
      float dx = 0.0f, dy = 0.0f, dz = 0.0f;
      const float* row = lds + (jl + N) * PW + tx + N;
#pragma unroll
      for (int l = 1; l <= N; ++l) {
        const float cl = a[l - 1];
        dx += cl * (row[l - 1] - row[-l]);
        dy += cl * (row[(l - 1) * PW] - row[-l * PW]);
        dz += cl * (q[r][N - 1 + l] - q[r][N - l]);
      }
      const float vx = Cvt<T>::ld(S.fld[s][1], c, S.fdec[1]) + (Cvt<T>::ld(S.mat[1], c, S.mdec[1]) * dtdh) * dx;
      const float vy = Cvt<T>::ld(S.fld[s][2], c, S.fdec[2]) + (Cvt<T>::ld(S.mat[2], c, S.mdec[2]) * dtdh) * dy;
      const float vz = Cvt<T>::ld(S.fld[s][3], c, S.fdec[3]) + (Cvt<T>::ld(S.mat[3], c, S.mdec[3]) * dtdh) * dz;
      Cvt<T>::st(S.fld[o][1], c, vx, S.fenc[1]);
      Cvt<T>::st(S.fld[o][2], c, vy, S.fenc[2]);
      Cvt<T>::st(S.fld[o][3], c, vz, S.fenc[3]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < YPT; ++r) {
#pragma unroll
      for (int m = 0; m < 2 * N - 1; ++m) q[r][m] = q[r][m + 1];
      q[r][2 * N - 1] = ldp(i, W.j0 + ty + 4 * r, k + 1 + N);
    }
  }
}

// Pressure: vx through x (LDS rows with x-halo), vy through y (LDS with y-halo rows), vz through
// z (register queue of planes k-N .. k+N-1).
template <int N, class T>
__device__ void interior_prs(const Sweep3DArgs& A, const Work3D& W, float* lds) {
  constexpr int TY = kTileY3, YPT = TY / 4, PWX = kTileW + 2 * N;
  const Shot3D& S = *A.shot;
  const G3 g{S.nx, S.ny, S.nz, (size_t)S.nx, (size_t)S.nx * S.ny};
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int s = A.parity, o = 1 - s;
  const int i = W.i0 + tx;
  const float* __restrict__ a = A.a;
  float* ldx = lds;                 // TY rows x PWX
  float* ldy = lds + TY * PWX;      // PHY rows x 64
  auto ldv = [&](int f, int ii, int jj, int kk) -> float {
    return g.in(ii, jj, kk) ? Cvt<T>::ld(S.fld[o][f], g.at(ii, jj, kk), S.fdec[f]) : 0.0f;
  };
  float q[YPT][2 * N];  // vz planes k-N .. k+N-1
#pragma unroll
  for (int r = 0; r < YPT; ++r)
#pragma unroll
    for (int m = 0; m < 2 * N; ++m) q[r][m] = ldv(3, i, W.j0 + ty + 4 * r, W.k0 - N + m);
  const float dtdh = S.dtdh;
  const int it = *A.step;
  for (int k = W.k0; k < W.k0 + W.hz; ++k) {
#pragma unroll
    for (int r = 0; r < YPT; ++r) {
      const int jl = ty + 4 * r;
      const int j = W.j0 + jl;
      ldx[jl * PWX + tx + N] = ldv(1, i, j, k);
      if (tx < N) {
        ldx[jl * PWX + tx] = ldv(1, W.i0 - N + tx, j, k);
        ldx[jl * PWX + kTileW + N + tx] = ldv(1, W.i0 + kTileW + tx, j, k);
      }
      ldy[(jl + N) * kTileW + tx] = ldv(2, i, j, k);
    }
    for (int hr = ty; hr < 2 * N; hr += 4) {
      const int jl = hr < N ? hr - N : TY + (hr - N);
      ldy[(jl + N) * kTileW + tx] = ldv(2, i, W.j0 + jl, k);
    }
    __syncthreads();
    const bool col_ok = tx < W.w;
#pragma unroll
    for (int r = 0; r < YPT; ++r) {
      const int jl = ty + 4 * r;
      if (!col_ok || jl >= W.hy) continue;
      const int j = W.j0 + jl;
      const size_t c = g.at(i, j, k);
      float dx = 0.0f, dy = 0.0f, dz = 0.0f;
      const float* rx = ldx + jl * PWX + tx + N;
      const float* ry = ldy + (jl + N) * kTileW + tx;
#pragma unroll
      for (int l = 1; l <= N; ++l) {
        const float cl = a[l - 1];
        dx += cl * (rx[l] - rx[-(l - 1)]);
        dy += cl * (ry[l * kTileW] - ry[-(l - 1) * kTileW]);
        dz += cl * (q[r][N + l - 1] - q[r][N - l]);
      }
      float p = Cvt<T>::ld(S.fld[s][0], c, S.fdec[0]) + (Cvt<T>::ld(S.mat[0], c, S.mdec[0]) * dtdh) * ((dx + dy) + dz);
      inject3d(S, c, it, A.nt, p);
      Cvt<T>::st(S.fld[o][0], c, p, S.fenc[0]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < YPT; ++r) {
#pragma unroll
      for (int m = 0; m < 2 * N - 1; ++m) q[r][m] = q[r][m + 1];
      q[r][2 * N - 1] = ldv(3, i, W.j0 + ty + 4 * r, k + N);
    }
  }
}

template <int PHASE, int N, class T>
__global__ __launch_bounds__(256) void sweep3d_kernel(const Sweep3DArgs A) {
  constexpr int LV = (kTileW + 2 * N) * (kTileY3 + 2 * N);
  constexpr int LP = kTileY3 * (kTileW + 2 * N) + (kTileY3 + 2 * N) * kTileW;
  __shared__ float lds[LV > LP ? LV : LP];
  if ((int)blockIdx.x >= A.n_work) return;
  const Work3D W = A.work[blockIdx.x];
  if (W.kind < 0) return;
  if (W.kind == 0) {
    frame_box<N, T>(A, W, PHASE);
  } else if (PHASE == 0) {
    interior_vel<N, T>(A, W, lds);
  } else {
    interior_prs<N, T>(A, W, lds);
  }
}

template <class T>
__global__ __launch_bounds__(256) void gather3d_kernel(const Gather3DArgs A) {
  const int it = *A.step;
  const Shot3D& S = *A.shot;
  if (it < A.nt) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < A.n_rec; r += gridDim.x * blockDim.x) {
      const size_t c = (size_t)A.rec_c[r];
      const size_t o = (size_t)r + (size_t)A.n_rec * (size_t)it;
      for (int k = 0; k < A.nslots; ++k) {
        const int f = A.slot_field[k];
        A.seis[k][o] = Cvt<T>::ld(S.fld[A.parity_out][f], c, S.fdec[f]);
      }
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

// Decode a stored array to Float32 (dst[c] = value of cell c).
template <class T>
__global__ void decode_kernel(const void* src, float* dst, size_t n, float dec) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
    dst[c] = Cvt<T>::ld(src, c, dec);
}

// Pack a slice of a padded field (interior part) into dst, decoded to Float32.
// plane 0: xz at y index `idx`; 1: xy at z index `idx`; 2: yz at x index `idx` (padded index).
template <class T>
__global__ void slice_kernel(const void* src, float* dst, int nx, int ny, int nz, int pad, int plane, int idx,
                             float dec) {
  const int ni = nx - 2 * pad, nj = ny - 2 * pad, nk = nz - 2 * pad;
  const int a_n = plane == 2 ? nj : ni;
  const int b_n = plane == 1 ? nj : nk;
  const size_t n = (size_t)a_n * b_n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int aa = (int)(e % a_n), bb = (int)(e / a_n);
    int i, j, k;
    if (plane == 0) {
      i = pad + aa; j = idx; k = pad + bb;
    } else if (plane == 1) {
      i = pad + aa; j = pad + bb; k = idx;
    } else {
      i = idx; j = pad + aa; k = pad + bb;
    }
    dst[e] = Cvt<T>::ld(src, (size_t)i + (size_t)j * nx + (size_t)k * nx * ny, dec);
  }
}

hipError_t launch_decode3d(int storage, const void* src, float* dst, size_t n, float dec, hipStream_t s) {
  const int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  switch (storage) {
    case FOMO_STORE_F32: hipLaunchKernelGGL(decode_kernel<float>, dim3(blocks), dim3(256), 0, s, src, dst, n, dec); break;
    case FOMO_STORE_F16: hipLaunchKernelGGL(decode_kernel<__half>, dim3(blocks), dim3(256), 0, s, src, dst, n, dec); break;
    case FOMO_STORE_BF16:
      hipLaunchKernelGGL(decode_kernel<__hip_bfloat16>, dim3(blocks), dim3(256), 0, s, src, dst, n, dec);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_slice3d(int storage, const void* src, float* dst, int nx, int ny, int nz, int pad, int plane,
                          int idx, float dec, hipStream_t s) {
  const int blocks = 1024;
  switch (storage) {
    case FOMO_STORE_F32:
      hipLaunchKernelGGL(slice_kernel<float>, dim3(blocks), dim3(256), 0, s, src, dst, nx, ny, nz, pad, plane, idx, dec);
      break;
    case FOMO_STORE_F16:
      hipLaunchKernelGGL(slice_kernel<__half>, dim3(blocks), dim3(256), 0, s, src, dst, nx, ny, nz, pad, plane, idx, dec);
      break;
    case FOMO_STORE_BF16:
      hipLaunchKernelGGL(slice_kernel<__hip_bfloat16>, dim3(blocks), dim3(256), 0, s, src, dst, nx, ny, nz, pad, plane,
                         idx, dec);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int PHASE, class T>
static hipError_t launch_n(int N, const Sweep3DArgs& a, hipStream_t s) {
  const dim3 blk(kTileW, 4);
  switch (N) {
    case 1: hipLaunchKernelGGL((sweep3d_kernel<PHASE, 1, T>), dim3(a.n_work), blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL((sweep3d_kernel<PHASE, 2, T>), dim3(a.n_work), blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL((sweep3d_kernel<PHASE, 3, T>), dim3(a.n_work), blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL((sweep3d_kernel<PHASE, 4, T>), dim3(a.n_work), blk, 0, s, a); break;
    case 5: hipLaunchKernelGGL((sweep3d_kernel<PHASE, 5, T>), dim3(a.n_work), blk, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class T>
static hipError_t launch_t(int phase, int N, const Sweep3DArgs& a, hipStream_t s) {
  return phase == 0 ? launch_n<0, T>(N, a, s) : launch_n<1, T>(N, a, s);
}

hipError_t launch_sweep3d(int storage, int phase, int N, const Sweep3DArgs& a, hipStream_t s) {
  if (a.n_work <= 0) return hipSuccess;
  switch (storage) {
    case FOMO_STORE_F32: return launch_t<float>(phase, N, a, s);
    case FOMO_STORE_F16: return launch_t<__half>(phase, N, a, s);
    case FOMO_STORE_BF16: return launch_t<__hip_bfloat16>(phase, N, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_gather3d(int storage, const Gather3DArgs& a, hipStream_t s) {
  const int blocks = a.n_rec > 0 ? std::min(1024, (a.n_rec + 255) / 256) : 1;
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
