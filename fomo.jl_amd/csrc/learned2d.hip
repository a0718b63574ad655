// learned2d.hip -- per-frame-cell MLP for the learned absorbing boundary: inference, training
// (forward + backward with deterministic gradient reduction) and a unit entry for gradient checks.
//
// Inference (matrix cores, BF16 operands) writes the two blend-weight corrections of every frame cell; the 2D frame tiles then
// blend f = a*f_pde + (1-a)*S with a = clamp(w(dist) + corr, 0, 1).  Training compares the
// student's new state at frame cells with a wide-boundary teacher run at the same physical cells
// (one-step loss L = sum ((f - f_teacher)/scale)^2), so that dL/dcorr = 2 (f - f_teacher)/scale^2
// * (f_pde - S) (inside the clamp), with f_pde - S written by the sweep.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <utility>

#include "learned2d.h"
#include "storage.h"

namespace fomo {

__device__ __forceinline__ float corrf(float z) { return kCorrMax * tanhf(z); }

// Load all weights into LDS as Float32 (BF16 storage, FP32 accumulate).
__device__ __forceinline__ void load_weights(const void* w_bf16, float* W) {
  const __hip_bfloat16* w = (const __hip_bfloat16*)w_bf16;
  for (int e = threadIdx.x; e < kMlpParams; e += blockDim.x) W[e] = __bfloat162float(w[e]);
  __syncthreads();
}

// Features of frame cell c: 5x5 patches (clamped to the grid) of p, vx, vz, p divided by its
// patch max-abs and vx, vz by theirs (no squares: precursor values ~1e-20 stay representable),
// then dist/H.
__device__ __forceinline__ void gather_features(const float* p, const float* vx, const float* vz, int nx, int nz,
                                                int H, int c, float* x) {
  const int i = c % nx, j = c / nx;
  float mp = 0.0f, mv = 0.0f;
#pragma unroll
  for (int dy = -2; dy <= 2; ++dy) {
    const int jj = min(max(j + dy, 0), nz - 1);
#pragma unroll
    for (int dx = -2; dx <= 2; ++dx) {
      const int ii = min(max(i + dx, 0), nx - 1);
      const int q = (dy + 2) * 5 + (dx + 2);
      const size_t g = (size_t)ii + (size_t)jj * (size_t)nx;
      const float a = p[g], b = vx[g], d = vz[g];
      x[q] = a;
      x[25 + q] = b;
      x[50 + q] = d;
      mp = fmaxf(mp, fabsf(a));
      mv = fmaxf(mv, fmaxf(fabsf(b), fabsf(d)));
    }
  }
  const float sp = mp + 1e-30f, sv = mv + 1e-30f;
#pragma unroll
  for (int q = 0; q < 25; ++q) x[q] = x[q] / sp;
#pragma unroll
  for (int q = 25; q < 75; ++q) x[q] = x[q] / sv;
  const int dist = min(min(i, nx - 1 - i), min(j, nz - 1 - j));
  x[75] = (float)dist / (float)H;
}

__device__ __forceinline__ void mlp_fwd(const float* W, const float* x, float* h1, float* h2, float* z) {
#pragma unroll 4
  for (int o = 0; o < kMlpH; ++o) {
    float acc = W[kOffB1 + o];
    const float* w = W + kOffW1 + o * kMlpIn;
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) acc += w[k] * x[k];
    h1[o] = fmaxf(acc, 0.0f);
  }
#pragma unroll 4
  for (int o = 0; o < kMlpH; ++o) {
    float acc = W[kOffB2 + o];
    const float* w = W + kOffW2 + o * kMlpH;
#pragma unroll
    for (int k = 0; k < kMlpH; ++k) acc += w[k] * h1[k];
    h2[o] = fmaxf(acc, 0.0f);
  }
#pragma unroll
  for (int o = 0; o < kMlpOut; ++o) {
    float acc = W[kOffB3 + o];
    const float* w = W + kOffW3 + o * kMlpH;
#pragma unroll
    for (int k = 0; k < kMlpH; ++k) acc += w[k] * h2[k];
    z[o] = acc;
  }
}

// Backward of one cell given dL/dcorr; leaves g_z, g_h2, g_h1 for the weight-gradient GEMMs.
// corr = kCorrMax tanh(z): dcorr/dz = kCorrMax (1 - (corr/kCorrMax)^2).
__device__ __forceinline__ void mlp_bwd(const float* W, const float* h1, const float* h2, const float* corr,
                                        const float* ga, float* gz, float* gh2, float* gh1) {
#pragma unroll
  for (int k = 0; k < kMlpOut; ++k) {
    const float t = corr[k] * (1.0f / kCorrMax);
    gz[k] = ga[k] * (kCorrMax * (1.0f - t * t));
  }
#pragma unroll 4
  for (int o = 0; o < kMlpH; ++o) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < kMlpOut; ++k) acc += W[kOffW3 + k * kMlpH + o] * gz[k];
    gh2[o] = h2[o] > 0.0f ? acc : 0.0f;
  }
#pragma unroll 4
  for (int o = 0; o < kMlpH; ++o) {
    float acc = 0.0f;
#pragma unroll 8
    for (int m = 0; m < kMlpH; ++m) acc += W[kOffW2 + m * kMlpH + o] * gh2[m];
    gh1[o] = h1[o] > 0.0f ? acc : 0.0f;
  }
}

// Inference on the matrix cores.  A wave takes 64 frame cells per pass: every lane gathers the 76
// features of one cell (FP32, scaled as above), rounds them to BF16 into LDS, and the three layers
// run as 32-cell x 32-unit tiles of v_mfma_f32_32x32x16_bf16 (K = 80 for layer 1, 76 features and
// 4 zeros; K = 32 for layers 2 and 3; layer 3 has 2 live output columns).  Weights stay in
// registers as BF16 B-fragments for the whole kernel (loaded from a copy kept in fragment order);
// accumulation is FP32; the hidden activations are kept in FP32 in LDS and rounded to BF16 as they
// are read into the next layer's A fragments.  Fragment layouts (32x32x16): A lane l holds row l%32,
// k = 8*(l/32) + 0..7 of a 16-wide k block; B lane l holds column l%32, the same k; D register r
// of lane l holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int kInferWaves = 4;
constexpr int kInferCells = 64;         // cells per wave pass
constexpr int kHidPitch = 40;           // words per cell row of the hidden-activation buffer

// A fragment of a hidden layer: 8 consecutive FP32 activations (32-byte aligned) rounded to BF16.
__device__ __forceinline__ bf16x8 hid_frag(const float* src) {
  const float4 lo = *(const float4*)src, hi = *(const float4*)(src + 4);
  bf16x8 a;
  a[0] = (__bf16)lo.x;
  a[1] = (__bf16)lo.y;
  a[2] = (__bf16)lo.z;
  a[3] = (__bf16)lo.w;
  a[4] = (__bf16)hi.x;
  a[5] = (__bf16)hi.y;
  a[6] = (__bf16)hi.z;
  a[7] = (__bf16)hi.w;
  return a;
}

// Feature k of a cell (k < 25 p, k < 50 vx, k < 75 vz at patch offset q = k mod 25, row-major
// 5 x 5 from the cell's patch corner; 75 the layer depth; 76.. zero) as a constant offset into the
// staged region (field-major, kInferStage cells per field, row pitch sw) and a scale selector.
struct FeatK {
  int field_off, dy, dx;  // field * kInferStage, patch row and column
  bool pscale, depth, zero;
};
__device__ constexpr FeatK feat_k(int k) {
  if (k > 75) return FeatK{0, 0, 0, false, false, true};
  if (k == 75) return FeatK{0, 0, 0, false, true, false};
  const int f = k < 25 ? 0 : (k < 50 ? 1 : 2), q = k - 25 * f;
  return FeatK{f * kInferStage, q / 5, q % 5, f == 0, false, false};
}

// The 8 BF16 features of an MFMA A fragment: features K0 .. K0+7 in the lower half-wave,
// K0+8 .. K0+15 in the upper one.  The half-wave selects an address and a scale, so each feature
// is one LDS read and no branch.  sc: 1/max|p|, 1/max|v|, depth of the lane's cell.
template <int K0>
__device__ __forceinline__ bf16x8 frag8(bool upper, const float* stg, int base, int sw, const float* sc) {
  bf16x8 a;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const FeatK l = feat_k(K0 + u), h = feat_k(K0 + 8 + u);
    const int off = upper ? h.field_off + h.dy * sw + h.dx : l.field_off + l.dy * sw + l.dx;
    const bool ps = upper ? h.pscale : l.pscale, dp = upper ? h.depth : l.depth, zr = upper ? h.zero : l.zero;
    float x = stg[base + off] * (ps ? sc[0] : sc[1]);
    x = dp ? sc[2] : x;
    x = zr ? 0.0f : x;
    a[u] = (__bf16)x;
  }
  return a;
}

// Inference over one tile of frame cells: the tile's patch region of p, vx, vz is staged in LDS
// once (each value serves up to 25 cells), the features of each MFMA A fragment are formed from
// it (the patch scale applied as a multiply by its reciprocal, within one FP32 ulp of the
// division and far below the BF16 rounding that follows), and the three layers run on BF16 MFMA
// with FP32 accumulation.
template <class T>
__global__ __launch_bounds__(64 * kInferWaves) void mlp_infer_kernel(const MlpArgs A) {
  static_assert(kInferWaves * kInferCells == kInferTileCells, "one tile cell per thread");
  // staged p, vx, vz of the tile's patch region; per-cell feature scales; hidden activations
  __shared__ float stg[3][kInferStage];
  __shared__ float scl[kInferTileCells][3];  // 1/max|p|, 1/max|v|, dist/H
  // hidden activations, FP32, one row per cell; the row pitch of 40 words puts the two half-waves'
  // stores of a D register in disjoint banks and the 16-byte A-fragment reads conflict-free
  __shared__ float hbuf[kInferWaves][kInferCells * kHidPitch];
  __shared__ float zout[kInferTileCells][kMlpOut];  // layer-3 outputs (pre-activation)
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int col = lane & 31, half = lane >> 5;
  const InferTile tl = A.tiles[blockIdx.x];
  const int nx = A.nx, nz = A.nz, SW = tl.w + 4, n_stage = SW * (tl.h + 4);
  // 1. stage the patch region (cells clamped to the grid, as the features define): every load is
  // issued before the first LDS write
  constexpr int KS = (kInferStage + 64 * kInferWaves - 1) / (64 * kInferWaves);
  float sv[KS][3];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int idx = min(tid + k * 64 * kInferWaves, n_stage - 1);
    const int r = idx / SW, cc = idx - r * SW;
    const int ii = min(max(tl.x0 - 2 + cc, 0), nx - 1), jj = min(max(tl.z0 - 2 + r, 0), nz - 1);
    const unsigned g = (unsigned)(ii + jj * nx);
    sv[k][0] = Cvt<T>::ld(A.p, g, A.dec[0]);
    sv[k][1] = Cvt<T>::ld(A.vx, g, A.dec[1]);
    sv[k][2] = Cvt<T>::ld(A.vz, g, A.dec[2]);
  }
  // the weights as B fragments: one 16-byte load per fragment, in fragment order (kFragW1 ..), and
  // the biases (layer 3 past its two units: a clamped load, then zero)
  const bf16x8* wf = (const bf16x8*)A.w_frag;
  bf16x8 b1[5], b2[2], b3[2];
#pragma unroll
  for (int kb = 0; kb < 5; ++kb) b1[kb] = wf[kFragW1 / 8 + kb * 64 + lane];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    b2[kb] = wf[kFragW2 / 8 + kb * 64 + lane];
    b3[kb] = wf[kFragW3 / 8 + kb * 64 + lane];
  }
  const __bf16* wb = (const __bf16*)A.w_frag + kFragB;
  const float bias1 = (float)wb[col], bias2 = (float)wb[kMlpH + col];
  const float b3v = (float)wb[2 * kMlpH + min(col, kMlpOut - 1)];
  const float bias3 = col < kMlpOut ? b3v : 0.0f;
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int idx = tid + k * 64 * kInferWaves;
    if (idx < n_stage) {
      stg[0][idx] = sv[k][0];
      stg[1][idx] = sv[k][1];
      stg[2][idx] = sv[k][2];
    }
  }
  __syncthreads();
  // 2. per-cell scales: max-abs of the 5x5 patches (p; vx and vz together) and the layer depth
  const int n_tc = tl.w * tl.h;
  {
    float mp = 0.0f, mv = 0.0f, dh = 0.0f;
    if (tid < n_tc) {
      const int li = tid % tl.w, lj = tid / tl.w;
#pragma unroll
      for (int dy = 0; dy < 5; ++dy)
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
          const int o = (lj + dy) * SW + li + dx;
          mp = fmaxf(mp, fabsf(stg[0][o]));
          mv = fmaxf(mv, fmaxf(fabsf(stg[1][o]), fabsf(stg[2][o])));
        }
      const int i = tl.x0 + li, j = tl.z0 + lj;
      const int dist = min(min(i, nx - 1 - i), min(j, nz - 1 - j));
      dh = (float)dist / (float)A.H;
    }
    scl[tid][0] = 1.0f / (mp + 1e-30f);
    scl[tid][1] = 1.0f / (mv + 1e-30f);
    scl[tid][2] = dh;
  }
  __syncthreads();
  float* Hf = hbuf[wv];
  // 3. layer 1 (features built from the staged region), per wave 64 cells in two 32-row groups
#pragma unroll
  for (int grp = 0; grp < 2; ++grp) {
    const int rb = grp * 32;
    const int t = wv * kInferCells + rb + col;
    const bool tv = t < n_tc;
    const int lj = tv ? t / tl.w : 0, li = tv ? t - lj * tl.w : 0;
    const int base = lj * SW + li;
    const float sc[3] = {tv ? scl[t][0] : 0.0f, tv ? scl[t][1] : 0.0f, tv ? scl[t][2] : 0.0f};
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    auto kblock = [&](auto kbc) {
      constexpr int kb = decltype(kbc)::value;
      const bf16x8 a = frag8<kb * 16>(half != 0, &stg[0][0], base, SW, sc);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1[kb], acc, 0, 0, 0);
    };
    kblock(std::integral_constant<int, 0>{});
    kblock(std::integral_constant<int, 1>{});
    kblock(std::integral_constant<int, 2>{});
    kblock(std::integral_constant<int, 3>{});
    kblock(std::integral_constant<int, 4>{});
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r / 4) + 4 * half + (r % 4);
      Hf[(rb + row) * kHidPitch + col] = fmaxf(acc[r] + bias1, 0.0f);
    }
  }
  __syncthreads();
  f32x16 h2[2];
#pragma unroll
  for (int grp = 0; grp < 2; ++grp) {
    const int rb = grp * 32;
#pragma unroll
    for (int r = 0; r < 16; ++r) h2[grp][r] = 0.0f;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      const bf16x8 a = hid_frag(Hf + (rb + col) * kHidPitch + kb * 16 + half * 8);
      h2[grp] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b2[kb], h2[grp], 0, 0, 0);
    }
  }
  __syncthreads();
#pragma unroll
  for (int grp = 0; grp < 2; ++grp)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r / 4) + 4 * half + (r % 4);
      Hf[(grp * 32 + row) * kHidPitch + col] = fmaxf(h2[grp][r] + bias2, 0.0f);
    }
  __syncthreads();
#pragma unroll
  for (int grp = 0; grp < 2; ++grp) {
    const int rb = grp * 32;
    f32x16 z;
#pragma unroll
    for (int r = 0; r < 16; ++r) z[r] = 0.0f;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      const bf16x8 a = hid_frag(Hf + (rb + col) * kHidPitch + kb * 16 + half * 8);
      z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b3[kb], z, 0, 0, 0);
    }
    if (col < kMlpOut) {
#pragma unroll
      for (int r = 0; r < 16; ++r) zout[wv * kInferCells + rb + 8 * (r / 4) + 4 * half + (r % 4)][col] = z[r] + bias3;
    }
  }
  __syncthreads();
  // 4. one cell per thread: the bounded corrections at the cell's padded index
  if (tid < n_tc) {
    const int lj = tid / tl.w, li = tid - lj * tl.w;
    const int c = tl.x0 + li + (tl.z0 + lj) * nx;
    G(A.alpha_v)[c] = corrf(zout[tid][0]);
    G(A.alpha_p)[c] = corrf(zout[tid][1]);
  }
}

// Per-chunk LDS layout for the weight-gradient GEMMs (32 cells per chunk, 64 threads).
constexpr int kChunk = 32;
constexpr int kTrainThreads = 64;
constexpr int kGradW = kMlpParams + 1;  // gradients + loss

struct ChunkLds {
  float x[kChunk][kMlpIn];
  float h1[kChunk][kMlpH];
  float h2[kChunk][kMlpH];
  float gh1[kChunk][kMlpH];
  float gh2[kChunk][kMlpH];
  float gz[kChunk][kMlpOut];
  float loss[kChunk];
};

// Accumulate gradient entry e over the chunk's cells in a fixed order.
__device__ __forceinline__ float grad_entry(const ChunkLds& L, int e) {
  float acc = 0.0f;
  if (e < kOffB1) {
    const int o = e / kMlpIn, k = e - o * kMlpIn;
    for (int q = 0; q < kChunk; ++q) acc += L.gh1[q][o] * L.x[q][k];
  } else if (e < kOffW2) {
    const int o = e - kOffB1;
    for (int q = 0; q < kChunk; ++q) acc += L.gh1[q][o];
  } else if (e < kOffB2) {
    const int m = (e - kOffW2) / kMlpH, o = (e - kOffW2) - m * kMlpH;
    for (int q = 0; q < kChunk; ++q) acc += L.gh2[q][m] * L.h1[q][o];
  } else if (e < kOffW3) {
    const int m = e - kOffB2;
    for (int q = 0; q < kChunk; ++q) acc += L.gh2[q][m];
  } else if (e < kOffB3) {
    const int k = (e - kOffW3) / kMlpH, o = (e - kOffW3) - k * kMlpH;
    for (int q = 0; q < kChunk; ++q) acc += L.gz[q][k] * L.h2[q][o];
  } else if (e < kMlpParams) {
    const int k = e - kOffB3;
    for (int q = 0; q < kChunk; ++q) acc += L.gz[q][k];
  } else if (e == kMlpParams) {
    for (int q = 0; q < kChunk; ++q) acc += L.loss[q];
  }
  return acc;
}

// One block walks chunks blockIdx.x, blockIdx.x + gridDim.x, ... of 32 cells (threads 0..31 take
// one cell each), then all 64 threads form the weight-gradient sums of the chunk in a fixed order.
// The per-block partial is accumulated into partial[blockIdx.x] (the host clears it).
template <bool FROM_FIELDS>
__device__ void train_blocks(const MlpTrainArgs* T, const void* w_bf16, const float* feats, const float* galpha,
                             int n, float* alpha_out, float* partial) {
  __shared__ float W[kMlpParams];
  __shared__ ChunkLds L;
  __shared__ float G[kGradW];
  for (int e = threadIdx.x; e < kGradW; e += blockDim.x) G[e] = 0.0f;
  load_weights(w_bf16, W);
  const int nchunks = (n + kChunk - 1) / kChunk;
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int q = threadIdx.x;
    if (q < kChunk) {
      const int t = ch * kChunk + q;
      float x[kMlpIn], h1[kMlpH], h2[kMlpH], z[kMlpOut], al[kMlpOut], ga[kMlpOut] = {0.0f, 0.0f};
      float gz[kMlpOut] = {0.0f, 0.0f}, gh1[kMlpH], gh2[kMlpH];
      float loss = 0.0f;
      if (t < n) {
        if (FROM_FIELDS) {
          const MlpArgs& A = T->m;
          const int c = A.cells[t];
          gather_features((const float*)A.p, (const float*)A.vx, (const float*)A.vz, A.nx, A.nz, A.H, c, x);
          const int i = c % A.nx, j = c / A.nx;
          const size_t tc = (size_t)(i + T->off_x) + (size_t)(j + T->off_z) * (size_t)T->tnx;
          const float rvx = (T->vx1[c] - T->tvx[tc]) * T->inv_scale_v;
          const float rvz = (T->vz1[c] - T->tvz[tc]) * T->inv_scale_v;
          const float rp = (T->p1[c] - T->tp[tc]) * T->inv_scale_p;
          ga[0] = 2.0f * T->inv_scale_v * (rvx * T->sens_vx[c] + rvz * T->sens_vz[c]);
          ga[1] = 2.0f * T->inv_scale_p * rp * T->sens_p[c];
          loss = rvx * rvx + rvz * rvz + rp * rp;
        } else {
#pragma unroll
          for (int k = 0; k < kMlpIn; ++k) x[k] = feats[(size_t)t * kMlpIn + k];
          ga[0] = galpha[2 * (size_t)t];
          ga[1] = galpha[2 * (size_t)t + 1];
        }
        mlp_fwd(W, x, h1, h2, z);
        al[0] = corrf(z[0]);
        al[1] = corrf(z[1]);
        if (!FROM_FIELDS && alpha_out) {
          alpha_out[2 * (size_t)t] = al[0];
          alpha_out[2 * (size_t)t + 1] = al[1];
        }
        mlp_bwd(W, h1, h2, al, ga, gz, gh2, gh1);
      } else {
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) x[k] = 0.0f;
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) h1[k] = h2[k] = gh1[k] = gh2[k] = 0.0f;
      }
#pragma unroll
      for (int k = 0; k < kMlpIn; ++k) L.x[q][k] = x[k];
#pragma unroll
      for (int k = 0; k < kMlpH; ++k) {
        L.h1[q][k] = h1[k];
        L.h2[q][k] = h2[k];
        L.gh1[q][k] = gh1[k];
        L.gh2[q][k] = gh2[k];
      }
      L.gz[q][0] = gz[0];
      L.gz[q][1] = gz[1];
      L.loss[q] = loss;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kGradW; e += blockDim.x) G[e] += grad_entry(L, e);
    __syncthreads();
  }
  for (int e = threadIdx.x; e < kGradW; e += blockDim.x) partial[(size_t)blockIdx.x * kGradW + e] += G[e];
}

__global__ __launch_bounds__(64) void mlp_train_kernel(const MlpTrainArgs T) {
  train_blocks<true>(&T, T.m.w_bf16, nullptr, nullptr, T.m.n_cells, nullptr, T.partial);
}

__global__ __launch_bounds__(64) void mlp_eval_kernel(const void* w, const float* feats, const float* galpha, int n,
                                                      float* alpha, float* partial) {
  train_blocks<false>(nullptr, w, feats, galpha, n, alpha, partial);
}

__global__ void reduce_partials_kernel(const float* partial, int blocks, int width, float* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= width) return;
  float acc = 0.0f;
  for (int b = 0; b < blocks; ++b) acc += partial[(size_t)b * width + e];
  out[e] = acc;
}

hipError_t launch_mlp_forward(const MlpArgs& a, hipStream_t s) {
  if (a.n_cells <= 0) return hipSuccess;
  const int blocks = a.n_tiles;
  if (blocks <= 0) return hipErrorInvalidValue;
  switch (a.storage) {
    case FOMO_STORE_F32: hipLaunchKernelGGL(mlp_infer_kernel<float>, dim3(blocks), dim3(64 * kInferWaves), 0, s, a); break;
    case FOMO_STORE_F16: hipLaunchKernelGGL(mlp_infer_kernel<__half>, dim3(blocks), dim3(64 * kInferWaves), 0, s, a); break;
    case FOMO_STORE_BF16:
      hipLaunchKernelGGL(mlp_infer_kernel<__hip_bfloat16>, dim3(blocks), dim3(64 * kInferWaves), 0, s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mlp_train(const MlpTrainArgs& a, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(mlp_train_kernel, dim3(blocks), dim3(kTrainThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_reduce_partials(const float* partial, int blocks, int width, float* out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((width + 255) / 256), dim3(256), 0, s, partial, blocks, width, out);
  return hipGetLastError();
}

hipError_t launch_mlp_eval(const void* w_bf16, const float* feats, const float* galpha, int n, float* alpha,
                           float* partial, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(mlp_eval_kernel, dim3(blocks), dim3(kTrainThreads), 0, s, w_bf16, feats, galpha, n, alpha, partial);
  return hipGetLastError();
}

}  // namespace fomo
