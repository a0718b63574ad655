// learned2d.h -- learned absorbing boundary for 2D acoustic runs: a per-frame-cell MLP that
// predicts the Higdon blend weights.
#pragma once
#include <hip/hip_runtime.h>

#include "fomo_common.h"

namespace fomo {

// MLP shape: 76 inputs (5x5 patches of p, vx, vz, each scaled by its patch max-abs, plus the depth
// in the layer dist/H), two hidden layers of 32 (ReLU), 2 outputs z.  The outputs are bounded
// corrections to the distance-ramp blend weights: corr = kCorrMax * tanh(z), and the frame blend
// uses clamp(w(dist) + corr, 0, 1) -- corr_v for vx and vz, corr_p for p.  Zero output weights
// give exactly the plain Higdon boundary.
constexpr int kMlpIn = 76, kMlpH = 32, kMlpOut = 2;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;   // W1 is [32][76] row-major
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;    // W2 [32][32]
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH;  // W3 [2][32]
constexpr int kMlpParams = kOffB3 + kMlpOut;      // 3586
constexpr float kCorrMax = 0.25f;

// The inference kernel's copy of the weights, in MFMA B-fragment order (v_mfma_f32_32x32x16_bf16):
// layer 1 as [5 k-blocks][64 lanes][8], layers 2 and 3 as [2][64][8] each, then the biases b1[32],
// b2[32], b3[2].  Lane l holds output unit l%32 and inputs k = 16 kb + 8 (l/32) + 0..7; inputs
// past 76 and layer-3 units past 2 hold zeros.  One 16-byte load per fragment and lane.
constexpr int kFragW1 = 0;
constexpr int kFragW2 = kFragW1 + 5 * 64 * 8;
constexpr int kFragW3 = kFragW2 + 2 * 64 * 8;
constexpr int kFragB = kFragW3 + 2 * 64 * 8;
constexpr int kFragParams = kFragB + 2 * kMlpH + kMlpOut;

// Inference tile: w x h frame cells from (x0, z0); w <= 64, w * h <= kInferTileCells and the
// staged patch region (w + 4) x (h + 4) fits kInferStage cells.
struct InferTile {
  int x0, z0, w, h;
};
constexpr int kInferTileCells = 256;
constexpr int kInferStage = 544;  // (64 + 4) x (4 + 4)

struct MlpArgs {
  const void* w_bf16;   // kMlpParams BFloat16 weights
  const void* w_frag;   // kFragParams BFloat16: the same weights in fragment order (inference)
  const int* cells;     // frame cells (padded index) of shot 0 (training)
  int n_cells;
  const InferTile* tiles;  // the same cells as inference tiles
  int n_tiles;
  int nx, nz, H;
  const void* p;        // state at time n (read buffer), stored as `storage` with scales dec
  const void* vx;
  const void* vz;
  float dec[3];
  int storage;          // FOMO_STORE_*: inference reads any storage, training needs Float32
  float* alpha_v;       // full-grid arrays of corrections, written at frame cells
  float* alpha_p;
};

struct MlpTrainArgs {
  MlpArgs m;
  // student after the step (time n+1) and the frame sensitivities f_pde - S written by the sweeps
  const float* p1;
  const float* vx1;
  const float* vz1;
  const float* sens_p;
  const float* sens_vx;
  const float* sens_vz;
  // teacher state at time n+1 and the offset of the student grid inside the teacher grid
  const float* tp;
  const float* tvx;
  const float* tvz;
  int tnx, off_x, off_z;
  float inv_scale_p, inv_scale_v;
  float* partial;       // [gridDim.x][kMlpParams + 1] gradient (+ loss) partial sums, accumulated
};

hipError_t launch_mlp_forward(const MlpArgs& a, hipStream_t s);
hipError_t launch_mlp_train(const MlpTrainArgs& a, int blocks, hipStream_t s);
hipError_t launch_reduce_partials(const float* partial, int blocks, int width, float* out, hipStream_t s);
// Unit entry: forward/backward on given features (n x 76) and output gradients (n x 2).
hipError_t launch_mlp_eval(const void* w_bf16, const float* feats, const float* galpha, int n, float* alpha,
                           float* partial, int blocks, hipStream_t s);

}  // namespace fomo
