// plan2d.h -- 2D plan object (one or many shots).
#pragma once
#include <vector>

#include "fomo_common.h"
#include "plan.h"

namespace fomo {

struct ShotMeta {
  int nxi, nzi, pad, x0, n_rec;
  long long seis_off;
};

struct Plan2D : PlanBase {
  int N = 0, nfld = 0, nslots = 0, tile_cfg = 1;
  int storage = FOMO_STORE_F32, elem = 4;  // storage of the fields
  int mat_storage = FOMO_STORE_F32;        // storage of the materials
  int sweep_storage = FOMO_STORE_F32;      // the requested code, which selects the sweep kernels
  float* scratch = nullptr;                // Float32 staging of one decoded field
  size_t scratch_n = 0;
  float coef[kMaxN] = {0};  // staggered first-derivative (scalar: centred second-derivative; coupled:
                            // centred first-derivative) weights
  float coef2[kMaxN] = {0}; // coupled: centred second-derivative weights
  float c0 = 0.0f;           // scalar, coupled: centre weight of the second derivative
  // coupled2d: the velocity launch covers the frame tiles and the interior strips (whose cells also
  // take their position half-step there); the position launch covers the frame tiles (d_frame).
  bool coupled = false;
  std::vector<Shot2D> shots;
  std::vector<ShotMeta> meta;
  Shot2D* d_shots = nullptr;
  Work2D* d_work[kNumTileCfgs] = {nullptr};
  int n_work[kNumTileCfgs] = {0};
  // acoustic2d / elastic2d: the interior's whole step in one launch (launch_fused2d) over
  // d_fwork[cfg] (interior strips of fused_rows(cfg) rows), the frame launches over the frame tiles
  // alone: d_frame, tiles of kFrameRowsF rows for launch_frame2d; elastic2d's stress frame d_frame8,
  // tiles of kFrameRows rows for launch_sweep2d (its 512-thread frame blocks run fewer waves per
  // CU: 8192^2 FD8 25.4 -> 28.4 us).  fusable: the plan allows it (at most two sources per column);
  // fused: the current tile shape is one of the fused family.
  bool fusable = false, fused = false;
  Work2D* d_fwork[kNumFusedCfgs] = {nullptr};
  int n_fwork[kNumFusedCfgs] = {0};
  Work2D* d_frame = nullptr;
  int n_frame = 0;
  Work2D* d_frame8 = nullptr;
  int n_frame8 = 0;
  int only = -1;  // enqueue_part: only the frame launch (1) or only the fused launch (0) when >= 0
  int n_entries = 0;
  int* rec_shot = nullptr;
  int* rec_c = nullptr;
  long long* rec_out = nullptr;
  int* rec_stride = nullptr;
  float* seis[3] = {nullptr};
  long long seis_total = 0;
  // learned boundary (shot 0 of an acoustic plan)
  bool learned = false, training = false;
  int* d_cells = nullptr;
  int n_cells = 0;
  void* d_itiles = nullptr;  // InferTile list over the same cells
  int n_itiles = 0;
  void* w_bf16 = nullptr;
  void* w_frag = nullptr;  // the same weights in MFMA fragment order
  float* alpha[2] = {nullptr, nullptr};
  float* sens[3] = {nullptr, nullptr, nullptr};
  float* partial = nullptr;
  float* grad = nullptr;
  int train_blocks = 0;

  ~Plan2D() override;
  hipError_t enqueue_part(int part, int s, int koff) override;
  bool gather_in_sweep() const override { return true; }
  int reset() override;
  int read_seis(int shot, float* const* dst) override;
  int seis_device(int shot, int slot, const float** ptr, int* n_rec) const override;
  int read_field(int shot, int which, float* out) override;
  int snap_buffer(int shot, const float** ptr, int* count, size_t* cells) const override {
    if (shot < 0 || shot >= (int)shots.size()) return set_error(FOMO_EINVAL, "shot index out of range");
    *ptr = shots[shot].snap;
    // coupled: the P slots, then the S slots (snap2 follows snap in one allocation)
    *count = shots[shot].snap ? shots[shot].snap_n * (coupled ? 2 : 1) : 0;
    *cells = (size_t)meta[shot].nxi * meta[shot].nzi;
    return FOMO_OK;
  }
  int dims(int shot, int* nx, int* ny, int* nz, int* pad) const override;
  int n_shots() const override { return (int)shots.size(); }
  // Launches of a step: velocity, pressure/stress (scalar: one); fused: velocity frame, interior
  // step, pressure frame.
  int n_launches() const override {
    return fused ? (d.eq == FOMO_SCALAR2D ? 2 : 3) : (d.eq == FOMO_SCALAR2D ? 1 : 2);
  }
  LaunchInfo launch_info(int k) const override;
  hipError_t enqueue_launch(int k, int s) override {
    if (!fused) return enqueue_part(k, s, 0);
    only = k == 1 ? 0 : 1;
    const hipError_t e = enqueue_part(k == 2 ? 1 : 0, s, 0);
    only = -1;
    return e;
  }
  int make_work(int cfg);
  void note_tuning(int source);  // record the layout in use (PlanBase::tune_*)
  int n_cfgs() const { return fused ? kNumFusedCfgs : kNumTileCfgs; }
  // Decode the current field `which` of a shot into `scratch` (asynchronous on the stream).
  int decode_field(int shot, int which);
  // Buffer holding the current state of field `which` of a shot.
  const void* current(int shot, int which) const;
  int autotune();
};

int create_plan2d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, int n_shots,
                  const fomo_shot* shots, Plan2D** out);
// The coupled2d medium on the device (coupled2d.hip): 8 new plan allocations in out.
int coupled_medium2d(PlanBase* p, const std::vector<float>& vpp, const std::vector<float>& vsp, int nx, int nz,
                     int fd_order, float dh, double sigma, float* out[8]);

}  // namespace fomo
