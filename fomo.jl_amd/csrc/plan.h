// plan.h -- host-side plan objects behind the opaque fomo_plan handle.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fomo_hip.h"

namespace fomo {

int set_error(int code, const std::string& msg);

#define FOMO_HIP_TRY(expr)                                                                    \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      return ::fomo::set_error(_e == hipErrorOutOfMemory ? FOMO_ENOMEM : FOMO_EHIP,           \
                               std::string(#expr) + ": " + hipGetErrorString(_e));            \
  } while (0)

// Steps per captured graph (even: a replay starts and ends on parity 0).
constexpr int kGraphSteps = 16;

// Common part of every plan: device allocations owned by the plan, its stream, the step
// counter, and the captured multi-step graph.
//
// Step indexing: kernels take the index of their step as (*step + koff), with koff a launch
// argument; the counter moves only by tick launches at the end of a run of steps.  A captured
// graph of kGraphSteps steps therefore holds every step's offset in its nodes and needs one
// counter update per replay instead of one per step.
struct PlanBase {
  fomo_desc d{};
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  int* step = nullptr;
  unsigned* ovf = nullptr;  // 16-bit storage overflow flag written by the sweeps (st_field)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long steps_done = 0;
  // The work layout in use (fomo_plan_tuning): family (0 separate sweeps, 1 a fused interior
  // launch), shape index, its extent (2D: strip rows; 3D: z-planes of a segment of the fused
  // interior, else of the sweeps), and where the choice came from (0 timed by this plan, 1 this
  // process's tuning cache, 2 the tuning file, 3 fixed by the descriptor or the environment).
  int tune_family = 0, tune_shape = 0, tune_extent = 0, tune_source = 3;

  virtual ~PlanBase();
  // Enqueue one part of the step with index *step + koff from state parity s: 0 velocity phase,
  // 1 pressure/stress phase, 2 receiver gather.  Parts 0 and 1 read only the parity-s buffers and
  // the step counter besides what they write, so repeating one is idempotent.
  virtual hipError_t enqueue_part(int part, int s, int koff) = 0;
  // True when the velocity phase also records the previous step's receivers; part 2 then only
  // records the latest state (once, at the end of a run of steps).
  virtual bool gather_in_sweep() const { return false; }
  // Enqueue m >= 1 steps from the current state and the counter tick; steps_done += m.  ev
  // (optional, 4 events) brackets the sweeps of the first step: ev[0] velocity ev[1] stress ev[2].
  hipError_t run_steps(int m, hipEvent_t* ev = nullptr);
  virtual int reset() = 0;
  virtual int read_seis(int shot, float* const* dst) = 0;
  // Device address of one slot's recorded traces of a shot (n_rec x nt, receiver index fastest)
  // and the shot's receiver count; fails for a slot the plan does not record.
  virtual int seis_device(int shot, int slot, const float** ptr, int* n_rec) const = 0;
  virtual int read_field(int shot, int which, float* out) = 0;
  // Snapshot buffer of a shot (count x cells of one snapshot, Float32); *count = 0 without one.
  virtual int snap_buffer(int shot, const float** ptr, int* count, size_t* cells) const = 0;
  virtual int dims(int shot, int* nx, int* ny, int* nz, int* pad) const = 0;
  virtual int n_shots() const = 0;
  // The kernel launches of one step, each enqueueable alone for per-kernel timing: launch k from
  // state parity s (idempotent, like a repeated part), its name, the cells it updates and the
  // algorithmic bytes per cell of the reference's per-phase traffic (16-bit storage: half).
  struct LaunchInfo {
    const char* name;
    long long cells;
    int bytes_per_cell;
  };
  virtual int n_launches() const = 0;
  virtual LaunchInfo launch_info(int k) const = 0;
  virtual hipError_t enqueue_launch(int k, int s) = 0;

  // Advance n steps (graph replays where they fit), then record the latest state's receivers when
  // the sweeps record them one step late.
  int advance(int n);
  int build_graph();
  // Waits for the stream; FOMO_EOVERFLOW when a 16-bit store overflowed since the last reset.
  int check_overflow();
  template <class T>
  int alloc(T** ptr, size_t count) {
    void* v = nullptr;
    const size_t bytes = (count ? count : 1) * sizeof(T);
    hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess)
      return set_error(e == hipErrorOutOfMemory ? FOMO_ENOMEM : FOMO_EHIP,
                       "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
    allocs.push_back(v);
    *ptr = (T*)v;
    return FOMO_OK;
  }
  template <class T>
  int upload(T** dst, const T* src, size_t count) {
    int rc = alloc(dst, count);
    if (rc) return rc;
    if (count) FOMO_HIP_TRY(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return FOMO_OK;
  }
};

// Host-side setup shared by the solvers (fomo_api.hip).
void pad2d(const float* in, int nxi, int nzi, int pad, int x0, int nx_model, std::vector<float>& out);
void buoyancy2d(const std::vector<float>& rho, int nx, int nz, std::vector<float>& bx, std::vector<float>& bz);
void elastic_params2d(const std::vector<float>& vp, const std::vector<float>& vs, const std::vector<float>& rho,
                      int nx, int nz, std::vector<float>& lam, std::vector<float>& lam2mu, std::vector<float>& muxz);
void weight_lut(int H, int pad, double off, float* out);
// Upload n Float32 values as the storage type (16-bit: RN(v * enc)) into a new plan allocation.
int upload_stored(PlanBase* p, int storage, void** dst, const float* h, size_t n, float enc);
// Max |v| over n values (0 if none).
float max_abs(const float* v, size_t n);
// Cerjan sponge factors by distance d = 0..pad-1 to the padded edge (src/boundary/sponge.jl:30-50).
void sponge_lut(int nbc, int pad, float factor, float* out);
float min_positive(const float* a, size_t n);

}  // namespace fomo

struct fomo_plan {
  fomo::PlanBase* impl;
};
