// learned_api.hip -- C-ABI of the learned absorbing boundary (2D acoustic, shot 0 of a plan).
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <vector>

#include "learned2d.h"
#include "plan.h"
#include "plan2d.h"

using namespace fomo;

static Plan2D* acoustic_plan(fomo_plan* p) {
  if (!p || !p->impl) return nullptr;
  Plan2D* q = dynamic_cast<Plan2D*>(p->impl);
  if (!q || q->d.eq != FOMO_ACOUSTIC2D) return nullptr;
  return q;
}

// The BF16 weights rearranged into the inference kernel's B-fragment order (kFragW1 ..).
static std::vector<__hip_bfloat16> weight_fragments(const std::vector<__hip_bfloat16>& w) {
  std::vector<__hip_bfloat16> f(kFragParams, __float2bfloat16(0.0f));
  for (int lane = 0; lane < 64; ++lane) {
    const int col = lane & 31, half = lane >> 5;
    for (int j = 0; j < 8; ++j) {
      for (int kb = 0; kb < 5; ++kb) {
        const int k = kb * 16 + half * 8 + j;
        if (k < kMlpIn) f[kFragW1 + (kb * 64 + lane) * 8 + j] = w[kOffW1 + col * kMlpIn + k];
      }
      for (int kb = 0; kb < 2; ++kb) {
        const int k = kb * 16 + half * 8 + j;
        f[kFragW2 + (kb * 64 + lane) * 8 + j] = w[kOffW2 + col * kMlpH + k];
        if (col < kMlpOut) f[kFragW3 + (kb * 64 + lane) * 8 + j] = w[kOffW3 + col * kMlpH + k];
      }
    }
  }
  for (int o = 0; o < kMlpH; ++o) {
    f[kFragB + o] = w[kOffB1 + o];
    f[kFragB + kMlpH + o] = w[kOffB2 + o];
  }
  for (int o = 0; o < kMlpOut; ++o) f[kFragB + 2 * kMlpH + o] = w[kOffB3 + o];
  return f;
}

static int upload_weights(Plan2D* q, const float* w) {
  std::vector<__hip_bfloat16> h(kMlpParams);
  for (int e = 0; e < kMlpParams; ++e) h[e] = __float2bfloat16(w[e]);
  const std::vector<__hip_bfloat16> f = weight_fragments(h);
  FOMO_HIP_TRY(hipMemcpyAsync(q->w_bf16, h.data(), kMlpParams * sizeof(__hip_bfloat16), hipMemcpyHostToDevice,
                              q->stream));
  FOMO_HIP_TRY(hipMemcpyAsync(q->w_frag, f.data(), kFragParams * sizeof(__hip_bfloat16), hipMemcpyHostToDevice,
                              q->stream));
  FOMO_HIP_TRY(hipStreamSynchronize(q->stream));
  return FOMO_OK;
}

extern "C" int fomo_learned_n_params(void) { return kMlpParams; }

extern "C" int fomo_learned_enable(fomo_plan* p, const float* weights, int32_t train) {
  Plan2D* q = acoustic_plan(p);
  if (!q) return set_error(FOMO_EINVAL, "the learned boundary needs an acoustic2d plan");
  if (q->shots.size() != 1) return set_error(FOMO_EUNSUP, "the learned boundary runs on single-shot plans");
  if (!weights) return set_error(FOMO_EINVAL, "weights must not be NULL");
  if (q->learned) return set_error(FOMO_EINVAL, "the learned boundary is already enabled");
  if (q->d.boundary != FOMO_BOUNDARY_HABC) return set_error(FOMO_EINVAL, "the learned boundary replaces the HABC");
  Shot2D& S = q->shots[0];
  const size_t n = (size_t)S.nx * S.nz;
  std::vector<int> cells;
  for (int j = 0; j < S.nz; ++j)
    for (int i = 0; i < S.nx; ++i) {
      const int d = std::min(std::min(i, S.nx - 1 - i), std::min(j, S.nz - 1 - j));
      if (d >= 1 && d <= S.H) cells.push_back(i + j * S.nx);
    }
  // Inference tiles over the four bands of the frame (cells with 1 <= dist <= H): top and bottom
  // bands (rows 1..H and nz-1-H..nz-2, columns 1..nx-2) in 64 x 4 tiles, left and right bands
  // (columns 1..H and nx-1-H..nx-2, rows H+1..nz-2-H) in tiles as tall as 256 cells allow.
  std::vector<InferTile> tiles;
  auto add_band = [&](int xa, int xb, int za, int zb) {  // inclusive-exclusive ranges
    for (int x = xa; x < xb; x += 64) {
      const int w = std::min(64, xb - x);
      const int hs = std::max(1, std::min(16, kInferTileCells / w));
      for (int z = za; z < zb; z += hs) tiles.push_back(InferTile{x, z, w, std::min(hs, zb - z)});
    }
  };
  {
    const int H = S.H, nx = S.nx, nz = S.nz;
    add_band(1, nx - 1, 1, std::min(H + 1, nz - 1));
    add_band(1, nx - 1, std::max(nz - 1 - H, H + 1), nz - 1);
    add_band(1, std::min(H + 1, nx - 1), H + 1, nz - 1 - H);
    add_band(std::max(nx - 1 - H, H + 1), nx - 1, H + 1, nz - 1 - H);
    size_t covered = 0;
    for (const auto& t : tiles) {
      if ((t.w + 4) * (t.h + 4) > kInferStage || t.w * t.h > kInferTileCells)
        return set_error(FOMO_EINVAL, "inference tile exceeds its staging area");
      covered += (size_t)t.w * t.h;
    }
    if (covered != cells.size()) return set_error(FOMO_EINVAL, "inference tiles do not cover the frame cells");
  }
  int rc;
  if ((rc = q->upload(&q->d_cells, cells.data(), cells.size()))) return rc;
  q->n_cells = (int)cells.size();
  {
    InferTile* dt = nullptr;
    if ((rc = q->upload(&dt, tiles.data(), tiles.size()))) return rc;
    q->d_itiles = dt;
    q->n_itiles = (int)tiles.size();
  }
  __hip_bfloat16* wb = nullptr;
  if ((rc = q->alloc(&wb, kMlpParams))) return rc;
  q->w_bf16 = wb;
  __hip_bfloat16* wf = nullptr;
  if ((rc = q->alloc(&wf, kFragParams))) return rc;
  q->w_frag = wf;
  for (int k = 0; k < 2; ++k)
    if ((rc = q->alloc(&q->alpha[k], n))) return rc;
  if (train) {
    for (int k = 0; k < 3; ++k)
      if ((rc = q->alloc(&q->sens[k], n))) return rc;
    q->train_blocks = 240;
    const size_t width = kMlpParams + 1;
    if ((rc = q->alloc(&q->partial, width * q->train_blocks)) || (rc = q->alloc(&q->grad, width))) return rc;
    FOMO_HIP_TRY(hipMemsetAsync(q->partial, 0, width * q->train_blocks * sizeof(float), q->stream));
  }
  if ((rc = upload_weights(q, weights))) return rc;
  S.alpha[0] = q->alpha[0];
  S.alpha[1] = q->alpha[1];
  for (int k = 0; k < 3; ++k) S.sens[k] = q->sens[k];
  FOMO_HIP_TRY(hipMemcpy(q->d_shots, q->shots.data(), q->shots.size() * sizeof(Shot2D), hipMemcpyHostToDevice));
  q->learned = true;
  q->training = train != 0;
  if (q->gexec) {
    (void)hipGraphExecDestroy(q->gexec);
    (void)hipGraphDestroy(q->graph);
    q->gexec = nullptr;
    q->graph = nullptr;
    if ((rc = q->build_graph())) return rc;
  }
  return FOMO_OK;
}

extern "C" int fomo_learned_set_weights(fomo_plan* p, const float* weights) {
  Plan2D* q = acoustic_plan(p);
  if (!q || !q->learned) return set_error(FOMO_EINVAL, "plan has no learned boundary");
  if (!weights) return set_error(FOMO_EINVAL, "weights must not be NULL");
  return upload_weights(q, weights);
}

// One training step: the teacher and the student each advance one step; then the one-step loss at
// the student's frame cells against the teacher's state at the same physical cells is
// differentiated through the MLP and its weight gradients are accumulated.  The student cell
// (i, j) is the teacher cell (i + off_x, j + off_z).  Asynchronous.
extern "C" int fomo_learned_train_step(fomo_plan* student, fomo_plan* teacher, int32_t off_x, int32_t off_z,
                                       float scale_p, float scale_v) {
  Plan2D* q = acoustic_plan(student);
  Plan2D* t = acoustic_plan(teacher);
  if (!q || !q->training) return set_error(FOMO_EINVAL, "student plan is not in training mode");
  if (!t || t->shots.size() != 1) return set_error(FOMO_EINVAL, "teacher must be a single-shot acoustic2d plan");
  if (q->sweep_storage != FOMO_STORE_F32 || t->sweep_storage != FOMO_STORE_F32)
    return set_error(FOMO_EINVAL,
                     "training needs Float32 storage of fields and materials in the student and teacher plans");
  const Shot2D& S = q->shots[0];
  const Shot2D& T = t->shots[0];
  if (off_x < 0 || off_z < 0 || S.nx + off_x > T.nx || S.nz + off_z > T.nz)
    return set_error(FOMO_EINVAL, "student grid does not fit in the teacher grid at this offset");
  if (!(scale_p > 0.0f) || !(scale_v > 0.0f)) return set_error(FOMO_EINVAL, "scales must be > 0");
  const int ts = (int)(t->steps_done % 2);
  FOMO_HIP_TRY(t->run_steps(1));
  hipEvent_t ev;
  FOMO_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  FOMO_HIP_TRY(hipEventRecord(ev, t->stream));
  const int s = (int)(q->steps_done % 2);
  FOMO_HIP_TRY(q->run_steps(1));
  FOMO_HIP_TRY(hipStreamWaitEvent(q->stream, ev, 0));
  MlpTrainArgs a{};
  a.m.w_bf16 = q->w_bf16;
  a.m.cells = q->d_cells;
  a.m.n_cells = q->n_cells;
  a.m.nx = S.nx;
  a.m.nz = S.nz;
  a.m.H = S.H;
  a.m.p = S.fld[s][0];
  a.m.vx = S.fld[s][1];
  a.m.vz = S.fld[s][2];
  for (int f = 0; f < 3; ++f) a.m.dec[f] = 1.0f;
  a.m.storage = FOMO_STORE_F32;
  a.p1 = (const float*)S.fld[1 - s][0];
  a.vx1 = (const float*)S.fld[1 - s][1];
  a.vz1 = (const float*)S.fld[1 - s][2];
  a.sens_vx = q->sens[0];
  a.sens_vz = q->sens[1];
  a.sens_p = q->sens[2];
  const int to = 1 - ts;
  a.tp = (const float*)T.fld[to][0];
  a.tvx = (const float*)T.fld[to][1];
  a.tvz = (const float*)T.fld[to][2];
  a.tnx = T.nx;
  a.off_x = off_x;
  a.off_z = off_z;
  a.inv_scale_p = 1.0f / scale_p;
  a.inv_scale_v = 1.0f / scale_v;
  a.partial = q->partial;
  FOMO_HIP_TRY(launch_mlp_train(a, q->train_blocks, q->stream));
  // The teacher's next steps overwrite the buffer this launch reads (T.fld[to] is the input of
  // its step after next): its stream waits for the training launch.
  FOMO_HIP_TRY(hipEventRecord(ev, q->stream));
  FOMO_HIP_TRY(hipStreamWaitEvent(t->stream, ev, 0));
  (void)hipEventDestroy(ev);
  return FOMO_OK;
}

// Sum of the accumulated weight gradients (and loss) since the last call; clears the accumulator.
extern "C" int fomo_learned_grad(fomo_plan* p, float* grad, double* loss) {
  Plan2D* q = acoustic_plan(p);
  if (!q || !q->training) return set_error(FOMO_EINVAL, "plan is not in training mode");
  const int width = kMlpParams + 1;
  FOMO_HIP_TRY(launch_reduce_partials(q->partial, q->train_blocks, width, q->grad, q->stream));
  std::vector<float> h(width);
  FOMO_HIP_TRY(hipMemcpyAsync(h.data(), q->grad, width * sizeof(float), hipMemcpyDeviceToHost, q->stream));
  FOMO_HIP_TRY(hipMemsetAsync(q->partial, 0, (size_t)width * q->train_blocks * sizeof(float), q->stream));
  FOMO_HIP_TRY(hipStreamSynchronize(q->stream));
  if (grad)
    for (int e = 0; e < kMlpParams; ++e) grad[e] = h[e];
  if (loss) *loss = h[kMlpParams];
  return FOMO_OK;
}

// Frame-cell list and current blend weights of a learned plan (n = number of frame cells).
extern "C" int fomo_learned_alpha(fomo_plan* p, int32_t* n, int32_t* cells, float* alpha_v, float* alpha_p) {
  Plan2D* q = acoustic_plan(p);
  if (!q || !q->learned) return set_error(FOMO_EINVAL, "plan has no learned boundary");
  FOMO_HIP_TRY(hipStreamSynchronize(q->stream));
  if (n) *n = q->n_cells;
  std::vector<int> c(q->n_cells);
  FOMO_HIP_TRY(hipMemcpy(c.data(), q->d_cells, c.size() * sizeof(int), hipMemcpyDeviceToHost));
  if (cells)
    for (int k = 0; k < q->n_cells; ++k) cells[k] = c[k];
  const size_t np = (size_t)q->shots[0].nx * q->shots[0].nz;
  std::vector<float> a(np);
  float* dst[2] = {alpha_v, alpha_p};
  for (int m = 0; m < 2; ++m) {
    if (!dst[m]) continue;
    FOMO_HIP_TRY(hipMemcpy(a.data(), q->alpha[m], np * sizeof(float), hipMemcpyDeviceToHost));
    for (int k = 0; k < q->n_cells; ++k) dst[m][k] = a[c[k]];
  }
  return FOMO_OK;
}

// Unit entry for gradient checks: the MLP forward and backward on given features (n x 76) and
// output gradients dL/dalpha (n x 2).  Writes alpha (n x 2) and the summed weight gradients.
extern "C" int fomo_mlp_eval(const float* weights, int32_t n, const float* feats, const float* galpha, float* alpha,
                             float* grad) {
  if (!weights || !feats || !galpha || n < 1) return set_error(FOMO_EINVAL, "invalid arguments");
  if (fomo_device_count() < 1) return set_error(FOMO_EHIP, "no HIP device visible");
  std::vector<__hip_bfloat16> wb(kMlpParams);
  for (int e = 0; e < kMlpParams; ++e) wb[e] = __float2bfloat16(weights[e]);
  const int blocks = 16, width = kMlpParams + 1;
  void* dw = nullptr;
  float *df = nullptr, *dg = nullptr, *da = nullptr, *dp = nullptr, *dout = nullptr;
  auto cleanup = [&]() {
    for (void* v : {dw, (void*)df, (void*)dg, (void*)da, (void*)dp, (void*)dout})
      if (v) (void)hipFree(v);
  };
  hipError_t e = hipMalloc(&dw, kMlpParams * sizeof(__hip_bfloat16));
  if (e == hipSuccess) e = hipMalloc(&df, (size_t)n * kMlpIn * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dg, (size_t)n * 2 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&da, (size_t)n * 2 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dp, (size_t)blocks * width * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dout, (size_t)width * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dw, wb.data(), kMlpParams * sizeof(__hip_bfloat16), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(df, feats, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dg, galpha, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dp, 0, (size_t)blocks * width * sizeof(float));
  if (e == hipSuccess) e = launch_mlp_eval(dw, df, dg, n, da, dp, blocks, nullptr);
  if (e == hipSuccess) e = launch_reduce_partials(dp, blocks, width, dout, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  std::vector<float> h(width);
  if (e == hipSuccess && alpha) e = hipMemcpy(alpha, da, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(h.data(), dout, width * sizeof(float), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return set_error(FOMO_EHIP, std::string("fomo_mlp_eval: ") + hipGetErrorString(e));
  if (grad)
    for (int k = 0; k < kMlpParams; ++k) grad[k] = h[k];
  return FOMO_OK;
}
