// plan2d.hip -- 2D plans: one or many shots (windows of one model) advanced by shared launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>

#include "fomo_common.h"
#include "plan.h"
#include "plan2d.h"
#include "learned2d.h"
#include "tune.h"

namespace fomo {

Plan2D::~Plan2D() = default;

// Work items of a tile shape: frame tiles (64 x kFrameRows) over the four pad bands and interior
// strips (64 x strip_rows) over the user's interior, for every shot, sorted by (shot, row, column).
static std::vector<Work2D> work_items(const std::vector<Shot2D>& shots, const std::vector<ShotMeta>& meta, int cfg,
                                      int frame_rows = kFrameRows) {
  std::vector<Work2D> items;
  const int zs = strip_rows(cfg);
  auto add_rect = [&](int s, int kind, int x0, int x1, int z0, int z1, int th) {
    for (int j = z0; j < z1; j += th)
      for (int i = x0; i < x1; i += kTileW)
        items.push_back(Work2D{s, kind, i, j, std::min(kTileW, x1 - i), std::min(th, z1 - j)});
  };
  for (size_t s = 0; s < shots.size(); ++s) {
    const int nx = shots[s].nx, nz = shots[s].nz, pad = meta[s].pad;
    const int xe = pad + meta[s].nxi, ze = pad + meta[s].nzi;
    add_rect((int)s, 0, 0, nx, 0, pad, frame_rows);       // top band
    add_rect((int)s, 0, 0, pad, pad, ze, frame_rows);     // left band
    add_rect((int)s, 1, pad, xe, pad, ze, zs);            // interior
    add_rect((int)s, 0, xe, nx, pad, ze, frame_rows);     // right band
    add_rect((int)s, 0, 0, nx, ze, nz, frame_rows);       // bottom band
  }
  std::stable_sort(items.begin(), items.end(), [](const Work2D& a, const Work2D& b) {
    if (a.shot != b.shot) return a.shot < b.shot;
    if (a.j0 != b.j0) return a.j0 < b.j0;
    return a.i0 < b.i0;
  });
  return items;
}

// Launch order of a list: 8 contiguous runs of about equal cost, one per XCD, interleaved so that
// block b (which the hardware places on XCD b % 8) takes the next item of run b % 8.  Slots past
// the end of a shorter run hold an empty item (kind -1).
static std::vector<Work2D> xcd_layout(const std::vector<Work2D>& items) {
  // A frame cell weighs 1.5 interior cells (ring recompute, blend).  Measured on elastic2d
  // 2000x800 FD8: 1.5 gives the fastest step (34.4 us; 1.0: 36.4, 2.0: 34.9, 3.0: 35.8, 4.0: 38.2),
  // and at 8192^2 and the 64-shot batch it matches 2.0.  FOMO_FRAME_COST overrides it.
  double fc = 1.5;
  if (const char* env = std::getenv("FOMO_FRAME_COST")) fc = std::min(64.0, std::max(0.5, std::atof(env)));  // NaN -> 0.5
  std::vector<double> cost(items.size());
  double total = 0.0;
  for (size_t k = 0; k < items.size(); ++k) {
    cost[k] = (double)items[k].w * items[k].h * (items[k].kind == 0 ? fc : 1.0);
    total += cost[k];
  }
  std::vector<std::vector<Work2D>> runs(8);
  double acc = 0.0;
  for (size_t k = 0; k < items.size(); ++k) {
    const int r = std::min(7, (int)(8.0 * (acc + 0.5 * cost[k]) / std::max(total, 1.0)));
    runs[r].push_back(items[k]);
    acc += cost[k];
  }
  size_t len = 0;
  for (auto& r : runs) len = std::max(len, r.size());
  std::vector<Work2D> flat(8 * len, Work2D{0, -1, 0, 0, 0, 0});
  for (int r = 0; r < 8; ++r)
    for (size_t k = 0; k < runs[r].size(); ++k) flat[k * 8 + r] = runs[r][k];
  return flat;
}

// Rows of one interior strip of the fused launch: 24, 32, 48, 96 or 176 (FOMO_FUSED2D_ROWS forces
// a length, rounded up to a multiple of 8).  Each strip recomputes velocities on about 2N rows
// past its ends and reads as many rows of its inputs again.  Strips alternate their march
// direction by row band, so that two strips read those rows at the same time and the second read
// finds them in the caches; short strips then cost little: at 8192^2 FD8 the acoustic2d step
// takes 0.576 / 0.575 / 0.578 / 0.582 ms at 16 / 24 / 32 / 40 rows, against 0.601 ms with every
// strip marching down at its best length (176), and elastic2d 1.009 / 1.002 / 1.013 / 1.045 ms at
// 24 / 32 / 48 / 96 rows against 1.042 ms (16-bit storage: fastest at 32-48 rows).  The autotuner
// times every length.
static int fused_rows(int cfg) {
  if (const char* env = std::getenv("FOMO_FUSED2D_ROWS")) {
    const int r = std::atoi(env);
    if (r > 0) return std::min(4096, (r + 7) / 8 * 8);
  }
  static const int rows[kNumFusedCfgs] = {24, 32, 48, 96, 176};
  return rows[cfg];
}

int Plan2D::make_work(int cfg) {
  if (fused ? d_fwork[cfg] != nullptr : d_work[cfg] != nullptr) return FOMO_OK;
  if (coupled && !d_frame) {
    std::vector<Work2D> frame;
    for (const Work2D& w : work_items(shots, meta, 0))
      if (w.kind == 0) frame.push_back(w);
    const std::vector<Work2D> ff = xcd_layout(frame);
    n_frame = (int)ff.size();
    const int rc = upload(&d_frame, ff.data(), ff.size());
    if (rc) return rc;
  }
  if (fused) {
    std::vector<Work2D> frame, inner;
    for (const Work2D& w : work_items(shots, meta, 0, kFrameRowsF))
      if (w.kind == 0) frame.push_back(w);
    const int rows = fused_rows(cfg), cols = fused2d_width(d.eq, N);
    for (size_t s = 0; s < shots.size(); ++s) {
      const int pad = meta[s].pad, xe = pad + meta[s].nxi, ze = pad + meta[s].nzi;
      // strips of about equal width: a window of 256 interior columns runs as two of 128 rather
      // than one of the full width and a sliver, whose block would march every row for a few columns
      const int nstrip = std::max(1, (meta[s].nxi + cols - 1) / cols);
      const int cw = (meta[s].nxi + nstrip - 1) / nstrip;
      // strips of odd row bands march up (kind 2), so that vertically adjacent strips read the rows
      // they share at the same time
      for (int j = pad, band = 0; j < ze; j += rows, ++band)
        for (int i = pad; i < xe; i += cw)
          inner.push_back(Work2D{(int)s, (band & 1) ? 2 : 1, i, j, std::min(cw, xe - i), std::min(rows, ze - j)});
    }
    if (!d_frame) {
      const std::vector<Work2D> ff = xcd_layout(frame);
      n_frame = (int)ff.size();
      const int rc = upload(&d_frame, ff.data(), ff.size());
      if (rc) return rc;
    }
    if (d.eq == FOMO_ELASTIC2D && !d_frame8) {
      std::vector<Work2D> frame8;
      for (const Work2D& w : work_items(shots, meta, 0))
        if (w.kind == 0) frame8.push_back(w);
      const std::vector<Work2D> ff = xcd_layout(frame8);
      n_frame8 = (int)ff.size();
      const int rc = upload(&d_frame8, ff.data(), ff.size());
      if (rc) return rc;
    }
    const std::vector<Work2D> fi = xcd_layout(inner);
    n_fwork[cfg] = (int)fi.size();
    return upload(&d_fwork[cfg], fi.data(), fi.size());
  }
  const std::vector<Work2D> flat = xcd_layout(work_items(shots, meta, cfg));
  n_work[cfg] = (int)flat.size();
  return upload(&d_work[cfg], flat.data(), flat.size());
}

hipError_t Plan2D::enqueue_part(int part, int s, int koff) {
  GatherArgs g{};
  g.shots = d_shots;
  g.rec_shot = rec_shot;
  g.rec_c = rec_c;
  g.rec_out = rec_out;
  g.rec_stride = rec_stride;
  g.nslots = nslots;
  for (int k = 0; k < nslots; ++k) g.seis[k] = seis[k];
  g.n_entries = n_entries;
  g.nt = d.nt;
  if (part == 2) return launch_gather2d(storage, g, step, koff, s, stream);
  if (part == 1 && d.eq == FOMO_SCALAR2D) return hipSuccess;  // one phase per step
  SweepArgs a{};
  a.shots = d_shots;
  a.work = d_work[tile_cfg];
  a.n_work = n_work[tile_cfg];
  a.parity = s;
  for (int l = 0; l < N; ++l) a.a[l] = coef[l];
  a.step = step;
  a.koff = koff;
  a.nt = d.nt;
  a.c0 = c0;
  for (int l = 0; l < N; ++l) a.a2[l] = coef2[l];
  if (part == 0) {
    a.g = g;
    a.n_gather = (n_entries + kGatherBlock - 1) / kGatherBlock;
  }
  if (fused) {
    // part 0: the frame's velocities (and the receivers of the previous step), then the interior's
    // whole step; part 1: the frame's pressure or stresses
    a.work = d_frame;
    a.n_work = n_frame;
    if (part == 1 && only == 0) return hipSuccess;
    if (part == 1 && d.eq == FOMO_ELASTIC2D) {
      a.work = d_frame8;
      a.n_work = n_frame8;
      return launch_sweep2d(sweep_storage, kElStr, N, tile_cfg, a, stream);
    }
    if (part == 1) return launch_frame2d(sweep_storage, kAcPrs, N, a, stream);
  } else if (part == 1 && coupled) {
    a.work = d_frame;  // the interior's position half-steps ran in the velocity launch
    a.n_work = n_frame;
    return launch_sweep2d(sweep_storage, kCpPos, N, tile_cfg, a, stream);
  } else if (part == 1) {
    return launch_sweep2d(sweep_storage, d.eq == FOMO_ACOUSTIC2D ? kAcPrs : kElStr, N, tile_cfg, a, stream);
  }
  if (learned && only != 0) {
    // blend-weight corrections of every frame cell from the state at time n
    MlpArgs m{};
    m.w_bf16 = w_bf16;
    m.w_frag = w_frag;
    m.cells = d_cells;
    m.n_cells = n_cells;
    m.tiles = (const InferTile*)d_itiles;
    m.n_tiles = n_itiles;
    m.nx = shots[0].nx;
    m.nz = shots[0].nz;
    m.H = shots[0].H;
    m.p = shots[0].fld[s][0];
    m.vx = shots[0].fld[s][1];
    m.vz = shots[0].fld[s][2];
    for (int f = 0; f < 3; ++f) m.dec[f] = shots[0].fdec[f];
    m.storage = storage;
    m.alpha_v = alpha[0];
    m.alpha_p = alpha[1];
    const hipError_t e = launch_mlp_forward(m, stream);
    if (e != hipSuccess) return e;
  }
  const int ph = d.eq == FOMO_ACOUSTIC2D ? kAcVel : (d.eq == FOMO_ELASTIC2D ? kElVel : (coupled ? kCpVel : kScal));
  if (!fused) return launch_sweep2d(sweep_storage, ph, N, tile_cfg, a, stream);
  if (only != 0) {
    const hipError_t e = launch_frame2d(sweep_storage, ph, N, a, stream);
    if (e != hipSuccess || only == 1) return e;
  }
  a.work = d_fwork[tile_cfg];
  a.n_work = n_fwork[tile_cfg];
  a.n_gather = 0;
  return launch_fused2d(sweep_storage, d.eq, N, a, stream);
}

PlanBase::LaunchInfo Plan2D::launch_info(int k) const {
  long long cells = 0;
  for (const auto& sh : shots) cells += (long long)sh.nx * sh.nz;
  // words per cell read + written: acoustic velocity p,vx,vz,bx,bz -> vx,vz (7) and pressure
  // vx,vz,p,kappa -> p (5); elastic velocity 9, stress 11; scalar u^n,u^(n-1),c2 -> u^(n+1) (4)
  int words = 0;
  const char* name = "";
  if (fused) {
    // the interior launch is credited with both phases' reads and writes (acoustic 7 + 5 words per
    // cell, elastic 9 + 11)
    long long inner = 0;
    for (const auto& m : meta) inner += (long long)m.nxi * m.nzi;
    if (d.eq == FOMO_SCALAR2D) {
      if (k == 0) return {"scalar2d frame", cells - inner, 4 * elem};
      return {"scalar2d interior step", inner, 4 * elem};
    }
    if (d.eq == FOMO_ELASTIC2D) {
      if (k == 0) return {"elastic2d velocity frame", cells - inner, 9 * elem};
      if (k == 1) return {"elastic2d interior step (velocity + stress)", inner, 20 * elem};
      return {"elastic2d stress frame", cells - inner, 11 * elem};
    }
    if (k == 0) return {"acoustic2d velocity frame", cells - inner, 7 * elem};
    if (k == 1) return {"acoustic2d interior step (velocity + pressure)", inner, 12 * elem};
    return {"acoustic2d pressure frame", cells - inner, 5 * elem};
  }
  if (coupled) {
    // velocity half-step: P, S, dP/dt, dS/dt and the eight medium arrays -> dP/dt, dS/dt (14 words);
    // position half-step: P, S, dP/dt, dS/dt -> P, S (6 words).  The velocity launch also takes
    // the interior's position half-steps; only its velocity words are credited to it.
    long long inner = 0;
    for (const auto& m : meta) inner += (long long)m.nxi * m.nzi;
    if (k == 0) return {"coupled2d velocity sweep (interior position half-step not credited)", cells, 14 * elem};
    return {"coupled2d position frame", cells - inner, 6 * elem};
  }
  if (d.eq == FOMO_ACOUSTIC2D) {
    words = k == 0 ? 7 : 5;
    name = k == 0 ? "acoustic2d velocity sweep" : "acoustic2d pressure sweep";
  } else if (d.eq == FOMO_ELASTIC2D) {
    words = k == 0 ? 9 : 11;
    name = k == 0 ? "elastic2d velocity sweep" : "elastic2d stress sweep";
  } else {
    words = 4;
    name = "scalar2d sweep";
  }
  return {name, cells, words * elem};
}

int Plan2D::reset() {
  for (auto& sh : shots) {
    const size_t n = (size_t)sh.nx * sh.nz;
    for (int b = 0; b < 2; ++b)
      for (int f = 0; f < nfld; ++f)
        if (sh.fld[b][f]) FOMO_HIP_TRY(hipMemsetAsync(sh.fld[b][f], 0, n * elem, stream));
  }
  for (int k = 0; k < nslots; ++k)
    if (seis_total) FOMO_HIP_TRY(hipMemsetAsync(seis[k], 0, (size_t)seis_total * sizeof(float), stream));
  // snapshots written before the reset (autotuning, an earlier run) must not survive it
  for (size_t k = 0; k < shots.size(); ++k)
    if (shots[k].snap)
      FOMO_HIP_TRY(hipMemsetAsync(shots[k].snap, 0,
                                  (size_t)shots[k].snap_n * (coupled ? 2 : 1) * meta[k].nxi * meta[k].nzi *
                                      sizeof(float),
                                  stream));
  FOMO_HIP_TRY(hipMemsetAsync(step, 0, sizeof(int), stream));
  if (ovf) FOMO_HIP_TRY(hipMemsetAsync(ovf, 0, sizeof(unsigned), stream));
  steps_done = 0;
  return FOMO_OK;
}

int Plan2D::read_seis(int shot, float* const* dst) {
  if (shot < 0 || shot >= (int)shots.size()) return set_error(FOMO_EINVAL, "shot index out of range");
  FOMO_HIP_TRY(hipStreamSynchronize(stream));
  const size_t cnt = (size_t)meta[shot].n_rec * d.nt;
  for (int k = 0; k < nslots; ++k)
    if (dst[k] && cnt)
      FOMO_HIP_TRY(hipMemcpy(dst[k], seis[k] + meta[shot].seis_off, cnt * sizeof(float), hipMemcpyDeviceToHost));
  return FOMO_OK;
}

int Plan2D::seis_device(int shot, int slot, const float** ptr, int* n_rec) const {
  if (shot < 0 || shot >= (int)shots.size()) return set_error(FOMO_EINVAL, "shot index out of range");
  if (slot < 0 || slot >= nslots || !seis[slot]) return set_error(FOMO_EINVAL, "trace slot not recorded");
  *ptr = seis[slot] + meta[shot].seis_off;
  *n_rec = meta[shot].n_rec;
  return FOMO_OK;
}

int Plan2D::decode_field(int shot, int which) {
  const Shot2D& sh = shots[shot];
  const size_t n = (size_t)sh.nx * sh.nz;
  if (scratch_n < n) {
    if (scratch) {
      FOMO_HIP_TRY(hipStreamSynchronize(stream));
      auto it = std::find(allocs.begin(), allocs.end(), (void*)scratch);
      if (it != allocs.end()) allocs.erase(it);
      (void)hipFree(scratch);
      scratch = nullptr;
      scratch_n = 0;
    }
    const int rc = alloc(&scratch, n);
    if (rc) return rc;
    scratch_n = n;
  }
  FOMO_HIP_TRY(launch_decode(storage, current(shot, which), scratch, n, sh.fdec[which], stream));
  return FOMO_OK;
}

const void* Plan2D::current(int shot, int which) const {
  const Shot2D& sh = shots[shot];
  if (sh.nbuf == 3) return sh.fld[0][steps_done % 3];  // scalar u^k
  return sh.fld[steps_done % 2][which];
}

int Plan2D::read_field(int shot, int which, float* out) {
  if (shot < 0 || shot >= (int)shots.size()) return set_error(FOMO_EINVAL, "shot index out of range");
  if (which < 0 || which >= nfld) return set_error(FOMO_EINVAL, "field index out of range");
  const Shot2D& sh = shots[shot];
  const size_t n = (size_t)sh.nx * sh.nz;
  const void* src = current(shot, which);
  if (storage != FOMO_STORE_F32) {
    const int rc = decode_field(shot, which);
    if (rc) return rc;
    src = scratch;
  }
  FOMO_HIP_TRY(hipStreamSynchronize(stream));
  FOMO_HIP_TRY(hipMemcpy(out, src, n * sizeof(float), hipMemcpyDeviceToHost));
  return FOMO_OK;
}

int Plan2D::dims(int shot, int* nx, int* ny, int* nz, int* pad) const {
  if (shot < 0 || shot >= (int)shots.size()) return set_error(FOMO_EINVAL, "shot index out of range");
  if (nx) *nx = shots[shot].nx;
  if (ny) *ny = 1;
  if (nz) *nz = shots[shot].nz;
  if (pad) *pad = meta[shot].pad;
  return FOMO_OK;
}

void Plan2D::note_tuning(int source) {
  tune_family = fused ? 1 : 0;
  tune_shape = tile_cfg;
  tune_extent = fused ? fused_rows(tile_cfg) : strip_rows(tile_cfg);
  tune_source = source;
}

// Times a few direct-launch steps per tile shape and keeps the fastest.  An acoustic2d plan that
// allows the fused interior launch times both families: the separate sweeps' strips (16, 32, 64
// rows) and the fused interior's (64 .. 256 rows); FOMO_2D_FUSED=1 / FOMO_2D_UNFUSED=1 keep one.
// (At 8192^2 the fused family is 12 % faster; on the 64-window batch, whose narrow windows leave
// lanes of the 249-column fused strips idle, the separate sweeps are 1 % faster.)  The choice is
// kept per (device and library build, equation, order, sweep storage, boundary, learned, families,
// shot count, padded size and window widths), in this process and in the tuning file (tune.h), so a
// later process with the same problem reuses it; FOMO_TILE_CFG forces a shape of the current family.
int Plan2D::autotune() {
  long long cells = 0;
  unsigned long long widths = 1469598103934665603ull;  // FNV-1a over the shots' padded widths
  for (auto& sh : shots) {
    cells += (long long)sh.nx * sh.nz;
    widths = (widths ^ (unsigned long long)sh.nx) * 1099511628211ull;
  }
  const bool only_fused = fusable && std::getenv("FOMO_2D_FUSED");
  const bool try_fused = fusable && !std::getenv("FOMO_2D_UNFUSED");
  const bool try_sweeps = !only_fused || !try_fused;
  const int fam = (try_fused ? 1 : 0) + (try_sweeps ? 2 : 0);
  if (const char* env = std::getenv("FOMO_TILE_CFG")) {
    const int c = std::atoi(env);
    if (c >= 0 && c < n_cfgs()) {
      tile_cfg = c;
      note_tuning(3);
      return make_work(c);
    }
  }
  const std::string key = tune_device_tag() + " 2d eq=" + std::to_string(d.eq) + " N=" + std::to_string(N) +
                          " storage=" + std::to_string(sweep_storage) + " boundary=" + std::to_string(d.boundary) +
                          " learned=" + std::to_string(learned ? 1 : 0) + " families=" + std::to_string(fam) +
                          " shots=" + std::to_string(shots.size()) + " cells=" + std::to_string(cells) +
                          " widths=" + std::to_string(widths);
  std::vector<int> rec;
  if (tune_lookup(key, rec) && rec.size() == 2 && (rec[0] == 0 || rec[0] == 1)) {
    const bool f = rec[0] == 1;
    const int c = rec[1];
    if ((f ? try_fused : try_sweeps) && c >= 0 && c < (f ? kNumFusedCfgs : kNumTileCfgs)) {
      fused = f;
      tile_cfg = c;
      note_tuning(tune_last_source());
      return make_work(tile_cfg);
    }
  }
  std::vector<std::pair<bool, int>> cand;
  if (try_sweeps)
    for (int c = 0; c < kNumTileCfgs; ++c) cand.push_back({false, c});
  if (try_fused)
    for (int c = 0; c < kNumFusedCfgs; ++c) cand.push_back({true, c});
  for (const auto& c : cand) {
    fused = c.first;
    const int rc = make_work(c.second);
    if (rc) return rc;
  }
  hipEvent_t e0, e1;
  FOMO_HIP_TRY(hipEventCreate(&e0));
  FOMO_HIP_TRY(hipEventCreate(&e1));
  // Every shape is timed in two interleaved rounds (so that no shape is favoured by running first
  // or last) over about 2^24 cell-updates each, and keeps its faster round.
  std::vector<float> best_ms(cand.size(), 1e30f);
  const int reps = (int)std::max<long long>(4, std::min<long long>(32, (1LL << 24) / std::max(cells, 1LL)));
  for (int round = 0; round < 2; ++round)
    for (size_t c = 0; c < cand.size(); ++c) {
      fused = cand[c].first;
      tile_cfg = cand[c].second;
      FOMO_HIP_TRY(run_steps(2));
      FOMO_HIP_TRY(hipEventRecord(e0, stream));
      FOMO_HIP_TRY(run_steps(reps));
      FOMO_HIP_TRY(hipEventRecord(e1, stream));
      FOMO_HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.0f;
      FOMO_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      best_ms[c] = std::min(best_ms[c], ms);
    }
  size_t best = 0;
  for (size_t c = 1; c < cand.size(); ++c)
    if (best_ms[c] < best_ms[best]) best = c;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  fused = cand[best].first;
  tile_cfg = cand[best].second;
  tune_store(key, {fused ? 1 : 0, tile_cfg});
  note_tuning(0);
  return reset();
}

static int check_pts(const char* what, int n, const int32_t* x, const int32_t* z, int nx, int nz) {
  if (n > 0 && (!x || !z)) return set_error(FOMO_EINVAL, std::string(what) + " arrays must not be NULL");
  for (int s = 0; s < n; ++s)
    if (x[s] < 1 || x[s] > nx || z[s] < 1 || z[s] > nz)
      return set_error(FOMO_EINVAL, std::string(what) + " index out of range: need 1 <= x <= " + std::to_string(nx) +
                                        ", 1 <= z <= " + std::to_string(nz));
  return FOMO_OK;
}

int create_plan2d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, int n_shots,
                  const fomo_shot* sh, Plan2D** out) {
  *out = nullptr;
  if (d->eq != FOMO_ACOUSTIC2D && d->eq != FOMO_ELASTIC2D && d->eq != FOMO_SCALAR2D && d->eq != FOMO_COUPLED2D)
    return set_error(FOMO_EINVAL, "not a 2D equation");
  const bool scalar = d->eq == FOMO_SCALAR2D, coupled = d->eq == FOMO_COUPLED2D;
  if (n_shots < 1) return set_error(FOMO_EINVAL, "at least one shot is required");
  if (!vp || (!scalar && !coupled && !rho) || ((d->eq == FOMO_ELASTIC2D || coupled) && !vs))
    return set_error(FOMO_EINVAL, "model arrays must not be NULL");
  if (!sh) return set_error(FOMO_EINVAL, "shot table is NULL");
  for (int s = 0; s < n_shots; ++s) {
    if (sh[s].x0 < 1 || sh[s].nx < 1 || sh[s].x0 + sh[s].nx - 1 > d->nx)
      return set_error(FOMO_EINVAL, "shot " + std::to_string(s) + ": window outside the model");
    int rc = check_pts("source", sh[s].n_src, sh[s].sx, sh[s].sz, sh[s].nx, d->nz);
    if (!rc) rc = check_pts("receiver", sh[s].n_rec, sh[s].rx, sh[s].rz, sh[s].nx, d->nz);
    if (rc) return rc;
    if (sh[s].n_src > 0 && !sh[s].wavelet) return set_error(FOMO_EINVAL, "wavelet must not be NULL");
  }
  if (d->storage < FOMO_STORE_F32 || d->storage > FOMO_STORE_F32_MAT16) return set_error(FOMO_EINVAL, "unknown storage");
  if (d->storage == FOMO_STORE_F32_MAT16 && scalar)
    return set_error(FOMO_EINVAL, "Float16 materials need a velocity-stress equation (acoustic2d, elastic2d)");
  if (coupled && d->storage != FOMO_STORE_F32) return set_error(FOMO_EUNSUP, "coupled2d runs with Float32 storage");
  if (coupled && d->boundary != FOMO_BOUNDARY_HABC)
    return set_error(FOMO_EUNSUP, "coupled2d uses the HABC (as in the reference)");
  Plan2D* p = new Plan2D();
  p->d = *d;
  const bool mat16 = d->storage == FOMO_STORE_F32_MAT16;
  p->storage = mat16 ? FOMO_STORE_F32 : d->storage;
  p->mat_storage = mat16 ? FOMO_STORE_F16 : d->storage;
  p->sweep_storage = d->storage;
  // acoustic2d and elastic2d may fuse the interior's two phases, scalar2d run its interior as a
  // row march (the autotuner decides; a forced tile shape is one of that family unless
  // FOMO_2D_UNFUSED=1)
  p->fusable = d->eq == FOMO_ACOUSTIC2D || d->eq == FOMO_ELASTIC2D || d->eq == FOMO_SCALAR2D;
  p->elem = p->storage == FOMO_STORE_F32 ? 4 : 2;
  p->N = d->fd_order / 2;
  p->coupled = coupled;
  if (scalar) {
    fomo_centered_d2(d->fd_order, &p->c0, p->coef);
  } else if (coupled) {
    fomo_centered_d1(d->fd_order, p->coef);
    fomo_centered_d2(d->fd_order, &p->c0, p->coef2);
  } else {
    fomo_fd_coefficients(d->fd_order, p->coef);
  }
  // scalar: the three u buffers; coupled: P, S, dP/dt, dS/dt
  p->nfld = d->eq == FOMO_ELASTIC2D ? 5 : (coupled ? 4 : 3);
  p->nslots = d->eq == FOMO_ACOUSTIC2D ? 3 : (scalar ? 1 : 2);
  auto bail = [&](int code) {
    delete p;
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(set_error(FOMO_EHIP, std::string("stream: ") + hipGetErrorString(e)));
  const int pad = d->nbc + p->N, H = pad - 1;
  const float dtx = d->dt / d->dh;
  // Cerjan sponge factors by distance, shared by every shot
  float* d_damp = nullptr;
  if (d->boundary == FOMO_BOUNDARY_SPONGE) {
    std::vector<float> lut(pad);
    sponge_lut(d->nbc, pad, d->sponge_factor, lut.data());
    int rc0 = p->upload(&d_damp, lut.data(), lut.size());
    if (rc0) return bail(rc0);
  }
  // Medium and HABC tables, shared between shots with the same window and v_ref.
  std::map<std::tuple<int, int, float>, int> medium_of;
  std::map<const float*, const float*> wav_dev;
  int rc = FOMO_OK;
  long long seis_off = 0;
  std::vector<int> rs, rcell, rstride;
  std::vector<long long> rout;
  if ((rc = p->alloc(&p->ovf, 1))) return bail(rc);
  for (int s = 0; s < n_shots; ++s) {
    const fomo_shot& S = sh[s];
    Shot2D x{};
    x.ovf = p->ovf;
    ShotMeta m{};
    m.nxi = S.nx;
    m.nzi = d->nz;
    m.pad = pad;
    m.x0 = S.x0;
    m.n_rec = S.n_rec;
    x.nx = S.nx + 2 * pad;
    x.nz = d->nz + 2 * pad;
    x.H = H;
    x.dtx = dtx;
    x.damp = d_damp;
    if (d->snap_interval > 0 && n_shots == 1 && d->nt / d->snap_interval > 0) {
      // snapshots of the interior, filled by the sweeps (no host copy inside the step loop);
      // coupled: the P slots, then the S slots
      x.snap_every = d->snap_interval;
      x.snap_n = d->nt / d->snap_interval;
      const size_t per = (size_t)x.snap_n * S.nx * d->nz, tot = per * (coupled ? 2 : 1);
      if ((rc = p->alloc(&x.snap, tot))) return bail(rc);
      if (coupled) x.snap2 = x.snap + per;
      if (hipMemset(x.snap, 0, tot * sizeof(float)) != hipSuccess)
        return bail(set_error(FOMO_EHIP, "snapshot buffer: hipMemset failed"));
    }
    const size_t n = (size_t)x.nx * x.nz;
    float v_ref = S.v_ref;
    if (!(v_ref > 0.0f)) {
      std::vector<float> crop((size_t)S.nx * d->nz);
      for (int j = 0; j < d->nz; ++j)
        std::memcpy(&crop[(size_t)j * S.nx], vp + (size_t)(S.x0 - 1) + (size_t)j * d->nx, S.nx * sizeof(float));
      v_ref = min_positive(crop.data(), crop.size());
      if (!(v_ref > 0.0f)) return bail(set_error(FOMO_EINVAL, "vp has no positive entry in shot window"));
    }
    float hc[3];
    fomo_habc_coefficients(v_ref, d->dt, d->dh, hc);
    x.q = x.q2 = hc[0];
    x.qt = x.qt2 = hc[1];
    x.qxt = x.qxt2 = hc[2];
    // window's vs, for the coupled S fields' reference velocity (coupled2d.jl:66)
    float v_ref_s = 0.0f;
    if (coupled) {
      v_ref_s = d->v_ref_s;
      if (!(v_ref_s > 0.0f)) {
        std::vector<float> crop((size_t)S.nx * d->nz);
        for (int j = 0; j < d->nz; ++j)
          std::memcpy(&crop[(size_t)j * S.nx], vs + (size_t)(S.x0 - 1) + (size_t)j * d->nx, S.nx * sizeof(float));
        v_ref_s = min_positive(crop.data(), crop.size());
        if (!(v_ref_s > 0.0f)) return bail(set_error(FOMO_EINVAL, "vs has no positive entry in shot window"));
      }
      fomo_habc_coefficients(v_ref_s, d->dt, d->dh, hc);
      x.q2 = hc[0];
      x.qt2 = hc[1];
      x.qxt2 = hc[2];
      // update_fields.jl:161-162: ih = 1/dh, ih2 = ih*ih
      x.dt = d->dt;
      x.ih = 1.0f / d->dh;
      x.ih2 = x.ih * x.ih;
    }
    // storage scales.  Sources into p or stresses: p / stresses in units of the shot's peak source
    // amplitude, velocities in that over the smallest impedance rho*vp of the window (cells with
    // vp, rho > 0).  Vertical force into vz: velocities in units of the peak amplitude, stresses
    // in that times the largest impedance.
    const bool scaled = p->storage != FOMO_STORE_F32;
    {
      float wmax = 0.0f, zmin = INFINITY, zmax = 0.0f;
      if (S.wavelet) wmax = max_abs(S.wavelet, (size_t)S.n_src * d->nt);
      for (int j = 0; j < d->nz; ++j)
        for (int i = 0; i < S.nx; ++i) {
          const size_t c = (size_t)(S.x0 - 1 + i) + (size_t)j * d->nx;
          if (!scalar && !coupled && vp[c] > 0.0f && rho[c] > 0.0f) {
            zmin = std::min(zmin, rho[c] * vp[c]);
            zmax = std::max(zmax, rho[c] * vp[c]);
          }
        }
      const float amp = scaled && wmax > 0.0f ? wmax : 1.0f;
      const bool has_z = std::isfinite(zmin) && zmin > 0.0f;
      const bool force = d->eq == FOMO_ELASTIC2D && d->source_type == FOMO_SRC_FZ;
      float sdec = amp, vdec = amp;
      if (scaled && has_z) {
        if (force)
          sdec = amp * zmax;
        else
          vdec = amp / zmin;
      }
      if (!scaled) sdec = vdec = 1.0f;
      for (int f = 0; f < 5; ++f) {
        const bool vel = d->eq == FOMO_ACOUSTIC2D ? (f == 1 || f == 2) : (f < 2);
        x.fdec[f] = vel ? vdec : sdec;
        x.fenc[f] = 1.0f / x.fdec[f];
        x.mdec[f] = 1.0f;
      }
    }
    const auto key = std::make_tuple(S.x0, S.nx, v_ref);
    auto mit = medium_of.find(key);
    if (mit != medium_of.end()) {
      const Shot2D& o = p->shots[mit->second];
      for (int k = 0; k < 8; ++k) {
        x.mat[k] = o.mat[k];
        x.mdec[k] = o.mdec[k];
      }
      x.wlut = o.wlut;
    } else {
      std::vector<float> vpp, rhop, bx, bz;
      pad2d(vp, S.nx, d->nz, pad, S.x0 - 1, d->nx, vpp);
      if (!scalar && !coupled) {
        pad2d(rho, S.nx, d->nz, pad, S.x0 - 1, d->nx, rhop);
        buoyancy2d(rhop, x.nx, x.nz, bx, bz);
      }
      void* dm[8] = {nullptr};
      // material m, stored with scale max |value| when 16-bit
      auto put = [&](int m, const std::vector<float>& h) -> int {
        const float mx = max_abs(h.data(), h.size());
        x.mdec[m] = p->mat_storage != FOMO_STORE_F32 && mx > 0.0f ? mx : 1.0f;
        return upload_stored(p, p->mat_storage, &dm[m], h.data(), h.size(), 1.0f / x.mdec[m]);
      };
      if (coupled) {
        // alpha, beta and the smoothed-medium derivatives, built on the device (coupled2d.hip)
        std::vector<float> vsp;
        pad2d(vs, S.nx, d->nz, pad, S.x0 - 1, d->nx, vsp);
        float* cm[8] = {nullptr};
        if ((rc = coupled_medium2d(p, vpp, vsp, x.nx, x.nz, d->fd_order, d->dh, d->smooth_sigma, cm))) return bail(rc);
        for (int k = 0; k < 8; ++k) {
          dm[k] = cm[k];
          x.mdec[k] = 1.0f;
        }
      } else if (scalar) {
        // c2 = (v * Float32(dt/dh))^2 on the edge-padded model, scalar2d.jl:126-128
        std::vector<float> c2(n);
        for (size_t c = 0; c < n; ++c) {
          const float t = vpp[c] * dtx;
          c2[c] = t * t;
        }
        if ((rc = put(0, c2))) return bail(rc);
      } else if (d->eq == FOMO_ACOUSTIC2D) {
        std::vector<float> kappa(n);
        for (size_t c = 0; c < n; ++c) kappa[c] = rhop[c] * (vpp[c] * vpp[c]);
        if ((rc = put(0, kappa)) || (rc = put(1, bx)) || (rc = put(2, bz))) return bail(rc);
      } else {
        std::vector<float> vsp, lam, l2m, muxz;
        pad2d(vs, S.nx, d->nz, pad, S.x0 - 1, d->nx, vsp);
        elastic_params2d(vpp, vsp, rhop, x.nx, x.nz, lam, l2m, muxz);
        if ((rc = put(0, lam)) || (rc = put(1, l2m)) || (rc = put(2, muxz)) || (rc = put(3, bx)) || (rc = put(4, bz)))
          return bail(rc);
      }
      for (int k = 0; k < 8; ++k) x.mat[k] = dm[k];
      std::vector<float> lut(4 * (size_t)(H + 1));
      weight_lut(H, pad, 0.0, lut.data());
      weight_lut(H, pad, 0.5, lut.data() + (H + 1));
      weight_lut(H, pad, 0.25, lut.data() + 2 * (H + 1));
      weight_lut(H, pad, 0.75, lut.data() + 3 * (H + 1));
      float* dl = nullptr;
      if ((rc = p->upload(&dl, lut.data(), lut.size()))) return bail(rc);
      x.wlut = dl;
      medium_of[key] = s;
    }
    x.nbuf = scalar ? 3 : 2;
    for (int b = 0; b < (scalar ? 1 : 2); ++b)
      for (int f = 0; f < p->nfld; ++f)
      {
        char* fp = nullptr;
        if ((rc = p->alloc(&fp, n * p->elem))) return bail(rc);
        x.fld[b][f] = fp;
      }
    // sources
    x.n_src = S.n_src;
    x.src_fz = d->source_type == FOMO_SRC_FZ ? 1 : 0;
    {
      std::vector<int> sc(std::max(S.n_src, 1), -1);
      for (int q = 0; q < S.n_src; ++q) sc[q] = (S.sx[q] - 1 + pad) + (S.sz[q] - 1 + pad) * x.nx;
      // the fused interior launch holds at most two sources per column of a strip in registers
      std::map<int, int> per_column;
      for (int q = 0; q < S.n_src; ++q)
        if (++per_column[S.sx[q]] > 2) p->fusable = false;
      int* dsc = nullptr;
      if ((rc = p->upload(&dsc, sc.data(), sc.size()))) return bail(rc);
      x.src_c = dsc;
      if (S.n_src > 0) {
        auto wit = wav_dev.find(S.wavelet);
        if (wit != wav_dev.end()) {
          x.wav = wit->second;
        } else {
          float* dw = nullptr;
          if ((rc = p->upload(&dw, S.wavelet, (size_t)S.n_src * d->nt))) return bail(rc);
          x.wav = dw;
          wav_dev[S.wavelet] = dw;
        }
      }
    }
    // receivers
    m.seis_off = seis_off;
    for (int r = 0; r < S.n_rec; ++r) {
      rs.push_back(s);
      rcell.push_back((S.rx[r] - 1 + pad) + (S.rz[r] - 1 + pad) * x.nx);
      rout.push_back(seis_off + r);
      rstride.push_back(S.n_rec);
    }
    seis_off += (long long)S.n_rec * d->nt;
    p->shots.push_back(x);
    p->meta.push_back(m);
  }
  p->seis_total = seis_off;
  p->n_entries = (int)rs.size();
  if ((rc = p->upload(&p->rec_shot, rs.data(), rs.size())) || (rc = p->upload(&p->rec_c, rcell.data(), rcell.size())) ||
      (rc = p->upload(&p->rec_out, rout.data(), rout.size())) ||
      (rc = p->upload(&p->rec_stride, rstride.data(), rstride.size())))
    return bail(rc);
  for (int k = 0; k < p->nslots; ++k)
    if ((rc = p->alloc(&p->seis[k], (size_t)std::max<long long>(seis_off, 1)))) return bail(rc);
  if ((rc = p->upload(&p->d_shots, p->shots.data(), p->shots.size()))) return bail(rc);
  if ((rc = p->alloc(&p->step, 1))) return bail(rc);
  if ((rc = p->reset())) return bail(rc);
  p->fused = p->fusable && !std::getenv("FOMO_2D_UNFUSED");
  if (d->tile_cfg > 0 && d->tile_cfg <= p->n_cfgs()) {
    p->tile_cfg = d->tile_cfg - 1;
    p->note_tuning(3);
    if ((rc = p->make_work(p->tile_cfg))) return bail(rc);
  } else if ((rc = p->autotune())) {
    return bail(rc);
  }
  if (d->use_graph && (rc = p->build_graph())) return bail(rc);
  e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return bail(set_error(FOMO_EHIP, std::string("setup: ") + hipGetErrorString(e)));
  *out = p;
  return FOMO_OK;
}

}  // namespace fomo
