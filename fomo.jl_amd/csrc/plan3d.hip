// plan3d.hip -- 3D plans (acoustic3d, elastic3d): setup (padding, media, HABC tables, storage
// scales), work list, step enqueue, snapshots and the acoustic3d / elastic3d one-shot entry points.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <type_traits>
#include <vector>

#include "engine3d.h"
#include "plan.h"
#include "tune.h"

namespace fomo {

static void parallel_for(long long n, const std::function<void(long long, long long)>& fn) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < (1 << 20) || hw == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  const long long chunk = (n + hw - 1) / hw;
  for (unsigned t = 0; t < hw; ++t) {
    const long long a = t * chunk, b = std::min(n, a + chunk);
    if (a < b) th.emplace_back(fn, a, b);
  }
  for (auto& x : th) x.join();
}

struct Plan3D : PlanBase {
  int N = 0, storage = 0, nxi = 0, nyi = 0, nzi = 0, pad = 0, elem = 4;
  int nf = 4;       // fields: acoustic 4, elastic 9
  int snap_f = 0;   // field of the snapshots: acoustic p, elastic vz
  float coef[kMaxN] = {0};
  Shot3D shot{};
  int xoff = 0;        // elements from an array's allocation to its base (row_pitch3d)
  size_t alloc_n = 0;  // elements of each array's allocation: shot.px * ny * nz
  Work3D* d_work[2] = {nullptr, nullptr};  // core items, frame items
  int n_work[2] = {0, 0};
  // acoustic3d: the interior's two phases in one launch (engine3d_fused.hip) over these items,
  // between the frame velocity and frame pressure launches; the core lists above are then unused
  bool fused = false;
  Work3D* d_fused = nullptr;
  int n_fused = 0;
  long long* rec_c = nullptr;
  float* seis[4] = {nullptr};
  int nslots = 3;
  float* scratch = nullptr;  // decode staging
  size_t scratch_n = 0;
  // Snapshots (single-shot runs with snap_interval > 0): written by the per-step gather launch
  // into this device buffer, graph replays included, and downloaded once after the step loop.
  float* snap = nullptr;
  int snap_n = 0, snap_plane = 0, snap_idx = 0;  // snap_idx: padded index of the slice
  size_t snap_cells = 0;                          // cells of one slice

  hipError_t enqueue_part(int part, int s, int koff) override {
    if (part == 2) {
      Gather3DArgs g{};
      g.shot = shot;
      g.rec_c = rec_c;
      const bool ac = d.eq == FOMO_ACOUSTIC3D;
      const int fields[4] = {ac ? 1 : 0, ac ? 2 : 1, ac ? 3 : 2, 0};  // vx, vy, vz (, p)
      for (int k = 0; k < nslots; ++k) {
        g.seis[k] = seis[k];
        g.slot_field[k] = fields[k];
      }
      g.nslots = nslots;
      g.n_rec = d.n_rec;
      g.parity_out = 1 - s;
      g.nt = d.nt;
      g.step = step;
      g.koff = koff;
      if (snap) {
        g.snap = snap;
        g.snap_every = d.snap_interval;
        g.snap_n = snap_n;
        g.snap_plane = snap_plane;
        g.snap_idx = snap_idx;
        g.snap_f = snap_f;
      }
      return launch_gather3d(storage, g, stream);
    }
    Sweep3DArgs a{};
    a.shot = shot;
    a.parity = s;
    for (int l = 0; l < N; ++l) a.a[l] = coef[l];
    a.step = step;
    a.koff = koff;
    a.nt = d.nt;
    Sweep3DArgs f = a;
    if (fused) {
      // part 0: frame velocities, then the interior's whole step; part 1: frame pressure
      f.work = d_work[1];
      f.n_work = only == 0 ? 0 : n_work[1];  // only: 0 the interior launch, 1 the frame launch
      a.work = d_work[0];
      a.n_work = 0;
      hipError_t e = launch_sweep3d(d.eq, storage, part, N, a, f, stream);
      if (e != hipSuccess || part != 0 || only == 1) return e;
      a.work = d_fused;
      a.n_work = n_fused;
      return launch_fused3d(storage, N, a, stream);
    }
    a.work = d_work[0];
    a.n_work = only == 1 ? 0 : n_work[0];
    f.work = d_work[1];
    f.n_work = only == 0 ? 0 : n_work[1];
    return launch_sweep3d(d.eq, storage, part, N, a, f, stream);
  }

  int reset() override {
    for (int b = 0; b < 2; ++b)
      for (int f = 0; f < nf; ++f)
        FOMO_HIP_TRY(hipMemsetAsync((char*)shot.fld[b][f] - (size_t)xoff * elem, 0, alloc_n * elem, stream));
    const size_t ns = (size_t)d.n_rec * d.nt;
    for (int k = 0; k < nslots; ++k)
      if (ns) FOMO_HIP_TRY(hipMemsetAsync(seis[k], 0, ns * sizeof(float), stream));
    if (snap) FOMO_HIP_TRY(hipMemsetAsync(snap, 0, (size_t)snap_n * snap_cells * sizeof(float), stream));
    if (ovf) FOMO_HIP_TRY(hipMemsetAsync(ovf, 0, sizeof(unsigned), stream));
    FOMO_HIP_TRY(hipMemsetAsync(step, 0, sizeof(int), stream));
    steps_done = 0;
    return FOMO_OK;
  }

  int read_seis(int sh, float* const* dst) override {
    if (sh != 0) return set_error(FOMO_EINVAL, "shot index out of range");
    FOMO_HIP_TRY(hipStreamSynchronize(stream));
    const size_t cnt = (size_t)d.n_rec * d.nt;
    for (int k = 0; k < nslots; ++k)
      if (dst[k] && cnt) FOMO_HIP_TRY(hipMemcpy(dst[k], seis[k], cnt * sizeof(float), hipMemcpyDeviceToHost));
    return FOMO_OK;
  }

  int seis_device(int sh, int slot, const float** ptr, int* n_rec) const override {
    if (sh != 0) return set_error(FOMO_EINVAL, "shot index out of range");
    if (slot < 0 || slot >= nslots || !seis[slot]) return set_error(FOMO_EINVAL, "trace slot not recorded");
    *ptr = seis[slot];
    *n_rec = d.n_rec;
    return FOMO_OK;
  }

  int ensure_scratch(size_t n) {
    if (scratch_n >= n) return FOMO_OK;
    if (scratch) {
      (void)hipStreamSynchronize(stream);
      auto it = std::find(allocs.begin(), allocs.end(), (void*)scratch);
      if (it != allocs.end()) allocs.erase(it);
      (void)hipFree(scratch);
      scratch = nullptr;
      scratch_n = 0;
    }
    int rc = alloc(&scratch, n);
    if (!rc) scratch_n = n;
    return rc;
  }

  int read_field(int sh, int which, float* out) override {
    if (sh != 0) return set_error(FOMO_EINVAL, "shot index out of range");
    if (which < 0 || which >= nf) return set_error(FOMO_EINVAL, "field index out of range");
    const size_t n = (size_t)shot.nx * shot.ny * shot.nz;
    int rc = ensure_scratch(n);
    if (rc) return rc;
    FOMO_HIP_TRY(launch_decode3d(storage, shot.fld[steps_done % 2][which], scratch, shot.nx, (size_t)shot.ny * shot.nz,
                                 shot.px, shot.fdec[which], stream));
    FOMO_HIP_TRY(hipMemcpyAsync(out, scratch, n * sizeof(float), hipMemcpyDeviceToHost, stream));
    FOMO_HIP_TRY(hipStreamSynchronize(stream));
    return FOMO_OK;
  }

  int dims(int sh, int* nx, int* ny, int* nz, int* pd) const override {
    if (sh != 0) return set_error(FOMO_EINVAL, "shot index out of range");
    if (nx) *nx = shot.nx;
    if (ny) *ny = shot.ny;
    if (nz) *nz = shot.nz;
    if (pd) *pd = pad;
    return FOMO_OK;
  }
  int n_shots() const override { return 1; }
  int snap_buffer(int sh, const float** ptr, int* count, size_t* cells) const override {
    if (sh != 0) return set_error(FOMO_EINVAL, "shot index out of range");
    *ptr = snap;
    *count = snap ? snap_n : 0;
    *cells = snap_cells;
    return FOMO_OK;
  }
  // Launches of a step: velocity core, velocity frame, pressure/stress core, pressure/stress frame;
  // fused: velocity frame, interior step, pressure frame.
  int only = -1;  // enqueue_part launches only the core (0) or the frame (1) items when >= 0
  long long cells_of[2] = {0, 0};
  int n_launches() const override { return fused ? 3 : 4; }
  LaunchInfo launch_info(int k) const override {
    if (fused) {
      // the interior launch is credited with both phases' reads and writes (10 + 6 words per cell)
      if (k == 0) return {"acoustic3d velocity frame", cells_of[1], 10 * elem};
      if (k == 1) return {"acoustic3d interior step (velocity + pressure)", cells_of[0], 16 * elem};
      return {"acoustic3d pressure frame", cells_of[1], 6 * elem};
    }
    static const char* names[2][4] = {
        {"acoustic3d velocity core", "acoustic3d velocity frame", "acoustic3d pressure core",
         "acoustic3d pressure frame"},
        {"elastic3d velocity core", "elastic3d velocity frame", "elastic3d stress core", "elastic3d stress frame"}};
    const bool ac = d.eq == FOMO_ACOUSTIC3D;
    // words per cell read + written: acoustic velocity p,vx,vy,vz,bx,by,bz -> vx,vy,vz (10),
    // pressure vx,vy,vz,p,kappa -> p (6); elastic velocity 15, stress 20
    const int words = ac ? (k < 2 ? 10 : 6) : (k < 2 ? 15 : 20);
    return {names[ac ? 0 : 1][k], cells_of[k & 1], words * elem};
  }
  hipError_t enqueue_launch(int k, int s) override {
    hipError_t e;
    if (fused) {
      only = k == 1 ? 0 : 1;
      e = enqueue_part(k == 2 ? 1 : 0, s, 0);
    } else {
      only = k & 1;
      e = enqueue_part(k >> 1, s, 0);
    }
    only = -1;
    return e;
  }
};

// Work lists over the cells 1..n-2 of each axis (the outer ring is never written).  The x-y
// plane is cut into core tiles (64 x rows over the user's interior, no x/y frame cell) and ring
// tiles (62 x (rows-2) over the pad bands).  Core tiles over the user's z range form the core
// list; ring tiles over all of z and core tiles over the top and bottom pad slabs form the frame
// list.  Each column is marched over z-segments of about `kseg` planes.  Each list is ordered by
// (z, y, x) and split into 8 cost-balanced contiguous runs, one per XCD (block b runs on XCD b % 8).
static int upload_work(Plan3D* p, int which, std::vector<Work3D>& items, int rows) {
  std::stable_sort(items.begin(), items.end(), [](const Work3D& a, const Work3D& b) {
    if (a.k0 != b.k0) return a.k0 < b.k0;
    if (a.j0 != b.j0) return a.j0 < b.j0;
    return a.i0 < b.i0;
  });
  std::vector<double> cost(items.size());
  double total = 0.0;
  for (size_t q = 0; q < items.size(); ++q) {
    cost[q] = (double)kTileW * rows * (items[q].hz + 2);
    total += cost[q];
  }
  std::vector<std::vector<Work3D>> runs(8);
  double acc = 0.0;
  for (size_t q = 0; q < items.size(); ++q) {
    const int r = std::min(7, (int)(8.0 * (acc + 0.5 * cost[q]) / std::max(total, 1.0)));
    runs[r].push_back(items[q]);
    acc += cost[q];
  }
  size_t len = 0;
  for (auto& r : runs) len = std::max(len, r.size());
  std::vector<Work3D> flat(8 * len, Work3D{-1, 0, 0, 0, 0, 0, 0, 0});
  for (int r = 0; r < 8; ++r)
    for (size_t q = 0; q < runs[r].size(); ++q) flat[q * 8 + r] = runs[r][q];
  if (flat.size() > (size_t)0x7fffffff) return set_error(FOMO_EUNSUP, "grid too large for one launch");
  if (which == 2) {
    p->n_fused = (int)flat.size();
    return p->upload(&p->d_fused, flat.data(), flat.size());
  }
  p->n_work[which] = (int)flat.size();
  return p->upload(&p->d_work[which], flat.data(), flat.size());
}

static int make_work3d(Plan3D* p, int kseg, int rows, int core_rows, int kseg_fused) {
  const int nx = p->shot.nx, ny = p->shot.ny, nz = p->shot.nz, pad = p->pad;
  std::vector<Work3D> core_fp, ring_fp;
  // rx / ry: the rectangle holds x / y frame cells (then the tile computes a ring in that direction)
  auto ring_rect = [&](int x0, int x1, int y0, int y1, int rx, int ry) {
    const int tw = kTileW - 2 * rx, th = rows - 2 * ry;
    for (int j = y0; j < y1; j += th)
      for (int i = x0; i < x1; i += tw)
        ring_fp.push_back(Work3D{1, i, j, 0, std::min(tw, x1 - i), std::min(th, y1 - j), 0, rx | (ry << 1)});
  };
  // core footprints for the core list (core_rows) and for the frame list's top/bottom slabs (rows)
  std::vector<Work3D> slab_fp;
  for (int j = pad; j < ny - pad; j += core_rows)
    for (int i = pad; i < nx - pad; i += kTileW)
      core_fp.push_back(Work3D{1, i, j, 0, std::min(kTileW, nx - pad - i), std::min(core_rows, ny - pad - j), 0, 0});
  for (int j = pad; j < ny - pad; j += rows)
    for (int i = pad; i < nx - pad; i += kTileW)
      slab_fp.push_back(Work3D{1, i, j, 0, std::min(kTileW, nx - pad - i), std::min(rows, ny - pad - j), 0, 0});
  ring_rect(1, nx - 1, 1, pad, 1, 1);
  ring_rect(1, pad, pad, ny - pad, 1, 0);
  ring_rect(nx - pad, nx - 1, pad, ny - pad, 1, 0);
  ring_rect(1, nx - 1, ny - pad, ny - 1, 1, 1);
  // z-segments of [z0, z1) of about kseg planes each
  auto segs = [&](int z0, int z1, const std::vector<Work3D>& fp, std::vector<Work3D>& out) {
    const int nk = z1 - z0;
    if (nk <= 0) return;
    const int nseg = std::max(1, (nk + kseg - 1) / kseg);
    const int hz = (nk + nseg - 1) / nseg;
    for (int k = z0; k < z1; k += hz)
      for (const Work3D& f : fp) {
        Work3D w = f;
        w.k0 = k;
        w.hz = std::min(hz, z1 - k);
        out.push_back(w);
      }
  };
  std::vector<Work3D> core, frame;
  segs(pad, nz - pad, core_fp, core);
  {
    // the top and bottom slabs are pad - 1 planes deep: one segment each (a second one would pay
    // the z-queue prologue again for a few planes)
    const int keep = kseg;
    kseg = std::max(kseg, pad);
    segs(1, pad, slab_fp, frame);
    segs(nz - pad, nz - 1, slab_fp, frame);
    kseg = keep;
  }
  segs(1, nz - 1, ring_fp, frame);
  int rc = upload_work(p, 0, core, core_rows);
  if (!rc) rc = upload_work(p, 1, frame, rows);
  if (!rc && p->fused) {
    // interior tiles of the fused launch: 64 x kFTY footprints over z-segments of about kseg_fused planes
    std::vector<Work3D> fp, items;
    for (int j = pad; j < ny - pad; j += kFTY)
      for (int i = pad; i < nx - pad; i += kTileW)
        fp.push_back(Work3D{1, i, j, 0, std::min(kTileW, nx - pad - i), std::min(kFTY, ny - pad - j), 0, 0});
    const int keep = kseg;
    kseg = kseg_fused;
    segs(pad, nz - pad, fp, items);
    kseg = keep;
    rc = upload_work(p, 2, items, kFTY);
  }
  return rc;
}

// Upload a padded array (rows of nx values, x fastest) as the storage type (16-bit: RN(v * enc))
// in the arrays' row layout; *dst receives the array base.  The pitch padding holds zeros.
template <class T>
static int upload_stored(Plan3D* p, void** dst, const std::vector<float>& h, float enc) {
  const Shot3D& S = p->shot;
  const long long rows = (long long)S.ny * S.nz;
  std::vector<T> tmp(p->alloc_n);
  parallel_for(rows, [&](long long a, long long b) {
    for (long long r = a; r < b; ++r) {
      const float* in = h.data() + (size_t)r * S.nx;
      T* out = tmp.data() + p->xoff + (size_t)r * S.px;
      for (int i = 0; i < S.nx; ++i) {
        if constexpr (std::is_same<T, float>::value)
          out[i] = in[i];
        else if constexpr (std::is_same<T, __half>::value)
          out[i] = __float2half_rn(in[i] * enc);
        else
          out[i] = __float2bfloat16(in[i] * enc);
      }
    }
  });
  T* d = nullptr;
  int rc = p->upload(&d, tmp.data(), tmp.size());
  *dst = d ? d + p->xoff : nullptr;
  return rc;
}

static int upload_any(Plan3D* p, void** dst, const std::vector<float>& h, float enc) {
  if (p->storage == FOMO_STORE_F16) return upload_stored<__half>(p, dst, h, enc);
  if (p->storage == FOMO_STORE_BF16) return upload_stored<__hip_bfloat16>(p, dst, h, enc);
  return upload_stored<float>(p, dst, h, enc);
}

}  // namespace fomo

using namespace fomo;

static int layout3d(Plan3D* p, const fomo_desc* d, int rows, int core_rows);

int fomo_plan_create3d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, const int32_t* sx,
                       const int32_t* sy, const int32_t* sz, const float* wavelet, const int32_t* rx,
                       const int32_t* ry, const int32_t* rz, fomo_plan** out) {
  *out = nullptr;
  const bool elastic = d->eq == FOMO_ELASTIC3D;
  if (d->eq != FOMO_ACOUSTIC3D && !elastic) return set_error(FOMO_EINVAL, "not a 3D equation");
  if (d->storage < FOMO_STORE_F32 || d->storage > FOMO_STORE_BF16) return set_error(FOMO_EINVAL, "unknown storage");
  if (!vp || !rho || (elastic && !vs)) return set_error(FOMO_EINVAL, "model arrays must not be NULL");
  if (elastic && d->record_p) return set_error(FOMO_EINVAL, "elastic3d has no pressure field to record");
  if (d->n_src > 0 && (!sx || !sy || !sz || !wavelet)) return set_error(FOMO_EINVAL, "source arrays must not be NULL");
  if (d->n_rec > 0 && (!rx || !ry || !rz)) return set_error(FOMO_EINVAL, "receiver arrays must not be NULL");
  for (int s = 0; s < d->n_src; ++s)
    if (sx[s] < 1 || sx[s] > d->nx || sy[s] < 1 || sy[s] > d->ny || sz[s] < 1 || sz[s] > d->nz)
      return set_error(FOMO_EINVAL, "source index out of range");
  for (int r = 0; r < d->n_rec; ++r)
    if (rx[r] < 1 || rx[r] > d->nx || ry[r] < 1 || ry[r] > d->ny || rz[r] < 1 || rz[r] > d->nz)
      return set_error(FOMO_EINVAL, "receiver index out of range");
  // snapshot slice: plane 0 xz at y, 1 xy at z, 2 yz at x (1-based interior index; 0 = middle)
  const int snap_idx1 = d->snap_index > 0 ? d->snap_index
                                          : (d->snap_plane == 0 ? d->ny / 2 : (d->snap_plane == 1 ? d->nz / 2 : d->nx / 2));
  if (d->snap_interval > 0) {
    if (d->snap_plane < 0 || d->snap_plane > 2) return set_error(FOMO_EINVAL, "snap_plane must be 0 (xz), 1 (xy) or 2 (yz)");
    const int axis = d->snap_plane == 0 ? d->ny : (d->snap_plane == 1 ? d->nz : d->nx);
    if (snap_idx1 < 1 || snap_idx1 > axis) return set_error(FOMO_EINVAL, "snap_index out of range");
  }
  Plan3D* p = new Plan3D();
  auto bail = [&](int code) {
    delete p;
    return code;
  };
  p->d = *d;
  p->N = d->fd_order / 2;
  p->storage = d->storage;
  p->elem = d->storage == FOMO_STORE_F32 ? 4 : 2;
  p->nf = elastic ? 9 : 4;
  p->snap_f = elastic ? 2 : 0;
  fomo_fd_coefficients(d->fd_order, p->coef);
  p->nxi = d->nx;
  p->nyi = d->ny;
  p->nzi = d->nz;
  p->pad = d->nbc + p->N;
  const int pad = p->pad;
  Shot3D& S = p->shot;
  S.nx = d->nx + 2 * pad;
  S.ny = d->ny + 2 * pad;
  S.nz = d->nz + 2 * pad;
  S.H = pad - 1;
  S.pad = pad;
  {
    const RowPitch rp = row_pitch3d(S.nx, pad, p->elem);
    S.px = rp.px;
    p->xoff = rp.xoff;
    p->alloc_n = (size_t)S.px * S.ny * S.nz;
  }
  S.dtdh = d->dt / d->dh;
  hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(set_error(FOMO_EHIP, std::string("stream: ") + hipGetErrorString(e)));
  const size_t sy_ = (size_t)S.nx, sz_ = (size_t)S.nx * S.ny, n = sz_ * S.nz;
  const size_t ixy = (size_t)d->nx * d->ny;
  // v_ref: min(vp > 0) over the interior unless given
  float v_ref = d->v_ref;
  if (!(v_ref > 0.0f)) v_ref = min_positive(vp, ixy * d->nz);
  float hc[3];
  fomo_habc_coefficients(v_ref, d->dt, d->dh, hc);
  S.q = hc[0];
  S.qt = hc[1];
  S.qxt = hc[2];
  const bool scaled = d->storage != FOMO_STORE_F32;
  int rc;
  // Upload one material array (index m) with its storage scale (max |value| when scaled).
  auto put_mat = [&](int m, const std::vector<float>& h) -> int {
    float mx = 0.0f;
    for (size_t c = 0; c < h.size(); ++c) mx = std::max(mx, std::fabs(h[c]));
    const float dec = scaled && mx > 0.0f ? mx : 1.0f;
    void* dm = nullptr;
    int r2 = upload_any(p, &dm, h, 1.0f / dec);
    if (r2) return r2;
    S.mat[m] = dm;
    S.mdec[m] = dec;
    return FOMO_OK;
  };
  // padding: pad_array_3d.jl:8-29 (clamp indexing)
  auto pad3 = [&](const float* in, std::vector<float>& o) {
    o.assign(n, 0.0f);
    parallel_for(S.nz, [&](long long k0, long long k1) {
      for (long long k = k0; k < k1; ++k) {
        const int kk = std::min(std::max((int)k - pad, 0), d->nz - 1);
        for (int j = 0; j < S.ny; ++j) {
          const int jj = std::min(std::max(j - pad, 0), d->ny - 1);
          for (int i = 0; i < S.nx; ++i) {
            const int ii = std::min(std::max(i - pad, 0), d->nx - 1);
            o[(size_t)i + j * sy_ + k * sz_] = in[(size_t)ii + (size_t)jj * d->nx + (size_t)kk * ixy];
          }
        }
      }
    });
  };
  std::vector<float> rhop;
  pad3(rho, rhop);
  const int mb = elastic ? 0 : 1;  // index of bx among the materials
  // acoustic3d with Float32 storage runs the interior's step in one launch (512^3 FD8: 3.12 against
  // 3.59 ms per step).  With 16-bit storage the fused loop's decodes, encodes and round trips make
  // it issue-bound (interior 2.29 ms against 1.08 ms for the two 16-bit core sweeps), so the
  // separate sweeps run; FOMO_3D_FUSED=1 forces the fused form, FOMO_3D_UNFUSED=1 the sweeps.
  p->fused = !elastic && !std::getenv("FOMO_3D_UNFUSED") &&
             (d->storage == FOMO_STORE_F32 || std::getenv("FOMO_3D_FUSED") != nullptr);
  if (!elastic) {
    // kappa = rho * vp^2 (acoustic3d/medium.jl:40-62)
    std::vector<float> vpp, kappa(n);
    pad3(vp, vpp);
    for (size_t c = 0; c < n; ++c) kappa[c] = rhop[c] * (vpp[c] * vpp[c]);
    if ((rc = put_mat(0, kappa))) return bail(rc);
  }
  // staggered buoyancy, pad_array_3d.jl:37-89
  auto hb = [](float r1, float r2) {
    if (r1 == 0.0f && r2 == 0.0f) return 0.0f;
    if (r1 == 0.0f) return 2.0f / r2;
    if (r2 == 0.0f) return 2.0f / r1;
    return 2.0f / (r1 + r2);
  };
  for (int ax = 0; ax < 3; ++ax) {
    std::vector<float> b(n, 0.0f);
    parallel_for(S.nz, [&](long long k0, long long k1) {
      for (long long k = k0; k < k1; ++k)
        for (int j = 0; j < S.ny; ++j)
          for (int i = 0; i < S.nx; ++i) {
            const size_t c = (size_t)i + j * sy_ + k * sz_;
            if (ax == 0) b[c] = i >= 1 ? hb(rhop[c - 1], rhop[c]) : 0.0f;
            else if (ax == 1) b[c] = j >= 1 ? hb(rhop[c - sy_], rhop[c]) : 0.0f;
            else b[c] = k + 1 < S.nz ? hb(rhop[c], rhop[c + sz_]) : 0.0f;
          }
    });
    // edge copies: buoy_vx[1,:,:] = buoy_vx[2,:,:]; buoy_vy[:,1,:] = buoy_vy[:,2,:]; buoy_vz[:,:,nz] = buoy_vz[:,:,nz-1]
    if (ax == 0)
      for (int k = 0; k < S.nz; ++k)
        for (int j = 0; j < S.ny; ++j) b[j * sy_ + k * sz_] = b[1 + j * sy_ + k * sz_];
    if (ax == 1)
      for (int k = 0; k < S.nz; ++k)
        for (int i = 0; i < S.nx; ++i) b[i + k * sz_] = b[i + sy_ + k * sz_];
    if (ax == 2)
      for (int j = 0; j < S.ny; ++j)
        for (int i = 0; i < S.nx; ++i) b[i + j * sy_ + (S.nz - 1) * sz_] = b[i + j * sy_ + (S.nz - 2) * sz_];
    if ((rc = put_mat(mb + ax, b))) return bail(rc);
  }
  if (elastic) {
    // elastic3d/medium.jl:76-123: mu = rho vs^2, lam = rho vp^2 - 2 mu, lam2mu = lam + 2 mu, and the
    // four-point harmonic means of mu at txy, txz, tyz (0 if any of the four is 0), edge-copied.
    std::vector<float> vpp, vsp;
    pad3(vp, vpp);
    pad3(vs, vsp);
    std::vector<float> mu(n), lam(n), l2m(n);
    for (size_t c = 0; c < n; ++c) {
      mu[c] = rhop[c] * (vsp[c] * vsp[c]);
      lam[c] = rhop[c] * (vpp[c] * vpp[c]) - 2.0f * mu[c];
      l2m[c] = lam[c] + 2.0f * mu[c];
    }
    std::vector<float>().swap(vpp);
    std::vector<float>().swap(vsp);
    if ((rc = put_mat(3, lam)) || (rc = put_mat(4, l2m))) return bail(rc);
    std::vector<float>().swap(lam);
    std::vector<float>().swap(l2m);
    auto h4 = [](float m1, float m2, float m3, float m4) {
      if (m1 == 0.0f || m2 == 0.0f || m3 == 0.0f || m4 == 0.0f) return 0.0f;
      return 4.0f / (((1.0f / m1 + 1.0f / m2) + 1.0f / m3) + 1.0f / m4);
    };
    const int NX = S.nx, NY = S.ny, NZ = S.nz;
    for (int which = 0; which < 3; ++which) {
      std::vector<float> m(n, 0.0f);
      parallel_for(NZ, [&](long long k0, long long k1) {
        for (long long k = k0; k < k1; ++k)
          for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
              const size_t c = (size_t)i + j * sy_ + k * sz_;
              if (which == 0 && i >= 1 && j >= 1)
                m[c] = h4(mu[c - 1 - sy_], mu[c - sy_], mu[c - 1], mu[c]);
              else if (which == 1 && i >= 1 && k + 1 < NZ)
                m[c] = h4(mu[c - 1], mu[c], mu[c - 1 + sz_], mu[c + sz_]);
              else if (which == 2 && j >= 1 && k + 1 < NZ)
                m[c] = h4(mu[c - sy_], mu[c], mu[c - sy_ + sz_], mu[c + sz_]);
            }
      });
      if (which == 0) {  // mu_txy[1,:,:] = mu_txy[2,:,:]; mu_txy[:,1,:] = mu_txy[:,2,:]
        for (int k = 0; k < NZ; ++k)
          for (int j = 0; j < NY; ++j) m[j * sy_ + k * sz_] = m[1 + j * sy_ + k * sz_];
        for (int k = 0; k < NZ; ++k)
          for (int i = 0; i < NX; ++i) m[i + k * sz_] = m[i + sy_ + k * sz_];
      } else if (which == 1) {  // mu_txz[1,:,:] = mu_txz[2,:,:]; mu_txz[:,:,nz] = mu_txz[:,:,nz-1]
        for (int k = 0; k < NZ; ++k)
          for (int j = 0; j < NY; ++j) m[j * sy_ + k * sz_] = m[1 + j * sy_ + k * sz_];
        for (int j = 0; j < NY; ++j)
          for (int i = 0; i < NX; ++i) m[i + j * sy_ + (NZ - 1) * sz_] = m[i + j * sy_ + (NZ - 2) * sz_];
      } else {  // mu_tyz[:,1,:] = mu_tyz[:,2,:]; mu_tyz[:,:,nz] = mu_tyz[:,:,nz-1]
        for (int k = 0; k < NZ; ++k)
          for (int i = 0; i < NX; ++i) m[i + k * sz_] = m[i + sy_ + k * sz_];
        for (int j = 0; j < NY; ++j)
          for (int i = 0; i < NX; ++i) m[i + j * sy_ + (NZ - 1) * sz_] = m[i + j * sy_ + (NZ - 2) * sz_];
      }
      if ((rc = put_mat(5 + which, m))) return bail(rc);
    }
  }
  std::vector<float>().swap(rhop);
  // field scales: p / stresses in units of the peak source amplitude, velocities in that over the
  // smallest impedance rho*vp (over cells with vp, rho > 0)
  float zmin = INFINITY, wmax = 0.0f;
  for (size_t c = 0; c < ixy * d->nz; ++c)
    if (vp[c] > 0.0f && rho[c] > 0.0f) zmin = std::min(zmin, rho[c] * vp[c]);
  for (size_t c = 0; c < (size_t)d->n_src * d->nt; ++c) wmax = std::max(wmax, std::fabs(wavelet[c]));
  const float sdec = scaled ? (wmax > 0.0f ? wmax : 1.0f) : 1.0f;
  const float vdec = scaled ? (std::isfinite(zmin) && zmin > 0.0f ? sdec / zmin : sdec) : 1.0f;
  for (int f = 0; f < p->nf; ++f) {
    const bool vel = elastic ? f < 3 : f >= 1;
    S.fdec[f] = vel ? vdec : sdec;
    S.fenc[f] = 1.0f / S.fdec[f];
  }
  for (int b = 0; b < 2; ++b)
    for (int f = 0; f < p->nf; ++f) {
      char* ptr = nullptr;
      if ((rc = p->alloc(&ptr, p->alloc_n * p->elem))) return bail(rc);
      S.fld[b][f] = ptr + (size_t)p->xoff * p->elem;
    }
  {
    std::vector<float> lut(4 * (size_t)(S.H + 1));
    weight_lut(S.H, pad, 0.0, lut.data());
    weight_lut(S.H, pad, 0.5, lut.data() + (S.H + 1));
    weight_lut(S.H, pad, 0.25, lut.data() + 2 * (S.H + 1));
    weight_lut(S.H, pad, 0.75, lut.data() + 3 * (S.H + 1));
    float* dl = nullptr;
    if ((rc = p->upload(&dl, lut.data(), lut.size()))) return bail(rc);
    S.wlut = dl;
  }
  {
    S.n_src = d->n_src;
    std::vector<long long> sc(std::max(d->n_src, 1), -1);
    // cells as element offsets from the array base (row pitch S.px)
    const long long dy = S.px, dz = (long long)S.px * S.ny;
    for (int q = 0; q < d->n_src; ++q)
      sc[q] = (long long)(sx[q] - 1 + pad) + (long long)(sy[q] - 1 + pad) * dy + (long long)(sz[q] - 1 + pad) * dz;
    long long* dsc = nullptr;
    if ((rc = p->upload(&dsc, sc.data(), sc.size()))) return bail(rc);
    S.src_c = dsc;
    float* dw = nullptr;
    if ((rc = p->upload(&dw, wavelet ? wavelet : (const float*)nullptr, wavelet ? (size_t)d->n_src * d->nt : 0)))
      return bail(rc);
    S.wav = dw;
    std::vector<long long> rcell(std::max(d->n_rec, 1), 0);
    for (int r = 0; r < d->n_rec; ++r)
      rcell[r] = (long long)(rx[r] - 1 + pad) + (long long)(ry[r] - 1 + pad) * dy + (long long)(rz[r] - 1 + pad) * dz;
    if ((rc = p->upload(&p->rec_c, rcell.data(), rcell.size()))) return bail(rc);
    p->nslots = d->record_p ? 4 : 3;
    for (int k = 0; k < p->nslots; ++k)
      if ((rc = p->alloc(&p->seis[k], (size_t)d->n_rec * d->nt))) return bail(rc);
  }
  if ((rc = p->alloc(&p->step, 1)) || (rc = p->alloc(&p->ovf, 1))) return bail(rc);
  p->shot.ovf = p->ovf;
  if (d->snap_interval > 0 && d->nt / d->snap_interval > 0) {
    p->snap_plane = d->snap_plane;
    p->snap_idx = snap_idx1 - 1 + pad;
    p->snap_cells = d->snap_plane == 0 ? (size_t)d->nx * d->nz
                    : (d->snap_plane == 1 ? (size_t)d->nx * d->ny : (size_t)d->ny * d->nz);
    p->snap_n = d->nt / d->snap_interval;
    if ((rc = p->alloc(&p->snap, p->snap_cells * p->snap_n))) return bail(rc);
  }
  const int rows = elastic ? kRowsEl3 : kRowsAc3, core_rows = elastic ? kRowsEl3Core : kRowsAc3Core;
  p->cells_of[0] = (long long)p->nxi * p->nyi * p->nzi;
  p->cells_of[1] = (long long)p->shot.nx * p->shot.ny * p->shot.nz - p->cells_of[0];
  if ((rc = p->reset())) return bail(rc);
  if ((rc = layout3d(p, d, rows, core_rows))) return bail(rc);
  if (d->use_graph && (rc = p->build_graph())) return bail(rc);
  e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return bail(set_error(FOMO_EHIP, std::string("setup: ") + hipGetErrorString(e)));
  *out = new fomo_plan{p};
  return FOMO_OK;
}

// Fused interior z-segments: one block per CU fits (1024 threads, 60 KB of LDS), so the columns are
// cut into the fewest segments that still give every CU a block; each segment pays 2N - 1 extra
// levels of vz, and tiles that march together share their halos through L2.
static int fused_kseg_default(const Plan3D* p) {
  const long long cols = (long long)((p->nxi + kTileW - 1) / kTileW) * ((p->nyi + kFTY - 1) / kFTY);
  const long long nseg = std::max(1LL, (256 + cols - 1) / std::max(1LL, cols));
  return (int)std::max(8LL, (p->nzi + nseg - 1) / nseg);
}

// The work layout of a 3D plan.  d->tile_cfg = k > 0 fixes the z-segments (sweeps and frame
// 16 << (k-1) planes, fused interior 32 << (k-1)); FOMO_FUSED_KSEG fixes the fused interior's.
// Otherwise the candidates are timed on the plan's own grid, as the 2D autotuner does (two
// interleaved rounds of direct-launch steps, each candidate's faster round kept): sweep and frame
// segments of 16, 32 and 64 planes, and, where the fused interior launch is allowed (acoustic3d,
// Float32 storage), that family with its default segment (fused_kseg_default), half and a quarter
// of it, against the separate sweeps (FOMO_3D_FUSED=1 keeps the fused family only).  The choice is kept per (device and library build,
// equation, order, storage, padded dims, families) in this process and in the tuning file
// (tune.h).  Segment lengths change no bits.
static int layout3d(Plan3D* p, const fomo_desc* d, int rows, int core_rows) {
  struct Cand {
    bool fused;
    int kcfg, kseg, kseg_fused;
    Work3D* work[2];
    int n_work[2];
    Work3D* fw;
    int n_fw;
  };
  const bool may_fuse = p->fused;
  auto build = [&](Cand& c) -> int {
    p->fused = c.fused;
    const int rc = make_work3d(p, c.kseg, rows, core_rows, c.kseg_fused);
    if (rc) return rc;
    c.work[0] = p->d_work[0];
    c.work[1] = p->d_work[1];
    c.n_work[0] = p->n_work[0];
    c.n_work[1] = p->n_work[1];
    c.fw = c.fused ? p->d_fused : nullptr;
    c.n_fw = c.fused ? p->n_fused : 0;
    return FOMO_OK;
  };
  auto select = [&](const Cand& c) {
    p->fused = c.fused;
    p->d_work[0] = c.work[0];
    p->d_work[1] = c.work[1];
    p->n_work[0] = c.n_work[0];
    p->n_work[1] = c.n_work[1];
    p->d_fused = c.fw;
    p->n_fused = c.n_fw;
    p->tune_family = c.fused ? 1 : 0;
    p->tune_shape = c.kcfg;
    p->tune_extent = c.fused ? c.kseg_fused : c.kseg;
  };
  const int kf_default = fused_kseg_default(p);
  const char* kf_env = std::getenv("FOMO_FUSED_KSEG");
  if (d->tile_cfg > 0 || kf_env) {
    Cand c{};
    c.fused = may_fuse;
    c.kcfg = d->tile_cfg > 0 ? d->tile_cfg - 1 : 1;
    c.kseg = 16 << c.kcfg;
    c.kseg_fused = kf_env ? std::max(8, std::atoi(kf_env)) : (d->tile_cfg > 0 ? 32 << (d->tile_cfg - 1) : kf_default);
    const int rc = build(c);
    if (rc) return rc;
    select(c);
    p->tune_source = 3;
    return FOMO_OK;
  }
  std::vector<Cand> cand;
  const bool only_fused = may_fuse && std::getenv("FOMO_3D_FUSED") != nullptr;
  if (!only_fused)
    for (int k = 0; k < 3; ++k) cand.push_back(Cand{false, k, 16 << k, 0, {nullptr, nullptr}, {0, 0}, nullptr, 0});
  if (may_fuse) {
    std::vector<int> kfs;
    for (int div : {1, 2, 4}) {
      const int v = std::max(8, kf_default / div);
      if (std::find(kfs.begin(), kfs.end(), v) == kfs.end()) kfs.push_back(v);
    }
    for (int k = 0; k < 3; ++k)
      for (int kf : kfs) cand.push_back(Cand{true, k, 16 << k, kf, {nullptr, nullptr}, {0, 0}, nullptr, 0});
  }
  const std::string key = tune_device_tag() + " 3d eq=" + std::to_string(d->eq) + " N=" + std::to_string(p->N) +
                          " storage=" + std::to_string(p->storage) + " dims=" + std::to_string(p->shot.nx) + "x" +
                          std::to_string(p->shot.ny) + "x" + std::to_string(p->shot.nz) +
                          " pad=" + std::to_string(p->pad) + " fuse=" + std::to_string(may_fuse ? (only_fused ? 2 : 1) : 0);
  std::vector<int> rec;
  if (tune_lookup(key, rec) && rec.size() == 3) {
    for (Cand& c : cand)
      if ((c.fused ? 1 : 0) == rec[0] && c.kcfg == rec[1] && (!c.fused || c.kseg_fused == rec[2])) {
        const int rc = build(c);
        if (rc) return rc;
        select(c);
        p->tune_source = tune_last_source();
        return FOMO_OK;
      }
  }
  for (Cand& c : cand) {
    const int rc = build(c);
    if (rc) return rc;
  }
  hipEvent_t e0, e1;
  FOMO_HIP_TRY(hipEventCreate(&e0));
  FOMO_HIP_TRY(hipEventCreate(&e1));
  const long long cells = (long long)p->shot.nx * p->shot.ny * p->shot.nz;
  const int reps = (int)std::max<long long>(2, std::min<long long>(16, (1LL << 24) / std::max(cells, 1LL)));
  std::vector<float> best_ms(cand.size(), 1e30f);
  hipError_t e = hipSuccess;
  for (int round = 0; round < 2 && e == hipSuccess; ++round)
    for (size_t k = 0; k < cand.size() && e == hipSuccess; ++k) {
      select(cand[k]);
      e = p->run_steps(1);
      if (e == hipSuccess) e = hipEventRecord(e0, p->stream);
      if (e == hipSuccess) e = p->run_steps(reps);
      if (e == hipSuccess) e = hipEventRecord(e1, p->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      best_ms[k] = std::min(best_ms[k], ms);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return set_error(FOMO_EHIP, std::string("3D autotune: ") + hipGetErrorString(e));
  size_t best = 0;
  for (size_t k = 1; k < cand.size(); ++k)
    if (best_ms[k] < best_ms[best]) best = k;
  select(cand[best]);
  p->tune_source = 0;
  tune_store(key, {cand[best].fused ? 1 : 0, cand[best].kcfg, cand[best].fused ? cand[best].kseg_fused : 0});
  return p->reset();
}

// Shared one-shot driver of acoustic3d and elastic3d: warm-up step, reset, timed loop with
// snapshots, seismogram download.
static int run3d(const fomo_desc* d, const float* vp, const float* vs, const float* rho, const int32_t* sx,
                 const int32_t* sy, const int32_t* sz, const float* wavelet, const int32_t* rx, const int32_t* ry,
                 const int32_t* rz, float* const* seis, float* snaps, double* kernel_time_s) {
  fomo_plan* h = nullptr;
  int rc = fomo_plan_create(d, vp, vs, rho, sx, sy, sz, wavelet, rx, ry, rz, &h);
  if (rc) return rc;
  Plan3D* p = static_cast<Plan3D*>(h->impl);
  auto finish = [&](int code) {
    fomo_plan_destroy(h);
    return code;
  };
  if (d->nt > 0) {
    if ((rc = p->advance(1)) || (rc = fomo_plan_sync(h)) || (rc = p->reset())) return finish(rc);
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return finish(set_error(FOMO_EHIP, "hipEventCreate failed"));
  (void)hipEventRecord(e0, p->stream);
  if ((rc = p->advance(d->nt))) return finish(rc);
  (void)hipEventRecord(e1, p->stream);
  if (hipEventSynchronize(e1) != hipSuccess) return finish(set_error(FOMO_EHIP, "step loop failed"));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (kernel_time_s) *kernel_time_s = ms * 1e-3;
  if (snaps && p->snap) {
    const hipError_t ce = hipMemcpy(snaps, p->snap, p->snap_cells * p->snap_n * sizeof(float), hipMemcpyDeviceToHost);
    if (ce != hipSuccess) return finish(set_error(FOMO_EHIP, std::string("snapshot download: ") + hipGetErrorString(ce)));
  }
  rc = p->read_seis(0, seis);
  if (!rc) rc = p->check_overflow();
  return finish(rc);
}

extern "C" int fomo_acoustic3d(const fomo_desc* d, const float* vp, const float* rho, const int32_t* sx,
                               const int32_t* sy, const int32_t* sz, const float* wavelet, const int32_t* rx,
                               const int32_t* ry, const int32_t* rz, float* seis_vx, float* seis_vy, float* seis_vz,
                               float* seis_p, float* snaps, double* kernel_time_s) {
  if (!d || d->eq != FOMO_ACOUSTIC3D) return set_error(FOMO_EINVAL, "descriptor eq must be FOMO_ACOUSTIC3D");
  fomo_desc dd = *d;
  dd.record_p = seis_p ? 1 : 0;
  float* dst[4] = {seis_vx, seis_vy, seis_vz, seis_p};
  return run3d(&dd, vp, nullptr, rho, sx, sy, sz, wavelet, rx, ry, rz, dst, snaps, kernel_time_s);
}

extern "C" int fomo_elastic3d(const fomo_desc* d, const float* vp, const float* vs, const float* rho,
                              const int32_t* sx, const int32_t* sy, const int32_t* sz, const float* wavelet,
                              const int32_t* rx, const int32_t* ry, const int32_t* rz, float* seis_vx,
                              float* seis_vy, float* seis_vz, float* snaps, double* kernel_time_s) {
  if (!d || d->eq != FOMO_ELASTIC3D) return set_error(FOMO_EINVAL, "descriptor eq must be FOMO_ELASTIC3D");
  fomo_desc dd = *d;
  dd.record_p = 0;
  float* dst[4] = {seis_vx, seis_vy, seis_vz, nullptr};
  return run3d(&dd, vp, vs, rho, sx, sy, sz, wavelet, rx, ry, rz, dst, snaps, kernel_time_s);
}
