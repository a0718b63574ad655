// engine3d.h -- 3D acoustic engine: shared declarations (host + device).
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "../../include/fomo_hip.h"
#include "fomo_common.h"

namespace fomo {

// Fields: 0 p, 1 vx, 2 vy, 3 vz (ping-pong).  Materials: 0 kappa, 1 bx, 2 by, 3 bz.  Storage is
// Float32, Float16 or BFloat16; arithmetic is always Float32.  A reduced-precision value v is
// stored as RN(v * enc) and read back as float(stored) * dec, one (dec, enc) pair per array.
struct Shot3D {
  void* fld[2][4];
  const void* mat[4];
  float fdec[4], fenc[4];
  float mdec[4];
  const float* wlut;  // HABC weight tables by distance: vx, vz, vy, tau (H+1 each)
  float dtdh, q, qt, qxt;
  int nx, ny, nz, H, pad;
  int n_src;
  const long long* src_c;
  const float* wav;
};

// kind 0: frame box (x up to 64, y up to 4, one plane), cell by cell; kind 1: interior column
// tile (x up to 64, y up to kTileY3) marched down z over hz planes.
struct Work3D {
  int kind, i0, j0, k0, w, hy, hz;
};

constexpr int kTileY3 = 16;

struct Sweep3DArgs {
  const Shot3D* shot;
  const Work3D* work;
  int n_work;
  int parity;
  float a[kMaxN];
  const int* step;
  int nt;
};

struct Gather3DArgs {
  const Shot3D* shot;
  const long long* rec_c;
  float* seis[4];  // slots: vx, vy, vz, p
  int slot_field[4];
  int nslots;
  int n_rec;
  int parity_out;
  int nt;
  int* step;
  unsigned* done;
};

// storage: FOMO_STORE_F32 / F16 / BF16.  phase 0: velocity, 1: pressure.
hipError_t launch_sweep3d(int storage, int phase, int N, const Sweep3DArgs& a, hipStream_t s);
hipError_t launch_gather3d(int storage, const Gather3DArgs& a, hipStream_t s);
hipError_t launch_decode3d(int storage, const void* src, float* dst, size_t n, float dec, hipStream_t s);
hipError_t launch_slice3d(int storage, const void* src, float* dst, int nx, int ny, int nz, int pad, int plane,
                          int idx, float dec, hipStream_t s);

}  // namespace fomo
