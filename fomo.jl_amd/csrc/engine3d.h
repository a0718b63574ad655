// engine3d.h -- 3D engine (acoustic3d, elastic3d): shared declarations (host + device).
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "../../include/fomo_hip.h"
#include "fomo_common.h"

namespace fomo {

// Fields (ping-pong):  acoustic3d 0 p, 1 vx, 2 vy, 3 vz;
//                      elastic3d  0 vx, 1 vy, 2 vz, 3 txx, 4 tyy, 5 tzz, 6 txy, 7 txz, 8 tyz.
// Materials:           acoustic3d 0 kappa, 1 bx, 2 by, 3 bz;
//                      elastic3d  0 bx, 1 by, 2 bz, 3 lam, 4 lam2mu, 5 mu_txy, 6 mu_txz, 7 mu_tyz.
// Storage is Float32, Float16 or BFloat16; arithmetic is always Float32.  A reduced-precision
// value v is stored as RN(v * enc) and read back as float(stored) * dec, one pair per array.
constexpr int kMaxF3 = 9, kMaxM3 = 8;

// Row layout of every 3D array: cell (i, j, k) of the padded nx x ny x nz grid sits at element
// i + j*px + k*px*ny from the array's base.  The pitch px is a multiple of 128 bytes and the base
// lies xoff elements into the allocation, so that column `pad` (the first interior column) of
// every row starts on a 128-byte boundary: the interior's tiles of 64 columns then read and write
// whole cache lines (512^3 FD8, nbc 32: the fused interior launch takes 1.66 ms with rows at the
// unaligned 584-cell pitch and 1.38 ms aligned, elastic3d's step 8.64 and 8.06 ms).
struct RowPitch {
  int xoff, px;
};
inline RowPitch row_pitch3d(int nx, int pad, int elem) {
  const int a = 128 / elem;
  const int xoff = (a - pad % a) % a;
  return {xoff, (nx + xoff + a - 1) / a * a};
}

struct Shot3D {
  void* fld[2][kMaxF3];  // array bases (see RowPitch)
  const void* mat[kMaxM3];
  float fdec[kMaxF3], fenc[kMaxF3];
  float mdec[kMaxM3];
  const float* wlut;  // HABC weight tables by distance: vx, vz, vy, tau (H+1 each)
  float dtdh, q, qt, qxt;
  int nx, ny, nz, H, pad;
  int px;  // row pitch in elements
  int n_src;
  const long long* src_c;
  const float* wav;
  unsigned* ovf;  // 16-bit storage: set to 1 when a stored value overflowed (see st_field)
};

// A column tile marched down z.  Output cells: x in [i0, i0+w), y in [j0, j0+hy), z in
// [k0, k0+hz).  ring bit 0 (x) / bit 1 (y): the tile also computes pre-correction values one
// cell beyond its outputs in x (w <= 62) / y (hy <= rows-2), which the Higdon blend of x / y frame
// cells reads; without the bit no output cell is an x / y frame cell (w <= 64, hy <= rows).  Core
// items (the user's interior in x, y and z) go to the core list, all others to the frame list.
// kind < 0: padding.
struct Work3D {
  int kind, i0, j0, k0, w, hy, hz, ring;
};

// Rows of a tile (thread rows x rows per thread): acoustic 8 x 2 (512-thread blocks; 2 rows per
// thread keep the per-thread state small enough for 4 waves per SIMD, 16 rows halve the share of
// y-halo reads), elastic 8 x 1 (its z-queues and prefetch registers per row are three times the
// acoustic ones: one row per thread keeps it near 128 VGPRs, 3-4 waves per SIMD, where two rows
// per thread needed 224-234 and ran 2 waves per SIMD, 9 % slower).
constexpr int kRowsAc3 = 16, kRowsEl3 = 8;
// Core tiles (user interior, no frame logic): acoustic 16 x 2 rows (1024-thread blocks, y-halo
// reads a quarter of the plane), elastic as its frame tiles.
constexpr int kRowsAc3Core = 32, kRowsEl3Core = 8;

// Fused acoustic3d interior tiles (engine3d_fused.hip): 64 x kFTY cells in x-y, 512 threads.
constexpr int kFTY = 16;

// Raw-add source injection into a cell of plane-step `it` (src/acquisition/inject_source_3d.jl:5-17).
__device__ __forceinline__ void inject3d(const Shot3D& S, size_t c, int it, int nt, float& v) {
  if (it >= nt) return;
  for (int q = 0; q < S.n_src; ++q)
    if ((size_t)G(S.src_c)[q] == c) v += G(S.wav)[(size_t)q + (size_t)S.n_src * (size_t)it];
}

// The shot descriptor travels by value in the kernel arguments (constant, so its pointers and
// scales stay in scalar registers or are reloaded from the argument segment).
struct Sweep3DArgs {
  Shot3D shot;
  const Work3D* work;
  int n_work;
  int parity;
  float a[kMaxN];
  const int* step;  // device step counter; the step index is *step + koff
  int koff;
  int nt;
};

struct Gather3DArgs {
  Shot3D shot;
  const long long* rec_c;
  float* seis[4];
  int slot_field[4];
  int nslots;
  int n_rec;
  int parity_out;  // buffer holding the state after the step
  int nt;
  const int* step;  // the state after step *step + koff goes into that column
  int koff;
  // Snapshots, written by the same launch after every snap_every-th step (none when snap is
  // NULL): the interior slice of field snap_f in plane snap_plane (0 xz, 1 xy, 2 yz) at padded
  // index snap_idx, decoded to Float32, snapshot k at snap + k * (cells of one slice).
  float* snap;
  int snap_every, snap_n, snap_plane, snap_idx, snap_f;
};

// eq: FOMO_ACOUSTIC3D / FOMO_ELASTIC3D; storage: FOMO_STORE_F32 / F16 / BF16; phase 0 velocity,
// 1 pressure or stress.  Two launches: the core items (user interior, no frame cell) and the
// frame items (everything else); they write disjoint cells and read only read-only buffers.
hipError_t launch_sweep3d(int eq, int storage, int phase, int N, const Sweep3DArgs& core, const Sweep3DArgs& frame,
                          hipStream_t s);
hipError_t launch_gather3d(int storage, const Gather3DArgs& a, hipStream_t s);
// acoustic3d: the whole step (new velocities, then new pressure) of the interior tiles in a.work;
// the frame velocity launch runs before it and the frame pressure launch after it.
hipError_t launch_fused3d(int storage, int N, const Sweep3DArgs& a, hipStream_t s);
// Float32 copy of a stored array without its row pitch (rows of nx values, px apart at src).
hipError_t launch_decode3d(int storage, const void* src, float* dst, int nx, size_t rows, int px, float dec,
                           hipStream_t s);

}  // namespace fomo
