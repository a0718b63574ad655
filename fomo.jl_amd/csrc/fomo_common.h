// fomo_common.h -- shared declarations of the gfx950 forward-modeling engine (host + device).
//
// Storage convention for every device field: padded grid, x fastest, 0-based (i, j[, k]);
// cell (i, j) of a 2D grid lives at i + j*nx.  The reference's 1-based (i_ref, j_ref) is
// (i+1, j+1).  Each time-dependent field has two buffers (ping-pong): after k steps the state
// of every field lives in buffer k % 2.
//
// A run is a table of *shots*.  A solo run is a table of one shot; a batch (the reference's
// acoustic2d_batch, or uneven per-shot windows) is a longer table.  Every kernel walks a flat
// work list of (shot, tile) pairs, so a shot's arithmetic does not depend on the other shots
// in the table: a batched shot is bit-identical to its solo run.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/fomo_hip.h"

namespace fomo {

constexpr int kMaxN = 5;  // fd_order/2 for fd_order <= 10

// Pointers tagged as device global memory.  A plain pointer taken from a table in memory is a
// generic pointer, and accesses through it compile to flat_* instructions, whose completion is
// counted together with LDS traffic (a wait for an LDS read then also waits for HBM); accesses
// through these compile to global_* instructions.
template <class T>
using gp = T __attribute__((address_space(1)))*;
template <class T>
__device__ __forceinline__ gp<T> G(T* p) {
  return (gp<T>)p;
}

// Sweep phases.
// kCpVel / kCpPos: the coupled P-S potential equations' "velocity" (dP/dt, dS/dt) and "position"
// (P, S) half-steps.
enum Phase : int { kAcVel = 0, kAcPrs = 1, kElVel = 2, kElStr = 3, kScal = 4, kCpVel = 5, kCpPos = 6 };

// One shot of a 2D run (device-resident table entry).  Velocity-stress equations keep each field
// in two buffers (ping-pong, fld[s][f]); the second-order scalar equation keeps u in three buffers
// fld[0][0..2] and the state after k steps, u^k, lives in fld[0][k % 3] (u^{k-1} in the previous
// one), so the leapfrog reads u^k and u^{k-1} and writes the third buffer.
// Fields and materials are stored as the plan's storage type (Float32, Float16 or BFloat16); a
// 16-bit value v is stored as RN(v * enc) and read back as float(stored) * dec.
struct Shot2D {
  void* fld[2][5];      // ping-pong fields. acoustic: p, vx, vz; elastic: vx, vz, txx, tzz, txz;
                        // coupled: P, S, dP/dt, dS/dt
  const void* mat[8];   // acoustic: kappa, bx, bz; elastic: lam, lam2mu, muxz, bx, bz; coupled: alpha,
                        // beta, d alpha/dx, d alpha/dz, d beta/dx, d beta/dz, lap alpha, lap beta
  float fdec[5], fenc[5], mdec[8];
  const float* wlut;    // HABC weights by distance to the edge: 4 tables of H+1 (vx, vz, vy, tau)
  float dtx;            // Float32(dt/dh)
  float q, qt, qxt;     // Higdon coefficients (coupled: of the P fields)
  float q2, qt2, qxt2;  // coupled: Higdon coefficients of the S fields
  float dt, ih, ih2;    // coupled: dt, 1/dh and (1/dh)^2 of the velocity half-step
  int nx, nz;           // padded dims
  int H;                // HABC frame width (pad - 1)
  int n_src;
  const int* src_c;     // padded cell index of each source
  const float* wav;     // n_src x nt
  int src_fz;           // elastic: 1 = vertical force into vz, 0 = explosive into txx, tzz
  // Learned boundary (acoustic only; NULL otherwise): per-cell corrections of the blend weights of
  // the velocity and pressure fields at frame cells, and, while training, the sensitivities
  // f_pde - S of vx, vz, p.
  const float* alpha[2];
  float* sens[3];
  // Cerjan sponge instead of the Higdon blend (NULL: HABC): damping factor by distance to the
  // padded edge, pad entries (1 at distances >= nbc).
  const float* damp;
  // Snapshots (NULL: none): after every snap_every-th step, the interior of p (acoustic) or vz
  // (elastic) as Float32 into snap[k] (x fastest), k = steps / snap_every - 1 < snap_n.
  float* snap;
  float* snap2;         // coupled: snapshots of S (snap holds P)
  int snap_every, snap_n;
  int nbuf;             // 2: ping-pong fields; 3: scalar u^{k+1}, u^k, u^{k-1} rotation
  unsigned* ovf;        // 16-bit storage: set to 1 when a stored value overflowed (see st_field)
};

// One unit of sweep work.  kind 0: a frame tile (at most kTileW x kFrameRows cells of the
// absorbing frame, blended on-chip); kind 1: an interior strip (at most kTileW columns of the
// user's interior, marched down z).
struct Work2D {
  int shot, kind, i0, j0, w, h;
};

// Receiver tables over all shots.  Slot k records field k (acoustic p, vx, vz; elastic vx, vz)
// of the state after step t into column t.
struct GatherArgs {
  const Shot2D* shots;
  const int* rec_shot;  // shot of each receiver entry
  const int* rec_c;     // padded cell index
  const long long* rec_out;  // offset of (receiver, step 0) in the seismogram buffer
  const int* rec_stride;     // step stride (n_rec of the shot)
  float* seis[3];       // slot buffers
  int nslots;
  int n_entries;
  int nt;
};
constexpr int kGatherBlock = 256;  // receiver entries per gather block

// Arguments of one sweep launch over all shots.  The step index is *step + koff.
struct SweepArgs {
  const Shot2D* shots;
  const Work2D* work;
  int n_work;
  int parity;           // state parity s: read buffers [s], write buffers [1-s]
  float a[kMaxN];       // staggered FD coefficients (coupled: centred first-derivative weights)
  float a2[kMaxN];      // coupled: centred second-derivative weights
  const int* step;      // device step counter
  int koff;             // offset of this launch's step from the counter
  int nt;
  float c0;             // scalar and coupled equations: centre coefficient of the second derivative
  // Velocity phase only: blocks n_work .. n_work + n_gather - 1 record the state held in buffer
  // [s] (the state after step it - 1) into column it - 1.
  GatherArgs g;
  int n_gather;
};

// Launchers (defined in engine2d.hip).  tile_cfg selects the interior strip length; storage is
// FOMO_STORE_F32 / F16 / BF16.
hipError_t launch_sweep2d(int storage, int phase, int N, int tile_cfg, const SweepArgs& a, hipStream_t s);
// acoustic2d / elastic2d (eq): the whole step (new velocities, then new pressure or stresses) of
// the interior strips in a.work (kind 1 marching down, 2 up; at most fused2d_width(eq, N) columns
// each); the frame velocity launch runs before it and the frame pressure / stress launch after it.
// scalar2d: the interior's leapfrog step; its frame launch (with the gather blocks) runs before it.
hipError_t launch_fused2d(int storage, int eq, int N, const SweepArgs& a, hipStream_t s);
// The frame launches beside it (phases kAcVel / kAcPrs / kElVel / kScal; elastic2d's stress frame
// runs launch_sweep2d over 8-row tiles): the frame tiles in a.work, of kFrameRowsF rows, and in the
// first phase of a step the gather blocks.
hipError_t launch_frame2d(int storage, int phase, int N, const SweepArgs& a, hipStream_t s);
int fused2d_width(int eq, int N);
// Record the state held in buffer [parity] into column (*step + koff - 1).
hipError_t launch_gather2d(int storage, const GatherArgs& g, const int* step, int koff, int parity, hipStream_t s);
// Decode n stored values (storage type, scale dec) to Float32.
hipError_t launch_decode(int storage, const void* src, float* dst, size_t n, float dec, hipStream_t s);
int strip_rows(int tile_cfg);   // rows (z) of one interior strip
constexpr int kTileW = 64;      // tile / strip width (x): one wave64 row
constexpr int kFrameRows = 8;   // rows of a frame tile
constexpr int kFrameRowsF = 16; // rows of a frame tile of the fused families' frame launches
constexpr int kNumTileCfgs = 3;
constexpr int kNumFusedCfgs = 5;  // strip lengths of the fused acoustic2d / elastic2d / scalar2d interior launch

}  // namespace fomo
