// fomo_common.h -- shared declarations of the gfx950 forward-modeling engine (host + device).
//
// Storage convention for every device field: padded grid, x fastest, 0-based (i, j[, k]);
// cell (i, j) of a 2D grid lives at i + j*nx.  The reference's 1-based (i_ref, j_ref) is
// (i+1, j+1).  Each time-dependent field has two buffers (ping-pong): after k steps the state
// of every field lives in buffer k % 2.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fomo {

constexpr int kMaxN = 5;  // fd_order/2 for fd_order <= 10

// Sweep phases.
enum Phase : int { kAcVel = 0, kAcPrs = 1, kElVel = 2, kElStr = 3 };

// One sweep (one phase of one time step) over a 2D padded grid.  Everything the sweep reads is
// either a read-only input of this phase or the phase's own fields at time n (old[]), and every
// write goes to out[] (time n+1): the sweep is a pure function of its inputs.
struct Sweep2DArgs {
  const float* rd[3];    // stencil inputs: AcVel {p}; AcPrs {vx,vz}; ElVel {txx,tzz,txz}; ElStr {vx,vz}
  const float* old[3];   // this phase's fields at time n
  float* out[3];         // this phase's fields at time n+1
  const float* mat[3];   // AcVel {bx,bz}; AcPrs {kappa}; ElVel {bx,bz}; ElStr {lam,lam2mu,muxz}
  const float* wlut[3];  // HABC blend weight as a function of the distance to the grid edge
  float a[kMaxN];        // staggered FD coefficients
  float dtx;             // Float32(dt/dh)
  float q, qt, qxt;      // Higdon coefficients
  int nx, nz;            // padded dims
  int H;                 // HABC frame width (pad - 1)
  // Raw-add injection, applied after the HABC to the fields in inj_mask.
  int inj_mask;
  int n_src;
  const int* src_c;      // padded cell index of each source
  const float* wav;      // n_src x nt, column-major
  const int* step;       // device step counter (0-based index of the step being taken)
  int nt;
};

// Receiver gather after the step, then the step counter is advanced by the last block.
struct RecordArgs {
  const float* f[4];
  float* seis[4];
  int nf;
  const int* rec_c;      // padded cell index of each receiver
  int n_rec;
  int nt;
  int* step;
  unsigned* done;
};

// Launchers (defined in sweep2d.hip).  Return hipSuccess or the launch error.
hipError_t launch_sweep2d(int phase, int N, const Sweep2DArgs& a, hipStream_t s);
hipError_t launch_record(const RecordArgs& a, hipStream_t s);

}  // namespace fomo
