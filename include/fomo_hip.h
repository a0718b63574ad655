/*
 * fomo_hip.h -- C-ABI of the MI355X (gfx950) forward-modeling engine.
 *
 * Plain C types only: host pointers to contiguous column-major (x fastest) Float32 / Int32
 * arrays, sizes as int32, a status code per call and a thread-local error string.  Every call
 * is blocking unless it says otherwise.  The library never keeps a caller's host pointer after
 * the call returns.
 *
 * Index conventions are the reference's: source / receiver indices are 1-based interior grid
 * indices (the engine adds the padding `pad = nbc + fd_order/2` itself, as
 * src/acquisition/source.jl:22-23 and receiver.jl:17-18 do); wavelets are (n_src x nt)
 * column-major; seismograms are (n_rec x nt) column-major, receiver fastest.
 *
 * Citations are to the Fomo.jl sources the entry point replaces.
 */
#ifndef FOMO_HIP_H
#define FOMO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
#define FOMO_OK 0
#define FOMO_EINVAL 1  /* invalid argument: the Julia binding raises ArgumentError */
#define FOMO_EHIP 2    /* HIP runtime error */
#define FOMO_ENOMEM 3  /* device or host allocation failed */
#define FOMO_EUNSUP 4  /* combination not supported by this build */
#define FOMO_EOVERFLOW 5 /* 16-bit storage: a wavefield value overflowed the storage format since the last
                          * reset (stored saturated to the largest finite value); reported by the calls
                          * that wait for the plan's stream (sync, timed advance, reads, one-shot runs) */

/* ---- equation / option enums -------------------------------------------------------------- */
#define FOMO_ACOUSTIC2D 1
#define FOMO_ELASTIC2D 2
#define FOMO_ACOUSTIC3D 3
#define FOMO_ELASTIC3D 4
#define FOMO_SCALAR2D 5  /* second-order constant-density scalar equation, u in three rotating buffers */
#define FOMO_COUPLED2D 6 /* coupled P-S potential equations (constant density), velocity-position split */

#define FOMO_SRC_EXPLOSIVE 0 /* elastic: wavelet added to txx and tzz (reference behaviour) */
#define FOMO_SRC_FZ 1        /* elastic: vertical force, wavelet added to vz after the velocity HABC */

#define FOMO_BOUNDARY_HABC 0   /* deterministic two-pass Higdon (reference default) */
#define FOMO_BOUNDARY_SPONGE 1 /* Cerjan sponge, src/boundary/sponge.jl (2D solo and batch runs) */

#define FOMO_STORE_F32 0
#define FOMO_STORE_F16 1
#define FOMO_STORE_BF16 2
#define FOMO_STORE_F32_MAT16 3 /* 2D acoustic/elastic: Float32 fields, Float16 materials */

/* Message of the last failing call on this thread ("" if none). */
const char *fomo_last_error(void);
/* Library version string. */
const char *fomo_version(void);
/* Number of visible HIP devices (0 when no GPU is present; never fails). */
int fomo_device_count(void);

/* HBM calibration (no counterpart in the reference; a measurement helper for the roofline): copies
 * `bytes` between two device buffers `iters` times with a float4 streaming kernel and writes the
 * average milliseconds per copy.  Copy bandwidth = 2 * bytes / time. */
int fomo_stream_copy(int64_t bytes, int32_t iters, double *ms_per_copy);
/* A sweep's access pattern without its stencil (no counterpart in the reference; a measurement
 * helper for the roofline): n_read arrays read and n_write written, `bytes_per_array` each, one
 * float per lane per access as the sweeps load, `iters` passes; writes the average milliseconds
 * per pass.  Bandwidth = (n_read + n_write) * bytes_per_array / time.  1 <= n_write <= 4,
 * n_write <= n_read <= 8. */
int fomo_stream_probe(int32_t n_read, int32_t n_write, int64_t bytes_per_array, int32_t iters, double *ms_per_pass);

/* ---- host-side setup helpers (no GPU needed) ---------------------------------------------- */

/* Staggered FD coefficients for fd_order in {2,4,6,8,10}; writes fd_order/2 floats.
 * Replaces get_fd_coefficients, src/utils/FD_utils.jl:16-19.  Returns FOMO_EINVAL otherwise. */
int fomo_fd_coefficients(int32_t fd_order, float *out);

/* Centred second-derivative weights for fd_order in {2,4,6,8,10}: the centre weight to *c0 and
 * fd_order/2 offset weights to d2.  Replaces get_centered_d2, src/utils/FD_centered.jl:59-62. */
int fomo_centered_d2(int32_t fd_order, float *c0, float *d2);

/* Centred first-derivative weights for fd_order in {2,4,6,8,10}: fd_order/2 floats to d1.
 * Replaces get_centered_d1, src/utils/FD_centered.jl:48-51. */
int fomo_centered_d1(int32_t fd_order, float *d1);

/* Ricker wavelet of length nt.  Replaces ricker_wavelet, src/utils/ricker_wavelet.jl:6-12. */
int fomo_ricker_wavelet(float f0, float dt, int32_t nt, float *out);

/* Higdon coefficients {q, qt, qxt} for r = v_ref*dt/dh.  Replaces the scalar part of init_habc,
 * src/boundary/habc/habc.jl:21-28. */
int fomo_habc_coefficients(float v_ref, float dt, float dh, float *out3);

/* ---- run descriptor ----------------------------------------------------------------------- */
typedef struct fomo_desc {
  int32_t eq;            /* FOMO_ACOUSTIC2D ... */
  int32_t nx, ny, nz;    /* interior grid (ny = 1 for 2D) */
  int32_t nbc;           /* absorbing layers; pad = nbc + fd_order/2 */
  int32_t fd_order;      /* 2,4,6,8,10 */
  int32_t nt;            /* time steps */
  float dh, dt;          /* grid spacing (m), time step (s) */
  float v_ref;           /* HABC reference velocity */
  int32_t n_src, n_rec;
  int32_t source_type;   /* FOMO_SRC_* (elastic only) */
  int32_t storage;       /* FOMO_STORE_*; FOMO_STORE_F32_MAT16 for acoustic2d/elastic2d only */
  int32_t snap_interval; /* 0 = no snapshots */
  int32_t snap_plane;    /* 3D snapshots: 0 = xz, 1 = xy, 2 = yz */
  int32_t snap_index;    /* 3D snapshots: 1-based interior index of the slice */
  int32_t use_graph;     /* 1: one hipGraph launch per pair of steps; 0: direct launches */
  int32_t record_p;      /* acoustic3d: also gather p (extension) */
  int32_t tile_cfg;      /* 0: autotune the tile shape; k > 0: fixed shape k-1 (2D sweeps: interior
                            strips of 16, 32, 64 rows; acoustic2d / elastic2d fused interior: 24, 32,
                            48, 96, 176 rows; 3D: z-segments of 16 << (k-1) planes, of the fused acoustic3d
                            interior 32 << (k-1)) */
  int32_t boundary;      /* FOMO_BOUNDARY_HABC or FOMO_BOUNDARY_SPONGE (2D) */
  float sponge_factor;   /* Cerjan factor (reference default 0.015); used with FOMO_BOUNDARY_SPONGE */
  double smooth_sigma;   /* coupled2d: Gaussian sigma in grid points of the smoothed alpha = vp^2, beta = vs^2
                            whose derivatives enter the coupling terms (the reference's default is 3.0;
                            <= 0: unsmoothed) */
  float v_ref_s;         /* coupled2d: HABC reference velocity of the S and dS/dt fields (<= 0: min(vs > 0));
                            v_ref is the P fields' */
  int32_t reserved[1];
} fomo_desc;
/* Size limit: cells are addressed with 32-bit offsets, so a padded 2D field (nx + 2 pad) x
 * (nz + 2 pad), or one padded z plane of a 3D field, holds at most 2^30 cells; larger grids are
 * refused with FOMO_EUNSUP (pad = nbc + fd_order / 2; for a batch, the whole model counts). */

/* One shot of a batch: a window of the model (full depth) with its own sources and receivers.
 * The shot is simulated exactly as a solo run on the cropped model would be: padding, medium,
 * HABC coefficients (from this window's v_ref) and arithmetic are per window. */
typedef struct fomo_shot {
  int32_t x0;              /* first interior column of the window in the model (1-based) */
  int32_t nx;              /* window width in interior columns */
  int32_t n_src;
  const int32_t *sx, *sz;  /* 1-based indices within the window */
  const float *wavelet;    /* n_src x nt, column-major */
  int32_t n_rec;
  const int32_t *rx, *rz;  /* 1-based indices within the window */
  float v_ref;             /* HABC reference velocity; <= 0 selects min(vp > 0) over the window */
} fomo_shot;

/* ---- one-shot drop-in entry points -------------------------------------------------------- */
/*
 * acoustic2d -- replaces Fomo.acoustic2d (src/equations/acoustic2d/acoustic2d.jl:47-192);
 * d->boundary selects :habc or :sponge (acoustic2d.jl:82-88, src/boundary/sponge.jl:30-78).  vp, rho: nx*nz.  wavelet: n_src x nt (already repeated per source and, if
 * requested, scaled by dt, as acoustic2d.jl:91-97 does).  Outputs (n_rec x nt each, any may be
 * NULL): seis_p, seis_vx, seis_vz.  snaps: (nt/snap_interval) x nx*nz interior p, or NULL.
 * kernel_time_s: hipEvent time of the step loop (setup, upload and download excluded).
 */
int fomo_acoustic2d(const fomo_desc *d, const float *vp, const float *rho, const int32_t *sx, const int32_t *sz,
                    const float *wavelet, const int32_t *rx, const int32_t *rz, float *seis_p, float *seis_vx,
                    float *seis_vz, float *snaps, double *kernel_time_s);

/*
 * elastic2d -- replaces Fomo.elastic2d (src/equations/elastic2d/elastic2d.jl:45-163);
 * d->boundary selects :habc or :sponge (elastic2d.jl:82-87, 180-195).  Outputs seis_vx, seis_vz (n_rec x nt); snaps of interior vz.
 */
int fomo_elastic2d(const fomo_desc *d, const float *vp, const float *vs, const float *rho, const int32_t *sx,
                   const int32_t *sz, const float *wavelet, const int32_t *rx, const int32_t *rz, float *seis_vx,
                   float *seis_vz, float *snaps, double *kernel_time_s);

/*
 * scalar2d -- replaces Fomo.scalar2d (src/equations/scalar2d/scalar2d.jl:113-189): second-order
 * constant-density leapfrog, u^{n+1} = 2u^n - u^{n-1} + c2 * lap (c2 = (v dt/dh)^2, centred
 * weights of fomo_centered_d2), the deterministic Higdon blend on u (f = u^{n+1}, f_old = u^n,
 * tau weights), raw-add injection and the u gather.  d->eq = FOMO_SCALAR2D, HABC only, no
 * snapshots.  wavelet: n_src x nt, already scaled as source_scale asks (scalar2d.jl:139-146).
 * Output seis_u (n_rec x nt).
 */
int fomo_scalar2d(const fomo_desc *d, const float *vp, const int32_t *sx, const int32_t *sz, const float *wavelet,
                  const int32_t *rx, const int32_t *rz, float *seis_u, double *kernel_time_s);

/*
 * coupled2d -- replaces Fomo.coupled2d (src/equations/coupled2d/coupled2d.jl:53-183): the coupled
 * P-S potential equations of constant-density elastic media as a velocity-position split (the
 * "velocity" half-step dP/dt += dt rhs_P, dS/dt += dt rhs_S with centred fd_order weights on the
 * regular grid, then P += dt dP/dt, S += dt dS/dt), a Higdon blend after each half-step (P fields
 * with d->v_ref, S fields with d->v_ref_s; vx weights after the velocity half-step, tau weights
 * after the position half-step), raw-add injection into P and the P, S gather.  The medium is
 * alpha = vp^2 and beta = vs^2 with the derivatives of their Gaussian-smoothed copies
 * (d->smooth_sigma), as medium.jl:50-117 builds it.  The Higdon blend is the deterministic two-pass
 * form (the reference applies its in-place single-field kernel, src/boundary/habc/kernels.jl:
 * 122-190).  d->eq = FOMO_COUPLED2D, HABC, Float32 storage.  vp, vs: nx*nz.  wavelet: n_src x nt.
 * Outputs (n_rec x nt, either may be NULL): seis_P, seis_S.  snaps_P, snaps_S:
 * (nt/snap_interval) x nx*nz interior P and S, or NULL.
 */
int fomo_coupled2d(const fomo_desc *d, const float *vp, const float *vs, const int32_t *sx, const int32_t *sz,
                   const float *wavelet, const int32_t *rx, const int32_t *rz, float *seis_P, float *seis_S,
                   float *snaps_P, float *snaps_S, double *kernel_time_s);

/*
 * acoustic3d -- replaces Fomo.acoustic3d (src/equations/acoustic3d/acoustic3d.jl:31-129).
 * vp, rho: nx*ny*nz (x fastest).  Sources (sx,sy,sz) with wavelet n_src x nt; receivers
 * (rx,ry,rz).  Outputs n_rec x nt: seis_vx, seis_vy, seis_vz, and seis_p when not NULL (an
 * extension: the reference records only velocities).  snaps: (nt/snap_interval) slices of p in
 * d->snap_plane (0 xz, 1 xy, 2 yz) at d->snap_index (1-based interior index; 0 = middle).
 * d->storage selects Float32, Float16 or BFloat16 storage of wavefields and materials
 * (arithmetic stays Float32).  The Higdon boundary is the deterministic two-pass form.
 */
int fomo_acoustic3d(const fomo_desc *d, const float *vp, const float *rho, const int32_t *sx, const int32_t *sy,
                    const int32_t *sz, const float *wavelet, const int32_t *rx, const int32_t *ry, const int32_t *rz,
                    float *seis_vx, float *seis_vy, float *seis_vz, float *seis_p, float *snaps,
                    double *kernel_time_s);

/*
 * elastic3d -- replaces Fomo.elastic3d (src/equations/elastic3d/elastic3d.jl:31-117).
 * vp, vs, rho: nx*ny*nz.  Explosive source: the wavelet row of each source is added to txx, tyy
 * and tzz after the stress HABC (elastic3d.jl:143-146).  Outputs n_rec x nt: seis_vx, seis_vy,
 * seis_vz.  snaps: slices of vz as in fomo_acoustic3d.  d->storage as in fomo_acoustic3d.
 */
int fomo_elastic3d(const fomo_desc *d, const float *vp, const float *vs, const float *rho, const int32_t *sx,
                   const int32_t *sy, const int32_t *sz, const float *wavelet, const int32_t *rx, const int32_t *ry,
                   const int32_t *rz, float *seis_vx, float *seis_vy, float *seis_vz, float *snaps,
                   double *kernel_time_s);

/*
 * acoustic2d_batch -- replaces Fomo.acoustic2d_batch (src/equations/acoustic2d/batch.jl:180-261):
 * n_shots shots over the shared model and receiver spread; d->n_src sources per shot.
 * sx, sz: n_src x n_shots (column-major, as init_batched_source stores them, batch_acq.jl:27-37);
 * wavelet: n_src x nt shared by all shots; outputs n_rec x nt x n_shots.  Each shot is
 * bit-identical to a solo fomo_acoustic2d run.
 */
int fomo_acoustic2d_batch(const fomo_desc *d, const float *vp, const float *rho, int32_t n_shots, const int32_t *sx,
                          const int32_t *sz, const float *wavelet, const int32_t *rx, const int32_t *rz,
                          float *seis_p, float *seis_vx, float *seis_vz, double *kernel_time_s);

/* elastic2d_batch -- replaces Fomo.elastic2d_batch (src/equations/elastic2d/batch.jl:197-261). */
int fomo_elastic2d_batch(const fomo_desc *d, const float *vp, const float *vs, const float *rho, int32_t n_shots,
                         const int32_t *sx, const int32_t *sz, const float *wavelet, const int32_t *rx,
                         const int32_t *rz, float *seis_vx, float *seis_vz, double *kernel_time_s);

/* ---- device-resident plans (bench, batch engine, tests) ----------------------------------- */
typedef struct fomo_plan fomo_plan;

/* Build a plan: pads the model, derives the medium and the HABC tables, uploads everything,
 * zeroes the state and (if d->use_graph) captures the step graph.  vs may be NULL for
 * acoustic equations; sy/ry NULL for 2D. */
int fomo_plan_create(const fomo_desc *d, const float *vp, const float *vs, const float *rho, const int32_t *sx,
                     const int32_t *sy, const int32_t *sz, const float *wavelet, const int32_t *rx,
                     const int32_t *ry, const int32_t *rz, fomo_plan **out);
/* Build a multi-shot plan: every shot is a window of the model (fomo_shot).  All shots advance
 * together, one launch per phase over all tiles of all shots.  An extension: the reference's
 * batch shares one grid between shots (fomo_acoustic2d_batch covers that form). */
int fomo_batch_create(const fomo_desc *d, const float *vp, const float *vs, const float *rho, int32_t n_shots,
                      const fomo_shot *shots, fomo_plan **out);
/* Number of shots of a plan. */
int fomo_plan_n_shots(const fomo_plan *p, int32_t *out);
/* Seismograms of one shot (n_rec_shot x nt each; NULL skips). */
int fomo_plan_read_shot_seis(fomo_plan *p, int32_t shot, float *s0, float *s1, float *s2, float *s3);
/* Current padded field `which` of one shot. */
int fomo_plan_read_shot_field(fomo_plan *p, int32_t shot, int32_t which, float *out);
/* Padded dims and pad width of one shot. */
int fomo_plan_shot_dims(const fomo_plan *p, int32_t shot, int32_t *nx, int32_t *ny, int32_t *nz, int32_t *pad);
/* Zero every field and seismogram, reset the step counter to 0.  Asynchronous. */
int fomo_plan_reset(fomo_plan *p);
/* Advance n steps (enqueue only; asynchronous on the plan's stream). */
int fomo_plan_advance(fomo_plan *p, int32_t n);
/* Advance n steps bracketed by hipEvents on the plan's stream; blocks; returns milliseconds. */
int fomo_plan_timed_advance(fomo_plan *p, int32_t n, double *ms);
int fomo_plan_sync(fomo_plan *p);
/* Per-kernel timing on the plan's stream with hipEvents: the velocity and the pressure/stress sweep
 * are each launched n times back to back from the current state (repeating a sweep rewrites the
 * same values); then n ordinary steps are timed.  Writes the average milliseconds per launch of
 * the velocity sweep and of the pressure/stress sweep, and the rest of a step's time (receiver
 * gather, step-counter updates, gaps between launches), to out3.  The state advances by n steps.
 * Blocks. */
int fomo_plan_kernel_times(fomo_plan *p, int32_t n, double *out3);
/* Per-launch timing (a measurement helper with no counterpart in the reference): the kernel
 * launches of one step (2D: velocity sweep, pressure/stress sweep; scalar2d: one sweep; 3D:
 * velocity core, velocity frame, pressure/stress core, pressure/stress frame), with each one's
 * name, the cells it updates and the algorithmic bytes per cell of its phase (the reference's
 * per-phase reads and writes; 16-bit storage halves them).  fomo_plan_launch_times launches each
 * one n times back to back from the current state (idempotent) between hipEvents on the plan's
 * stream and writes the average milliseconds per launch, ms[k] for k < count.  Blocks. */
int fomo_plan_launch_count(const fomo_plan *p, int32_t *n);
int fomo_plan_launch_info(const fomo_plan *p, int32_t k, const char **name, int64_t *cells, int32_t *bytes_per_cell);
int fomo_plan_launch_times(fomo_plan *p, int32_t n, double *ms);
/* The work layout a plan runs with (no counterpart in the reference; the autotuner's result):
 * out4[0] family (0 separate sweeps, 1 a fused interior launch), out4[1] shape index (the
 * descriptor's tile_cfg - 1), out4[2] its extent (2D: rows of an interior strip; 3D: z-planes of a
 * segment of the fused interior launch, else of the sweeps), out4[3] where it came from (0 timed
 * by this plan, 1 an earlier plan of this process, 2 the tuning file, 3 fixed by d->tile_cfg or
 * the environment).  Choices are kept per device, library build and problem in the file named by
 * FOMO_TUNE_CACHE (default $XDG_CACHE_HOME/fomo_hip/tune.txt or ~/.cache/fomo_hip/tune.txt;
 * "off" disables the file).  Layouts change no bits. */
int fomo_plan_tuning(const fomo_plan *p, int32_t *out4);
/* Steps taken since the last reset. */
int fomo_plan_steps_done(const fomo_plan *p, int32_t *out);
/* Copy seismograms to host (n_rec x nt each; NULL skips).  Slots: acoustic {p, vx, vz},
 * elastic2d {vx, vz, -}, scalar2d {u}, coupled2d {P, S}, acoustic3d {vx, vy, vz, p}. */
int fomo_plan_read_seis(fomo_plan *p, float *s0, float *s1, float *s2, float *s3);
/* FP64 per-receiver sums over one trace slot of two plans' recorded traces, time samples
 * t0 <= t < t1: num[r] = sum_t (a[r,t] - b[r,t])^2 and den[r] = sum_t b[r,t]^2, with b = 0 when pb
 * is NULL (then num[r] is the energy of trace r of a).  The shots must have equal receiver counts
 * and nt.  Reduced on the GPU in a fixed order, so repeated calls give the same bits.  Replaces the
 * host-side norms of the reference's residual checks: test/verify_patch.jl:86-87
 * (||h - ref|| / ||ref||, the reflection measure) and the per-trace relative L2 of its solver
 * comparisons.  Blocks. */
int fomo_trace_misfit(fomo_plan *pa, int32_t shot_a, fomo_plan *pb, int32_t shot_b, int32_t slot, int32_t t0,
                      int32_t t1, double *num, double *den);
/* fomo_trace_misfit without allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device or pinned memory (tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and bits as fomo_trace_misfit; the partial sums live in a static device
 * array of 2^21 doubles (16 MiB, held by every GPU the library runs on until the process ends), so
 * (chunks + 1) x 2 x n_rec must not exceed 2^21, with chunks = min(64, ceil((t1 - t0) / 256)) (for
 * example 65,536 receivers over 4000 samples); larger calls fail with FOMO_EINVAL.  The results are
 * copied straight into num and den.  Calls that overlap on one GPU share its array and race.
 * Blocks. */
int fomo_trace_misfit_noalloc(fomo_plan *pa, int32_t shot_a, fomo_plan *pb, int32_t shot_b, int32_t slot,
                              int32_t t0, int32_t t1, double *num, double *den);
/* Copy the current padded field number `which` (order: acoustic2d p,vx,vz; elastic2d
 * vx,vz,txx,tzz,txz; coupled2d P,S,dP/dt,dS/dt; acoustic3d p,vx,vy,vz; elastic3d
 * vx,vy,vz,txx,tyy,tzz,txy,txz,tyz) to host as a dense padded Float32 array, x fastest (decoded from
 * the storage precision when reduced-precision storage is used; 3D plans keep rows at a 128-byte
 * pitch on the device, which the copy drops). */
int fomo_plan_read_field(fomo_plan *p, int32_t which, float *out);
/* Snapshots of a plan created with d->snap_interval > 0 (2D single-shot and 3D plans): writes the
 * number of snapshot slots to *count and, when out is not NULL, copies every slot (count x the
 * interior cells of one snapshot, as the one-shot entry points return them; coupled2d: the
 * nt/snap_interval slots of P, then as many of S).  Slots of steps not
 * yet taken since the last reset are zero. */
int fomo_plan_read_snaps(fomo_plan *p, int32_t shot, int32_t *count, float *out);
/* Padded grid dims and pad width. */
int fomo_plan_dims(const fomo_plan *p, int32_t *nx, int32_t *ny, int32_t *nz, int32_t *pad);
int fomo_plan_destroy(fomo_plan *p);

/* ---- learned absorbing boundary (extension; acoustic2d single-shot plans) ----------------- */
/* The Higdon blend weight of every frame cell comes from a small MLP run per frame cell
 * (76 inputs: max-abs-scaled 5x5 patches of p, vx, vz and dist/H; 32-32 ReLU; 2 sigmoid outputs:
 * the weights of the velocity fields and of p).  Weights are stored as BFloat16, arithmetic is
 * Float32.  The reference has no counterpart; its distance-ramp weights (habc.jl:30-34) remain
 * the default and the fallback. */
int fomo_learned_n_params(void);
/* Attach the learned boundary to plan p with weights (fomo_learned_n_params() floats); train != 0
 * also allocates the training buffers. */
int fomo_learned_enable(fomo_plan *p, const float *weights, int32_t train);
int fomo_learned_set_weights(fomo_plan *p, const float *weights);
/* Advance teacher and student one step each and accumulate the gradient of the one-step loss
 * sum(((f_student - f_teacher)/scale)^2) over the student's frame cells (student cell (i,j) =
 * teacher cell (i+off_x, j+off_z)).  Asynchronous. */
int fomo_learned_train_step(fomo_plan *student, fomo_plan *teacher, int32_t off_x, int32_t off_z, float scale_p,
                            float scale_v);
/* Accumulated weight gradients (n_params floats) and loss since the last call; clears them. */
int fomo_learned_grad(fomo_plan *p, float *grad, double *loss);
/* Frame cells (padded indices) and their current blend weights. */
int fomo_learned_alpha(fomo_plan *p, int32_t *n, int32_t *cells, float *alpha_v, float *alpha_p);
/* MLP forward + backward on given features (n x 76, row-major) and dL/dalpha (n x 2): writes
 * alpha (n x 2) and the summed weight gradients (for gradient checks). */
int fomo_mlp_eval(const float *weights, int32_t n, const float *feats, const float *galpha, float *alpha,
                  float *grad);

#ifdef __cplusplus
}
#endif
#endif /* FOMO_HIP_H */
