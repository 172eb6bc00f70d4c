/*
 * fomo_mi355x.h -- C ABI of the MI355X (gfx950) time-stepping path of the
 * Fomo_julia 2D elastic solver.  Plain pointers, sizes and opaque handles;
 * no GPU or framework types cross this boundary, so Julia binds it with
 * `ccall` (the reference's only FFI) and Python with ctypes.
 *
 * Conventions
 *  - Every function returns int32: FOMO_OK (0) or a negative FOMO_ERR_* code;
 *    fomo_ctx_last_error() returns the message of the last failure on that
 *    context (owned by the context, valid until the next call on it).
 *  - Arrays follow the reference's Julia layout: a field is (nx, nz)
 *    column-major Float32 (element (i, j), 1-based, at (i-1) + (j-1)*nx);
 *    HABC weights are (nz, nx) (src/utils/init.jl:193-195); a gather is
 *    [nt x n_rec] column-major (src/simulation/shot_manager.jl:17-28).
 *  - Grid indices passed in (source, receivers) are 1-based, as in Julia.
 *  - Host pointers are borrowed for the duration of the call only.
 *  - A context owns one HIP stream; all work is queued on it in call order.
 *    Runs on one context are driven by one host thread at a time.  Calls that
 *    allocate, free or capture graphs, and the pool calls, hold one library
 *    lock, so other threads may create and destroy models, wavefields and
 *    shots while that thread runs (the serving loop does).
 *  - Device memory: one model (medium) is shared by any number of
 *    wavefields and shots; a shot holds its wavelet, receivers and gather.
 */
#ifndef FOMO_MI355X_H
#define FOMO_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FOMO_OK 0
#define FOMO_ERR_ARG (-1)         /* invalid argument or size */
#define FOMO_ERR_HIP (-2)         /* HIP runtime failure (message has the HIP error) */
#define FOMO_ERR_UNSUPPORTED (-3) /* valid for the reference, not supported here */
#define FOMO_ERR_STATE (-4)       /* object used with a mismatched model/context */

/* Source types: MultiShotConfig.source_type (src/simulation/shot_manager.jl:99-102). */
#define FOMO_SRC_PRESSURE 0 /* txx += w, tzz += w   (src/kernels/source_receiver.jl:18-30) */
#define FOMO_SRC_FORCE_Z 1  /* vz += w*dt/rho_vz    (src/backend/cpu.jl:180-181) */
#define FOMO_SRC_FORCE_X 2  /* vx += w*dt/rho_vx    (src/backend/cpu.jl:182-183) */

/* Receiver types: Receivers.type (src/core/Structures.jl:131). */
#define FOMO_REC_VZ 0
#define FOMO_REC_VX 1
#define FOMO_REC_P 2 /* (txx+tzz)/2, the CPU path's :p (source_receiver.jl:72-73) */

/* Field ids for download/upload: Wavefield's components (Structures.jl:18-32). */
#define FOMO_FIELD_VX 0
#define FOMO_FIELD_VZ 1
#define FOMO_FIELD_TXX 2
#define FOMO_FIELD_TZZ 3
#define FOMO_FIELD_TXZ 4
#define FOMO_FIELD_VX_OLD 5
#define FOMO_FIELD_VZ_OLD 6
#define FOMO_FIELD_TXX_OLD 7
#define FOMO_FIELD_TZZ_OLD 8
#define FOMO_FIELD_TXZ_OLD 9

/* Storage of the wavefields and medium coefficients; stencil arithmetic is
 * FP32 in every case (no counterpart in the reference, which is FP32 only).
 *   F32   everything FP32 (the reference's arithmetic, bit for bit)
 *   BF16  coefficients and fields BF16
 *   F16   coefficients and fields FP16
 *   F16C  coefficients FP16, fields FP32
 *   F24   coefficients FP32, fields rounded to a 16-bit significand (FP32's
 *         exponent range, no scale): 3 bytes per field cell, a plane of each
 *         value's high 16 bits and a plane of its next 8 bits
 * Fields are rounded to nearest (ties to even) whenever a stage writes them.
 * FP16 arrays carry a power-of-two scale (exact): per coefficient array from
 * its maximum, per field from a bound of the injected wavelet, so values stay
 * inside FP16's range.  Coefficients are quantized once, with a deterministic
 * per-cell dither (round up when the remainder exceeds a hash of the cell), so
 * their rounding errors average out over a wavelength instead of biasing a
 * homogeneous medium.  Reduced storage needs fd_order 4, 8 or 12 and the fused
 * step (step mode 1); per-stage calls, the two-pass step and the *_old fields
 * return FOMO_ERR_UNSUPPORTED.  A wavefield's FP16 state carries the scales
 * of the shot that produced it: continuing it with a shot of other scales
 * returns FOMO_ERR_STATE (reset it first). */
#define FOMO_STORAGE_F32 0
#define FOMO_STORAGE_BF16 1
#define FOMO_STORAGE_F16 2
#define FOMO_STORAGE_F16C 3
#define FOMO_STORAGE_F24 4

typedef struct fomo_ctx fomo_ctx;
typedef struct fomo_model fomo_model;
typedef struct fomo_wavefield fomo_wavefield;
typedef struct fomo_shot fomo_shot;
typedef struct fomo_batch fomo_batch;

/* The scalars of Medium, HABCConfig and SimParams that the step needs. */
typedef struct fomo_model_desc {
  int32_t nx, nz;       /* padded grid: Medium.nx, Medium.nz (Structures.jl:67-68) */
  int32_t pad;          /* Medium.pad = nbc + fd_order/2 (init.jl:72-73) */
  int32_t nbc;          /* HABCConfig.nbc (Structures.jl:95); 1 <= nbc <= 63 */
  int32_t fd_order;     /* SimParams.fd_order (Structures.jl:148); even, 2..16 */
  int32_t free_surface; /* Medium.is_free_surface (Structures.jl:75) */
  float dx, dz;         /* Medium.dx, Medium.dz */
  float dt, dtx, dtz;   /* SimParams.dt, .dtx, .dtz (Structures.jl:144-147) */
  float qx, qz, qt_x, qt_z, qxt; /* HABCConfig (Structures.jl:96-100) */
  int32_t storage;      /* FOMO_STORAGE_* */
} fomo_model_desc;

/* ---- context ------------------------------------------------------------ */
/* One HIP device + stream.  Replaces CUDA.jl's implicit device/stream of the
 * CUDABackend (src/backends/backend.jl:20,68-79). */
int32_t fomo_ctx_create(int32_t device, fomo_ctx** out);
int32_t fomo_ctx_destroy(fomo_ctx* ctx);
const char* fomo_ctx_last_error(const fomo_ctx* ctx);
/* synchronize(::CUDABackend) (src/backends/backend.jl:78-79). */
int32_t fomo_ctx_synchronize(fomo_ctx* ctx);
/* The hipStream_t all work of this context is queued on (for event timing). */
void* fomo_ctx_stream(fomo_ctx* ctx);

/* ---- model: Medium + HABCConfig + SimParams + FD coefficients ----------- */
/* Replaces to_device of the Medium arrays (init.jl:90-98), of the HABC
 * weights (init.jl:197-202) and of the coefficient vector (run.jl:93).
 * fd_coeffs: fd_order/2 values (get_fd_coefficients, init.jl:26-29).
 * rho_vx..mu_txz: (nx, nz) arrays of Medium (Structures.jl:78-82).
 * w_vx, w_vz, w_tau: (nz, nx) HABC weights; they must be the distance ramps
 * init_habc builds (init.jl:188-195), else FOMO_ERR_UNSUPPORTED.  Passing
 * NULL for all three builds those ramps from nbc directly. */
int32_t fomo_model_create(fomo_ctx* ctx, const fomo_model_desc* desc, const float* fd_coeffs,
                          const float* rho_vx, const float* rho_vz, const float* lam,
                          const float* mu_txx, const float* mu_txz, const float* w_vx,
                          const float* w_vz, const float* w_tau, fomo_model** out);
/* The same model built on the device from the unpadded (nx_in, nz_in) vp, vs,
 * rho arrays of init_medium's arguments (init.jl:68-99 with _pad_array and
 * _compute_staggered_params, init.jl:119-158): padding, staggered averages
 * and coefficients are computed by a kernel, so only the three unpadded
 * arrays cross the bus.  desc->nx/nz must be nx_in/nz_in + 2*pad.  HABC
 * weights are the init_habc ramps of desc->nbc.  Identical results to
 * fomo_model_create on init_medium's arrays. */
int32_t fomo_model_create_from_velocity(fomo_ctx* ctx, const fomo_model_desc* desc,
                                        const float* fd_coeffs, const float* vp, const float* vs,
                                        const float* rho, int32_t nx_in, int32_t nz_in,
                                        fomo_model** out);
int32_t fomo_model_destroy(fomo_model* m);
/* Bytes of device memory the model holds. */
int64_t fomo_model_device_bytes(const fomo_model* m);

/* ---- wavefield ---------------------------------------------------------- */
/* Wavefield(nx, nz, backend) (Structures.jl:39-55): 10 zeroed fields. */
int32_t fomo_wavefield_create(fomo_model* m, fomo_wavefield** out);
int32_t fomo_wavefield_destroy(fomo_wavefield* w);
/* reset!(backend, W) (src/kernels/source_receiver.jl:115-129). */
int32_t fomo_wavefield_reset(fomo_wavefield* w);
/* Copy one field (FOMO_FIELD_*) to / from a host (nx, nz) array. */
int32_t fomo_wavefield_download(fomo_wavefield* w, int32_t field, float* host);
int32_t fomo_wavefield_upload(fomo_wavefield* w, int32_t field, const float* host);
int64_t fomo_wavefield_device_bytes(const fomo_wavefield* w);

/* ---- shot: Source + Receivers + gather ----------------------------------- */
/* _create_source / _create_receivers (shot_manager.jl:196-222): source cell
 * (src_i, src_j) as _coord_to_index returns it, wavelet of n_wavelet samples
 * (steps beyond it inject nothing, source_receiver.jl:19-21), nt steps of
 * gather for n_rec receivers at (rec_i[r], rec_j[r]) of type rec_type[r]
 * (NULL = all FOMO_REC_VZ).  The gather starts zeroed. */
int32_t fomo_shot_create(fomo_model* m, int32_t src_i, int32_t src_j, int32_t src_type,
                         const float* wavelet, int32_t n_wavelet, int32_t nt,
                         const int32_t* rec_i, const int32_t* rec_j, const int32_t* rec_type,
                         int32_t n_rec, fomo_shot** out);
int32_t fomo_shot_destroy(fomo_shot* s);
/* The power-of-two store scales (vx vz txx tzz txz) FP16 storage uses for this
 * shot's wavefield: a value x is stored as x * scale.  All 1 for FP32 and BF16
 * storage.  No counterpart in the reference. */
int32_t fomo_shot_field_scales(const fomo_shot* s, float* out5);
/* fill!(rec.data, 0f0) (shot_manager.jl:54). */
int32_t fomo_shot_reset(fomo_shot* s);
/* _get_gather (shot_manager.jl:71-72): [nt x n_rec] column-major into out. */
int32_t fomo_shot_gather(fomo_shot* s, float* out);
/* No counterpart in the reference.  fomo_shot_gather for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same [nt x n_rec] result, with the
 * transpose done in chunks of time steps through a fixed 16 MiB device array
 * of the library (one per GPU, held until the process ends), so the call
 * allocates nothing.  n_rec above 65536 (64 steps per chunk) returns
 * FOMO_ERR_ARG.  All calls on one GPU share that array: calls that overlap in
 * time on one GPU (other threads, other contexts of the device) race. */
int32_t fomo_shot_gather_noalloc(fomo_shot* s, float* out);

/* ---- the time step ------------------------------------------------------ */
/* time_step!(backend, W, M, H, a, src, rec, k, params) (time_stepper.jl:39-63), k 1-based. */
int32_t fomo_time_step(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int32_t k);
/* run_time_loop! (time_stepper.jl:97-128) for steps k_first .. k_first+n_steps-1,
 * replayed as captured HIP graphs of whole steps; asynchronous (call
 * fomo_ctx_synchronize). */
int32_t fomo_run_time_loop(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int32_t k_first,
                           int32_t n_steps);
/* Captures the graphs fomo_run_time_loop replays for (m, w, s) without
 * running a step, so that a later run (a timed one, or the first of a
 * serving loop) does not capture inside it.  Part of run_time_loop!'s job
 * (time_stepper.jl:97-128); the reference has no separate call for it. */
int32_t fomo_run_time_loop_prepare(fomo_model* m, fomo_wavefield* w, fomo_shot* s);

/* ---- wavefield frames: the on_step video callback without leaving the graphs
 * MultiFieldRecorder (video_recorder.jl:113-148,207-228) called as on_step
 * (time_stepper.jl:112-119) takes, on every step k with k % skip == 0, the
 * frame of each field (_extract_frame!, video_recorder.jl:260-268):
 *   frame[i_d, j_d] = raw(min(i_d*downsample, nx-1), min(j_d*downsample, nz-1))
 * (0-based; raw over the whole padded grid) for i_d < cld(nx_frame, downsample),
 * j_d < cld(nz_frame, downsample), with raw = (txx + tzz) * 0.5 (P),
 * sqrt(vx^2 + vz^2) (VEL), vx, vz (_get_field_data, :270-283).  A frame store
 * keeps `capacity` such frames in device memory, [slot][field][j_d][i_d]
 * (x fastest), so a run records them inside its captured graphs with no host
 * round trip per frame. */
#define FOMO_FRAME_P 0
#define FOMO_FRAME_VEL 1
#define FOMO_FRAME_VX 2
#define FOMO_FRAME_VZ 3
typedef struct fomo_frames fomo_frames;
/* VideoConfig(fields, skip, downsample) (video_recorder.jl:68-81) + MultiFieldRecorder(nx, nz, ...)
 * (:113-127, 164-200), without the plotting. */
int32_t fomo_frames_create(fomo_ctx* ctx, int32_t nx_frame, int32_t nz_frame, const int32_t* fields,
                           int32_t n_fields, int32_t skip, int32_t downsample, int32_t capacity,
                           fomo_frames** out);
int32_t fomo_frames_destroy(fomo_frames* f);
/* Frame size per field: cld(nx_frame, downsample), cld(nz_frame, downsample). */
int32_t fomo_frames_shape(const fomo_frames* f, int32_t* nx_d, int32_t* nz_d);
int64_t fomo_frames_device_bytes(const fomo_frames* f);
/* Slots first .. first+count-1 into host ([count][n_fields][nz_d][nx_d]); synchronous. */
int32_t fomo_frames_download(fomo_frames* f, int32_t first, int32_t count, float* host);
/* fomo_run_time_loop that, after every step k with k % skip == 0, writes the
 * frames of the state after step k into slot (k/skip - 1) mod capacity (a
 * ring: download before a later frame reuses the slot):
 * run_time_loop!(...; on_step = recorder) (time_stepper.jl:112-119). */
int32_t fomo_run_time_loop_frames(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int32_t k_first,
                                  int32_t n_steps, fomo_frames* f);
/* The frames of the wavefield's current state into `slot` (_SingleRecorder
 * called on W, video_recorder.jl:207-228). */
int32_t fomo_wavefield_frame(fomo_wavefield* w, fomo_frames* f, int32_t slot);

/* Per-stage entry points, each one kernel of the step (Elastic2D.jl:55-58). */
int32_t fomo_backup_boundary(fomo_model* m, fomo_wavefield* w);             /* boundary.jl:45-56 */
int32_t fomo_inject_source(fomo_model* m, fomo_wavefield* w, fomo_shot* s,
                           int32_t k);                                      /* source_receiver.jl:18-30 */
int32_t fomo_update_velocity(fomo_model* m, fomo_wavefield* w);             /* velocity.jl:21-46 */
int32_t fomo_apply_habc_velocity(fomo_model* m, fomo_wavefield* w);         /* boundary.jl:139-144 */
int32_t fomo_update_stress(fomo_model* m, fomo_wavefield* w);               /* stress.jl:18-47 */
int32_t fomo_apply_habc_stress(fomo_model* m, fomo_wavefield* w);           /* boundary.jl:151-157 */
int32_t fomo_apply_free_surface(fomo_model* m, fomo_wavefield* w);          /* boundary.jl:168-181 */
int32_t fomo_record_receivers(fomo_model* m, fomo_wavefield* w, fomo_shot* s,
                              int32_t k);                                   /* source_receiver.jl:65-77 */

/* ---- batches: many shots (uneven grids allowed) in one launch per kernel ---
 * Replaces the sequential shot loop of run_shots! (shot_manager.jl:145-180):
 * shot b runs on wavefields[b] with models[b]; every model must have the same
 * fd_order.  Each step of every shot is one launch per kernel for the whole
 * batch (a work list over all shots' tiles). */
int32_t fomo_batch_create(fomo_ctx* ctx, int32_t n, fomo_model* const* models,
                          fomo_wavefield* const* wavefields, fomo_shot* const* shots,
                          fomo_batch** out);
int32_t fomo_batch_destroy(fomo_batch* b);
/* Steps k_first .. k_first+n_steps-1 for every shot of the batch; a shot
 * steps while k <= its nt and is skipped after that (its wavefield keeps the
 * state of its own last step).  The shots that step in a run must be at the
 * same buffer parity (FOMO_ERR_STATE otherwise): true for every batch whose
 * runs stay within every shot's nt.  In the whole-grid step a batch of two
 * or more shots runs as two halves on two streams (the context's stream and
 * its second stream), joined before the call's work ends on the context's
 * stream. */
int32_t fomo_batch_run(fomo_batch* b, int32_t k_first, int32_t n_steps);

/* ---- pools: continuous batching for a serving loop ----------------------
 * No counterpart in the reference.  A pool runs the fused step for a changing
 * set of shots (members) of one fd_order and storage type, uneven grids
 * allowed, up to `capacity` at once.  A member joins with fomo_pool_add
 * (it starts at step 1 on the next run; reset its wavefield and shot first),
 * advances with every fomo_pool_run, and leaves with fomo_pool_remove
 * (which waits for the stream).  Members past their nt are skipped until
 * removed.  Members need dtx == dtz.  Runs are asynchronous and take an even
 * number of steps; one captured graph per pool serves every membership.
 * Each pool has its own stream, so the runs of different pools (other FD
 * orders or storage types) proceed side by side on the GPU; wait for a pool
 * with fomo_pool_synchronize (fomo_ctx_synchronize does not).
 * Ordering and threads: a joining member's zeroing and earlier work, queued
 * on the context's stream, complete before its first pool step.  Pool calls
 * hold the library lock that every allocating or freeing call takes, so
 * other threads may create and destroy objects meanwhile; a model, wavefield
 * or shot that is in a pool cannot be destroyed (FOMO_ERR_STATE) until it is
 * removed or the pool is destroyed. */
typedef struct fomo_pool fomo_pool;
int32_t fomo_pool_create(fomo_ctx* ctx, int32_t fd_order, int32_t storage, int32_t capacity,
                         fomo_pool** out);
int32_t fomo_pool_destroy(fomo_pool* p);
int32_t fomo_pool_add(fomo_pool* p, fomo_model* m, fomo_wavefield* w, fomo_shot* s);
int32_t fomo_pool_remove(fomo_pool* p, fomo_shot* s);
int32_t fomo_pool_run(fomo_pool* p, int32_t n_steps);
/* Number of members. */
int32_t fomo_pool_size(const fomo_pool* p);
/* Waits until the pool's queued runs are done. */
int32_t fomo_pool_synchronize(fomo_pool* p);
/* The member's next step (1-based): it is done when this exceeds its nt. */
int32_t fomo_pool_next_step(const fomo_pool* p, const fomo_shot* s, int32_t* k_next);

/* ---- tuning ------------------------------------------------------------- */
/* Step implementation for runs created after the call on this context:
 * 1 (default) = fused: velocity and stress in one pass over ping-pong state
 *   buffers (the read buffer doubles as the strip backup), then the velocity
 *   HABC, the stress cells beside the HABC strips, the stress HABC, free
 *   surface and receivers;
 * 0 = two-pass in place: velocity pass (with strip backup and source), HABC,
 *   stress pass, HABC, free surface and receivers;
 * 2 = split fused: the fused pass over the interior tiles (whose cells lie
 *   2M + 2 cells inside the band frame the strip kernels write) runs on the
 *   context's stream while, on a second stream of the context, the fused pass
 *   over the tiles near the edges and then the HABC, band stress and free
 *   surface run; the two are joined before the next step.  Batches only
 *   (pools run mode 1);
 * 3 = double: pairs of steps.  One pass over the interior tiles advances
 *   them two steps (the first step's values stay on chip), while beside it
 *   two split-step chains advance the rest: step k outside the interior
 *   shrunk by 2M into a third state buffer, then step k + 1 outside the
 *   interior.  FP32 storage with dtx == dtz, fd_order 4 or 8,
 *   nbc >= fd_order / 2, and no receiver inside the shrunk interior; other
 *   runs, runs with a frame store whose frame steps are not all second
 *   steps of pairs (an odd skip, or an even first step), and an odd last
 *   step take mode 2.  The third buffer (5 fields) is
 *   allocated with the first such run and counted in
 *   fomo_wavefield_device_bytes.
 * All give identical results. */
int32_t fomo_ctx_set_step_mode(fomo_ctx* ctx, int32_t mode);
/* Split step shape (mode 2): rows per edge tile (1..256, default 16) and the
 * most interior blocks per compute unit (0 = as many as fit), which leaves
 * room on the GPU for the strip kernels that run beside them.  Applies to
 * runs created after the call.  No counterpart in the reference. */
int32_t fomo_ctx_set_split(fomo_ctx* ctx, int32_t edge_rows, int32_t interior_blocks_per_cu);
/* Stencil tile: 64*cols_per_lane columns (cols_per_lane in {1,2,4}) by
 * rows_per_tile rows per wavefront (0 = chosen so that the tiles fill the
 * GPU's resident wavefronts in whole rounds); steps_per_graph = whole steps
 * per captured graph (even, 2..256), or 0 to launch every step's kernels
 * directly, without graphs.  Applies to runs created after the call on
 * this context. */
int32_t fomo_ctx_set_tiling(fomo_ctx* ctx, int32_t cols_per_lane, int32_t rows_per_tile,
                            int32_t steps_per_graph);
/* Queue `reps` launches of one kernel for (m, w, s) on the context's stream,
 * for timing: 0 = velocity pass, 1 = stress pass (two-pass step, with strip
 * backup and injection), 2 = HABC velocity, 3 = HABC stress (in place),
 * 4 = free surface + record, 5 = fused velocity+stress pass, 6 = band
 * stress, 7 = HABC velocity, 8 = HABC stress (fused-step buffer roles),
 * 9 = fused pass over the split step's edge tiles, 10 = over its interior
 * tiles (mode 2's geometry); for a double-step run (mode 3): 11 = the
 * two-step pass over the interior (steps k and k + 1), 12 = the step-k edge
 * tiles, 13 = the step-(k+1) edge tiles. */
int32_t fomo_launch_stage(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int32_t stage,
                          int32_t k, int32_t reps);

#ifdef __cplusplus
}
#endif
#endif /* FOMO_MI355X_H */
