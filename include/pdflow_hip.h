/* C-ABI of the MI355X PD-Flow library (libpdflow_hip.so).
 *
 * This is the boundary the reference's host code crosses: every entry point replaces one
 * CSF_cuda host method or one "bridge" free function of pdflow_cudalib.h:125-174, keeping
 * its meaning and its place in the call sequence of scene_flow_impair.cpp:80-168.  The
 * C++ drop-in class (include/pdflow_cudalib.h) is a thin wrapper over these functions.
 *
 * Conventions (the reference returns void and ignores cudaError_t; here):
 *   - every function returns PDF_OK (0) or a PDF_ERR_* code; pdf_last_error() returns
 *     the message of the last failure on that context;
 *   - all planes are FP32, column-major, element (v, u) at v + u*rows (scene_flow_impair.cpp:232);
 *   - host-pointer functions are blocking, like the reference's cudaMemcpy; device-pointer
 *     functions are asynchronous on the context's stream;
 *   - stage functions enqueue kernels on the context's stream and return without waiting.
 * No torch, HIP or CUDA types appear in these signatures; a stream is passed as void*.
 */
#ifndef PDFLOW_HIP_H
#define PDFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "pdflow_fields.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdf_ctx pdf_ctx;

enum pdf_status {
    PDF_OK = 0,
    PDF_ERR_ARG = 1,      /* bad argument (null pointer, size out of range, bad field id) */
    PDF_ERR_STATE = 2,    /* call out of sequence (e.g. a stage before pdf_allocate_level) */
    PDF_ERR_HIP = 3,      /* a HIP runtime call failed */
    PDF_ERR_NOMEM = 4     /* device allocation failed */
};

/* ---- lifecycle ------------------------------------------------------------------ */

/* CSF_cuda::readParameters + CSF_cuda::allocateDevMemory (pdflow_cudalib.cu:176-193,
 * :29-72).  rows x cols is the finest solve size, the input frames are
 * (640/cam_mode) x (480/cam_mode).  All device memory is allocated here, once; the
 * per-level "allocation" calls below only select sizes. */
int pdf_create(pdf_ctx **out, unsigned rows, unsigned cols, float lambda_i, float lambda_d,
               float mu, const float *g_mask25, unsigned ctf_levels, unsigned cam_mode,
               float fovh, float fovv);

/* CSF_cuda::freeDeviceMemory (pdflow_cudalib.cu:249-288). */
int pdf_destroy(pdf_ctx *ctx);

/* No counterpart in the reference (which never checks an allocation, .cu:36-37): a process-wide
 * cap on the bytes of one arena (a pdf_create solver's, a pdf_batch_create batch's); 0 = none,
 * the default.  A create whose arena would exceed it returns PDF_ERR_NOMEM without any device
 * call, the same status as a failed device allocation.  Arenas that exist are not affected. */
int pdf_set_arena_limit(size_t bytes);

/* Runs the context's work on a caller-owned hipStream_t (NULL = the library's own). */
int pdf_set_stream(pdf_ctx *ctx, void *hip_stream);
int pdf_synchronize(pdf_ctx *ctx);
const char *pdf_last_error(const pdf_ctx *ctx);
/* message of the last pdf_create (or pdf_solve_pair_noalloc) failure on this thread */
const char *pdf_last_create_error(void);

/* ---- frame ingest and pyramid ------------------------------------------------------ */

/* CSF_cuda::copyNewFrames (pdflow_cudalib.cu:195-222): uploads one frame (host planes
 * I, Z of (640/cam_mode) x (480/cam_mode)) and makes the previous frame "old". */
int pdf_copy_new_frames(pdf_ctx *ctx, const float *I, const float *Z);
/* Same, from device-resident planes (asynchronous device-to-device copy). */
int pdf_copy_new_frames_device(pdf_ctx *ctx, const float *dI, const float *dZ);

/* GaussianPyramidBridge (pdflow_cudalib.cu:1172-1184): builds `levels` levels of the
 * "new" pyramid. */
int pdf_gaussian_pyramid(pdf_ctx *ctx, unsigned levels, unsigned cam_mode);

/* ---- per-level solve stages ---------------------------------------------------------- */

/* CSF_cuda::allocateMemoryNewLevel (pdflow_cudalib.cu:74-162): selects the level size
 * and pyramid slot; no allocation happens (the planes are preallocated). */
int pdf_allocate_level(pdf_ctx *ctx, unsigned rows_i, unsigned cols_i, unsigned level,
                       unsigned level_image);
/* CSF_cuda::freeLevelVariables (pdflow_cudalib.cu:290-326): ends the level. */
int pdf_free_level(pdf_ctx *ctx);

int pdf_assign_zeros(pdf_ctx *ctx);        /* AssignZerosBridge    .cu:1186-1189 */
int pdf_upsample(pdf_ctx *ctx);            /* UpsampleBridge       .cu:1191-1195 */
int pdf_rij(pdf_ctx *ctx);                 /* RijBridge            .cu:1207-1210 */
int pdf_image_gradients(pdf_ctx *ctx);     /* ImageGradientsBridge .cu:1197-1200 */
int pdf_warping(pdf_ctx *ctx);             /* WarpingBridge        .cu:1202-1205 */
int pdf_mu_and_step_sizes(pdf_ctx *ctx);   /* MuAndStepSizesBridge .cu:1212-1215 */
int pdf_gradient(pdf_ctx *ctx);            /* GradientBridge       .cu:1232-1235 */
int pdf_dual_variables(pdf_ctx *ctx);      /* DualVariablesBridge  .cu:1217-1220 */
int pdf_divergence(pdf_ctx *ctx);          /* DivergenceBridge     .cu:1227-1230 */
int pdf_primal_variables(pdf_ctx *ctx);    /* PrimalVariablesBridge .cu:1222-1225 */

/* n iterations of the primal-dual loop of scene_flow_impair.cpp:138-144
 * (Gradient, DualVariables, Divergence, PrimalVariables) as one fused kernel per
 * iteration; same results as the four bridges.  Step sizes and gradients are
 * recomputed in registers, so the sigma_*, tau_*, grad* and divp* planes are not
 * written by this call. */
int pdf_pd_iterations(pdf_ctx *ctx, unsigned n);

/* Iteration-kernel selection (no reference counterpart; the reference launches four kernels
 * per iteration): levels of at least min_pixels pixels run k iterations per launch in the
 * temporally blocked kernel (k in {2,3,4,5,6,8,10,12}); k <= 1 disables it.  Results are
 * identical for every setting.  Default: from 60x80 up, 12 iterations per launch on levels of at
 * most 60x80 pixels, 10 up to 120x160 and 6 above (batches: k = 4 on every level,
 * pdf_batch_set_iteration_kernel); calling this function replaces the per-level choice by k. */
int pdf_set_iteration_kernel(pdf_ctx *ctx, int k, unsigned min_pixels);
/* Region shape of the temporally blocked kernel (no reference counterpart): 0 = 64 rows x 32
 * columns per 1024-thread workgroup, 1 = 32 x 32 per 512-thread workgroup (default: two
 * workgroups share a CU, so one's loads overlap the other's arithmetic).  Results are identical
 * for every shape. */
int pdf_set_tb_shape(pdf_ctx *ctx, int shape);

/* Solve-path options (no reference counterpart).  fused_setup: pdf_solve runs each level's
 * set-up in four fused launches (default) instead of the reference's eleven bridges; the
 * fused path leaves the sigma_*, tau_*, *_upsamp, grad* and divp* planes unwritten.
 * level_kernel: levels of at most 2048 pixels run all their iterations in one launch
 * (default on).  Results are identical for every setting. */
int pdf_set_solve_options(pdf_ctx *ctx, int fused_setup, int level_kernel);

int pdf_filter(pdf_ctx *ctx);              /* FilterBridge         .cu:1237-1241 */
int pdf_motion_field(pdf_ctx *ctx);        /* MotionFieldBridge    .cu:1243-1246 */

/* ---- results ------------------------------------------------------------------------ */

/* CSF_cuda::copyMotionField (pdflow_cudalib.cu:239-244): rows_i*cols_i values each. */
int pdf_copy_motion_field(pdf_ctx *ctx, float *dx, float *dy, float *dz);
/* CSF_cuda::copyAllSolutions (pdflow_cudalib.cu:224-237): flow plus the level's
 * pyramid planes of both frames. */
int pdf_copy_all_solutions(pdf_ctx *ctx, float *dx, float *dy, float *dz, float *depth,
                           float *depth_old, float *colour, float *colour_old, float *xx,
                           float *xx_old, float *yy, float *yy_old);
/* device pointers of the current dx, dy, dz planes (valid until the next solve stage) */
int pdf_motion_field_device(pdf_ctx *ctx, const float **dx, const float **dy,
                            const float **dz);

/* ---- whole solve ------------------------------------------------------------------- */

/* The coarse-to-fine loop of scene_flow_impair.cpp:95-168 (PD_flow_opencv::
 * solveSceneFlowGPU) for the current old/new pyramids, with num_max_iter[ctf_levels]
 * iterations per level, coarse first (NULL = the reference schedule: 100 at the finest
 * level, 15 fewer per coarser level).  The finest level's flow is left in dx/dy/dz. */
int pdf_solve(pdf_ctx *ctx, const unsigned *num_max_iter);

/* The solve started at level start_level (coarse-first index) from a warm start instead of the
 * upsampled coarser solution (no reference counterpart; the step before the solve of
 * scene_flow_impair.cpp:95-168).  warm_device holds three consecutive device planes of that
 * level's size, column-major: du and dv in pixels of that level, dw in metres.  The level
 * starts with du_prev, dv_prev = du, dv, dw_new = dw_acc = dw and all duals at 0; finer levels
 * proceed as in pdf_solve.  Coarser levels are skipped. */
int pdf_solve_warm(pdf_ctx *ctx, const unsigned *num_max_iter, unsigned start_level,
                   const float *warm_device);

/* Same as pdf_solve, but replays a HIP graph captured on first use (one graph per
 * schedule and old/new parity). */
int pdf_solve_graph(pdf_ctx *ctx, const unsigned *num_max_iter);

/* The whole pair of scene_flow_impair.cpp:218-277 + :95-168 from device planes:
 * ingest frame 1, pyramid, ingest frame 2, pyramid, solve. */
int pdf_solve_pair_device(pdf_ctx *ctx, const float *dI1, const float *dZ1, const float *dI2,
                          const float *dZ2, const unsigned *num_max_iter, int use_graph);

/* No counterpart in the reference: pdf_solve_pair_device without a context, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Takes pdf_create's parameters (fovv is unused by every kernel), the
 * pair's four input planes and three output planes of rows*cols floats (dx, dy, dz: the
 * finest level's flow), all device memory of the current device, and runs the same kernels as
 * pdf_solve_pair_device on a default context (same results, bit for bit) on hip_stream (a
 * hipStream_t of the current device; NULL = the null stream), asynchronously.  It calls no
 * hipMalloc, hipFree or other allocating routine: its working planes are a device array of
 * 80 MiB that the library holds on every GPU it runs on until the process ends, enough for the
 * largest pair the reference takes (640x480 frames); the first call on a device looks up that
 * array's address.  Every call on one GPU uses the same array, so calls that overlap on one
 * GPU (on different streams or from different threads) race and must be serialised by the
 * caller.  Failures return PDF_ERR_* with the message in pdf_last_create_error(). */
int pdf_solve_pair_noalloc(unsigned rows, unsigned cols, float lambda_i, float lambda_d, float mu,
                           const float *g_mask25, unsigned ctf_levels, unsigned cam_mode, float fovh,
                           const float *dI1, const float *dZ1, const float *dI2, const float *dZ2,
                           const unsigned *num_max_iter, float *dx, float *dy, float *dz,
                           void *hip_stream);

/* Streaming: the live viewer's loop (main_scene_flow_visualization.cpp:114-130,
 * scene_flow_visualization.cpp:112-208): ingest one new frame from device planes, build its
 * pyramid (the previous frame's becomes "old"), and, from the second frame on, solve for the
 * flow from the previous frame to this one.  *solved (optional) tells whether a solve ran. */
int pdf_stream_frame_device(pdf_ctx *ctx, const float *dI, const float *dZ,
                            const unsigned *num_max_iter, int use_graph, int *solved);

/* ---- inspection (no reference counterpart: test and tooling access) ----------------- */

/* rows_i*cols_i values of one plane of the current level (blocking) */
int pdf_get_field(pdf_ctx *ctx, int field, float *host_out);
int pdf_set_field(pdf_ctx *ctx, int field, const float *host_in);
/* one pyramid plane; (640/cam_mode >> level) x (480/cam_mode >> level) values */
int pdf_get_pyramid(pdf_ctx *ctx, int kind, unsigned level, float *host_out);
int pdf_level_dims(const pdf_ctx *ctx, unsigned *rows_i, unsigned *cols_i);
/* tanf(0.5f*fovh) as used by the pyramid and motion-field stages */
float pdf_tan_half_fovh(const pdf_ctx *ctx);

/* Kernel timing of the fused iteration: when enabled, every pdf_pd_iterations launch
 * is bracketed by HIP events on the context's stream; pdf_kernel_timing returns the
 * number of timed launches, their total time (ms), the pixel-iterations they did and the pixels
 * of the levels they covered (a launch of k iterations over P pixels counts k*P and P: each
 * launch streams its level's planes through HBM once, whatever k).  Outputs may be NULL. */
int pdf_enable_kernel_timing(pdf_ctx *ctx, int enable);
int pdf_kernel_timing(pdf_ctx *ctx, uint64_t *launches, double *total_ms,
                      uint64_t *pixel_iters, uint64_t *pixels);

/* ---- batched ragged solve (no reference counterpart: the reference solves one pair per
 * CSF_cuda object) -------------------------------------------------------------------- */

typedef struct pdf_batch pdf_batch;

/* A batch of npairs independent pairs.  Pair i is solved at rows[i] x rows[i]*4/3 from input
 * frames of (640/cam_mode[i]) x (480/cam_mode[i]) (cam_mode NULL: 480/rows[i], the input at
 * the solve size), with the reference schedule for its level count (num_max_iter NULL) or
 * num_max_iter[i*6 + level], coarse first (at most 100000 per level).  lambda_i, lambda_d, mu, the mask and fovh are
 * shared (readParameters, pdflow_cudalib.cu:176-193).  All memory is allocated here. */
int pdf_batch_create(pdf_batch **out, unsigned npairs, const unsigned *rows, const unsigned *cam_mode,
                     float lambda_i, float lambda_d, float mu, const float *g_mask25, float fovh,
                     const unsigned *num_max_iter);
int pdf_batch_destroy(pdf_batch *b);
int pdf_batch_size(const pdf_batch *b, unsigned *npairs);
const char *pdf_batch_last_error(const pdf_batch *b);   /* NULL b: last create failure */
/* iteration kernels for the batch (as pdf_set_iteration_kernel; level_kernel as in
 * pdf_set_solve_options); invalidates the recorded plan and graph */
int pdf_batch_set_iteration_kernel(pdf_batch *b, int k, unsigned min_pixels, int level_kernel);
/* region shape of the temporally blocked kernel for the batch (as pdf_set_tb_shape) */
int pdf_batch_set_tb_shape(pdf_batch *b, int shape);
/* Autotuner: times the batch's own solve (one captured solve, then two graph replays between HIP
 * events) for each region shape and iterations per launch K in {3, 4, 5, 6} with the batch's
 * set-up tile, then each set-up tile (pdf_batch_set_setup_tile) with the fastest of those; keeps
 * the fastest configuration and returns its shape, K and time (any output may be NULL; *ms = one
 * solve; pdf_batch_setup_tile gives the tile).  Results are identical for every choice.  With
 * cache_path, a text file of "<key>\t<shape> <K> <ms> <tile>" lines keyed by device and batch
 * composition: a hit applies the cached choice without timing (a line without the tile applies the
 * default tile), a miss appends the new line; lines that do not parse to a supported shape, K and
 * tile are skipped.  A cache_path that exists but is not a readable regular file (a directory, no
 * read permission) applies the defaults (32 x 32, K = 4, tile 32 x 8) without timing and returns
 * PDF_OK with *ms = 0.  The batch's inputs must be set (it is solved several times). */
int pdf_batch_autotune(pdf_batch *b, const char *cache_path, int *shape, int *k, float *ms);
/* No counterpart in the reference (its kernels have one launch shape, pdflow_cudalib.h:28-29): the
 * tile of the batch's set-up stencils that stage a neighbourhood through LDS, the level-init
 * upsampling and the median filter.  Rows x columns, one pixel per thread: 0 = 32x8 (default),
 * 1 = 16x16, 2 = 8x32 (4 waves), 3 = 32x4, 4 = 16x8, 5 = 8x16 (2 waves), 6 = 8x8 (1 wave).  Results
 * do not depend on it; changing it invalidates the recorded plan. */
int pdf_batch_set_setup_tile(pdf_batch *b, int tile);
int pdf_batch_setup_tile(const pdf_batch *b, int *tile);
/* device pointers of pair's four input planes (the caller may fill them directly).  They are
 * also level 0 of the pair's colour and depth pyramids, so new inputs take effect there at once:
 * a continuation or re-solve after new inputs would mix them with the pyramid levels the last
 * solve built.  Set new inputs, then solve. */
int pdf_batch_inputs_device(pdf_batch *b, unsigned pair, float **I1, float **Z1, float **I2,
                            float **Z2);
/* Warm start of one pair (as pdf_solve_warm): the pair's solve starts at start_level (0 =
 * normal coarse-to-fine) from the three planes (du, dv, dw at that level, column-major)
 * returned in *warm_device, which the caller fills before pdf_batch_solve.  Changing the
 * start level invalidates the recorded plan. */
int pdf_batch_set_start_level(pdf_batch *b, unsigned pair, unsigned start_level, float **warm_device);
/* per-pair iterations per level, coarse first, at most 100000 each (invalidates the plan when
 * they change) */
int pdf_batch_set_schedule(pdf_batch *b, unsigned pair, const unsigned *num_max_iter);
/* blocking upload of one pair's frames from host planes (into the planes of
 * pdf_batch_inputs_device) */
int pdf_batch_set_inputs(pdf_batch *b, unsigned pair, const float *I1, const float *Z1,
                         const float *I2, const float *Z2);
/* solves every pair (two pyramids + coarse-to-fine solve each); asynchronous on the batch's
 * stream; use_graph replays a HIP graph captured on first use */
int pdf_batch_solve(pdf_batch *b, int use_graph);
int pdf_batch_synchronize(pdf_batch *b);
/* No counterpart in the reference, which runs one frame pair at a time: orders two batches'
 * streams without a host wait.  Work queued on b after this call starts only once everything
 * queued on other before it has finished (an event recorded on other's stream, which b's stream
 * waits on).  With two batches in flight, queue the next batch's solve after waiting on the
 * current batch right behind its solve: the solves then run one after the other at full width,
 * and only the current batch's rule, continuation and re-solves overlap the next solve.  b ==
 * other does nothing. */
int pdf_batch_wait(pdf_batch *b, pdf_batch *other);
/* pair's finest-level flow planes (rows x cols, column-major) on the device / copied out */
int pdf_batch_flow_device(pdf_batch *b, unsigned pair, const float **dx, const float **dy,
                          const float **dz, unsigned *rows, unsigned *cols);
int pdf_batch_copy_flow(pdf_batch *b, unsigned pair, float *dx, float *dy, float *dz);
/* No counterpart in the reference (its copyNewFrames, pdflow_cudalib.cu:195-222, copies one frame
 * and blocks): queues the copy of every pair's four input planes from host memory on the batch's
 * stream and returns at once.  host holds, pair after pair in batch order, I1, Z1, I2, Z2 of
 * rows x cols floats each (column-major); host_floats must equal that total.  The host memory must
 * stay unchanged until the stream has passed the copy (pdf_batch_synchronize, or any call that
 * waits on the stream); from page-locked memory (pdf_host_alloc) the copy overlaps kernels on
 * other streams.  Queue it after the previous solve's results are used and before the next solve,
 * as for pdf_batch_inputs_device. */
int pdf_batch_upload_inputs_async(pdf_batch *b, const float *host, size_t host_floats);
/* No counterpart in the reference (its copyMotionField, pdflow_cudalib.cu:239-244, blocks): queues
 * the copy of every pair's finest-level flow, dx, dy, dz of rows x cols floats each, pair after
 * pair, into host memory on the batch's stream and returns at once; the values are there once the
 * stream has passed the copy. */
int pdf_batch_download_flows_async(pdf_batch *b, float *host, size_t host_floats);
/* number of kernel launches one batched solve makes */
int pdf_batch_launch_count(const pdf_batch *b, unsigned *launches);
/* HIP-event timing of the iteration launches, as pdf_enable_kernel_timing (solves run
 * eagerly while enabled); enable = 2 times only the temporally blocked launches */
int pdf_batch_enable_kernel_timing(pdf_batch *b, int enable);
int pdf_batch_kernel_timing(pdf_batch *b, uint64_t *launches, double *total_ms, uint64_t *pixel_iters,
                            uint64_t *pixels);
/* planes of pair's last solve on the device: a field of include/pdflow_fields.h (those the
 * batched path keeps, at the finest level; the dual carry planes pd_l and p*_l hold the last
 * coarser level's duals, since nothing reads them after the finest level) or a pyramid plane
 * (kind, level) */
int pdf_batch_field_device(pdf_batch *b, unsigned pair, int field, const float **ptr);
int pdf_batch_pyramid_device(pdf_batch *b, unsigned pair, int kind, unsigned level, const float **ptr,
                             unsigned *rows, unsigned *cols);

/* ---- learned warm start and its fallback (include/pdflow_cnn.h; no reference
 * counterpart) ------------------------------------------------------------------------ */

struct pdf_net;
/* The listed pairs (all if pairs is NULL) start at their level of the network's size (h x w
 * of pdf_net_create), initialised from one batched forward pass of the network over those
 * levels' I1, Z1, I2, Z2 inside the solve (same graph); each keeps its own schedule
 * (pdf_batch_set_schedule).  net NULL switches every pair back to the plain start.  The
 * network must outlive the batch's use of it and hold at least count images. */
int pdf_batch_set_learned(pdf_batch *b, struct pdf_net *net, const unsigned *pairs, unsigned count);
/* Per pair of the last solve: the finest level's energy on the last primal iterate,
 *   E = sum |dct + dcu du + dcv dv| + mu_uv |ddt + ddu du + ddv dv - dw|
 *       + lambda_i (|grad_r(du+du_prev)| + |grad_r(dv+dv_prev)|) + lambda_d |grad_r dw|
 * (the terms of the prox .cu:906-923, the depth dual .cu:836-846, the TV duals .cu:849-885
 * with grad_r of .cu:982-1013), and the warped-intensity residual: mean |I2(x + u) - I1(x)|
 * over pixels with frame-1 depth, u the final flow (bilinear warp of .cu:654-682).
 * Either output may be NULL. */
int pdf_batch_metrics(pdf_batch *b, float *energy, float *residual);
/* No counterpart in the reference: pdf_batch_metrics for the listed pairs only; energy[j] and
 * residual[j] belong to pairs[j] (for re-checking the pairs a continuation touched) */
int pdf_batch_metrics_pairs(pdf_batch *b, const unsigned *pairs, unsigned count, float *energy, float *residual);
/* No counterpart in the reference: per listed pair (all pairs if pairs is NULL, in pair order),
 * pdf_batch_metrics' energy and residual and the zero-motion residual, mean |I2(x) - I1(x)| over
 * the same pixels with frame-1 depth: the intensity change a flow has to explain.  Any output may
 * be NULL. */
int pdf_batch_quality(pdf_batch *b, const unsigned *pairs, unsigned count, float *energy, float *residual,
                      float *residual0);
/* No counterpart in the reference: the learned path's acceptance rule that needs no plain solve.
 * Every pair whose finest-level energy exceeds energy_per_pixel_max times its pixel count, or whose
 * residual exceeds residual_ratio_max times its zero-motion residual (pdf_batch_quality), is
 * re-solved by the plain path (pdf_batch_resolve_plain); NaN values count as exceeding.  The bounds
 * are calibrated on the plain solver's results over training pairs (pdflow/calibrate.py).
 * flags[i] = 1 for the re-solved pairs (flags may be NULL), *count their number. */
int pdf_batch_fallback_bounds(pdf_batch *b, float energy_per_pixel_max, float residual_ratio_max,
                              unsigned char *flags, unsigned *count);
/* per pair: mean 3-D end-point error (metres) between this batch's flow and ref's over pixels
 * with frame-1 depth (pairs of equal sizes) */
int pdf_batch_epe(pdf_batch *b, const pdf_batch *ref, float *epe);
/* re-solves the listed pairs in place with the plain path: level 0, the reference schedule for
 * the pair's level count (whatever pdf_batch_set_schedule set), the pyramids already built */
int pdf_batch_resolve_plain(pdf_batch *b, const unsigned *pairs, unsigned count);
/* No counterpart in the reference: runs extra_iters more primal-dual iterations of the listed
 * pairs (each at most once) at their finest level, continuing from the state the last solve,
 * continuation or re-solve left, then that level's median filter and motion field again.  The
 * flow equals that of a solve whose finest level ran its iterations plus extra_iters.  For the
 * learned path's pairs that miss the fallback rule by little (an adaptive iteration budget). */
int pdf_batch_continue(pdf_batch *b, const unsigned *pairs, unsigned count, unsigned extra_iters);
/* fallback rule: every pair whose energy or residual exceeds (1 + tol) times the reference
 * value (e.g. the plain solver's, from pdf_batch_metrics of a plain batch) is re-solved by
 * the plain path; flags[i] = 1 for those (flags may be NULL), *count their number */
int pdf_batch_fallback(pdf_batch *b, const float *energy_ref, const float *residual_ref, float tol,
                       unsigned char *flags, unsigned *count);
/* training samples from a solved batch: for every pair, its h x w level's network input
 * (x_dev + pair*h*w*4, NHWC as the network takes it), the label (t_dev + pair*h*w*3: the
 * final flow averaged over each pixel's block, du, dv in pixels of that level, 10*dw) and
 * the loss mask (m_dev + pair*h*w) */
int pdf_batch_learn_samples(pdf_batch *b, unsigned h, unsigned w, float *x_dev, float *t_dev, float *m_dev);

/* ---- device memory helpers for callers without a HIP binding ------------------------ */
int pdf_device_alloc(void **ptr, size_t bytes);
int pdf_device_free(void *ptr);
int pdf_memcpy_h2d(void *dst, const void *src, size_t bytes);
int pdf_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* page-locked host memory (hipHostMalloc), for copies that overlap kernels */
int pdf_host_alloc(void **ptr, size_t bytes);
int pdf_host_free(void *ptr);
/* blocking device-to-device copy by a plain streaming kernel (bytes a multiple of 16, 16-byte
 * aligned pointers): every byte read once and written once, the calibration point for HBM byte
 * counters */
int pdf_device_copy(void *dst, const void *src, size_t bytes);

/* No counterpart in the reference: checks the iteration kernels' short exact arithmetic
 * sequences against the compiler's IEEE operations on every float whose bit pattern lies in
 * [lo, hi] (blocking).  which = 0: square root (valid for 1 <= x <= +inf), 1: reciprocal
 * (1 <= x < 2^64), 2: reciprocal square root (x >= 1), 3: 1/x and 4: 0.5/x (2^-94 <= x <= 2^124).  *mismatches = inputs whose results
 * differ in any bit, *first_mismatch = the smallest such bit pattern (0xffffffff if none). */
int pdf_check_exact_sequence(int which, uint32_t lo, uint32_t hi, uint64_t *mismatches, uint32_t *first_mismatch);
int pdf_device_count(int *count);

#ifdef __cplusplus
}
#endif
#endif
