/*
 * relion_coarse.h -- C ABI of the MI355X (gfx950) coarse diff2 search.
 *
 * Drop-in for RELION's coarse-pass squared-difference call:
 *
 *   diff2[o*T + t] (+)= sum_p corr_p/2 * | P_o(p) - X_p * exp(i*(x_p*tx_t + y_p*ty_t)) |^2
 *
 * where P_o(p) is the trilinear central-slice sample of a padded
 * half-complex Fourier volume at the pixel rotated by orientation o.
 *
 * Every pointer named d_* / device below is device memory owned by the
 * caller; host pointers are named host_*.  All calls are stream-ordered on
 * the hipStream_t passed as `stream` (NULL = default stream), never
 * synchronise the host, and return an rc_status; rc_last_error() gives the
 * message of the last failure on the calling thread.  Nothing here aborts
 * or raises a signal.
 */
#ifndef RELION_COARSE_H
#define RELION_COARSE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RC_OK = 0,
    RC_ERR_INVALID = 1,      /* bad argument or inconsistent dimensions      */
    RC_ERR_UNSUPPORTED = 2,  /* a mode this build does not implement          */
    RC_ERR_NO_MEMORY = 3,    /* device allocation failed                      */
    RC_ERR_DEVICE = 4,       /* HIP runtime / launch failure                  */
    RC_ERR_LIMIT = 5,        /* size beyond a documented limit                */
    RC_ERR_IO = 6            /* file missing or truncated                     */
} rc_status;

/* Arithmetic of the cross term Re(P . conj(w X e^{i phi})). */
typedef enum {
    RC_FORM_FP32 = 0,        /* FP32 MFMA (v_mfma_f32_16x16x4_f32), FP32 accumulate       */
    RC_FORM_BF16X3 = 1,      /* BF16 hi/lo split, 3 MFMAs per product, FP32 accumulate    */
    RC_FORM_FP16 = 2,        /* FP16 with per-image / per-volume power-of-two max-abs
                                scaling, 1 MFMA per product, FP32 accumulate              */
    RC_FORM_FP32_DIRECT = 3  /* FP32 vector ALU, direct |P - Xs|^2 form (no expansion)   */
} rc_form;

/* ---------------------------------------------------------------------- */
/* Projector: device-resident padded Fourier volume                        */
/* ---------------------------------------------------------------------- */
typedef struct rc_projector rc_projector;

/* Replaces AccProjector::setMdlDim + AccProjector::initMdl
 * (acc_projector_impl.h:10-118, 120-216).  host_real/host_imag hold
 * xdim*ydim*zdim floats, x fastest then y then z; zdim == 1 means 2D and is
 * rejected (RC_ERR_UNSUPPORTED): only 3D references are on this path.
 * The volume is copied to the current device with the calling thread's
 * device; the call synchronises on that copy. */
int rc_projector_create(int xdim, int ydim, int zdim, int inity, int initz, int maxr,
                        float padding_factor, const float *host_real, const float *host_imag,
                        rc_projector **out);

/* Replaces AccProjector::constructFromFile (acc_projector_impl.h:300-356):
 * int32 xdim, ydim, zdim, inity, initz, maxr; float32 padding; then the
 * real and the imaginary planes.  Missing/short file -> RC_ERR_IO. */
int rc_projector_create_from_file(const char *path, rc_projector **out);

/* Replaces AccProjector::clear / ~AccProjector (acc_projector_impl.h:246-297). */
int rc_projector_destroy(rc_projector *p);

/* Volume dimensions of a projector (any pointer may be NULL). */
int rc_projector_dims(const rc_projector *p, int *xdim, int *ydim, int *zdim, int *inity,
                      int *initz, int *maxr, float *padding_factor);

/* By-value view for one image geometry.  Replaces
 * AccProjectorKernel::makeKernel (acc_projectorkernel_impl.h:410-433):
 * maxR = min(model maxr, imgMaxR), maxR2_padded = (int)(maxR^2 * pad^2). */
typedef struct {
    const rc_projector *proj;
    int imgX, imgY, imgZ;
    int maxR, maxR2_padded;
    float padding_factor;
} rc_projector_kernel;

int rc_projector_make_kernel(const rc_projector *p, int imgX, int imgY, int imgZ, int imgMaxR,
                             rc_projector_kernel *out);

/* ---------------------------------------------------------------------- */
/* Single-image entry points (one particle, one class)                     */
/* ---------------------------------------------------------------------- */

/* Replaces runDiff2KernelCoarse (acc_helper_functions_impl.h:15-31).
 * Unfiltered image: image_size pixels whose (x, y) follow from the pixel
 * index (x = i % imgX, y = i / imgX, y -= imgY if y > maxR; diff2.cuh:182-188).
 * d_diff2s[o*T+t] is accumulated (+=).  trans_z and local_sqrtXi2 are
 * accepted and unused, as in the reference's 2D-data path.  do_CC != 0
 * writes the cross-correlation score -sum corr Re(P conj Xs) /
 * sqrt(sum corr |P|^2) instead (diff2.cuh:691-809, overwriting, with that
 * kernel's pixel-coordinate rule).  data_is_3D != 0 -> RC_ERR_UNSUPPORTED.
 *
 * Pixels: the search runs over the pixels that a rotation can bring inside
 * the projector's mask, (x^2 + y^2) pad^2 <= (maxR2_padded + 1)(1 + 2^-12)
 * in float32 (the coordinates of the pixel rule above), moved first; every
 * other pixel projects to 0 for every rotation matrix, so it adds only
 * corr/2 |X|^2 (diff2) or nothing (CC), which the call adds directly.  This
 * holds for orthonormal eulers (rotations), as RELION's are.
 *
 * Work split: the pixel axis is cut into slices (rc_coarse_split) so that
 * one image fills the device; each slice writes float32 partial sums and a
 * second kernel adds the slices in a fixed order in float64.  The slice count
 * depends on the shape and the image geometry only, so a shape always gives
 * bit-identical results.
 *
 * Scratch: the reference signature has no workspace, so the call keeps one
 * scratch buffer per (device, stream) -- calls on different streams (e.g.
 * RELION's host threads, each with its own stream) never share one, and
 * calls on one stream are ordered by it.  The first call on a stream, and a
 * call needing more than that stream has used before, allocates (and frees
 * the smaller buffer after synchronising that stream): such calls must not
 * be captured into a hipGraph.  All other calls neither allocate nor
 * synchronise.  rc_scratch_release frees a stream's buffer.
 *
 * Projection cache: the scratch of a stream also keeps the projections of
 * the last call (orientation_num x image_size complex floats, at most 1 GiB
 * per stream; larger searches gather every call).  A call with the same
 * projector, eulers pointer, sizes and image geometry reuses them after a
 * device-side check that the eulers' contents are bit-identical to those
 * the projections were made from (no host synchronisation); anything else
 * makes them again.  The results are bit-identical either way. */
int rc_run_diff2_kernel_coarse(const rc_projector_kernel *projector,
                               const float *d_trans_x, const float *d_trans_y,
                               const float *d_trans_z, const float *d_corr_img,
                               const float *d_fimg_real, const float *d_fimg_imag,
                               const float *d_eulers, float *d_diff2s, float local_sqrtXi2,
                               unsigned long orientation_num, unsigned long translation_num,
                               unsigned long image_size, void *stream, int do_CC,
                               int data_is_3D);

/* rc_run_diff2_kernel_coarse without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops).  Same arguments, semantics, kernels
 * and results (bit-identical) as rc_run_diff2_kernel_coarse; its scratch is
 * a 24 MiB device array inside the library (one per GPU, held until the
 * process ends), so shapes whose workspace (rc_coarse_workspace_bytes with
 * the FP32 form) exceeds 24 MiB are rejected with RC_ERR_LIMIT before any
 * launch.  Calls that overlap on one GPU -- any two streams, or host threads
 * -- share that array and race: serialise them. */
int rc_run_diff2_kernel_coarse_noalloc(const rc_projector_kernel *projector,
                                       const float *d_trans_x, const float *d_trans_y,
                                       const float *d_trans_z, const float *d_corr_img,
                                       const float *d_fimg_real, const float *d_fimg_imag,
                                       const float *d_eulers, float *d_diff2s,
                                       float local_sqrtXi2, unsigned long orientation_num,
                                       unsigned long translation_num, unsigned long image_size,
                                       void *stream, int do_CC, int data_is_3D);

/* Frees the scratch rc_run_diff2_kernel_coarse keeps for `stream` on the
 * current device (synchronises that stream first).  No counterpart in the
 * reference, whose kernels keep no state. */
int rc_scratch_release(void *stream);

/* Replaces CoarseMatrixKernel*::get_workspace_size_bytes
 * (coarse_splitImg_BCF_PMO_TMDB_NFUSE.cuh:69-88): image staging plus the
 * split-K partial sums (ksplit * orientation_num * (translation_num + 1)
 * floats when rc_coarse_split gives ksplit > 1). */
size_t rc_coarse_workspace_bytes(int translation_num, int orientation_num, int image_size,
                                 int full_image_size, int form);

/* Slices of the pixel axis the single-image calls use for a shape (the
 * counterpart of the split-K count of coarse_scheduler.cuh:211-213, here
 * min(chunks, max(1, floor(512 / tiles))) bounded by 32 MiB of partials, with no empty
 * slice): ksplit slices of kc_per_split 32-pixel chunks.  Diagnostic. */
int rc_coarse_split(int translation_num, int orientation_num, int image_size, int form,
                    int *ksplit, int *kc_per_split);

/* Replaces CoarseMatrixKernel*::run (coarse_splitImg_BCF_PMO_TMDB_NFUSE.cuh:90-131).
 * d_real/d_imag/d_corr: the first image_size pixels are inside the
 * projector radius; pixels image_size..full_image_size-1 (when
 * d_coor_xy != NULL and full_image_size > image_size) only add
 * corr/2*|X|^2.  d_coor_xy: interleaved float (x, y) per pixel, or NULL to
 * derive coordinates from the pixel index.
 * d_diff2s_dest[o*T+t] = (d_diff2s ? d_diff2s[o*T+t] : 0) + score; the two
 * pointers may alias (in-place accumulation, as the reference harness uses).
 * The workspace needs no reset between calls. */
int rc_coarse_run(int translation_num, int orientation_num, int image_size, int full_image_size,
                  int form, const float *d_eulers, const float *d_trans_x,
                  const float *d_trans_y, const float *d_real, const float *d_imag,
                  const rc_projector_kernel *projector, const float *d_corr,
                  const float *d_diff2s, float *d_diff2s_dest, void *stream,
                  const float *d_coor_xy, void *d_workspace, size_t workspace_bytes);

/* ---------------------------------------------------------------------- */
/* In-radius partition (FilterCoarseParamImage, main.cu:88-151)            */
/* ---------------------------------------------------------------------- */
/* Pixel i of an unfiltered half-plane image has x = i % imgX, y = i / imgX,
 * y -= imgY if y > maxR (diff2.cuh:182-188), and is inside when the float
 * (x*x + y*y) * pad * pad <= maxR2_padded (the view's geometry).  The
 * filtered order keeps inside pixels first in index order and stores the
 * outside pixels from the last position down.
 *
 * Host form: host_order[j] = source pixel of filtered position j (or NULL),
 * *n_inside = number of inside pixels (the `npix` of a batch).  Only the
 * view's geometry is read (its projector may be NULL). */
int rc_filter_coarse_order(const rc_projector_kernel *view, int image_size, int *host_order,
                           int *n_inside);

/* Device form for a batch, stream-ordered, no allocation: image b reads
 * d_fimg[b][2*image_size] (real block, then imaginary block, as the
 * reference's fimg) and d_corr[b][image_size]; it writes d_real, d_imag,
 * d_corr_out [b][image_size] in the filtered order, d_coor_xy[2*image_size]
 * (x, y) of each filtered position once (or NULL), and d_filtered_size[b]
 * (or NULL).  The outputs are the d_real / d_imag / d_corr / d_coor_xy of an
 * rc_batch_desc with full_npix = image_size and npix = *n_inside. */
int rc_filter_coarse_images(const rc_projector_kernel *view, int image_size, int n_images,
                            const float *d_fimg, const float *d_corr, float *d_real,
                            float *d_imag, float *d_corr_out, float *d_coor_xy,
                            int *d_filtered_size, void *stream);

/* ---------------------------------------------------------------------- */
/* Batched entry points (no reference counterpart: the reference calls the */
/* single-image form once per particle)                                    */
/* ---------------------------------------------------------------------- */
typedef struct {
    int n_particles;
    int n_classes;
    int orientation_num;        /* O, shared by all classes                      */
    int npix;                   /* in-radius pixels (projected) per image        */
    int full_npix;              /* total pixels per image (>= npix)              */
    int form;                   /* rc_form (not RC_FORM_FP32_DIRECT)             */
    const rc_projector_kernel *projectors;   /* host array [n_classes]          */
    const int *host_trans_offset;            /* host [n_particles+1], CSR       */
    const int *d_trans_offset;               /* device copy of the same         */
    const float *d_trans_x, *d_trans_y;      /* device [trans_offset[P]]         */
    const float *d_eulers;                   /* device [O*9]                     */
    const float *d_real, *d_imag, *d_corr;   /* device [n_particles*full_npix]   */
    const float *d_coor_xy;                  /* device [full_npix*2] (x, y)      */
    const float *d_diff2_init;               /* device, same layout as diff2, or NULL */
    float *d_diff2;     /* device out: particle i at n_classes*O*trans_offset[i],
                           then [class][orientation][translation]                */
    double *d_best_diff2;    /* device out [n_particles] (float64) or NULL        */
    int64_t *d_best_index;   /* device out [n_particles] flat (c*O+o)*T_i+t, or NULL */
    int do_cc;               /* 1: cross-correlation score (diff2.cuh:691-809) in
                                d_diff2 (written; d_diff2_init ignored); min = best */
} rc_batch_desc;

typedef struct rc_plan rc_plan;

/* Validate a batch shape, allocate the plan's device tables and workspace
 * (the only allocation of the batched path) and pick tile shapes. */
int rc_plan_create(const rc_batch_desc *desc, rc_plan **out);

/* Run the batch.  Pointers in desc may differ from plan creation; the
 * shape (particles, classes, O, npix, full_npix, form, per-particle
 * translation counts, projector dimensions) must match.  No allocation, no
 * host synchronisation: safe to capture in a hipGraph. */
int rc_plan_execute(rc_plan *plan, const rc_batch_desc *desc, void *stream);

int rc_plan_destroy(rc_plan *plan);

/* Bytes of device memory the plan holds. */
size_t rc_plan_device_bytes(const rc_plan *plan);

/* Tile shape the plan picked: variant id and the (rows x columns) tile. */
int rc_plan_tile(const rc_plan *plan, int *variant, int *tile_m, int *tile_n);

/* Per-launch timing of the cross-term kernel: allocate n_slots pairs of
 * hipEvents recorded on the execute stream around each cross-term launch
 * (slot = launch count mod n_slots); n_slots = 0 turns timing off.
 * rc_plan_timing_read waits for the slots and returns their milliseconds
 * (up to n values, oldest first) and the number written. */
int rc_plan_timing(rc_plan *plan, int n_slots);
int rc_plan_timing_read(rc_plan *plan, float *ms, int n, int *n_written);

/* hipGraph: capture one rc_plan_execute of desc on `stream` into a graph
 * held by the plan (replacing an earlier one), then replay it with
 * rc_plan_replay on any stream.  The captured graph holds the pointers of
 * desc; replays read the buffers' current contents.  Timing slots are not
 * recorded inside the graph. */
int rc_plan_capture(rc_plan *plan, const rc_batch_desc *desc, void *stream);
int rc_plan_replay(rc_plan *plan, void *stream);

/* Autotuner: time the cross-term variants of desc's form on desc's shape
 * (one plan per variant in turn, HIP events on `stream`): a screen of one
 * timed launch per variant, then two sweeps of two untimed and `iters`
 * timed launches over the variants within 25 % of the fastest screen, each
 * keeping its better median.  The fastest is kept in a process-wide table
 * keyed by the shape (particles, classes, orientations, pixels, form, row
 * groups) that later rc_plan_create calls consult, and returned.
 * Allocates and synchronises the stream; not for graph capture. */
int rc_autotune(const rc_batch_desc *desc, void *stream, int iters, int *best_variant,
                float *best_ms);

/* Force the cross-term kernel variant of later rc_plan_create calls
 * (-1 = automatic).  Process-wide; the autotuner does not use it (its
 * sweep names each variant to its own plans). */
int rc_set_variant_override(int variant);

/* Number of cross-term kernel variants and their tile shapes. */
int rc_variant_count(int form);
int rc_variant_info(int form, int variant, int *tile_m, int *tile_n, int *fused_gather);

/* Human-readable description of a variant ("cross_kernel_wide<bf16x3>
 * 256x256 ..."), naming the kernel template that a profiler lists; NULL for
 * an unknown variant.  Diagnostic only, no counterpart in the reference. */
const char *rc_variant_name(int form, int variant);

/* ---------------------------------------------------------------------- */
/* Replay of the reference's coarse-particle dumps                         */
/* ---------------------------------------------------------------------- */
/* Host copy of one `<iter>/coarse_data<d>00_1.dat` file, read as
 * CreateCoarseParamFromFile reads it (main.cu:153-203): int32 xdim, ydim,
 * zdim, rmax, translation_num, orientation_num, image_size; then float32
 * eulers[9*O], trans[2*T] (x block, y block), fimg[2*npix] (real block,
 * imag block), corr[npix], diff2s[O*T].  (The reference harness passes
 * eulers as the translation arrays, main.cu:234-235; trans_x/trans_y here
 * are the file's own translation block.) */
typedef struct {
    int xdim, ydim, zdim, rmax;
    int translation_num, orientation_num, image_size;
    float *eulers;       /* [9*O]  */
    float *trans_x;      /* [T]    */
    float *trans_y;      /* [T]    */
    float *real, *imag;  /* [npix] */
    float *corr;         /* [npix] */
    float *diff2s;       /* [O*T]  initial values */
} rc_coarse_data;

/* Missing or short file -> RC_ERR_IO; the arrays are malloc'd host memory
 * released by rc_coarse_data_free. */
int rc_coarse_data_load(const char *path, rc_coarse_data *out);
void rc_coarse_data_free(rc_coarse_data *d);

/* ---------------------------------------------------------------------- */
/* Serving engine (no reference counterpart): continuous batching of       */
/* coarse-search requests on one GPU                                       */
/* ---------------------------------------------------------------------- */
/* One worker thread owns the device.  Queued requests are admitted in
 * arrival order as soon as one of max_inflight slots is free and the
 * request's device memory (inputs, plan buffers, diff2 output) fits in the
 * HBM budget; each slot runs its request on its own stream, so admitted
 * requests overlap on the device.  Finished requests return their memory at
 * once (or at rc_server_release when keep_diff2 is set).  Projector volumes
 * stay device-resident: requests reference rc_projector objects that the
 * caller keeps alive for the server's lifetime.  Results are deterministic:
 * a request's diff2 is bit-identical alone and in any mix. */
typedef struct rc_server rc_server;

typedef struct {
    size_t hbm_budget_bytes;     /* device arena for admitted requests          */
    size_t pinned_budget_bytes;  /* pinned host staging for queued requests     */
    int max_inflight;            /* requests on the device at once, 1..64       */
    int form;                    /* rc_form of the cross term                    */
} rc_server_config;

typedef struct {
    int n_particles, n_classes, orientation_num, npix, full_npix;
    const rc_projector_kernel *projectors;   /* host [n_classes]              */
    const int *trans_offset;                 /* host [n_particles+1], CSR     */
    const float *trans_x, *trans_y;          /* host [trans_offset[P]]        */
    const float *eulers;                     /* host [O*9]                    */
    const float *real, *imag, *corr;         /* host [n_particles*full_npix]  */
    const float *coor_xy;                    /* host [full_npix*2]            */
    int keep_diff2;                          /* keep diff2 on the device until
                                                rc_server_release            */
} rc_request;

typedef struct {
    uint64_t completed;
    int queued, inflight, peak_inflight;
    size_t hbm_used, hbm_peak, hbm_budget;
    double busy_s;               /* first admission to last completion          */
} rc_server_counters;

int rc_server_create(const rc_server_config *cfg, rc_server **out);
/* Copies the request's host inputs into pinned staging and queues it;
 * returns at once.  RC_ERR_NO_MEMORY when the staging is full,
 * RC_ERR_LIMIT when the request can never fit the HBM budget. */
int rc_server_submit(rc_server *s, const rc_request *req, uint64_t *ticket);
/* Blocks until the request is done; copies the per-particle float64 minimum
 * and flat argmin (c*O+o)*T_i+t; latency = submission to completion. */
int rc_server_wait(rc_server *s, uint64_t ticket, double *best, int64_t *best_index,
                   double *latency_ms);
/* keep_diff2 requests: copy the diff2 block ([particle][class][o][t]) to host. */
int rc_server_diff2(rc_server *s, uint64_t ticket, float *host_out, size_t count);
int rc_server_release(rc_server *s, uint64_t ticket);
int rc_server_stats(rc_server *s, rc_server_counters *out);
int rc_server_destroy(rc_server *s);

/* ------------------------------------------------------------------------ */
/* Fine-pass diff2 over job lists.  Mirrors diff2_fine (utilities.h:213-235)
 * and cuda_kernel_diff2_fine (diff2.cuh:540-682) for a 3D reference and 2D
 * data: job b (b < todo_blocks) scores orientation d_rot_idx[d_job_idx[b]]
 * against translations d_trans_idx[d_job_idx[b]] + j, j < d_job_num[b],
 * writing d_diff2s[d_job_idx[b] + j] = sum_init + sum_p corr_p/2 |P - X_s|^2
 * over image_size unfiltered pixels (written, not accumulated).  Pixel
 * coordinates follow that kernel (rows between the half-planes parked at
 * x = maxR).  A job whose orientation or translations fall outside
 * orientation_num / translation_num is skipped.  d_diff2s must hold
 * d_job_idx[b] + d_job_num[b] entries for every job, as in the reference.
 * todo_blocks above 2^24 - 8 (one 256-thread workgroup per job):
 * RC_ERR_LIMIT.  data_is_3D: RC_ERR_UNSUPPORTED. */
int rc_run_diff2_kernel_fine(const rc_projector_kernel *projector, const float *d_eulers,
                             const float *d_real, const float *d_imag, const float *d_trans_x,
                             const float *d_trans_y, const float *d_trans_z, const float *d_corr,
                             float *d_diff2s, unsigned long image_size, float sum_init,
                             unsigned long orientation_num, unsigned long translation_num,
                             unsigned long todo_blocks, const unsigned long *d_rot_idx,
                             const unsigned long *d_trans_idx, const unsigned long *d_job_idx,
                             const unsigned long *d_job_num, int data_is_3D, void *stream);

/* ------------------------------------------------------------------------ */
/* Coarse weights and significance: the step after the search.  Replaces the
 * coarse weights exponent, exponentiate and significance-threshold steps of
 * RELION's coarse pass (kernel_weights_exponent_coarse and
 * kernel_exponentiate, utilities.h:405-459; array_over_threshold and
 * find_threshold_idx_in_cumulative, cuda_device_utils.cuh:238-267).
 *
 * Per particle i and sample k = (c*O + o)*T_i + t of a batch's diff2:
 *   lw_k = lp_orient[c][o] + lp_offset[c][t] + min_i - diff2_k   (float64),
 *   w_k  = exp(lw_k - max lw)                                      (float),
 * w_k = 0 where a log prior is -inf or diff2_k < (float)min_i;
 *   sum_i = sum_k w_k (float64);  with the weights sorted ascending s[0..n-1]
 *   and c[j] = s[0] + ... + s[j],  m = #{j : c[j] <= (1 - fraction) * sum_i},
 *   significant_weight_i = s[min(m, n-1)], and sample k is significant when
 *   w_k >= significant_weight_i.
 * Deterministic; one workgroup per particle; no workspace. */
typedef struct rc_weights_desc {
    int n_particles, n_classes, orientation_num;
    const int *host_trans_offset;     /* host [n_particles+1] (shape checks)     */
    const int *d_trans_offset;        /* device copy                             */
    const float *d_diff2;             /* device, rc_batch_desc.d_diff2 layout     */
    const double *d_min_diff2;        /* device [n_particles] (the plan's best)  */
    const float *d_log_prior_orient;  /* device [n_classes*O] per particle, or NULL */
    long long log_prior_orient_stride;/* elements between particles (0: shared)  */
    const float *d_log_prior_offset;  /* device, per particle [n_classes][T_i] at
                                         n_classes*trans_offset[i], or NULL       */
    double adaptive_fraction;         /* in (0, 1], RELION's adaptive_fraction    */
    float *d_weights;                 /* device out, diff2 layout (may alias d_diff2) */
    double *d_sum_weight;             /* device out [n_particles]                */
    float *d_significant_weight;      /* device out [n_particles]                */
    int *d_n_significant;             /* device out [n_particles]                */
    unsigned char *d_significant;     /* device out, diff2 layout, 1 = significant, or NULL */
    double *d_max_log_weight;         /* device out [n_particles] (max lw), or NULL */
} rc_weights_desc;

int rc_coarse_weights(const rc_weights_desc *desc, void *stream);

/* ---------------------------------------------------------------------- */
const char *rc_last_error(void);
const char *rc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RELION_COARSE_H */
