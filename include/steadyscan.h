/*
 * steadyscan.h -- C ABI of the MI355X-native align_group_to_template path.
 *
 * Drop-in for the MATLAB mex `align_group_to_template`
 * (mex/align_group_to_template.cpp:21 -> MexGateway<T>::execute,
 * mex/mexgateway.cpp:373-445).  The 26 mex arguments map onto
 * ss_group_in (arrays, args 1-13) and ss_settings (scalars, args 0, 14-25);
 * the 6 outputs used by steady_scan.m:239-242 (plus the diagnostics of
 * mexgateway.cpp:266-354) map onto ss_group_out.
 *
 * Conventions
 *  - Every matrix is MATLAB column-major: element (r, c) of an R x C matrix
 *    is at [c * R + r].  A mex wrapper passes mxGetData() pointers unchanged.
 *  - All buffers are owned by the caller.  The handle owns device state
 *    (workspaces) across calls; it never resets the device
 *    (unlike task.cpp:190).
 *  - Errors are returned as ss_status, never thrown or longjmp'ed
 *    (the mex raised GPUMC:* errors through mexErrMsgIdAndTxt,
 *    mex/reporter/reporter.cpp:25-27).  ss_last_error() has the message.
 *  - A handle is used by one host thread at a time; different handles are
 *    independent.
 *  - Large host-side copies inside a call (host frames into the pinned
 *    staging of ss_align_group / ss_align_batch, the widening of
 *    ss_reconstruct's results) run on up to 16 short-lived host threads
 *    that end before the call returns.
 */
#ifndef STEADYSCAN_H_
#define STEADYSCAN_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SS_ABI_VERSION 1

typedef struct ss_handle ss_handle;

/* Status codes (GPUMC:* report ids of mex/reporter/reporter.cpp:3-37). */
typedef enum {
    SS_OK = 0,
    SS_ERR_INVALID_ARGUMENT = 1, /* GPUMC:InvalidArgument, GPUMC:NotEnoughInputs */
    SS_ERR_OUT_OF_MEMORY = 2,    /* "unable to allocate device memory" (task.cpp:116-122) */
    SS_ERR_DEVICE = 3,           /* GPUMC:CudaError equivalents: a HIP call failed */
    SS_ERR_NUMERICAL = 4,        /* a line search that cannot terminate (non-finite step) */
    SS_ERR_NOT_AVAILABLE = 5     /* no usable gfx950 device */
} ss_status;

/* Stop criterion, same values as StopCriterion (interface/stopcriterion.h:4-10). */
typedef enum {
    SS_STOP_CORRELATION_INCREASE = 0,
    SS_STOP_HALT_CORRELATION = 1,
    SS_STOP_MOVEMENT = 2,
    SS_STOP_ITERATION_LIMIT = 3,
    SS_STOP_ERROR = 4
} ss_stop_criterion;

/* Element type of the frame data (arg 8, bmData). */
typedef enum { SS_F64 = 0, SS_F32 = 1, SS_U16 = 2, SS_I16 = 3 } ss_dtype;

/* Scalar arguments: args 14-25 of the mex (mexgateway.cpp:116-174) plus the
 * precision letter of arg 0 (configuration.h:10-11). */
typedef struct {
    int32_t max_iterations;                  /* arg 14 max_iter                          */
    double move_threshold;                   /* arg 15 move_thresh                       */
    double halt_correlation;                 /* arg 16 halt_correlation                  */
    double correlation_increase_threshold;   /* arg 17                                   */
    double filter_strength;                  /* arg 18                                   */
    int32_t max_filter_iterations;           /* arg 19                                   */
    double filter_correlation_threshold;     /* arg 20 filter_deactivation_threshold     */
    double lambda;                           /* arg 21: 0 = reference behaviour (unused there);
                                                >0 adds lambda*(delta_{k+1}-delta_k)^2 per axis */
    double linesearch_reduction_multiplier;  /* arg 22                                   */
    int32_t solver_id;                       /* arg 23: 0 ILU(0) (x-then-y order), 1/2 exact */
    int32_t filter_id;                       /* arg 24: 0 none, 1 smearing, 2 range      */
    int32_t device_id;                       /* arg 25 gpu_device_id                     */
    char precision;                          /* arg 0: 'D' FP64 or 'S' FP32 pixel arithmetic
                                                (and the mask epsilon, helper/eps.h)        */
    int32_t linesearch_candidates;           /* fused candidates in a half-iteration's first
                                                line-search round: 1, 2 or 4 (0 = 2); later
                                                rounds use 4.  Results do not depend on it. */
    int32_t template_bf16;                   /* 1: mixed precision, precision 'S' only: the
                                                template's gradients (the Jacobian factors)
                                                are stored as BF16 pairs, residuals sample the
                                                FP32 template, sums are FP32 per thread and
                                                FP64 across threads, and every knot update
                                                gets one FP64 iterative-refinement step.
                                                0 (default): the precision's own width.    */
    int32_t save_frame_correlations;         /* flag 'c' (configuration.cpp:106-108): frame
                                                correlations after every iteration, on the
                                                state the global correlation leaves
                                                (task.cpp:286-306); frame_corr then holds the
                                                last iteration's (task.cpp:392-394).  0: off */
} ss_settings;

/* Fills the defaults of default_alignment_settings (steady_scan.m:836-867). */
void ss_settings_default(ss_settings *s);

/* Array arguments 1-13 of the mex (mexgateway.cpp:45-114).  frame_h and
 * n_blocks are the size of base_x (mexgateway.cpp:176-177). */
typedef struct {
    int32_t frame_h;             /* rows of every block matrix                     */
    int32_t n_blocks;            /* knot intervals per frame (nknots - 1)           */
    int32_t n_frames;            /* arg 12 nsetframes                               */
    int32_t group_size;          /* arg 11 (2 * n_blocks from steady_scan.m:195)    */
    int32_t template_h, template_w;
    const double *template_image;/* arg 1: template_h x template_w                  */
    const double *xgrad;         /* arg 2: [diff(template,1,2) NaN]                 */
    const double *ygrad;         /* arg 3: [diff(template,1,1); NaN]                */
    const double *frac;          /* arg 4: tfrac block matrix, frame_h x n_blocks   */
    const double *base_x;        /* arg 5: basex - 1                                */
    const double *base_y;        /* arg 6: basey - 1                                */
    const uint8_t *base_mask;    /* arg 7                                           */
    ss_dtype data_type;          /* element type of data                            */
    const void *data;            /* arg 8: frame_h x (n_blocks * n_frames)          */
    int32_t data_on_device;      /* nonzero: data is a device pointer (HBM resident)*/
    const double *init_x;        /* arg 9: est_dx_knots, (n_blocks+1) x >= n_frames */
    const double *init_y;        /* arg 10                                          */
    const int32_t *min_points;   /* arg 13: n_blocks + 1 (truncated toward zero)    */
    int32_t template_id;         /* > 0: use a resident template (args 1-3 ignored) */
    int32_t geometry_id;         /* > 0: use resident geometry (args 4-7 ignored)   */
} ss_group_in;

/* Outputs 1-7 of the mex (mexgateway.cpp:320-354) and the diagnostics 8-15.
 * NULL pointers skip an output. */
typedef struct {
    double *knots_x;             /* out 1: (n_blocks+1) x n_frames                  */
    double *knots_y;             /* out 2                                           */
    double *frame_corr;          /* out 3: n_frames                                 */
    int32_t iterations;          /* out 4                                           */
    uint8_t *mask;               /* out 5: frame_h x (n_blocks * n_frames)          */
    double *residrms;            /* out 6: n_frames                                 */
    double device_init_time;     /* out 7: seconds                                  */
    double *iteration_corr;      /* out 8 ('C'): max_iterations                     */
    double *all_params;          /* out 15: 2*(n_blocks*n_frames+1) x max_iterations */
    int32_t stop_criterion;      /* ss_stop_criterion ('E' flag message)            */
    int32_t half_iterations;     /* executed half-iterations                        */
    int32_t linesearch_rounds;   /* executed fused line-search rounds               */
    int32_t linesearch_tries;    /* sequential tries those rounds resolved          */
    int32_t mask_on_device;      /* nonzero: mask is a device pointer; the kernels
                                    write it in place (no copy to the host)         */
    double *frame_corr_iter;     /* out 11 ('c'): n_frames x max_iterations, written
                                    when settings.save_frame_correlations is set    */
    double *iteration_errvals;   /* out 12 ('e'): 2*max_iterations x n_frames, rows
                                    2k and 2k+1 the two shifts' errvals after every
                                    iteration k+1 that continues (task.cpp:362-370) */
    double *iteration_times;     /* out 9 ('t'/'T'): max_iterations, seconds of each
                                    iteration by the device clock, from the end of the
                                    previous one (the first from its first assembly);
                                    in a batch, the ticks' time shared with its groups */
    double *raw_params;          /* out 13 ('P'): 2*(n_blocks*n_frames+1), the returned
                                    knots in the device layout, x then y (task.cpp:216-218) */
    double *rg_params;           /* out 14: 2*(n_blocks*n_frames+1) x 2*max_iterations,
                                    the knots after each half-iteration's line search,
                                    column shift + 2*(iteration-1)
                                    (timedmainiteration.cpp:124)                      */
    double *function_times;      /* out 10 ('t'/'T'): 14 slots of seconds, the
                                    reference's functionTimes order (timedmainiteration
                                    .cpp:50-126, task.cpp:295, 355) with the fused stages
                                    in slot 1 (parameters, difference image, gradients,
                                    normal equations), 7 (solve, filter, expansion), 9
                                    (line search) and 10 (correlations, final statistics);
                                    device clock between the group's stage ends        */
} ss_group_out;

/* Handle lifecycle.  Replaces CudaHelper<T>::initDevice (cudahelper.cu:8-27)
 * and Task<T>::initialize / freePreAllocatedMemory (task.cpp:47-197). */
ss_status ss_create(int32_t device_id, ss_handle **out);
void ss_destroy(ss_handle *h);
const char *ss_last_error(const ss_handle *h);

/* Resident inputs.  A template (args 1-3) or a scan geometry (args 4-7)
 * uploaded once is converted to the device layout and kept in HBM until
 * released; ss_group_in refers to it by id.  No reference counterpart: the
 * mex re-uploads every argument per call (task_allocate.cu:290-302). */
ss_status ss_upload_template(ss_handle *h, int32_t template_h, int32_t template_w,
                             const double *template_image, const double *xgrad,
                             const double *ygrad, int32_t *id);
ss_status ss_upload_geometry(ss_handle *h, int32_t frame_h, int32_t n_blocks, const double *frac,
                             const double *base_x, const double *base_y,
                             const uint8_t *base_mask, int32_t *id);
ss_status ss_release(ss_handle *h, int32_t id);

/* One align_group_to_template call (MexGateway<T>::execute,
 * mexgateway.cpp:373-445): Task::alignFrames + getParameters. */
ss_status ss_align_group(ss_handle *h, const ss_group_in *in, const ss_settings *s,
                         ss_group_out *out);

/* Several independent calls in one submission (groups may differ in frame
 * count, size and knot spacing).  Results equal n separate ss_align_group
 * calls.  No reference counterpart: steady_scan.m:216-287 loops over sets. */
ss_status ss_align_batch(ss_handle *h, int32_t n, const ss_group_in *in,
                         const ss_settings *s, ss_group_out *out);

/* ---- kernel timing ---- */

/* Kernel classes timed by ss_profile_* (HIP events on the handle's stream). */
typedef enum {
    SS_K_ASSEMBLE = 0,   /* fused warp / residual / J^T J / J^T d pass       */
    SS_K_SOLVE = 1,      /* FP64 block-tridiagonal solve + filter + expand   */
    SS_K_LS_ROUND = 2,   /* fused line-search candidates pass                */
    SS_K_LS_DECIDE = 3,  /* sequential resolution of the candidates          */
    SS_K_CORR = 4,       /* global correlation pass + reduction              */
    SS_K_FINAL = 5,      /* final errval / frame correlation / mask pass     */
    SS_K_COUNT = 6
} ss_kernel_class;

/* Enables (1) or disables (0) per-launch HIP event timing and clears the
 * counters.  No reference counterpart (the 't' flag's host timers,
 * timedmainiteration.cpp:46-145, synchronised the device per stage). */
ss_status ss_profile_enable(ss_handle *h, int32_t on);
/* Launch counts, summed kernel milliseconds and summed algorithmic frame
 * bytes (frame pixels read x element size) per kernel class. */
ss_status ss_profile_read(ss_handle *h, int64_t *launches, double *ms, double *frame_bytes);

/* ---- autotuning of the pixel kernels ----
 * No reference counterpart (the reference fixes THREADS_PER_BLOCK 256,
 * cudaincludes.h:36).  The assembly, line-search and correlation kernels are
 * built in three launch variants each: (template samples a thread keeps in
 * flight, minimum waves per SIMD).  By default each batch picks them by its
 * size: variant 0, measured fastest on batches of 512x512 groups, for a
 * batch of more than 8 block columns per CU; for a smaller one (a single
 * group, say) the variants measured fastest there (more waves per SIMD for
 * the FP64 assembly and correlation, more samples in flight for the FP32
 * line search).  With tuning on, the first time a
 * handle sees a group shape (precision, mixed precision, frame element type,
 * frame_h, n_blocks, lines per run, geometry form, gradient form, frame-count
 * class floor(log2(n_frames))), it runs that group's first tick twice per
 * variant from the group's initial state (a warm-up pass, then a pass timed
 * with HIP events) and keeps the fastest variant per kernel for the shape
 * (ss_align_group, ss_align_batch and server admission alike).  Every
 * variant computes the same results bit for bit, so tuning changes times
 * only.  A batch whose groups use different variants launches each kernel
 * once per variant, which can cost more than the tuning gains. */
typedef struct {
    char precision;              /* 'D' or 'S'                                      */
    int32_t mixed;               /* template_bf16                                   */
    int32_t data_type;           /* ss_dtype                                        */
    int32_t frame_h, n_blocks;
    int32_t frames;              /* frame count of the group the shape was tuned on */
    int32_t chosen[3];           /* variant per kernel: assembly, line search, correlation */
    double trial_us[3][3];       /* [kernel][variant]: kernel time in the trial tick */
    int32_t samples[3][3];       /* [kernel][variant]: the variants' parameters      */
    int32_t waves[3][3];
} ss_tuning_entry;
/* -2 (default): variants by the size of each batch; 0, 1 or 2: use that
 * variant for every kernel, no tuning; -1: tune each new group shape.  A
 * group keeps the mode in force when it was submitted. */
ss_status ss_tuning_mode(ss_handle *h, int32_t mode);
/* Copies up to cap of the handle's tuned shapes to out; returns how many
 * there are. */
int32_t ss_tuning_entries(const ss_handle *h, ss_tuning_entry *out, int32_t cap);

/* ---- serving: continuous batching of groups ----
 * No reference counterpart: the mex aligns one group per call and resets
 * the device afterwards (task.cpp:164-197).  A server keeps a running batch
 * on the handle's device: each step advances every resident group by one
 * tick (about one Gauss-Newton iteration), retires the groups that finished
 * (their outputs are written) and admits queued groups while the group-slot
 * count and the HBM budget allow.  A group's result is the same alone and in
 * any mix (every group's arithmetic is independent and deterministic).
 * A server is used by one thread at a time, like its handle. */
typedef struct ss_server ss_server;
typedef struct {
    int64_t hbm_budget_bytes;    /* device memory for admitted groups (0: 64 GiB) */
    int32_t max_groups;          /* groups resident at once (0: 256)            */
    int32_t reserved;            /* must be 0                                   */
    int64_t staging_limit_bytes; /* page-locked host staging the server may hold
                                    (0: no limit); groups whose staging would
                                    exceed it complete with SS_ERR_OUT_OF_MEMORY */
} ss_server_config;

ss_status ss_server_create(ss_handle *h, const ss_server_config *cfg, ss_server **out);
/* Frees the server's device memory; queued and resident groups are dropped. */
void ss_server_destroy(ss_server *s);
/* Queues one group (the arguments of ss_align_group).  The structs are
 * copied; the arrays they point to, device frames included, and the out
 * buffers must stay valid until the ticket is reported by
 * ss_server_completed. */
ss_status ss_server_submit(ss_server *s, const ss_group_in *in, const ss_settings *st,
                           ss_group_out *out, int64_t *ticket);
/* Admits, advances one tick, retires.  Reports the resident and queued
 * group counts after the step.  A queued group that cannot be admitted (a bad
 * request, one that can never fit the budget, a failed set-up) completes
 * with its error and gives back its memory; the step itself fails only when
 * the device does. */
ss_status ss_server_step(ss_server *s, int32_t *n_resident, int32_t *n_queued);
/* Moves up to cap finished tickets (with their status) to the caller;
 * returns how many.  messages (may be NULL) receives each ticket's error
 * message ("" on success), owned by the server and valid until the next
 * ss_server_completed call. */
int32_t ss_server_completed(ss_server *s, int64_t *tickets, ss_status *status, const char **messages,
                            int32_t cap);
/* Device bytes reserved by the server's pool and held by resident groups. */
ss_status ss_server_memory(const ss_server *s, int64_t *reserved, int64_t *in_use);

/* ---- host-side geometry: init_settings + image2blockmatrix ---- */

/* Scan geometry (settings fields of steady_scan.m:748-828, 836-867). */
typedef struct {
    int32_t nlines, linewidth;      /* raw frame size (H x W)                     */
    int32_t template_h, template_w;
    int32_t nknots;                 /* settings.nknots                            */
    int32_t cropping[4];            /* [left right bottom top]                    */
    double subrec[4];               /* [left right bottom top], if has_subrec     */
    int32_t has_subrec;
    int32_t bidirectional;          /* settings.isbidirectional                   */
    double min_blockpoints_ratio;   /* settings.min_blockpoints_ratio (0.3)       */
} ss_geometry_desc;

/* Sizes of the geometry: cropped frame hc x wc and block-matrix frame_h. */
ss_status ss_geometry_size(const ss_geometry_desc *g, int32_t *hc, int32_t *wc,
                           int32_t *frame_h);

/* init_settings (steady_scan.m:787-827) and the block-matrix packing of
 * estimate_displacements (steady_scan.m:166-175).  Image-shaped outputs are
 * hc x wc; block outputs frame_h x (nknots-1); min_points has nknots entries.
 * Any output pointer may be NULL. */
ss_status ss_geometry_init(const ss_geometry_desc *g, int32_t *prevknotind, double *tfrac,
                           double *basex, double *basey, double *frac_bm, double *base_x_bm,
                           double *base_y_bm, uint8_t *base_mask_bm, int32_t *min_points);

/* image2blockmatrix of a stack of frames (steady_scan.m:229-234, 601-613):
 * frames are nlines x linewidth x n (raw, uncropped); out is
 * frame_h x ((nknots-1) * n) of the same element type. */
ss_status ss_pack_frames(const ss_geometry_desc *g, ss_dtype type, const void *frames,
                         int32_t n, void *out);

/* ---- the steady_scan driver: estimate_displacements (steady_scan.m:93-288) ---- */

/* Driver settings (default_alignment_settings, steady_scan.m:836-867). */
typedef struct {
    int32_t frames_per_set;      /* settings.frames_per_set (500)                    */
    int32_t min_frames_per_set;  /* settings.min_frames_per_set (100); see below     */
    double corr_thresh;          /* settings.corr_thresh (0.75)                      */
    double minblockfrac;         /* settings.minblockfrac (0.25)                     */
    int32_t frame_groupsize;     /* settings.gpu_frame_groupsize (2)                 */
    double pregauss;             /* settings.pregauss (0): prefilter Gaussian sigma  */
    int32_t mednorm;             /* settings.mednorm (0): divide frames by median    */
    int32_t pad;
} ss_driver_settings;

void ss_driver_settings_default(ss_driver_settings *d);

/* Per-frame results of estimate_displacements; caller-owned, NULL skips an
 * output except the knots.  Frames that are not estimated (the stack's last
 * frame, and failed frames' corr) hold NaN, as in the reference. */
typedef struct {
    double *dx_knots;            /* nknots x n_frames                                */
    double *dy_knots;            /* nknots x n_frames                                */
    double *corr;                /* n_frames                                          */
    double *iter_used;           /* n_frames                                          */
    double *residrms;            /* n_frames                                          */
    uint8_t *failed;             /* n_frames                                          */
} ss_displacements_out;

/* estimate_displacements: frames (nlines x linewidth x n_frames, raw and
 * uncropped, the frame_list already applied) are split into sets starting
 * every frames_per_set frames, each extended by the next set's first frame
 * (steady_scan.m:97-99, 216; min_frames_per_set only rewrites set_size,
 * which estimate_displacements does not read), every set is aligned against the
 * template (template_h x template_w, column-major) and the per-frame outputs
 * are unpacked with the reference's success tests (steady_scan.m:246-284).
 * All sets go to the device in one batched submission.  est_dx/est_dy
 * (nknots x n_frames) may be NULL for zero estimates; as in the reference,
 * every set reads their first columns.  Each set's frames are prefiltered
 * (pregauss, mednorm; steady_scan.m:228) before packing, as ss_prefilter does;
 * the template is used as given (steady_scan.m:94 prefilters it before the
 * call; ss_prefilter does that step). */
ss_status ss_estimate_displacements(ss_handle *h, const ss_geometry_desc *geometry,
                                    const ss_driver_settings *driver, const ss_settings *s,
                                    const double *template_image, ss_dtype type, const void *frames,
                                    int32_t n_frames, const double *est_dx, const double *est_dy,
                                    ss_displacements_out *out);

/* ---- call capture and replay (the 'F' flag's ROM file) ----
 * Replaces MexSimulator<T>::saveRomToFile / loadRomFromFile
 * (mex/mexsimulator/mexsimulator.cpp:52-98), which dump the host struct
 * image; this format is self-describing and little-endian (header "SSROM001",
 * byte-order mark, tagged column-major arrays, settings field by field). */
typedef struct ss_rom ss_rom;
/* Writes the arrays and settings of one call (host arrays only). */
ss_status ss_rom_write(const char *path, const ss_group_in *in, const ss_settings *s);
/* Reads a ROM; on success in/s point into *rom, valid until ss_rom_free. */
ss_status ss_rom_read(const char *path, ss_rom **rom, ss_group_in *in, ss_settings *s);
void ss_rom_free(ss_rom *rom);

/* ---- corrected frames: resample_frame / pushgrid (steady_scan.m:562-599) ---- */
typedef struct {
    const ss_geometry_desc *geometry;
    int32_t n_frames;
    ss_dtype data_type;
    const void *frames;          /* nlines x linewidth x n_frames (raw, uncropped) */
    int32_t data_on_device;
    const double *knots_x;       /* nknots x n_frames                              */
    const double *knots_y;
    int32_t out_on_device;       /* ss_resample only: out is a device pointer      */
} ss_resample_in;

/* Forward bilinear splat of every frame into template space, normalised by
 * the splatted weights: out is template_h x template_w x n_frames float32,
 * NaN where nothing landed.  Repeated targets keep the last contribution,
 * as MATLAB's indexed assignment does.  The frames run on the handle's
 * stream in launches of as many frames as fit in the handle's work memory
 * (at least 256 MiB); host frames and host output are staged there. */
ss_status ss_resample(ss_handle *h, const ss_resample_in *in, float *out);

/* ss_resample without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Frames and output must be
 * device memory (data_on_device and out_on_device set); knots stay host
 * arrays.  Each pixel's geometry is computed inside the kernels, and the knots
 * and winner tables live in a static device work area of 64 MiB per device,
 * held from library load until the process ends.  A launch takes as many
 * frames as fit there: one frame needs 16 bytes per template pixel, so
 * templates up to about 4.1 million pixels are accepted and larger ones are
 * rejected with SS_ERR_INVALID_ARGUMENT.  All calls on one device share that
 * area: calls that overlap on one GPU (from several threads or handles) race
 * and must be serialised by the caller.  Results equal ss_resample's bit for
 * bit. */
ss_status ss_resample_noalloc(ss_handle *h, const ss_resample_in *in, float *out);

/* ---- corrected frames: invert_bilinear_sampling (steady_scan.m:290-404) ---- */
typedef struct {
    double regfac;          /* settings.ibs_regfac (3e-3)                          */
    double pixel_aspect;    /* settings.ibs_template_pixel_aspect_ratio (1)        */
    int32_t inpaint;        /* settings.ibs_inpaint: must be 0; inpainting is
                               ss_fill_gaps on the result                         */
    int32_t max_iterations; /* Gauss-Seidel iterations allowed per frame; 0 = 100000.
                               The reference has no bound; a frame that needs more
                               fails the call with SS_ERR_NUMERICAL               */
    int32_t chunk_frames;   /* frames reconstructed together on the device; 0 = as
                               many as fit in work_bytes                          */
    int32_t outputs_on_device; /* out and mask are device pointers                */
    int64_t work_bytes;     /* device work memory the handle may keep for this
                               call; 0 = half of the device memory free or
                               already held by the handle (about 53 MB per
                               512x512 frame)                                     */
    int32_t chain_frames;   /* 0: the frames form one chain, as one call of the
                               reference (the energy-check interval starts at 1
                               and carries from frame to frame).  k > 0: the
                               interval restarts at 1 on frames 0, k, 2k, ...,
                               as if every k frames were a separate call (the
                               driver's sets, steady_scan.m:124-134), so the
                               sets of a stack run as one launch               */
    int32_t reserved;       /* must be 0                                          */
} ss_ibs_settings;

void ss_ibs_settings_default(ss_ibs_settings *s);

/* Replaces invert_bilinear_sampling(data, settings, corrected_data_size,
 * frame_list, dx_knots, dy_knots) (steady_scan.m:290, called at :133).
 * in describes the frames as for ss_resample (raw nlines x linewidth frames,
 * cropped here by the geometry's cropping; knots nknots x n_frames, one
 * column per frame, i.e. dx_knots(:, frame_list)).  Frames of any element
 * type are taken as the double values they hold.  Outputs, template_h x
 * template_w x n_frames: out (float32, NaN outside the mask) = data_corrected,
 * mask = mask; iterations (may be NULL) the Gauss-Seidel iterations of each
 * frame, 0 for a frame with no constraint.  The energy-check interval starts
 * at 1 for the call and carries from frame to frame as in the reference. */
ss_status ss_invert_bilinear_sampling(ss_handle *h, const ss_resample_in *in,
                                      const ss_ibs_settings *s, float *out, uint8_t *mask,
                                      int32_t *iterations);

/* ---- frame preparation and gap filling ---- */

/* Replaces prefilter_foralignment(data, gsd, mednorm) (steady_scan.m:625-643):
 * frames height x width x n_frames (column-major, element type `type`) are
 * blurred by the reference's circular Gaussian on the grid padded by
 * L = ceil(sqrt(2 gsd^2 log(10000 / (sqrt(2 pi) gsd)))) pixels (gsd =
 * pregauss > 0), then divided by each frame's median (mednorm != 0).  The
 * result keeps the element type, as MATLAB keeps the class: integer frames are
 * rounded half away from zero and saturated.  Median normalisation of integer
 * frames is not supported (SS_ERR_INVALID_ARGUMENT).  in == out is allowed;
 * on_device: both are device pointers.  The handle keeps the work memory
 * (frames are processed in chunks within half of the free device memory). */
ss_status ss_prefilter(ss_handle *h, ss_dtype type, const void *in, void *out, int32_t height,
                       int32_t width, int32_t n_frames, double pregauss, int32_t mednorm,
                       int32_t on_device);

/* Replaces imgapfill_leastsquares(a) (steady_scan.m:525-560), the inpainting
 * of invert_bilinear_sampling's ibs_inpaint: the NaN pixels of each frame are
 * set to the minimiser of the summed squared adjacent differences with the
 * other pixels fixed (a frame without any known pixel is left as it is).
 * type is SS_F32 or SS_F64; the solve runs in FP64 (relative residual 1e-13)
 * and rounds to the type.  in == out is allowed; on_device: device pointers.
 * The handle keeps the work memory (about 16 MB per 512 x 512 frame; frames
 * are processed in chunks within half of the free device memory). */
ss_status ss_fill_gaps(ss_handle *h, ss_dtype type, const void *in, void *out, int32_t height,
                       int32_t width, int32_t n_frames, int32_t on_device);

/* The reconstruction loop of steady_scan (steady_scan.m:124-134, and
 * :136-154 for each extra channel): frames are taken set by set (sets start
 * every frames_per_set frames, without overlap), cropped, prefiltered
 * (pregauss, mednorm) and reconstructed by invert_bilinear_sampling with the
 * frames' knots (nknots x n_frames; a frame with NaN knots, such as the
 * stack's last frame, comes out NaN), the energy-check interval restarting
 * at each set.  With ibs->inpaint, each set's result is gap-filled
 * (ss_fill_gaps) in single precision for SS_F32 frames and in double
 * otherwise, as the reference fills in the frames' class.  data_corrected
 * (double) and mask are template_h x template_w x n_frames.  The result is
 * that of one invert_bilinear_sampling call per set; the sets themselves run
 * together, up to about 4096 frames per launch (ibs->chain_frames is set to
 * frames_per_set). */
ss_status ss_reconstruct(ss_handle *h, const ss_geometry_desc *geometry, const ss_driver_settings *driver,
                         const ss_ibs_settings *ibs, ss_dtype type, const void *frames, int32_t n_frames,
                         const double *dx_knots, const double *dy_knots, double *data_corrected,
                         uint8_t *mask);

#ifdef __cplusplus
}
#endif
#endif
