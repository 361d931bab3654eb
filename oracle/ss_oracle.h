/*
 * ss_oracle.h -- FP64 CPU restatement of SteadyScan's align_group_to_template
 * path.  TEST INFRASTRUCTURE ONLY: this library is the parity checker for the
 * HIP implementation.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.  Built a
 * second time with ORC_REAL=float (liboracle_f32.so), it is the benchmark's
 * FP32 CPU baseline; only the FP64 build checks anything.
 *
 * Every array crossing this interface is MATLAB column-major, exactly as the
 * reference mex receives it from steady_scan.m (steady_scan.m:239-242):
 * element (r, c) of an R x C matrix lives at [c * R + r].
 */
#ifndef SS_ORACLE_H_
#define SS_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Stop criteria, same order as interface/stopcriterion.h:4-10. */
enum {
    ORC_STOP_CORRELATION_INCREASE = 0,
    ORC_STOP_HALT_CORRELATION = 1,
    ORC_STOP_MOVEMENT = 2,
    ORC_STOP_ITERATION_LIMIT = 3,
    ORC_STOP_ERROR = 4
};

/* Status codes returned by every oracle entry point. */
enum {
    ORC_OK = 0,
    ORC_INVALID_ARGUMENT = 1,
    ORC_OUT_OF_MEMORY = 2
};

/* Inputs 1..13 of the mex call (mex/mexgateway.cpp:45-114). */
typedef struct {
    int32_t frame_h;        /* rows of the block matrices (mxGetM(prhs[5]))   */
    int32_t n_blocks;       /* columns of basex  (mxGetN(prhs[5]))            */
    int32_t n_frames;       /* arg 12                                          */
    int32_t template_h, template_w;
    int32_t group_size;     /* arg 11 (= 2 * n_blocks from steady_scan.m:195)  */
    const double *template_image; /* template_h x template_w                   */
    const double *xgrad, *ygrad;  /* template_h x template_w                   */
    const double *frac, *base_x, *base_y; /* frame_h x n_blocks                */
    const uint8_t *base_mask;             /* frame_h x n_blocks                */
    const double *data;           /* frame_h x (n_blocks * n_frames)           */
    const double *init_x, *init_y;/* (n_blocks + 1) x n_frames (ld n_blocks+1) */
    const int32_t *min_points;    /* n_blocks + 1, already truncated to int    */
} orc_in;

/* Inputs 14..24 (mex/mexgateway.cpp:116-169) plus the precision flag. */
typedef struct {
    int32_t max_iterations;
    double move_threshold;
    double halt_correlation;
    double correlation_increase_threshold;
    double filter_strength;
    int32_t n_filter_iterations;
    double filter_correlation_threshold;
    double lambda;                 /* smoothness extension; 0 = reference   */
    double linesearch_reduction_multiplier;
    int32_t solver_id;             /* 0 ILU(0), 1/2 exact                   */
    int32_t filter_id;             /* 0 none, 1 smearing, 2 range           */
    char precision;                /* 'D' or 'S': selects the mask epsilon  */
} orc_settings;

typedef struct {
    double *knots_x, *knots_y;     /* (n_blocks+1) x n_frames                */
    double *frame_corr;            /* n_frames                                */
    double *residrms;              /* n_frames                                */
    uint8_t *mask;                 /* frame_h x (n_blocks*n_frames) or NULL   */
    double *iteration_corr;        /* max_iterations or NULL                  */
    double *all_params;            /* 2*(n_blocks*n_frames+1) x max_iterations or NULL */
    int32_t iterations;
    int32_t stop_criterion;
    int32_t half_iterations;       /* executed half-iterations               */
    int32_t linesearch_tries;      /* sum over frames of executed tries       */
    /* flag 'c' (configuration.cpp:106-108): n_frames x max_iterations, the
     * frame correlations after every iteration (task.cpp:286-306); non-NULL
     * also makes frame_corr the last iteration's values (task.cpp:392-394) */
    double *frame_corr_iter;
    /* flag 'e' (configuration.cpp:145-147): 2*max_iterations x n_frames,
     * both shifts' errvals after every iteration that continues (task.cpp:362-370) */
    double *iteration_errvals;
    /* output 14: 2*(n_blocks*n_frames+1) x 2*max_iterations, the parameters
     * after each half-iteration's line search, column shift + 2*(iteration-1)
     * (timedmainiteration.cpp:124) */
    double *rg_params;
} orc_out;

/* Whole call: MexGateway<double>::execute (mex/mexgateway.cpp:373-445). */
int orc_align_group(const orc_in *in, const orc_settings *s, orc_out *out, int n_threads);

/* ---- stage-level hooks on a persistent state (for localized parity) ---- */
typedef struct orc_state orc_state;
int  orc_state_create(const orc_in *in, const orc_settings *s, orc_state **st);
void orc_state_destroy(orc_state *st);
/* Copies of the internal p (2*nParamPoints), deltaP, errval (2*n_frames). */
void orc_state_get(const orc_state *st, double *p, double *delta_p, double *errval);
void orc_state_set_p(orc_state *st, const double *p);
void orc_state_set_delta(orc_state *st, const double *delta_p);
/* One half-iteration up to (and including) the solve, filter and expand.
 * Returns nParamsUsed; optional outputs (NULL to skip):
 *   used[nParamPoints], jtr_full[2*nParamPoints] (uncompacted J^T d),
 *   h_full[nParamPoints * 6]: per knot {Hxx, Hyy, Hxy, Oxx, Oyy, Oxy} where
 *   O couples knot c with knot c+1 (zero where not inserted), errval[2*nF],
 *   delta_new[2*nParamPoints] expanded step. */
int orc_half_iteration_assemble(orc_state *st, int shift, uint8_t *used, double *jtr_full,
                                double *h_full, double *delta_new);
/* Per-pixel state (frame_h x n_blocks*n_frames, column-major): mask, x, y,
 * differenceImage; any pointer may be NULL. */
void orc_state_pixels(const orc_state *st, uint8_t *mask, double *x, double *y, double *d);
/* The line search of that half-iteration; returns the `improved` flag. */
int orc_half_iteration_linesearch(orc_state *st, int shift);
/* Global correlation statistic of the current state (task_correlation.cu:63-99). */
double orc_global_correlation(orc_state *st);
/* Per-frame Pearson correlation of the current state's warped template and
 * image over its masked pixels (hd_frameCorrelation, task_correlation.cu:101-183);
 * corr has n_frames entries. */
void orc_frame_correlation(orc_state *st, double *corr);

/* ---- small pure functions ---- */
/* ParameterOrganizer::organize2D1 / organize1D2 (parameterorganizer.cpp:30-99). */
void orc_organize_in(const double *px, const double *py, int n_frames, int n_blocks, double *p);
void orc_organize_out(const double *p, int n_frames, int n_blocks, double *px, double *py);
/* Exact solve of the reference's compacted COO system (eigensolve.cpp:5-41). */
int orc_solve_coo(int n, int nnz, const int *row, const int *col, const double *val,
                  const double *rhs, double *x, int solver_id);
/* SmearingFilter / RangeFilter apply (smearingfilter.cpp:24-33, rangefilter.cpp:24-40). */
void orc_filter_apply(int filter_id, double strength, double *v, int n);

/* ---- geometry (init_settings, steady_scan.m:748-828) ---- */
typedef struct {
    int32_t nlines, linewidth;      /* raw frame size                          */
    int32_t template_h, template_w;
    int32_t nknots;
    int32_t cropping[4];            /* [left right bottom top]                 */
    double subrec[4];               /* [left right bottom top]                 */
    int32_t has_subrec;
    int32_t bidirectional;
    double min_blockpoints_ratio;
} orc_geometry_in;

/* Outputs sized to the cropped frame Hc x Wc (column-major):
 * prevknotind (1-based), tfrac, basex, basey (1-based template coordinates). */
int orc_geometry(const orc_geometry_in *g, int32_t *prevknotind, double *tfrac,
                 double *basex, double *basey, int32_t *hc, int32_t *wc);
/* image2blockmatrix (steady_scan.m:601-613): returns frameH (max points per
 * block) when y == NULL; otherwise fills y (frameH x nblocks) and mask. */
int orc_image2blockmatrix(const double *x, const int32_t *blockind, int npix, int nblocks,
                          double *y, uint8_t *bmmask);

/* resample_frame + pushgrid (steady_scan.m:562-599), MATLAB semantics:
 * single-precision accumulators, last-write-wins indexed assignment. */
void orc_resample_frame(const double *data, int npix, const double *dx_knots,
                        const double *dy_knots, const double *basex, const double *basey,
                        const int32_t *prevknotind, const double *tfrac,
                        int template_h, int template_w, float *out);

#ifdef __cplusplus
}
#endif
#endif

/* invert_bilinear_sampling (steady_scan.m:290-404) on one call's frames:
 * frames Hc x Wc x nframes (cropped, double), knots nknots x nframes,
 * image-shaped geometry (Hc x Wc), template th x tw; out th x tw x nframes
 * (NaN outside the mask), maskout likewise; iters per frame (may be NULL). */
int orc_invert_bilinear_sampling(const double *frames, int Hc, int Wc, int nframes, const double *dxk,
                                 const double *dyk, int nknots, const int32_t *prevknotind,
                                 const double *tfrac, const double *basex, const double *basey, int th,
                                 int tw, double regfac, double aspect, float *out, uint8_t *maskout,
                                 int32_t *iters);
int orc_ibs_neighbour(long i0, int dir, long nrows, long npix, long *j0);
/* The steps of invert_bilinear_sampling, exposed for the tests: one frame's
 * masked system (stencil qm np x 9: diagonal, then N NE E SE S SW W NW),
 * the stopping energy, calc_interactions, one Gauss-Seidel iteration and
 * the energy-check interval that follows a frame of iter iterations. */
int orc_ibs_system(const double *data, int npd, const double *dx, const double *dy,
                   const int32_t *prevknotind, const double *tfrac, const double *basex,
                   const double *basey, int th, int tw, double regfac, double aspect, double *qm,
                   double *wy, float *iv, uint8_t *mask);
float orc_ibs_energy(const double *qm, const double *wy, const float *x, const uint8_t *mask, int th,
                     int tw);
void orc_ibs_interactions(const double *qm, const double *wy, const uint8_t *mask, int th, int tw,
                          float *Lk, float *Rm);
void orc_ibs_sweep(const float *Lk, const float *Rm, int th, int tw, float *x);
int orc_ibs_check_interval(int iter);
