/*
 * steadyscan_synth.h -- seeded synthetic acquisitions for tests and benchmarks.
 *
 * No counterpart in the reference (it ships no data).  Frames are a template
 * bilinearly sampled at (column + dx(line), line + dy(line)) with Poisson
 * noise, where dx, dy is a per-scan-line displacement trace.  The generator
 * uses its own RNG (xoshiro256** seeded by splitmix64), Box-Muller normals
 * and Poisson samplers (Knuth below mean 10, Hormann's PTRS above), so the
 * same seed gives the same data on every machine with the same libm.
 * All matrices are MATLAB column-major.
 */
#ifndef STEADYSCAN_SYNTH_H_
#define STEADYSCAN_SYNTH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t h, w;                       /* template size                       */
    int32_t n_blobs;                    /* isotropic Gaussian blobs            */
    double sigma_lo, sigma_hi;          /* blob sigma ~ U[lo, hi] px           */
    double amp_lo, amp_hi;              /* blob amplitude ~ U[lo, hi]          */
    double offset;                      /* constant added                      */
    int32_t n_background;               /* broad Gaussians (low frequency)     */
    double background_sigma;
    double background_amp_lo, background_amp_hi;
    double target_mean;                 /* > 0: rescale to this mean           */
    uint64_t seed;
} ss_synth_template_desc;

typedef struct {
    int32_t n_frames, nlines;           /* one displacement per scan line      */
    double frame_rate;                  /* frames per second                   */
    double walk_sigma;                  /* random-walk step sigma (px/line)    */
    double walk_revert;                 /* mean reversion per line (0 = none)  */
    double sin_amp_x, sin_amp_y;        /* sinusoid amplitudes (px)            */
    double sin_freq;                    /* Hz                                  */
    double sin_phase_x, sin_phase_y;    /* radians                             */
    double jump_prob;                   /* per line                            */
    double jump_size;                   /* px, random sign                     */
    uint64_t seed;
} ss_synth_motion_desc;

/* Template image h x w. Returns 0 on success. */
int ss_synth_template(const ss_synth_template_desc *d, double *out);

/* Displacement traces dx, dy: n_frames * nlines values each (frame-major). */
int ss_synth_motion(const ss_synth_motion_desc *d, double *dx, double *dy);

/* Frames h x w x n_frames of type out_type (0 f64, 1 f32, 2 u16, 3 i16):
 * template sampled bilinearly (clamped at the border) at
 * (j + dx[f*h+i], i + dy[f*h+i]), then Poisson noise if poisson != 0. */
int ss_synth_frames(const double *tmpl, int32_t th, int32_t tw, int32_t h, int32_t w,
                    int32_t n_frames, const double *dx, const double *dy, int32_t poisson,
                    uint64_t seed, int32_t out_type, void *out, int32_t n_threads);

#ifdef __cplusplus
}
#endif
#endif
