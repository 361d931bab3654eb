// ss_device.h -- device-side view of one alignment group (shared by the HIP
// kernels and the host engine).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ss {

// Per-column partial sums written by the assembly pass (FP64).
// Column c covers block b = c % nb of frame f = c / nb and holds:
enum ColSum {
    CS_D2 = 0,   // sum d^2 over masked pixels
    CS_CNT,      // masked pixel count
    CS_AXX, CS_AYY, CS_AXY,   // sum w w' (1-f)^2   -> left-knot diagonal
    CS_BXX, CS_BYY, CS_BXY,   // sum w w' f^2       -> right-knot diagonal
    CS_CXX, CS_CYY, CS_CXY,   // sum w w' f (1-f)   -> left/right coupling
    CS_JXA, CS_JYA,           // sum d w (1-f)      -> left-knot J^T d
    CS_JXB, CS_JYB,           // sum d w f          -> right-knot J^T d
    CS_QXX, CS_QYY, CS_QXY,   // sum w w' fn^2 with fn = frac of block (b+1) % nb
    CS_N
};

// Per-column sums of the correlation passes (FP64).
enum CorrSum {
    CR_SUMI = 0,  // sum of image over all pixels
    CR_SUMD,      // sum of the swapped difference image over all pixels
    CR_NMASK,     // masked pixels
    CR_SA, CR_SAA, CR_SB, CR_SBB, CR_SAB, CR_NS,  // moments over the leading masked run
    CR_N
};

// Per-column sums of the final pass (FP64).
enum FinalSum { FS_D2 = 0, FS_CNT, FS_SA, FS_SAA, FS_SB, FS_SBB, FS_SAB, FS_N };

constexpr int kMaxCand = 8;   // fused line-search candidates per round

// Device status block read by the host after a stage.
struct Status {
    double corr;          // global correlation of the last correlation pass
    int32_t need_more;    // some frame still needs another line-search round
    int32_t improved;     // `improved` of the last completed line search
    int32_t tries;        // sequential tries resolved (accumulated)
    int32_t active;       // frames still unresolved after the last decide
};

enum ImgType { IMG_F32 = 0, IMG_U16 = 1, IMG_I16 = 2 };

struct GroupDev {
    // sizes (mexgateway.cpp:176-198)
    int frameH, nb, nF, fframeW, N, th, tw, groupSize;
    int fracUniform;       // all frac columns equal: block c-1 diagonal term uses own frac
    float xmax, ymax;      // template_w - 1 - eps, template_h - 1 - eps
    double rmult;          // line-search reduction multiplier
    // inputs
    const void *img;       // frameH x fframeW, column-contiguous
    int imgType;
    const float4 *geo;     // frameH x nb: {base_x, base_y, frac, base_mask}
    const float *fnext;    // frameH x nb: frac of block (b+1) % nb (only if !fracUniform)
    const float *tmpl, *xg, *yg;  // template and gradients, row-major
    const int *minPts;     // nb
    // state
    double *p, *pLast, *p0, *d0, *deltaP;
    double *colsum, *lssum, *errval, *corrsum, *finalsum;
    double *ws;            // solver workspace
    uint8_t *used;         // nParamPoints
    int *lsActive, *lsNextK, *lsMk, *lsDk;
    uint8_t *maskOut;      // optional final mask, frameH x fframeW column-major
    double *resid, *fcorr; // final per-frame outputs
    Status *status;
};

}  // namespace ss
