// ss_device.h -- device-side view of a batch of alignment groups (shared by
// the HIP kernels and the host engine).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ss_tuning.h"

namespace ss {

// Per-column partial sums written by the assembly pass (FP64).
// Column c covers block b = c % nb of frame f = c / nb and holds:
enum ColSum {
    CS_D2 = 0,                // sum d^2 over masked pixels
    CS_SUMI,                  // sum of the image over all pixels
    CS_CNT,                   // masked pixel count
    CS_AXX, CS_AYY, CS_AXY,   // sum w w' (1-f)^2   -> left-knot diagonal
    CS_BXX, CS_BYY, CS_BXY,   // sum w w' f^2       -> right-knot diagonal
    CS_CXX, CS_CYY, CS_CXY,   // sum w w' f (1-f)   -> left/right coupling
    CS_JXA, CS_JYA,           // sum d w (1-f)      -> left-knot J^T d
    CS_JXB, CS_JYB,           // sum d w f          -> right-knot J^T d
    CS_QXX, CS_QYY, CS_QXY,   // sum w w' fn^2 with fn = frac of block (b+1) % nb
    CS_N
};

// Per-column sums of the correlation pass (FP64).
enum CorrSum {
    CR_SUMD = 0,  // sum of the swapped difference image over all pixels
    CR_NMASK,     // masked pixels
    CR_SA, CR_SAA, CR_SB, CR_SBB, CR_SAB, CR_NS,  // moments over the leading masked run
    CR_N
};

// Per-column sums of the final pass (FP64).
enum FinalSum { FS_D2 = 0, FS_CNT, FS_SA, FS_SAA, FS_SB, FS_SBB, FS_SAB, FS_N };
// output 10: the reference's 14 function-time slots and the ones the fused
// stages fill (timedmainiteration.cpp:50-126, task.cpp:295, 355)
constexpr int kFuncSlots = 14;
enum FuncSlot { FT_ASSEMBLY = 1, FT_SOLVE = 7, FT_LINESEARCH = 9, FT_CORRELATION = 10 };
// per-iteration frame correlation sums (flag 'c'), kept in finalsum's slots
// while the group is in its correlation phase
enum FrameCorrSum { FC_CNT = 0, FC_SA, FC_SAA, FC_SB, FC_SBB, FC_SAB, FC_N };
static_assert((int)FC_N <= (int)FS_N, "frame correlation sums share the final sums' slots");

constexpr int kMaxCand = 4;  // fused line-search candidates per round

// Phases of the per-group state machine (Task<T>::alignFrames +
// MainIteration<T>::execute).
enum Phase {
    PH_ASSEMBLE = 0,  // next: assembly pass of half-iteration `shift`
    PH_LS = 1,        // line-search rounds of the current half-iteration
    PH_CORR = 2,      // global correlation + stop tests
    PH_FINAL = 3,     // final errval / frame correlation / mask
    PH_DONE = 4,
    PH_ERROR = 5
};

enum ImgType { IMG_F32 = 0, IMG_U16 = 1, IMG_I16 = 2, IMG_F64 = 3, IMG_NTYPES = 4 };

// Affine scan geometry of a block, in the block's (column j, line i) pixel
// coordinates: base_x = gx0 + gxj j + gxi i, base_y = gy0 + gyb b + gyj j +
// gyi i, frac = f0 + fj j + fi i.
template <typename R>
struct AffineGeo {
    R gx0, gxj, gxi, gy0, gyb, gyj, gyi, f0, fj, fi;
};

// Static description of one group (written once per call by the host).
struct GroupDev {
    // sizes (mexgateway.cpp:176-198)
    int frameH, nb, nF, fframeW, N, th, tw, groupSize;
    int itemStart;        // first work item (= column) of this group in the batch
    int var[TK_COUNT];    // launch variant of the assembly, line-search and correlation kernels
    int varBySize;        // var follows the size of the batch the group runs in
    int frameStart;       // first frame item of this group in the batch
    // settings
    double xmax, ymax;    // template_w - 1 - eps, template_h - 1 - eps
    int fp64;             // pixel arithmetic in FP64 (precision 'D') or FP32 ('S')
    double rmult, thr, halt, cit, fct, strength, lambda;
    int solverId;         // 0: ILU(0); 1, 2: exact
    int maxIter, K1, Kn, maxRounds;
    // geometry: affine fast path or per-pixel table
    int affine, L;
    int runLen;           // rows a thread walks in order (a block's lines of one image column)
    AffineGeo<float> afF;
    AffineGeo<double> afD;
    int fracUniform;
    // per-pixel table, frameH x nb: {base_x, base_y, frac, base_mask}, and the
    // frac of block (b+1) % nb (only if !fracUniform); one of the two widths
    const float4 *geo;
    const double4 *geoD;
    const float *fnext;
    const double *fnextD;
    // template: pairs (T[t], T[t+1]) row-major; gradient pairs unless derived
    const float2 *tp, *gxv, *gyh;
    const double2 *tpD, *gxvD, *gyhD;
    const uint32_t *gxvB, *gyhB;  // BF16 gradient pairs (mixed precision, FP32 arithmetic)
    int gradBF16;         // Jacobian gradients from the BF16 pairs
    int refine;           // one FP64 iterative-refinement step of the knot update
    int gradFD;           // gradients are forward differences of the template
    // frame data, frameH x fframeW column-contiguous
    const void *img;
    int imgType, vec;
    const int *minPts;
    // state
    double *p, *pLast, *p0, *d0, *deltaP, *errval;
    double *colsum, *lssum, *corrsum, *finalsum, *ws;
    uint8_t *used;
    int *lsActive, *lsNextK, *lsMk, *lsDk;
    uint8_t *maskOut;
    double *resid, *fcorr;
    double *allParams, *iterCorr;  // optional diagnostics (or null)
    // flag 'c': nF x maxIter frame correlations after every iteration, taken
    // on the state the global correlation leaves (or null); flag 'e':
    // 2*maxIter x nF errvals of both shifts after every continuing iteration
    double *fcIter, *iterErr;
    // output 14: knots after each half-iteration; output 9: seconds per
    // iteration by the device clock (clockHz ticks per second)
    double *rgParams, *iterTimes;
    double clockHz;
    double *funcTimes;  // output 10: kFuncSlots seconds by stage (or null)
};

// Dynamic control state of one group (read by the host after each tick).
struct GroupCtl {
    int phase, shift, iteration, improved, iterImproved;
    int solverFilter, filterID, nFilterIterations;
    int retv, error;
    int halfIterations, rounds, tries, roundsThisLS;
    int slotRounds, K;    // rounds run in the tick's current line-search slots, candidates per frame
    int lsFramePasses;    // frames evaluated over all line-search rounds
    int pad;
    double correlation, corr;
    unsigned long long tIter;  // device clock at the end of the last iteration (0: not started)
    unsigned long long tStage; // device clock at the end of the group's last stage (output 10)
};

struct Batch {
    int nGroups, nItems, nFrames;
    GroupDev *groups;
    GroupCtl *ctl;
    int *itemGroup;       // work item (column) -> group
    int *frameGroup;      // frame item -> group
};

struct LaunchBatch {
    Batch b;
    int hasType[2 * IMG_NTYPES];  // [imgType + IMG_NTYPES * fp64]
    int hasVar[TK_COUNT][2 * IMG_NTYPES];  // bit v: some group of the type uses variant v
    int hasFrameCorr;     // some group saves per-iteration frame correlations (flag 'c')
    int itemGrid, frameGrid;
    // line-search rounds after a tick's first slot: workgroups per column
    // (2: each takes two of the four candidates), and their grid
    int lsSplit, lsSplitGrid;
};

// Stages of one host submission, in order; see launch_stage.
enum Stage { ST_ASSEMBLE = 0, ST_SOLVE, ST_LS, ST_LS_DECIDE, ST_CORR, ST_CORR_FINISH, ST_FINAL, ST_FINAL_FINISH };

// Line-search rounds per half-iteration in a tick.  The per-group
// bookkeeping of the rounds (k_ls_finish) runs once, after the last slot.
constexpr int kLsSlots = 3;
// A tick entry: the stage, and for ST_LS and ST_LS_DECIDE the round's slot
// (0 .. kLsSlots - 1) above bit 8.
constexpr int tick_entry(int stage, int slot = 0) { return stage | slot << 8; }
constexpr int entry_stage(int e) { return e & 0xff; }
constexpr int entry_slot(int e) { return e >> 8; }

hipError_t launch_stage(const LaunchBatch &L, int entry, hipStream_t s);
size_t solve_workspace_bytes(int n);

}  // namespace ss
