/*
 * ss_oracle.c -- FP64 CPU restatement of SteadyScan's align_group_to_template.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the HIP library;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * The code follows the reference's stage structure and loops one to one
 * (same per-pixel state arrays x, y, mask, differenceImage, wxgrad, wygrad;
 * same COO assembly; same host loop).  Where the reference races or reads
 * uninitialised memory, the pinned definitions below are used:
 *   - line search (task_linesearch.cu:35-106): candidate k of frame f uses
 *     p0 + r^k * delta for every knot it touches (its own knots and frame
 *     f+1's first knot, "lockstep"); it commits only its own knots.  The
 *     group's last knot is never written in pnew and stays 1e6
 *     (task_allocate.cu:309).
 *   - paramsUsed[nParamPoints-1] is never written on the untimed path
 *     (task_calcerrval.cu:33,55): pinned false.
 *   - groupActive[2*nGroups] (odd frame counts, task_calcerrval.cu:29-31):
 *     pinned active; all groups are active (flag 'M' is not supported).
 *   - the static Eigen matrix reuse (eigensolve.cpp:7-33): rebuilt per solve.
 *   - the filter `continue` path cannot run past max_iterations (task.cpp:324).
 *   - the line-search movement test abs() is the floating-point magnitude.
 * Eigen::SimplicialLDLT is replaced by an exact banded LDL^T of the same
 * matrix in interleaved (x_k, y_k) order (exact solves agree to rounding).
 *
 * The file also restates invert_bilinear_sampling (steady_scan.m:290-404)
 * with its helpers.  Its summation orders are fixed here where MATLAB's
 * sparse() and BLAS leave them open: constraint sums over data pixels in
 * ascending index; the stopping energy in 1024 partial sums, each over a
 * row segment (the layout is described at orc_ibs_energy), added by a fixed
 * tree; (Q x)_i over neighbours in ascending index.  Those rounding-level
 * orders are parity unpinned against the reference (no MATLAB here);
 * everything else is pinned by the dense-definition tests.
 */
#include "ss_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX2R(i, j, ld) ((size_t)(i) * (size_t)(ld) + (size_t)(j))

/* ORC_PLANT = n > 0 compiles deliberate mistake n into a copy of this file.
 * Only tests/test_oracle_closed_forms.py builds such copies (into a temporary
 * directory), to show that its closed-form checks reject each mistake; the
 * oracle itself is built with ORC_PLANT = 0 and contains none of them.
 *   1 bidirectional scan flips the even lines instead of the odd ones
 *   2 integer-spacing knot times taken from the cropped lines
 *   3 linspace divides by n instead of n - 1
 *   4 ILU(0) eliminates rows whose columns were not sorted
 *   5 ILU(0) divides the multiplier by the row's own pivot
 *   6 frame correlation means taken over all pixels of the frame
 *   7 frame correlation of the residual instead of the warped template
 *   8 the global correlation swaps every pixel back
 *   9 per-iteration errvals store the two shifts in each other's rows */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

/* Pixel arithmetic type: double in the oracle.  A copy built with
 * ORC_REAL=float (oracle/liboracle_f32.so) is the FP32 CPU baseline that the
 * benchmark times: template, frames, warp, residuals and their per-column
 * sums in float, as the reference's precision 'S' (T = float) keeps them.
 * Only the double build is pinned and used as the checker. */
#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

struct orc_state {
    int frameH, frameW, fframeW, nParamPoints, nParams, nFrames;
    int templateH, templateW, groupSize, nGroups;
    double eps;
    /* row-major copies, as MexConverter's RowMajor targets (mexconverter.cpp:223-285) */
    real *tmpl, *xg, *yg, *frac, *baseX, *baseY, *image;
    uint8_t *baseMask;
    int *minPoints;
    /* DeviceMemory<T> state (interface/memory.h:86-163) */
    double *p, *pLast, *deltaP, *newDeltaP, *jtr;
    real *x, *y, *d, *wx, *wy;
    uint8_t *mask, *paramsUsed, *blocksPresent;
    int *subOff;
    int nParamsUsed;
    double *Hval;
    int *Hrow, *Hcol;
    double *errval, *corr;
    double *hdbg; /* per-knot {Hxx,Hyy,Hxy,Oxx,Oyy,Oxy} of the last assembly */
    /* settings and filter lifecycle */
    orc_settings s;
    int solverFilter; /* filter object held by the solver: 0 none, 1 smearing, 2 range */
    int filterID;     /* mem.filterID */
    int nFilterIterations;
    int nthreads;
    long lsTries;
    int halfIterations;
};

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* ------------------------------------------------------------------ */
/* ParameterOrganizer (interface/parameterorganizer/parameterorganizer.cpp) */
/* ------------------------------------------------------------------ */

/* organize2D1(px, py) (parameterorganizer.cpp:30-55): shared knots are the
 * average of frame f's last and frame f+1's first estimate. */
void orc_organize_in(const double *px, const double *py, int n_frames, int n_blocks, double *p) {
    int frameW = n_blocks;
    int nParamPoints = n_frames * n_blocks + 1;
    for (int frame = 0; frame < n_frames; frame++) {
        int j = frame * frameW + (frame != 0);
        for (int i = (frame != 0); i < frameW; i++, j++) {
            p[j] = px[frame * (frameW + 1) + i];
            p[j + nParamPoints] = py[frame * (frameW + 1) + i];
        }
        if (frame < n_frames - 1) {
            p[j] = (px[frame * (frameW + 1) + frameW] + px[(frame + 1) * (frameW + 1)]) / 2;
            p[j + nParamPoints] = (py[frame * (frameW + 1) + frameW] + py[(frame + 1) * (frameW + 1)]) / 2;
        } else {
            p[j] = px[frame * (frameW + 1) + frameW];
            p[j + nParamPoints] = py[frame * (frameW + 1) + frameW];
        }
    }
}

/* organize1D2(p, px, py) (parameterorganizer.cpp:79-99): a frame's last knot
 * repeats p[f*nb + nb - 1], not the shared end knot. */
void orc_organize_out(const double *p, int n_frames, int n_blocks, double *px, double *py) {
    int frameW = n_blocks, nParamPoints = n_frames * n_blocks + 1, ct = 0;
    for (int f = 0; f < n_frames; f++) {
        int fi = f * frameW, i;
        for (i = fi; i < fi + frameW; i++) {
            px[ct] = p[i];
            py[ct] = p[i + nParamPoints];
            ct++;
        }
        px[ct] = p[i - 1];
        py[ct] = p[i - 1 + nParamPoints];
        ct++;
    }
}

/* ------------------------------------------------------------------ */
/* filters (task/filter/smearingfilter.cpp:24-33, rangefilter.cpp:24-40) */
/* ------------------------------------------------------------------ */
void orc_filter_apply(int filter_id, double strength, double *v, int n) {
    if (filter_id == 1) {
        for (int i = 0; i < n; i++) {
            double pv = (i == 0) ? 0 : v[i - 1];
            v[i] = pv * strength + v[i] * (1 - strength);
        }
    } else if (filter_id == 2) {
        for (int i = 1; i < n - 1; i++) {
            double tc = (v[i - 1] + v[i + 1]) / 2;
            double diff = (tc - v[i - 1]) * strength;
            if (v[i - 1] < v[i + 1]) {
                if (v[i] > v[i + 1] + diff || v[i] < v[i - 1] - diff) v[i] = tc;
            } else {
                if (v[i] < v[i + 1] + diff || v[i] > v[i - 1] - diff) v[i] = tc;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* solvers                                                              */
/* ------------------------------------------------------------------ */

/* Exact symmetric solve of the compacted system (replaces the
 * Eigen::SimplicialLDLT of eigensolve.cpp:35-38).  The compacted index space
 * is [x of used knots..., y of used knots...]; it is permuted to interleaved
 * (x_k, y_k) order, where the matrix is banded, and factored as LDL^T. */
static int solve_exact(int n, int nnz, const int *row, const int *col, const double *val,
                       const double *rhs, double *xout) {
    if (n == 0) return 0;
    int u = n / 2;
    int *perm = (int *)xcalloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; i++) perm[i] = (i < u) ? 2 * i : 2 * (i - u) + 1;
    if (n % 2) { /* cannot happen: x and y share paramsUsed; keep identity */
        for (int i = 0; i < n; i++) perm[i] = i;
    }
    int bw = 0;
    for (int k = 0; k < nnz; k++) {
        if (row[k] == INT_MAX) continue;
        int a = perm[row[k]], b = perm[col[k]];
        int w = a > b ? a - b : b - a;
        if (w > bw) bw = w;
    }
    int ld = bw + 1; /* L[i][bw - (i - j)] for j in [i-bw, i] */
    double *A = (double *)xcalloc((size_t)n * ld, sizeof(double));
    for (int k = 0; k < nnz; k++) {
        if (row[k] == INT_MAX) continue;
        int a = perm[row[k]], b = perm[col[k]];
        if (a >= b) A[(size_t)a * ld + bw - (a - b)] = val[k];
    }
    double *D = (double *)xcalloc((size_t)n, sizeof(double));
    for (int j = 0; j < n; j++) {
        double dj = A[(size_t)j * ld + bw];
        int k0 = j - bw < 0 ? 0 : j - bw;
        for (int k = k0; k < j; k++) {
            double ljk = A[(size_t)j * ld + bw - (j - k)];
            dj -= ljk * ljk * D[k];
        }
        D[j] = dj;
        int i1 = j + bw < n - 1 ? j + bw : n - 1;
        for (int i = j + 1; i <= i1; i++) {
            double aij = A[(size_t)i * ld + bw - (i - j)];
            int kk0 = i - bw < 0 ? 0 : i - bw;
            for (int k = kk0; k < j; k++)
                aij -= A[(size_t)i * ld + bw - (i - k)] * A[(size_t)j * ld + bw - (j - k)] * D[k];
            A[(size_t)i * ld + bw - (i - j)] = aij / dj;
        }
    }
    double *z = (double *)xcalloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; i++) z[perm[i]] = rhs[i];
    for (int i = 0; i < n; i++) {
        int k0 = i - bw < 0 ? 0 : i - bw;
        for (int k = k0; k < i; k++) z[i] -= A[(size_t)i * ld + bw - (i - k)] * z[k];
    }
    for (int i = 0; i < n; i++) z[i] /= D[i];
    for (int i = n - 1; i >= 0; i--) {
        int i1 = i + bw < n - 1 ? i + bw : n - 1;
        for (int k = i + 1; k <= i1; k++) z[i] -= A[(size_t)k * ld + bw - (k - i)] * z[k];
    }
    for (int i = 0; i < n; i++) xout[i] = z[perm[i]];
    free(perm);
    free(A);
    free(D);
    free(z);
    return 0;
}

/* Incomplete LU with zero fill on the compacted x-then-y CSR pattern,
 * followed by unit-lower and upper triangular solves: the arithmetic of
 * cusparse csrilu02 + csrsv2 (cusparsesolver.cu:119-158). */
static int solve_ilu0(int n, int nnz, const int *row, const int *col, const double *val,
                      const double *rhs, double *xout) {
    if (n == 0) return 0;
    int *cnt = (int *)xcalloc((size_t)n + 1, sizeof(int));
    int m = 0;
    for (int k = 0; k < nnz; k++)
        if (row[k] != INT_MAX) { cnt[row[k] + 1]++; m++; }
    for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    int *ci = (int *)xcalloc((size_t)m, sizeof(int));
    double *cv = (double *)xcalloc((size_t)m, sizeof(double));
    int *fill = (int *)xcalloc((size_t)n, sizeof(int));
    for (int k = 0; k < nnz; k++) {
        if (row[k] == INT_MAX) continue;
        int r = row[k], pos = cnt[r] + fill[r]++;
        ci[pos] = col[k];
        cv[pos] = val[k];
    }
    /* sort each row by column (d_sortColumnIndices, task_filldiagonals.cu:12-33) */
    for (int r = 0; r < n && ORC_PLANT != 4; r++) {
        for (int a = cnt[r] + 1; a < cnt[r + 1]; a++) {
            int j = a;
            while (j > cnt[r] && ci[j - 1] > ci[j]) {
                int t = ci[j - 1]; ci[j - 1] = ci[j]; ci[j] = t;
                double tv = cv[j - 1]; cv[j - 1] = cv[j]; cv[j] = tv;
                j--;
            }
        }
    }
    int *diag = (int *)xcalloc((size_t)n, sizeof(int));
    int *where = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) where[i] = -1;
    for (int r = 0; r < n; r++) {
        diag[r] = -1;
        for (int a = cnt[r]; a < cnt[r + 1]; a++)
            if (ci[a] == r) diag[r] = a;
    }
    for (int i = 0; i < n; i++) {
        for (int a = cnt[i]; a < cnt[i + 1]; a++) where[ci[a]] = a;
        for (int a = cnt[i]; a < cnt[i + 1]; a++) {
            int k = ci[a];
            if (k >= i) break;
            const int pk = ORC_PLANT == 5 ? i : k;
            double ukk = diag[pk] >= 0 ? cv[diag[pk]] : 0.0;
            cv[a] = cv[a] / ukk;
            double lik = cv[a];
            for (int b = cnt[k]; b < cnt[k + 1]; b++) {
                int j = ci[b];
                if (j <= k) continue;
                if (where[j] >= 0) cv[where[j]] -= lik * cv[b];
            }
        }
        for (int a = cnt[i]; a < cnt[i + 1]; a++) where[ci[a]] = -1;
    }
    double *z = (double *)xcalloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; i++) {
        double s = rhs[i];
        for (int a = cnt[i]; a < cnt[i + 1] && ci[a] < i; a++) s -= cv[a] * z[ci[a]];
        z[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = z[i];
        for (int a = cnt[i]; a < cnt[i + 1]; a++)
            if (ci[a] > i) s -= cv[a] * xout[ci[a]];
        xout[i] = s / (diag[i] >= 0 ? cv[diag[i]] : 0.0);
    }
    free(cnt); free(ci); free(cv); free(fill); free(diag); free(where); free(z);
    return 0;
}

int orc_solve_coo(int n, int nnz, const int *row, const int *col, const double *val,
                  const double *rhs, double *x, int solver_id) {
    if (solver_id == 0) return solve_ilu0(n, nnz, row, col, val, rhs, x);
    return solve_exact(n, nnz, row, col, val, rhs, x);
}

/* ------------------------------------------------------------------ */
/* per-pixel stages                                                     */
/* ------------------------------------------------------------------ */

/* d_applyParameters for one block column (task_applyparameters.cuh:15-63).
 * Returns 1 when the column counts as present. */
static int apply_column(orc_state *st, int idx, double px, double px1, double py, double py1) {
    int idxm = idx % st->frameW;
    double templateW = (double)(st->templateW - 1);
    double templateH = (double)(st->templateH - 1);
    int pxCount = 0;
    for (int row = 0; row < st->frameH; row++) {
        size_t ipxm = IDX2R(row, idxm, st->frameW);
        size_t ipx = IDX2R(row, idx, st->fframeW);
        real frac = st->frac[ipxm];
        real x = st->baseX[ipxm] + (real)px * (1 - frac) + (real)px1 * frac;
        real y = st->baseY[ipxm] + (real)py * (1 - frac) + (real)py1 * frac;
        int m = st->baseMask[ipxm] && x >= 0 && x < templateW - st->eps && y >= 0 &&
                y < templateH - st->eps;
        pxCount += m;
        st->mask[ipx] = (uint8_t)m;
        st->x[ipx] = x;
        st->y[ipx] = y;
    }
    return pxCount >= st->minPoints[idxm];
}

/* d_createDifferenceImage for one column (task_differenceimage.cuh:13-32). */
static void diff_column(orc_state *st, int idx) {
    int W = st->templateW;
    for (int row = 0; row < st->frameH; row++) {
        size_t rid = IDX2R(row, idx, st->fframeW);
        if (st->mask[rid]) {
            int iX = (int)st->x[rid];
            int iY = (int)st->y[rid];
            size_t iXY = IDX2R(iY, iX, W);
            real xFrac = st->x[rid] - iX;
            real yFrac = st->y[rid] - iY;
            const real *T = st->tmpl;
            real wI = (1 - xFrac) * ((1 - yFrac) * T[iXY] + yFrac * T[iXY + W]) +
                        xFrac * ((1 - yFrac) * T[iXY + 1] + yFrac * T[iXY + W + 1]);
            st->d[rid] = st->image[rid] - wI;
        } else {
            st->d[rid] = 0;
        }
    }
}

/* d_calcBaseErrval summed over a frame's columns (task_calcerrval.cu:7-19). */
static double frame_errval(orc_state *st, int f) {
    double e = 0;
    long cnt = 0;
    for (int t = 0; t < st->frameW; t++) {
        int idx = f * st->frameW + t;
        real ec = 0;
        int cc = 0;
        for (int row = 0; row < st->frameH; row++) {
            size_t i = IDX2R(row, idx, st->fframeW);
            ec += st->d[i] * st->d[i];
            cc += st->mask[i] ? 1 : 0;
        }
        e += ec;
        cnt += cc;
    }
    return e / (double)cnt;
}

/* hd_applyParameters over all columns with knots p (task_applyparameters.cu:5-16). */
static void apply_all(orc_state *st, const double *p) {
    int N = st->nParamPoints;
#pragma omp parallel for num_threads(st->nthreads) schedule(static)
    for (int idx = 0; idx < st->fframeW; idx++) {
        if (apply_column(st, idx, p[idx], p[idx + 1], p[idx + N], p[idx + N + 1]))
            st->blocksPresent[idx] = 1;
    }
}

/* d_calcAllGroupErrval (task_calcerrval.cu:22-56).  All groups are active. */
static void calc_group_errvals(orc_state *st, int shift) {
    int N = st->nParamPoints;
#pragma omp parallel for num_threads(st->nthreads) schedule(static)
    for (int idx = 0; idx < st->fframeW; idx++) {
        /* d_createParameterMask (task_createparametermask.cuh:12-20) */
        if (idx == 0) st->paramsUsed[idx] = st->blocksPresent[idx];
        else if (idx == N - 1) st->paramsUsed[idx] = st->blocksPresent[idx - 1];
        else st->paramsUsed[idx] = st->blocksPresent[idx] | st->blocksPresent[idx - 1];
        diff_column(st, idx);
    }
    for (int f = 0; f < st->nFrames; f++) st->errval[shift * st->nFrames + f] = frame_errval(st, f);
}

/* d_updateGradients (task_updategradients.cu:4-35). */
static void update_gradients(orc_state *st) {
    int W = st->templateW;
#pragma omp parallel for num_threads(st->nthreads) schedule(static)
    for (int idx = 0; idx < st->fframeW; idx++) {
        for (int row = 0; row < st->frameH; row++) {
            size_t i = IDX2R(row, idx, st->fframeW);
            if (st->mask[i]) {
                int iX = (int)st->x[i];
                int iY = (int)st->y[i];
                size_t t = IDX2R(iY, iX, W);
                real xFrac = st->x[i] - (real)iX;
                real yFrac = st->y[i] - (real)iY;
                st->wx[i] = (1 - yFrac) * st->xg[t] + yFrac * st->xg[t + W];
                st->wy[i] = (1 - xFrac) * st->yg[t] + xFrac * st->yg[t + 1];
            } else {
                st->wx[i] = 0;
                st->wy[i] = 0;
            }
        }
    }
}

/* hd_generateSubSparseOffsets (task_generatesubsparseoffset.cu:9-27). */
static void sub_sparse_offsets(orc_state *st) {
    int used = 0, off = 0;
    for (int i = 0; i < st->nParams; i++) {
        if (st->paramsUsed[i % st->nParamPoints]) used++;
        else off++;
        st->subOff[i] = off;
    }
    st->nParamsUsed = used;
}

/* d_insertValue (task_filldiagonals.cu:36-45). */
static void insert_value(orc_state *st, int nzId, double value, int row, int col) {
    int N = st->nParamPoints;
    int mrow = row >= N ? row - N : row;
    int mcol = col >= N ? col - N : col;
    if (st->paramsUsed[mrow] && st->paramsUsed[mcol]) {
        st->Hval[nzId] = value;
        st->Hrow[nzId] = row - st->subOff[row];
        st->Hcol[nzId] = col - st->subOff[col];
    }
}

/* hd_fillDiagonals / d_fillDiagonals (task_filldiagonals.cu:47-159, 173-182),
 * including the block c-1 diagonal term that reads frac column idx % frameW. */
static void fill_diagonals(orc_state *st) {
    int N = st->nParamPoints, nnz = st->nParams * 6;
    for (int i = 0; i < nnz; i++) st->Hrow[i] = INT_MAX;
#pragma omp parallel for num_threads(st->nthreads) schedule(static)
    for (int idx = 0; idx < N; idx++) {
        double *dbg = st->hdbg + (size_t)idx * 6;
        for (int k = 0; k < 6; k++) dbg[k] = 0;
        if (!st->paramsUsed[idx]) continue;
        real offx = 0, offy = 0, offxy = 0, diagx = 0, diagy = 0, diagxy = 0;
        int lastParam = (idx == st->fframeW);
        int firstParam = (idx == 0);
        long idxy = idx - lastParam;
        long idxy_1 = idx - 1;
        int idxm = idx % st->frameW;
        for (int row = 0; row < st->frameH; row++) {
            real frac = st->frac[IDX2R(row, idxm, st->frameW)];
            real fracsq = frac * frac;
            real compsq = (1 - frac) * (1 - frac);
            real fcprod = frac - fracsq;
            real wxgradsq = st->wx[idxy] * st->wx[idxy];
            real wygradsq = st->wy[idxy] * st->wy[idxy];
            real wxygrad = st->wx[idxy] * st->wy[idxy];
            offx += wxgradsq * fcprod;
            offy += wygradsq * fcprod;
            offxy += wxygrad * fcprod;
            switch (firstParam + (lastParam << 1)) {
            case 0:
                diagx += wxgradsq * compsq + st->wx[idxy_1] * st->wx[idxy_1] * fracsq;
                diagy += wygradsq * compsq + st->wy[idxy_1] * st->wy[idxy_1] * fracsq;
                diagxy += wxygrad * compsq + st->wx[idxy_1] * st->wy[idxy_1] * fracsq;
                break;
            case 1:
                diagx += wxgradsq * compsq;
                diagy += wygradsq * compsq;
                diagxy += wxygrad * compsq;
                break;
            case 2:
                diagx += wxgradsq * fracsq;
                diagy += wygradsq * fracsq;
                diagxy += wxygrad * fracsq;
                break;
            }
            idxy += st->fframeW;
            idxy_1 += st->fframeW;
        }
        dbg[0] = diagx; dbg[1] = diagy; dbg[2] = diagxy;
        if (!lastParam) {
            if (st->paramsUsed[idx + 1]) { dbg[3] = offx; dbg[4] = offy; dbg[5] = offxy; }
            int nz = idx;
            insert_value(st, nz, diagx, idx, idx); nz += N;
            insert_value(st, nz, offx, idx + 1, idx); nz += N;
            insert_value(st, nz, offx, idx, idx + 1); nz += N;
            insert_value(st, nz, diagy, idx + N, idx + N); nz += N;
            insert_value(st, nz, offy, idx + N + 1, idx + N); nz += N;
            insert_value(st, nz, offy, idx + N, idx + N + 1); nz += N;
            insert_value(st, nz, diagxy, idx, idx + N); nz += N;
            insert_value(st, nz, offxy, idx, idx + N + 1); nz += N;
            insert_value(st, nz, offxy, idx + 1, idx + N); nz += N;
            insert_value(st, nz, diagxy, idx + N, idx); nz += N;
            insert_value(st, nz, offxy, idx + N + 1, idx); nz += N;
            insert_value(st, nz, offxy, idx + N, idx + 1);
        } else {
            int nz = idx;
            insert_value(st, nz, diagx, idx, idx); nz += N;
            insert_value(st, nz, diagy, idx + N, idx + N); nz += N;
            insert_value(st, nz, diagxy, idx + N, idx); nz += N;
            insert_value(st, nz, diagxy, idx, idx + N);
        }
    }
}

/* d_fillJTr (task_filljtr.cu:4-58). */
static void fill_jtr(orc_state *st) {
    int N = st->nParamPoints;
#pragma omp parallel for num_threads(st->nthreads) schedule(static)
    for (int idx0 = 0; idx0 < N; idx0++) {
        int idx = idx0;
        if (!st->paramsUsed[idx]) continue;
        real sumx = 0, sumy = 0, sumx2 = 0, sumy2 = 0;
        int idxm = idx % st->frameW;
        if (idx != st->fframeW) {
            for (int row = 0; row < st->frameH; row++) {
                size_t idxy = IDX2R(row, idx, st->fframeW);
                real frac = st->frac[IDX2R(row, idxm, st->frameW)];
                real fdi = st->d[idxy] * (1 - frac);
                sumx += fdi * st->wx[idxy];
                sumy += fdi * st->wy[idxy];
            }
        }
        if (idx != 0) {
            if (idxm == 0) idxm = st->frameW - 1;
            else idxm -= 1;
            for (int row = 0; row < st->frameH; row++) {
                size_t idxy = IDX2R(row, idx - 1, st->fframeW);
                real frac = st->frac[IDX2R(row, idxm, st->frameW)];
                real fdi = st->d[idxy] * frac;
                sumx2 += fdi * st->wx[idxy];
                sumy2 += fdi * st->wy[idxy];
            }
        }
        st->jtr[idx - st->subOff[idx]] = sumx + sumx2;
        idx += N;
        st->jtr[idx - st->subOff[idx]] = sumy + sumy2;
    }
}

/* Smoothness extension (lambda, unused by the reference: steady_scan.m:207-208):
 * adds lambda * (delta_{c+1} - delta_c)^2 on each axis for adjacent used
 * knots.  The COO slots are those written by d_fillDiagonals: slot c + k*N
 * holds entry k of knot c (task_filldiagonals.cu:109-144). */
static void add_lambda(orc_state *st) {
    double lam = st->s.lambda;
    if (lam == 0) return;
    int N = st->nParamPoints;
    for (int c = 0; c + 1 < st->fframeW; c++) {
        if (!st->paramsUsed[c] || !st->paramsUsed[c + 1]) continue;
        st->Hval[c] += lam;              /* (x_c, x_c)         */
        st->Hval[c + 1] += lam;          /* (x_c+1, x_c+1)     */
        st->Hval[c + N] -= lam;          /* (x_c+1, x_c)       */
        st->Hval[c + 2 * N] -= lam;      /* (x_c, x_c+1)       */
        st->Hval[c + 3 * N] += lam;      /* (y_c, y_c)         */
        st->Hval[c + 1 + 3 * N] += lam;  /* (y_c+1, y_c+1)     */
        st->Hval[c + 4 * N] -= lam;      /* (y_c+1, y_c)       */
        st->Hval[c + 5 * N] -= lam;      /* (y_c, y_c+1)       */
    }
}

/* SolverEigen::solve (eigensolver.cu:38-58) or SolverCuSparse::solve: exact or
 * ILU(0) solve followed by the solver's filter on the x and y halves. */
static void solve_and_filter(orc_state *st) {
    int n = st->nParamsUsed;
    add_lambda(st);
    orc_solve_coo(n, st->nParams * 6, st->Hrow, st->Hcol, st->Hval, st->jtr, st->newDeltaP,
                  st->s.solver_id);
    if (st->solverFilter) {
        orc_filter_apply(st->solverFilter, st->s.filter_strength, st->newDeltaP, n / 2);
        orc_filter_apply(st->solverFilter, st->s.filter_strength, st->newDeltaP + n / 2, n / 2);
    }
}

/* d_expandResults (task_expandresults.cu:10-41). */
static void expand_results(orc_state *st, int shift) {
    int N = st->nParamPoints, G = st->groupSize;
    for (int idx = 0; idx < st->nParams; idx++) {
        int j = idx >= N ? idx - N : idx;
        int isFixed = ((((idx + 1) - (shift * (G >> 1))) % G) == 0);
        if (isFixed) continue; /* keep the previous delta */
        if (st->paramsUsed[j]) st->deltaP[idx] = st->newDeltaP[idx - st->subOff[idx]];
        else st->deltaP[idx] = 0;
    }
}

/* One frame of d_linesearch (task_linesearch.cu:35-106), with the pinned
 * lockstep semantics for the shared right knot.  p0/d0 are snapshots taken
 * at line-search start. */
static long linesearch_frame(orc_state *st, int shift, int f, const double *p0, const double *d0) {
    int nb = st->frameW, N = st->nParamPoints, base = f * nb;
    int last = (f == st->nFrames - 1);
    double r = st->s.linesearch_reduction_multiplier, thr = st->s.move_threshold;
    double drx = last ? 0 : d0[base + nb], dry = last ? 0 : d0[base + nb + N];
    double *kx = (double *)malloc(sizeof(double) * (size_t)(nb + 1));
    double *ky = (double *)malloc(sizeof(double) * (size_t)(nb + 1));
    long tries = 0;
    for (;;) {
        /* d_addDeltaP for own knots; right knot from frame f+1 (lockstep) or 1e6 */
        for (int t = 0; t < nb; t++) {
            kx[t] = p0[base + t] + st->deltaP[base + t];
            ky[t] = p0[base + t + N] + st->deltaP[base + t + N];
        }
        kx[nb] = last ? 1e6 : p0[base + nb] + drx;
        ky[nb] = last ? 1e6 : p0[base + nb + N] + dry;
        int present = 0;
        for (int t = 0; t < nb; t++) {
            if (apply_column(st, base + t, kx[t], kx[t + 1], ky[t], ky[t + 1])) {
                st->blocksPresent[base + t] = 1;
                present = 1;
            }
        }
        tries++;
        if (!present) {
            int moved = 0;
            for (int t = 0; t < nb; t++) {
                st->deltaP[base + t] *= r;
                st->deltaP[base + t + N] *= r;
            }
            drx *= r;
            dry *= r;
            for (int t = 0; t < nb; t++)
                if (fabs(st->deltaP[base + t]) > thr || fabs(st->deltaP[base + t + N]) > thr) moved = 1;
            if (!moved) break;
            continue;
        }
        for (int t = 0; t < nb; t++) diff_column(st, base + t);
        double errvalNew = frame_errval(st, f);
        if (errvalNew < st->errval[shift * st->nFrames + f]) {
            for (int t = 0; t < nb; t++) {
                st->p[base + t] = kx[t];
                st->p[base + t + N] = ky[t];
            }
            if (f == 0) st->errval[shift * st->nFrames] = errvalNew;
            break;
        }
        int moved = 0;
        for (int t = 0; t < nb; t++)
            if (fabs(st->deltaP[base + t]) > thr || fabs(st->deltaP[base + t + N]) > thr) moved = 1;
        if (!moved) break;
        for (int t = 0; t < nb; t++) {
            st->deltaP[base + t] *= r;
            st->deltaP[base + t + N] *= r;
        }
        drx *= r;
        dry *= r;
    }
    free(kx);
    free(ky);
    return tries;
}

/* hd_linesearch (task_linesearch.cu:108-126). */
static int linesearch(orc_state *st, int shift) {
    double *p0 = (double *)malloc(sizeof(double) * (size_t)st->nParams);
    double *d0 = (double *)malloc(sizeof(double) * (size_t)st->nParams);
    memcpy(p0, st->p, sizeof(double) * (size_t)st->nParams);
    memcpy(d0, st->deltaP, sizeof(double) * (size_t)st->nParams);
    long tries = 0;
#pragma omp parallel for num_threads(st->nthreads) schedule(dynamic, 1) reduction(+ : tries)
    for (int f = 0; f < st->nFrames; f++) tries += linesearch_frame(st, shift, f, p0, d0);
    st->lsTries += tries;
    free(p0);
    free(d0);
    double evSum = 0, evSum1 = 0;
    for (int i = 0; i < st->nFrames; i++) {
        evSum += st->errval[i];
        evSum1 += st->errval[i + st->nFrames];
    }
    return (shift && evSum > evSum1) || (!shift && evSum < evSum1);
}

/* hd_globalCorrelation (task_correlation.cu:63-99), including the mean swap,
 * the leading-masked-run walk and the partial swap-back. */
double orc_global_correlation(orc_state *st) {
    size_t numPixel = (size_t)st->frameH * st->fframeW;
    double *qcImg = (double *)malloc(sizeof(double) * numPixel);
    for (size_t i = 0; i < numPixel; i++) qcImg[i] = st->mask[i] ? st->image[i] : 0;
    for (size_t i = 0; i < numPixel; i++)
        if (st->mask[i]) st->d[i] = st->image[i] - st->d[i];
    double divSumA = 0, divSumB = 0;
    long n = 0;
    for (size_t i = 0; i < numPixel; i++) {
        divSumA += st->image[i];
        divSumB += st->d[i];
        n += st->mask[i] ? 1 : 0;
    }
    divSumA /= (double)n;
    divSumB /= (double)n;
    double assq = 0, bssq = 0, q = 0;
    for (int idx = 0; idx < st->fframeW; idx++) {
        double a = 0, b = 0, qq = 0;
        size_t pxId = (size_t)idx;
        for (int i = 0; i < st->frameH; i++) {
            if (st->mask[pxId]) {
                double va = st->d[pxId] - divSumA;
                a += va * va;
                double vb = qcImg[pxId] - divSumB;
                b += vb * vb;
                qq += va * vb;
                pxId += (size_t)st->fframeW;
            }
        }
        assq += a;
        bssq += b;
        q += qq;
    }
    double corr = q / sqrt(assq * bssq);
    size_t swapBack = (size_t)((st->fframeW + 255) / 256) * 256;
    if (swapBack > numPixel || ORC_PLANT == 8) swapBack = numPixel;
    for (size_t i = 0; i < swapBack; i++)
        if (st->mask[i]) st->d[i] = st->image[i] - st->d[i];
    free(qcImg);
    return corr;
}

/* hd_frameCorrelation (task_correlation.cu:101-183). */
static void frame_correlation(orc_state *st) {
    size_t numPixel = (size_t)st->frameH * st->fframeW;
    for (size_t i = 0; i < numPixel && ORC_PLANT != 7; i++)
        if (st->mask[i]) st->d[i] = st->image[i] - st->d[i];
    for (int f = 0; f < st->nFrames; f++) {
        double sa = 0, sb = 0;
        long sc = 0;
        for (int t = 0; t < st->frameW; t++) {
            int idx = f * st->frameW + t;
            for (int i = 0; i < st->frameH; i++) {
                size_t px = IDX2R(i, idx, st->fframeW);
                if (st->mask[px]) {
                    sc++;
                    sb += st->image[px];
                    sa += st->d[px];
                }
            }
        }
        const double cnt = ORC_PLANT == 6 ? (double)st->frameH * st->frameW : (double)sc;
        double mA = sa / cnt, mB = sb / cnt;
        double assq = 0, bssq = 0, q = 0;
        for (int t = 0; t < st->frameW; t++) {
            int idx = f * st->frameW + t;
            for (int i = 0; i < st->frameH; i++) {
                size_t px = IDX2R(i, idx, st->fframeW);
                double on = st->mask[px] ? 1.0 : 0.0;
                double va = (st->d[px] - mA) * on;
                double vb = (st->image[px] - mB) * on;
                assq += va * va;
                bssq += vb * vb;
                q += va * vb;
            }
        }
        st->corr[f] = q / sqrt(assq * bssq);
    }
    for (size_t i = 0; i < numPixel && ORC_PLANT != 7; i++)
        if (st->mask[i]) st->d[i] = st->image[i] - st->d[i];
}

void orc_frame_correlation(orc_state *st, double *corr) {
    frame_correlation(st);
    if (corr) memcpy(corr, st->corr, sizeof(double) * (size_t)st->nFrames);
}

/* ------------------------------------------------------------------ */
/* state construction                                                   */
/* ------------------------------------------------------------------ */
static int validate(const orc_in *in, const orc_settings *s) {
    if (!in || !s) return ORC_INVALID_ARGUMENT;
    if (in->frame_h < 1 || in->n_blocks < 1 || in->n_frames < 1) return ORC_INVALID_ARGUMENT;
    if (in->template_h < 2 || in->template_w < 2) return ORC_INVALID_ARGUMENT;
    if (in->group_size < 2 || in->group_size > 1024) return ORC_INVALID_ARGUMENT;
    if (s->max_iterations < 1) return ORC_INVALID_ARGUMENT;
    if (s->solver_id < 0 || s->solver_id > 2) return ORC_INVALID_ARGUMENT;
    int fsel = s->n_filter_iterations != 0 ? s->filter_id : 0;
    if (fsel < 0 || fsel > 2) return ORC_INVALID_ARGUMENT;
    return ORC_OK;
}

void orc_state_destroy(orc_state *st) {
    if (!st) return;
    void *ptrs[] = {st->tmpl, st->xg, st->yg, st->frac, st->baseX, st->baseY, st->image,
                    st->baseMask, st->minPoints, st->p, st->pLast, st->deltaP, st->newDeltaP,
                    st->jtr, st->x, st->y, st->d, st->wx, st->wy, st->mask, st->paramsUsed,
                    st->blocksPresent, st->subOff, st->Hval, st->Hrow, st->Hcol, st->errval,
                    st->corr, st->hdbg};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(st);
}

int orc_state_create(const orc_in *in, const orc_settings *s, orc_state **out) {
    int rc = validate(in, s);
    if (rc) return rc;
    orc_state *st = (orc_state *)xcalloc(1, sizeof(orc_state));
    if (!st) return ORC_OUT_OF_MEMORY;
    st->s = *s;
    st->nthreads = 1;
    /* derived sizes (mexgateway.cpp:176-198) */
    st->frameH = in->frame_h;
    st->frameW = in->n_blocks;
    st->nFrames = in->n_frames;
    st->fframeW = st->frameW * st->nFrames;
    st->nParamPoints = st->fframeW + 1;
    st->nParams = 2 * st->nParamPoints;
    st->templateH = in->template_h;
    st->templateW = in->template_w;
    st->groupSize = in->group_size;
    st->nGroups = st->fframeW / st->groupSize;
    st->eps = (s->precision == 'S') ? (double)FLT_EPSILON : DBL_EPSILON;
    size_t tpix = (size_t)st->templateH * st->templateW;
    size_t gpix = (size_t)st->frameH * st->frameW;
    size_t npix = (size_t)st->frameH * st->fframeW;
    st->tmpl = (real *)malloc(sizeof(real) * tpix);
    st->xg = (real *)malloc(sizeof(real) * tpix);
    st->yg = (real *)malloc(sizeof(real) * tpix);
    st->frac = (real *)malloc(sizeof(real) * gpix);
    st->baseX = (real *)malloc(sizeof(real) * gpix);
    st->baseY = (real *)malloc(sizeof(real) * gpix);
    st->baseMask = (uint8_t *)malloc(gpix);
    st->image = (real *)malloc(sizeof(real) * npix);
    st->minPoints = (int *)malloc(sizeof(int) * (size_t)(st->frameW + 1));
    st->p = (double *)xcalloc((size_t)st->nParams, sizeof(double));
    st->pLast = (double *)xcalloc((size_t)st->nParams, sizeof(double));
    st->deltaP = (double *)xcalloc((size_t)st->nParams, sizeof(double));
    st->newDeltaP = (double *)xcalloc((size_t)st->nParams, sizeof(double));
    st->jtr = (double *)xcalloc((size_t)st->nParams, sizeof(double));
    st->x = (real *)xcalloc(npix, sizeof(real));
    st->y = (real *)xcalloc(npix, sizeof(real));
    st->d = (real *)xcalloc(npix, sizeof(real));
    st->wx = (real *)xcalloc(npix, sizeof(real));
    st->wy = (real *)xcalloc(npix, sizeof(real));
    st->mask = (uint8_t *)xcalloc(npix, 1);
    st->paramsUsed = (uint8_t *)xcalloc((size_t)st->nParamPoints, 1);
    st->blocksPresent = (uint8_t *)xcalloc((size_t)st->nParamPoints, 1);
    st->subOff = (int *)xcalloc((size_t)st->nParams, sizeof(int));
    st->Hval = (double *)xcalloc((size_t)st->nParams * 6, sizeof(double));
    st->Hrow = (int *)xcalloc((size_t)st->nParams * 6, sizeof(int));
    st->Hcol = (int *)xcalloc((size_t)st->nParams * 6, sizeof(int));
    st->errval = (double *)xcalloc((size_t)st->nFrames * 2, sizeof(double));
    st->corr = (double *)xcalloc((size_t)st->nFrames, sizeof(double));
    st->hdbg = (double *)xcalloc((size_t)st->nParamPoints * 6, sizeof(double));
    if (!st->tmpl || !st->xg || !st->yg || !st->frac || !st->baseX || !st->baseY ||
        !st->baseMask || !st->image || !st->minPoints || !st->x || !st->y || !st->d ||
        !st->wx || !st->wy || !st->mask || !st->Hval || !st->Hrow || !st->Hcol) {
        orc_state_destroy(st);
        return ORC_OUT_OF_MEMORY;
    }
    /* column-major (MATLAB) -> row-major (mexconverter.cpp:223-285) */
    for (int iy = 0; iy < st->templateH; iy++)
        for (int ix = 0; ix < st->templateW; ix++) {
            size_t cm = (size_t)ix * st->templateH + iy, rm = IDX2R(iy, ix, st->templateW);
            st->tmpl[rm] = in->template_image[cm];
            st->xg[rm] = in->xgrad[cm];
            st->yg[rm] = in->ygrad[cm];
        }
    for (int r = 0; r < st->frameH; r++)
        for (int b = 0; b < st->frameW; b++) {
            size_t cm = (size_t)b * st->frameH + r, rm = IDX2R(r, b, st->frameW);
            st->frac[rm] = in->frac[cm];
            st->baseX[rm] = in->base_x[cm];
            st->baseY[rm] = in->base_y[cm];
            st->baseMask[rm] = in->base_mask[cm] ? 1 : 0;
        }
    for (int c = 0; c < st->fframeW; c++)
        for (int r = 0; r < st->frameH; r++)
            st->image[IDX2R(r, c, st->fframeW)] = in->data[(size_t)c * st->frameH + r];
    for (int b = 0; b <= st->frameW; b++) st->minPoints[b] = in->min_points[b];
    /* filter and solver selection (task.cpp:54-96) */
    st->filterID = s->filter_id;
    st->nFilterIterations = s->n_filter_iterations;
    switch (s->n_filter_iterations != 0 ? s->filter_id : 0) {
    case 0: st->solverFilter = 0; st->nFilterIterations = 0; break;
    case 1: st->solverFilter = 1; break;
    case 2: st->solverFilter = 2; break;
    }
    /* init values (task_allocate.cu:124-184) */
    orc_organize_in(in->init_x, in->init_y, st->nFrames, st->frameW, st->p);
    for (int i = 0; i < st->nFrames * 2; i++) st->errval[i] = 1e6;
    *out = st;
    return ORC_OK;
}

void orc_state_get(const orc_state *st, double *p, double *delta_p, double *errval) {
    if (p) memcpy(p, st->p, sizeof(double) * (size_t)st->nParams);
    if (delta_p) memcpy(delta_p, st->deltaP, sizeof(double) * (size_t)st->nParams);
    if (errval) memcpy(errval, st->errval, sizeof(double) * (size_t)st->nFrames * 2);
}
void orc_state_pixels(const orc_state *st, uint8_t *mask, double *x, double *y, double *d) {
    for (int c = 0; c < st->fframeW; c++)
        for (int r = 0; r < st->frameH; r++) {
            size_t cm = (size_t)c * st->frameH + r, rm = IDX2R(r, c, st->fframeW);
            if (mask) mask[cm] = st->mask[rm];
            if (x) x[cm] = st->x[rm];
            if (y) y[cm] = st->y[rm];
            if (d) d[cm] = st->d[rm];
        }
}
void orc_state_set_p(orc_state *st, const double *p) {
    memcpy(st->p, p, sizeof(double) * (size_t)st->nParams);
}
void orc_state_set_delta(orc_state *st, const double *delta_p) {
    memcpy(st->deltaP, delta_p, sizeof(double) * (size_t)st->nParams);
}

/* The half-iteration stages up to the expand (mainiteration.cpp:66-98). */
static void half_iteration_front(orc_state *st, int shift) {
    memset(st->blocksPresent, 0, (size_t)st->nParamPoints);
    apply_all(st, st->p);
    calc_group_errvals(st, shift);
    update_gradients(st);
    sub_sparse_offsets(st);
    fill_diagonals(st);
    fill_jtr(st);
}

int orc_half_iteration_assemble(orc_state *st, int shift, uint8_t *used, double *jtr_full,
                                double *h_full, double *delta_new) {
    int N = st->nParamPoints;
    half_iteration_front(st, shift);
    if (used) memcpy(used, st->paramsUsed, (size_t)N);
    if (jtr_full) {
        for (int i = 0; i < st->nParams; i++) {
            int j = i >= N ? i - N : i;
            jtr_full[i] = st->paramsUsed[j] ? st->jtr[i - st->subOff[i]] : 0.0;
        }
    }
    if (h_full) memcpy(h_full, st->hdbg, sizeof(double) * (size_t)N * 6);
    solve_and_filter(st);
    expand_results(st, shift);
    if (delta_new) memcpy(delta_new, st->deltaP, sizeof(double) * (size_t)st->nParams);
    return st->nParamsUsed;
}

int orc_half_iteration_linesearch(orc_state *st, int shift) { return linesearch(st, shift); }

/* MainIteration::execute (mainiteration.cpp:61-119); rg, if not NULL,
 * receives p after each half-iteration (timedmainiteration.cpp:124). */
static int main_iteration(orc_state *st, int iteration, double *rg) {
    int improved = 1;
    for (int shift = 0; shift < 2; shift++) {
        half_iteration_front(st, shift);
        if (st->nFilterIterations != 0 && iteration > st->nFilterIterations) {
            st->solverFilter = 0;
            st->filterID = 0;
            st->nFilterIterations = 0;
        }
        solve_and_filter(st);
        expand_results(st, shift);
        improved &= linesearch(st, shift);
        st->halfIterations++;
        if (rg && iteration <= st->s.max_iterations)
            memcpy(rg + (size_t)st->nParams * (size_t)(shift + 2 * (iteration - 1)), st->p,
                   sizeof(double) * (size_t)st->nParams);
    }
    if (!improved) {
        if (st->nFilterIterations != 0) {
            st->solverFilter = 0;
            st->nFilterIterations = 0;
            return main_iteration(st, iteration, rg);
        }
        return 0;
    }
    return 1;
}

/* Task::alignFrames + getParameters (task.cpp:201-219, 259-412). */
int orc_align_group(const orc_in *in, const orc_settings *s, orc_out *out, int n_threads) {
    orc_state *st = NULL;
    int rc = orc_state_create(in, s, &st);
    if (rc) return rc;
    st->nthreads = n_threads > 0 ? n_threads : 1;
    double correlation = 0, corr = 0;
    int iteration = 0, retv = ORC_STOP_ERROR;
    for (;;) {
        iteration++;
        int improved = main_iteration(st, iteration, out->rg_params);
        if (out->all_params && iteration <= s->max_iterations)
            memcpy(out->all_params + (size_t)st->nParams * (iteration - 1), st->p,
                   sizeof(double) * (size_t)st->nParams);
        corr = orc_global_correlation(st);
        if (out->frame_corr_iter) {
            /* on the state the global correlation left (its partial swap-back) */
            frame_correlation(st);
            if (iteration <= s->max_iterations)
                memcpy(out->frame_corr_iter + (size_t)st->nFrames * (iteration - 1), st->corr,
                       sizeof(double) * (size_t)st->nFrames);
        }
        if (out->iteration_corr && iteration <= s->max_iterations) out->iteration_corr[iteration - 1] = corr;
        if (corr - correlation < s->correlation_increase_threshold) {
            retv = ORC_STOP_CORRELATION_INCREASE;
            break;
        } else if (corr - correlation < s->filter_correlation_threshold) {
            if (st->filterID != 0) {
                st->solverFilter = 0;
                st->nFilterIterations = 0;
                st->filterID = 0;
                if (iteration >= s->max_iterations) { /* pinned: no run past max_iterations */
                    retv = ORC_STOP_ITERATION_LIMIT;
                    break;
                }
                continue;
            }
        }
        if (!improved) {
            retv = ORC_STOP_MOVEMENT;
            break;
        } else if (s->max_iterations == iteration) {
            retv = ORC_STOP_ITERATION_LIMIT;
            break;
        } else if (correlation >= s->halt_correlation) {
            retv = ORC_STOP_HALT_CORRELATION;
            break;
        }
        correlation = corr;
        memcpy(st->pLast, st->p, sizeof(double) * (size_t)st->nParams);
        if (out->iteration_errvals && iteration <= s->max_iterations) {
            const size_t ld = 2 * (size_t)s->max_iterations, s0 = ORC_PLANT == 9 ? 1 : 0;
            for (int f = 0; f < st->nFrames; f++) {
                out->iteration_errvals[(size_t)f * ld + 2 * (size_t)(iteration - 1) + s0] = st->errval[f];
                out->iteration_errvals[(size_t)f * ld + 2 * (size_t)(iteration - 1) + 1 - s0] = st->errval[f + st->nFrames];
            }
        }
    }
    if (iteration == 1) memcpy(st->pLast, st->p, sizeof(double) * (size_t)st->nParams);
    calc_group_errvals(st, 0);
    if (!out->frame_corr_iter) frame_correlation(st);
    out->iterations = iteration;
    out->stop_criterion = retv;
    out->half_iterations = st->halfIterations;
    out->linesearch_tries = (int32_t)st->lsTries;
    for (int f = 0; f < st->nFrames; f++) {
        if (out->frame_corr) out->frame_corr[f] = st->corr[f];
        if (out->residrms) out->residrms[f] = sqrt(st->errval[f]);
    }
    if (out->mask) {
        for (int c = 0; c < st->fframeW; c++)
            for (int r = 0; r < st->frameH; r++)
                out->mask[(size_t)c * st->frameH + r] = st->mask[IDX2R(r, c, st->fframeW)];
    }
    if (out->knots_x && out->knots_y)
        orc_organize_out(st->pLast, st->nFrames, st->frameW, out->knots_x, out->knots_y);
    orc_state_destroy(st);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* geometry: init_settings (steady_scan.m:748-828)                     */
/* ------------------------------------------------------------------ */

/* histc bin (MATLAB): k with edges[k] <= v < edges[k+1], v == edges[end] -> last. 1-based; 0 if outside. */
static int histc_bin(double v, const double *edges, int n) {
    if (!(v >= edges[0]) || v > edges[n - 1]) return 0;
    if (v == edges[n - 1]) return n;
    int lo = 0, hi = n - 1; /* edges[lo] <= v < edges[hi] */
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (edges[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo + 1;
}

/* MATLAB linspace(d1, d2, n): d1 + (0:n-1).*(d2-d1)/(n-1), ends forced to d1, d2. */
static double linspace_at(double d1, double d2, int n, int i) {
    if (n == 1) return d2;
    if (i == 0) return d1;
    if (i == n - 1) return d2;
    return d1 + ((double)i * (d2 - d1)) / (double)(ORC_PLANT == 3 ? n : n - 1);
}

int orc_geometry(const orc_geometry_in *g, int32_t *prevknotind, double *tfrac, double *basex,
                 double *basey, int32_t *hc_out, int32_t *wc_out) {
    int H = g->nlines, W = g->linewidth;
    const int32_t *cr = g->cropping; /* [left right bottom top] */
    int Hc = H - (cr[2] + cr[3]), Wc = W - (cr[0] + cr[1]);
    if (Hc <= 0 || Wc <= 0 || g->nknots < 2) return ORC_INVALID_ARGUMENT;
    if (cr[2] + cr[3] != 0) return ORC_INVALID_ARGUMENT; /* default basey is mis-sized (steady_scan.m:799,822-824) */
    if (hc_out) *hc_out = Hc;
    if (wc_out) *wc_out = Wc;
    if (!prevknotind) return ORC_OK;
    double sr[4];
    if (g->has_subrec) memcpy(sr, g->subrec, sizeof(sr));
    else if (Hc == g->template_h && Wc == g->template_w) {
        sr[0] = 1 - cr[0]; sr[1] = g->template_w + cr[1];
        sr[2] = 1 - cr[2]; sr[3] = g->template_h + cr[3];
    } else {
        sr[0] = 1; sr[1] = g->template_w; sr[2] = 1; sr[3] = g->template_h;
    }
    double npixels = (double)H * W;
    double dwell = 1 / npixels;
    /* pixelscantime(i,j), 0-based i, j (steady_scan.m:792-795) */
#define PST(i, j) (dwell / 2 + (double)((size_t)(i) * W + (size_t)((g->bidirectional && ((i) % 2 == (ORC_PLANT == 1 ? 0 : 1))) ? (W - 1 - (j)) : (j))) / npixels)
    int nk = g->nknots;
    double *tk = (double *)malloc(sizeof(double) * (size_t)nk);
    double lpk = (double)Hc / (nk - 1);
    if (lpk == floor(lpk)) {
        int L = (int)lpk;
        for (int k = 0; k < nk - 1; k++) {
            int line = k * L; /* knotlines index the uncropped pixelscantime (steady_scan.m:802-803) */
            const int j0 = ORC_PLANT == 2 ? cr[0] : 0, j1 = ORC_PLANT == 2 ? W - cr[1] : W;
            double mn = PST(line, j0);
            for (int j = j0 + 1; j < j1; j++) { double v = PST(line, j); if (v < mn) mn = v; }
            tk[k] = mn - dwell / 2;
        }
        tk[nk - 1] = 1;
        tk[0] = 0;
    } else {
        double mn = 1e300, mx = -1e300;
        for (int i = 0; i < Hc; i++)
            for (int j = 0; j < Wc; j++) {
                double v = PST(i + cr[2], j + cr[0]);
                if (v < mn) mn = v;
                if (v > mx) mx = v;
            }
        for (int k = 0; k < nk; k++) tk[k] = linspace_at(mn - dwell / 2, mx + dwell / 2, nk, k);
        tk[0] = 0;
        tk[nk - 1] = 1;
    }
    int bad = 0;
    for (int j = 0; j < Wc; j++)
        for (int i = 0; i < Hc; i++) {
            double v = PST(i + cr[2], j + cr[0]);
            int k = histc_bin(v, tk, nk);
            size_t o = (size_t)j * Hc + i;
            prevknotind[o] = k;
            if (k <= 0 || k >= nk) { bad = 1; tfrac[o] = 0; }
            else tfrac[o] = (v - tk[k - 1]) / (tk[k] - tk[k - 1]);
            /* meshgrid(linspace over linewidth, linspace over cropped nlines), then cropped columns */
            basex[o] = linspace_at(0, 1, W, j + cr[0]) * (sr[1] - sr[0]) + sr[0];
            basey[o] = linspace_at(0, 1, Hc, i) * (sr[3] - sr[2]) + sr[2];
        }
#undef PST
    free(tk);
    return bad ? ORC_INVALID_ARGUMENT : ORC_OK;
}

int orc_image2blockmatrix(const double *x, const int32_t *blockind, int npix, int nblocks,
                          double *y, uint8_t *bmmask) {
    int *ppb = (int *)xcalloc((size_t)nblocks, sizeof(int));
    for (int i = 0; i < npix; i++)
        if (blockind[i] >= 1 && blockind[i] <= nblocks) ppb[blockind[i] - 1]++;
    int frameH = 0;
    for (int b = 0; b < nblocks; b++)
        if (ppb[b] > frameH) frameH = ppb[b];
    if (y) {
        int *fill = (int *)xcalloc((size_t)nblocks, sizeof(int));
        for (size_t i = 0; i < (size_t)frameH * nblocks; i++) {
            y[i] = 0;
            if (bmmask) bmmask[i] = 0;
        }
        for (int i = 0; i < npix; i++) {
            int b = blockind[i] - 1;
            if (b < 0 || b >= nblocks) continue;
            size_t o = (size_t)b * frameH + fill[b]++;
            y[o] = x ? x[i] : 0;
            if (bmmask) bmmask[o] = 1;
        }
        free(fill);
    }
    free(ppb);
    return frameH;
}

/* resample_frame + pushgrid (steady_scan.m:562-599).  MATLAB's indexed
 * assignment z(ind) = z(ind) + v keeps the last value for repeated indices;
 * z and n are single precision. */
void orc_resample_frame(const double *data, int npix, const double *dx_knots,
                        const double *dy_knots, const double *basex, const double *basey,
                        const int32_t *prevknotind, const double *tfrac, int template_h,
                        int template_w, float *out) {
    size_t tp = (size_t)template_h * template_w;
    float *z = (float *)xcalloc(tp, sizeof(float));
    float *nn = (float *)xcalloc(tp, sizeof(float));
    float *zo = (float *)xcalloc(tp, sizeof(float));
    int nr = template_h, ncl = template_w;
    int *sel = (int *)malloc(sizeof(int) * (size_t)npix);
    double *xf = (double *)malloc(sizeof(double) * (size_t)npix);
    double *yf = (double *)malloc(sizeof(double) * (size_t)npix);
    long *mi = (long *)malloc(sizeof(long) * (size_t)npix);
    int m = 0;
    for (int i = 0; i < npix; i++) {
        int k = prevknotind[i] - 1;
        double tc = 1 - tfrac[i];
        double x = basex[i] + dx_knots[k] * tc + dx_knots[k + 1] * tfrac[i];
        double y = basey[i] + dy_knots[k] * tc + dy_knots[k + 1] * tfrac[i];
        if (x >= 1 && x < ncl && y >= 1 && y < nr) {
            double xi = floor(x), yi = floor(y);
            sel[m] = i;
            xf[m] = x - xi;
            yf[m] = y - yi;
            mi[m] = (long)(yi - 1) + (long)nr * (long)(xi - 1); /* 0-based matind */
            m++;
        }
    }
    /* four statements for z, then four for n (steady_scan.m:589-598) */
    long offz[4] = {0, 1, nr, (long)nr + 1};
    for (int s = 0; s < 4; s++) {
        memcpy(zo, z, sizeof(float) * tp);
        for (int q = 0; q < m; q++) {
            double w = s == 0 ? (1 - xf[q]) * (1 - yf[q]) : s == 1 ? (1 - xf[q]) * yf[q]
                     : s == 2 ? xf[q] * (1 - yf[q]) : xf[q] * yf[q];
            long t = mi[q] + offz[s];
            z[t] = zo[t] + (float)(w * data[sel[q]]);
        }
    }
    long offn[4] = {1, nr, 0, (long)nr + 1};
    for (int s = 0; s < 4; s++) {
        memcpy(zo, nn, sizeof(float) * tp);
        for (int q = 0; q < m; q++) {
            double w = s == 0 ? (1 - xf[q]) * yf[q] : s == 1 ? xf[q] * (1 - yf[q])
                     : s == 2 ? (1 - xf[q]) * (1 - yf[q]) : xf[q] * yf[q];
            long t = mi[q] + offn[s];
            nn[t] = zo[t] + (float)w;
        }
    }
    for (size_t t = 0; t < tp; t++) out[t] = z[t] / nn[t];
    free(z); free(nn); free(zo); free(sel); free(xf); free(yf); free(mi);
}

/* ------------------------------------------------------------------ */
/* invert_bilinear_sampling (steady_scan.m:290-404)                     */
/* ------------------------------------------------------------------ */
/* Neighbour slots of a template pixel, in the order of calc_indices'
 * find_* functions (steady_scan.m:645-723): N NE E SE S SW W NW. */
enum { IB_N, IB_NE, IB_E, IB_SE, IB_S, IB_SW, IB_W, IB_NW, IB_NB };

/* 0-based neighbour index and validity of pixel i (column-major, nrows rows,
 * npix pixels) as find_* computes them on 1-based indices.  find_SE's
 * "I < npixels" drops the one legitimate SE neighbour equal to the last
 * pixel; find_NW has no top-row test (its wrapped neighbour is never
 * coupled in Q, so its coefficient is zero either way). */
int orc_ibs_neighbour(long i0, int dir, long nrows, long npix, long *j0) {
    long i = i0 + 1, I = 1;
    int M = 0;
    switch (dir) {
    case IB_N: I = i - 1; M = (I % nrows) != 0; break;
    case IB_NE: I = i + nrows - 1; M = (I % nrows) != 0 && I < npix; break;
    case IB_E: I = i + nrows; M = I <= npix; break;
    case IB_SE: I = i + nrows + 1; M = (i % nrows) != 0 && I < npix; break;
    case IB_S: I = i + 1; M = (i % nrows) != 0; break;
    case IB_SW: I = i - nrows + 1; M = (i % nrows) != 0 && I > 0; break;
    case IB_W: I = i - nrows; M = I > 0; break;
    default: I = i - nrows - 1; M = I > 0; break;
    }
    if (!M) I = 1;
    *j0 = I - 1;
    return M;
}

/* Offset of neighbour slot dir as (drow, dcol). */
static void ib_offset(int dir, int *dr, int *dc) {
    static const int r[IB_NB] = {-1, -1, 0, 1, 1, 1, 0, -1};
    static const int c[IB_NB] = {0, 1, 1, 1, 0, -1, -1, -1};
    *dr = r[dir];
    *dc = c[dir];
}

static int ib_dir(int dr, int dc) {
    for (int d = 0; d < IB_NB; d++) {
        int a, b;
        ib_offset(d, &a, &b);
        if (a == dr && b == dc) return d;
    }
    return -1;
}

/* Q as a 9-point stencil per pixel: q[0] the diagonal, q[1 + dir] the
 * coupling to the neighbour in slot dir. */
#define IBQ(p, k) q[(size_t)(p) * 9 + (k)]

/* One frame's masked system (steady_scan.m:327-356).  qm (np x 9) is the
 * stencil of Q = ww + regmat with the rows and columns outside the mask
 * zeroed; wy (np) the constraint right-hand side, not yet masked; iv (np)
 * the resampled frame (resample_frame); mask (np).  Returns whether any
 * pixel had positive summed coefficients (the reference skips the frame
 * otherwise).  ww and wy are sums over the data pixels in ascending index
 * order; MATLAB's sparse() fixes no order, so that rounding is unpinned. */
int orc_ibs_system(const double *data, int npd_, const double *dx, const double *dy,
                   const int32_t *prevknotind, const double *tfrac, const double *basex,
                   const double *basey, int th, int tw, double regfac, double aspect, double *qm,
                   double *wy, float *iv, uint8_t *mask) {
    const long npd = npd_, np = (long)th * tw;
    const double eps = DBL_EPSILON, vfac = 1.0 / aspect, v2 = vfac * vfac;
    double *q = (double *)xcalloc((size_t)np * 9, sizeof(double));
    memset(wy, 0, sizeof(double) * (size_t)np);
    /* bilinear_map_linear_constraints (steady_scan.m:438-499) */
    for (long p = 0; p < npd; p++) {
        int k = prevknotind[p] - 1;
        double tc = 1 - tfrac[p];
        double xx = basex[p] + dx[k] * tc + dx[k + 1] * tfrac[p];
        double yy = basey[p] + dy[k] * tc + dy[k + 1] * tfrac[p];
        if (!((xx > 1 + eps) && (xx < tw - eps) && (yy > 1 + eps) && (yy < th - eps))) continue;
        double xi = floor(xx), yi = floor(yy);
        double fx = xx - xi, fy = yy - yi, cx = 1 - fx, cy = 1 - fy;
        long c0 = (long)(yi - 1) + (long)th * (long)(xi - 1);
        long cn[4] = {c0, c0 + 1, c0 + th, c0 + th + 1};
        int cr[4] = {0, 1, 0, 1}, cc[4] = {0, 0, 1, 1};
        double w[4] = {cx * cy, cx * fy, fx * cy, fx * fy};
        for (int a = 0; a < 4; a++) wy[cn[a]] += w[a] * data[p];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                double v = w[a] * w[b];
                if (a == b) IBQ(cn[a], 0) += v;
                else IBQ(cn[a], 1 + ib_dir(cr[b] - cr[a], cc[b] - cc[a])) += v;
            }
    }
    /* mask: summed coefficients (sum(ww)) positive and a resampled value;
     * any(mask) is tested before the resampled value masks pixels out
     * (steady_scan.m:333-338) */
    orc_resample_frame(data, (int)npd, dx, dy, basex, basey, prevknotind, tfrac, th, tw, iv);
    int any = 0;
    for (long i = 0; i < np; i++) {
        double s = 0;
        for (int k = 0; k < 9; k++) s += IBQ(i, k);
        mask[i] = s > 0;
        any |= mask[i];
        if (isnan(iv[i])) mask[i] = 0;
    }
    /* Q = ww + regmat (imageadjdiffquadmat, steady_scan.m:406-436): G'G has
     * -1 (horizontal) and -vfac^2 (vertical) off the diagonal and the sum of
     * the squared coefficients, W E N S, on it */
    for (long i = 0; i < np; i++) {
        long r = i % th, c = i / th;
        double dreg = 0;
        if (c > 0) { IBQ(i, 1 + IB_W) += -regfac; dreg += 1; }
        if (c + 1 < tw) { IBQ(i, 1 + IB_E) += -regfac; dreg += 1; }
        if (r > 0) { IBQ(i, 1 + IB_N) += -regfac * v2; dreg += v2; }
        if (r + 1 < th) { IBQ(i, 1 + IB_S) += -regfac * v2; dreg += v2; }
        IBQ(i, 0) += regfac * dreg;
    }
    /* rows and columns outside the mask zeroed */
    for (long i = 0; i < np; i++) {
        for (int k = 0; k < 9; k++) qm[i * 9 + k] = 0;
        if (!mask[i]) continue;
        qm[i * 9] = IBQ(i, 0);
        long r = i % th, c = i / th;
        for (int d = 0; d < IB_NB; d++) {
            int dr, dc;
            ib_offset(d, &dr, &dc);
            long rr = r + dr, cc2 = c + dc;
            if (rr < 0 || rr >= th || cc2 < 0 || cc2 >= tw) continue;
            long j = rr + cc2 * th;
            if (mask[j]) qm[i * 9 + 1 + d] = IBQ(i, 1 + d);
        }
    }
    free(q);
    return any;
}

/* The stopping energy single(double(x)' * Qe * double(x) - 2 * double(Le * x))
 * over the mask (steady_scan.m:370, 383).  The reference's order of summation
 * is that of MATLAB's sparse product and BLAS dot; this restatement fixes
 * one that parallel hardware can follow: (Q x)_i sums x_j Q_ij over the
 * neighbours j in ascending linear index (the order of a column of the
 * sparse matrix); the pixel terms x_i (Q x)_i (double) and Le_i x_i (single)
 * are summed in ORC_IBS_LANES partial sums, then added by a fixed binary
 * tree (t += t + 512, then 256, ... 1).  The partials come in 16 groups of
 * 64.  Rows are cut into blocks of 62 (nrb of them) and columns into
 * nseg = 16 / gcd(nrb, 16) segments (segment s: columns s tw / nseg up to
 * (s + 1) tw / nseg); unit u is row block u mod nrb by segment u / nrb.
 * Group g takes the units g, g + 16, ... in turn, and within a unit partial
 * 64 g + l (l = 1 .. 62) takes row 62 rb + l - 1 over the segment's columns
 * left to right; partials 64 g and 64 g + 63 stay zero. */
#define ORC_IBS_LANES 1024
#define ORC_IBS_GROUP 64
#define ORC_IBS_ROWS 62
float orc_ibs_energy(const double *qm, const double *wy, const float *x, const uint8_t *mask, int th,
                     int tw) {
    static const int ord[9][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {0, 0},
                                  {1, 0},   {-1, 1}, {0, 1},  {1, 1}};
    const int groups = ORC_IBS_LANES / ORC_IBS_GROUP;
    const int nrb = (th + ORC_IBS_ROWS - 1) / ORC_IBS_ROWS;
    int a = nrb, b = groups;
    while (b) {
        int t = a % b;
        a = b;
        b = t;
    }
    const int nseg = groups / a;
    double pq[ORC_IBS_LANES];
    float pl[ORC_IBS_LANES];
    for (int t = 0; t < ORC_IBS_LANES; t++) {
        pq[t] = 0;
        pl[t] = 0;
    }
    for (int g = 0; g < groups; g++)
        for (int u = g; u < nrb * nseg; u += groups) {
            const int rb = u % nrb, sg = u / nrb;
            const long c0 = (long)sg * tw / nseg, c1 = (long)(sg + 1) * tw / nseg;
            for (int l = 1; l <= ORC_IBS_ROWS; l++) {
                const long r = (long)rb * ORC_IBS_ROWS + l - 1;
                if (r >= th) continue;
                const int t = g * ORC_IBS_GROUP + l;
                for (long c = c0; c < c1; c++) {
                    const long i = r + c * (long)th;
                    if (!mask[i]) continue;
                    double qx = 0;
                    for (int k = 0; k < 9; k++) {
                        long rr = r + ord[k][0], cc = c + ord[k][1];
                        if (rr < 0 || rr >= th || cc < 0 || cc >= tw) continue;
                        long j = rr + cc * (long)th;
                        if (!mask[j]) continue;
                        int slot = k == 4 ? 0 : 1 + ib_dir(ord[k][0], ord[k][1]);
                        qx += (double)x[j] * qm[i * 9 + slot];
                    }
                    pq[t] += (double)x[i] * qx;
                    pl[t] += (float)wy[i] * x[i];
                }
            }
        }
    for (int off = ORC_IBS_LANES / 2; off >= 1; off >>= 1)
        for (int t = 0; t < off; t++) {
            pq[t] += pq[t + off];
            pl[t] += pl[t + off];
        }
    return (float)(pq[0] - 2.0 * (double)pl[0]);
}

/* calc_interactions (steady_scan.m:501-523) on one frame's masked system:
 * R = single(diag), Rm = single(single(Q_ij) / R) times the find_* mask,
 * L = single(wy) / R with NaN and Inf set to zero (steady_scan.m:358-362). */
void orc_ibs_interactions(const double *qm, const double *wy, const uint8_t *mask, int th, int tw,
                          float *Lk, float *Rm) {
    const long np = (long)th * tw;
    for (long i = 0; i < np; i++) {
        float R = (float)qm[i * 9];
        float Lf = (float)(mask[i] ? wy[i] : 0.0);
        float L = Lf / R;
        if (isnan(L) || isinf(L)) L = 0;
        Lk[i] = L;
        for (int d = 0; d < IB_NB; d++) {
            float v = 0;
            if (R != 0) {
                long j;
                int m = orc_ibs_neighbour(i, d, th, np, &j);
                /* Q(i, Imat) at the (possibly wrapped) neighbour index */
                long r = i % th, c = i / th, rj = j % th, cj = j / th;
                double qij = 0;
                if (labs(rj - r) <= 1 && labs(cj - c) <= 1 && j != i)
                    qij = qm[i * 9 + 1 + ib_dir((int)(rj - r), (int)(cj - c))];
                v = ((float)qij / R) * (m ? 1.0f : 0.0f);
            }
            Rm[i * IB_NB + d] = v;
        }
    }
}

/* One Gauss-Seidel iteration (steady_scan.m:374-377): the colours A B C D
 * (1-based odd/odd, odd/even, even/odd, even/even row/column) in turn, each
 * pixel x_i = L_i - sum over the slots N..NW of x(nbr) * Rm, in single. */
void orc_ibs_sweep(const float *Lk, const float *Rm, int th, int tw, float *x) {
    const long np = (long)th * tw;
    for (int color = 0; color < 4; color++) {
        int rpar = color >> 1, cpar = color & 1;
        for (long c = cpar; c < tw; c += 2)
            for (long r = rpar; r < th; r += 2) {
                long i = r + c * (long)th;
                float s = 0;
                for (int d = 0; d < IB_NB; d++) {
                    long j;
                    orc_ibs_neighbour(i, d, th, np, &j);
                    s += x[j] * Rm[i * IB_NB + d];
                }
                x[i] = Lk[i] - s;
            }
    }
}

int orc_invert_bilinear_sampling(const double *frames, int Hc, int Wc, int nframes, const double *dxk,
                                 const double *dyk, int nknots, const int32_t *prevknotind,
                                 const double *tfrac, const double *basex, const double *basey, int th,
                                 int tw, double regfac, double aspect, float *out, uint8_t *maskout,
                                 int32_t *iters) {
    const long npd = (long)Hc * Wc, np = (long)th * tw;
    double *wy = (double *)xcalloc((size_t)np, sizeof(double));
    double *qm = (double *)xcalloc((size_t)np * 9, sizeof(double));
    uint8_t *mask = (uint8_t *)xcalloc((size_t)np, 1);
    float *iv = (float *)xcalloc((size_t)np, sizeof(float));
    float *x = (float *)xcalloc((size_t)np, sizeof(float));
    float *xo = (float *)xcalloc((size_t)np, sizeof(float));
    float *Lk = (float *)xcalloc((size_t)np, sizeof(float));
    float *Rm = (float *)xcalloc((size_t)np * IB_NB, sizeof(float));
    int n_iter_error = 1;
    for (int f = 0; f < nframes; f++) {
        const double *data = frames + (size_t)f * npd;
        const double *dx = dxk + (size_t)f * nknots, *dy = dyk + (size_t)f * nknots;
        int any = orc_ibs_system(data, (int)npd, dx, dy, prevknotind, tfrac, basex, basey, th, tw, regfac,
                                 aspect, qm, wy, iv, mask);
        float *of = out + (size_t)f * np;
        uint8_t *mf = maskout + (size_t)f * np;
        if (!any) {
            for (long i = 0; i < np; i++) {
                of[i] = NAN;
                mf[i] = 0;
            }
            if (iters) iters[f] = 0;
            continue;
        }
        orc_ibs_interactions(qm, wy, mask, th, tw, Lk, Rm);
        /* initial x from the resampled frame, NaN -> 0; xo starts at zero */
        for (long i = 0; i < np; i++) {
            x[i] = isnan(iv[i]) ? 0.0f : iv[i];
            xo[i] = 0.0f;
        }
        float eo = orc_ibs_energy(qm, wy, x, mask, th, tw);
        int iter = 0;
        for (;;) {
            orc_ibs_sweep(Lk, Rm, th, tw, x);
            iter++;
            if (iter % n_iter_error == 0) {
                float e = orc_ibs_energy(qm, wy, x, mask, th, tw);
                if (eo <= e) break;
                eo = e;
                memcpy(xo, x, sizeof(float) * (size_t)np);
            }
        }
        /* the next frame checks the energy every int32(max(1, min(10,
         * sqrt(single(iter))))) iterations: round half away from zero */
        n_iter_error = orc_ibs_check_interval(iter);
        if (iters) iters[f] = iter;
        for (long i = 0; i < np; i++) {
            of[i] = mask[i] ? xo[i] : NAN;
            mf[i] = mask[i];
        }
    }
    free(wy); free(qm); free(mask); free(iv); free(x); free(xo); free(Lk); free(Rm);
    return 0;
}

int orc_ibs_check_interval(int iter) {
    float sq = sqrtf((float)iter);
    float v = sq < 10.0f ? sq : 10.0f;
    if (v < 1.0f) v = 1.0f;
    return (int)lroundf(v);
}
