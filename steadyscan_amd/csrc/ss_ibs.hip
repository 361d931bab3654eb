// ss_ibs.hip -- corrected frames by inverting the bilinear sampling
// (invert_bilinear_sampling, steady_scan.m:290-404, with its helpers
// bilinear_map_linear_constraints :438-499, imageadjdiffquadmat :406-436,
// calc_interactions :501-523, calc_indices and find_* :645-723, and
// resample_frame/pushgrid :562-599).
//
// Per frame the reference builds the normal equations of "the template,
// bilinearly sampled at every data pixel's displaced position, reproduces the
// data" plus an adjacent-difference regulariser, masks them, and runs a
// 4-colour Gauss-Seidel loop in single precision from the resampled frame,
// stopping when an energy check every n iterations finds no decrease.  n is
// 1 for the call's first frame and max(1, min(10, round(sqrt(iter)))) of the
// previous frame afterwards, which chains the frames.
//
// GPU form, for a chunk of frames at once:
//   points     one thread per data pixel: displaced position; the cell of
//              its constraint (counted per cell) and the pushgrid winners
//              (largest pixel index per target and corner, as MATLAB's
//              indexed assignment keeps the last value);
//   scan/fill  per-cell lists of data pixels (counting sort), each list
//              sorted by pixel index, so that
//   assemble   one thread per template pixel gathers the <= 4 cells it is a
//              corner of, merged in ascending pixel index: the constraint
//              sums come out in the same order as a sequential pass over the
//              data pixels, bit for bit, without floating-point atomics;
//   interact   the masked stencil, calc_interactions' R, L and Rm;
//   gs         one workgroup per frame runs the Gauss-Seidel loop with the
//              energy summed in a fixed order (per template row over
//              ascending columns, then over rows) and tracks the stopping
//              rule for all ten possible check intervals at once, keeping a
//              snapshot per interval; a frame whose interval is already
//              known (published by the previous frame's workgroup) drops
//              the others;
//   resolve    on the host: walks the frames in order, picks each frame's
//              interval from its predecessor, and
//   output     writes the chosen snapshot (NaN outside the mask).
// Everything is deterministic; all floating-point code is compiled without
// contraction so that it rounds as the sequential restatement does.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "ss_host.h"
#include "steadyscan.h"

namespace {

constexpr int kGsThreads = 1024;
constexpr int kNSched = 10;              // check intervals 1..10
constexpr int kNSlots = kNSched + 1;     // snapshots: one per interval + one spare
constexpr int kDefaultCap = 100000;      // Gauss-Seidel iterations per frame

// Neighbour slots in the order of calc_indices' find_* list.
enum { NB_N, NB_NE, NB_E, NB_SE, NB_S, NB_SW, NB_W, NB_NW, NB_COUNT };
// Stencil entry of corner b as seen from corner a of a cell: corners are
// (row, col) offsets (0,0) (1,0) (0,1) (1,1), i.e. matind_corners' 0 1 h h+1.
__constant__ int kCornerDir[4][4] = {{-1, NB_S, NB_E, NB_SE},
                                     {NB_N, -1, NB_NE, NB_E},
                                     {NB_W, NB_SW, -1, NB_S},
                                     {NB_NW, NB_W, NB_N, -1}};
__constant__ int kSlotDr[NB_COUNT] = {-1, -1, 0, 1, 1, 1, 0, -1};
__constant__ int kSlotDc[NB_COUNT] = {0, 1, 1, 1, 0, -1, -1, -1};

// Symmetric half of the masked stencil kept for the energy, structure of
// arrays: diagonal, then the couplings to S, NE, E, SE.
enum { QS_D, QS_S, QS_NE, QS_E, QS_SE, QS_COUNT };

struct PixGeo {
    double bx, by, tf;
    int k;  // prevknotind - 1
    int pad;
};

// One chunk of frames on the device; every per-frame array is indexed
// [frame][...] with the strides below.
struct Chunk {
    int B, th, tw, np, npd, hc, nl, cl, cb, nk, cap;
    double regfac, v2;
    size_t rawStride;         // elements per raw frame
    const PixGeo *geo;        // npd
    const double *kx, *ky;    // B x nk
    int *cell;                // B x npd: constraint cell or -1
    int *cnt;                 // B x np: per-cell count, then fill cursor
    int *start;               // B x (np + 1)
    int *list;                // B x npd
    int *win;                 // B x 4 x np: pushgrid winners
    double *qraw;             // B x np x 9: Q = ww + regmat before masking
    double *wy;               // B x np
    float *iv;                // B x np
    uint8_t *mask;            // B x np
    int *any;                 // B
    double *qs;               // B x QS_COUNT x np
    float *lf, *lk;           // B x np
    float *rm;                // B x np x 8
    float *x;                 // B x np
    float *snap;              // B x kNSlots x np
    double *rowq;             // B x th
    float *rowl;              // B x th
    int *tab;                 // B x 3 x kNSched: iterations, slot, stopped (+1 known)
    int *pub;                 // B + 1: each frame's check interval once known, else 0
};

// int32(max(1, min(10, sqrt(single(iter))))) (steady_scan.m:394): sqrt is
// never within float rounding of a half-integer here, so the rounded value is
// the k with k^2 - k < iter <= k^2 + k.
__host__ __device__ inline int check_interval(int iter) {
    int k = 1;
    while (k < 10 && iter > k * k + k) k++;
    return k;
}

// find_N .. find_NW on the 1-based index i0 + 1 (steady_scan.m:677-723): the
// neighbour's 0-based index, and 0 with a false mask where the find function
// rejects it (find_SE's strict "< npixels" included; find_NW tests no row).
__device__ inline long neighbour(long i0, int d, long nrows, long npix, bool *m) {
    const long i = i0 + 1;
    long I;
    bool M;
    switch (d) {
    case NB_N: I = i - 1; M = (I % nrows) != 0; break;
    case NB_NE: I = i + nrows - 1; M = (I % nrows) != 0 && I < npix; break;
    case NB_E: I = i + nrows; M = I <= npix; break;
    case NB_SE: I = i + nrows + 1; M = (i % nrows) != 0 && I < npix; break;
    case NB_S: I = i + 1; M = (i % nrows) != 0; break;
    case NB_SW: I = i - nrows + 1; M = (i % nrows) != 0 && I > 0; break;
    case NB_W: I = i - nrows; M = I > 0; break;
    default: I = i - nrows - 1; M = I > 0; break;
    }
    *m = M;
    return M ? I - 1 : 0;
}

__device__ inline int slot_of(int dr, int dc) {
    for (int d = 0; d < NB_COUNT; d++)
        if (kSlotDr[d] == dr && kSlotDc[d] == dc) return d;
    return -1;
}

__device__ inline void position(const PixGeo &q, const double *kx, const double *ky, double *x,
                                double *y) {
#pragma clang fp contract(off)
    const double tc = 1.0 - q.tf;
    *x = q.bx + kx[q.k] * tc + kx[q.k + 1] * q.tf;
    *y = q.by + ky[q.k] * tc + ky[q.k + 1] * q.tf;
}

template <typename TD>
__device__ inline double pixel_value(const Chunk &C, const TD *data, int f, int p) {
    const int ic = p % C.hc, jc = p / C.hc;  // cropped (row, column)
    return (double)data[(size_t)f * C.rawStride + (size_t)(jc + C.cl) * C.nl + (ic + C.cb)];
}

// ---------------------------------------------------------------------------
// points: constraint cells and pushgrid winners
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_points(Chunk C) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (p >= C.npd) return;
    double x, y;
    position(C.geo[p], C.kx + (size_t)f * C.nk, C.ky + (size_t)f * C.nk, &x, &y);
    const double th = C.th, tw = C.tw;
    int cell = -1;
    // bilinear_map_linear_constraints' mask (steady_scan.m:455)
    if ((x > 1.0 + DBL_EPSILON) && (x < tw - DBL_EPSILON) && (y > 1.0 + DBL_EPSILON) &&
        (y < th - DBL_EPSILON)) {
        cell = ((int)floor(y) - 1) + C.th * ((int)floor(x) - 1);
        atomicAdd(&C.cnt[(size_t)f * C.np + cell], 1);
    }
    C.cell[(size_t)f * C.npd + p] = cell;
    // pushgrid's mask (steady_scan.m:580)
    if (x >= 1.0 && x < tw && y >= 1.0 && y < th) {
        const int m0 = ((int)floor(y) - 1) + C.th * ((int)floor(x) - 1);
        int *w = C.win + (size_t)f * 4 * C.np;
        atomicMax(&w[m0], p);
        atomicMax(&w[C.np + m0 + 1], p);
        atomicMax(&w[2 * (size_t)C.np + m0 + C.th], p);
        atomicMax(&w[3 * (size_t)C.np + m0 + C.th + 1], p);
    }
}

// Exclusive scan of the per-cell counts (one workgroup per frame); the
// counts are cleared for use as fill cursors.
__global__ __launch_bounds__(1024) void k_scan(Chunk C) {
    __shared__ int part[1024];
    const int f = blockIdx.x, t = threadIdx.x;
    int *cnt = C.cnt + (size_t)f * C.np;
    int *start = C.start + (size_t)f * (C.np + 1);
    const int per = (C.np + 1023) / 1024;
    const int lo = t * per < C.np ? t * per : C.np;
    const int hi = lo + per < C.np ? lo + per : C.np;
    int s = 0;
    for (int i = lo; i < hi; i++) s += cnt[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int k = 0; k < 1024; k++) {
            const int v = part[k];
            part[k] = acc;
            acc += v;
        }
        start[C.np] = acc;
    }
    __syncthreads();
    int acc = part[t];
    for (int i = lo; i < hi; i++) {
        start[i] = acc;
        acc += cnt[i];
        cnt[i] = 0;
    }
}

__global__ __launch_bounds__(256) void k_fill(Chunk C) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (p >= C.npd) return;
    const int cell = C.cell[(size_t)f * C.npd + p];
    if (cell < 0) return;
    const int slot = atomicAdd(&C.cnt[(size_t)f * C.np + cell], 1);
    C.list[(size_t)f * C.npd + C.start[(size_t)f * (C.np + 1) + cell] + slot] = p;
}

// Sort each cell's list by pixel index: insertion sort for short lists,
// heapsort (in place, bounded) for long ones.
__device__ inline void sift_down(int *a, int root, int n) {
    while (true) {
        int c = 2 * root + 1;
        if (c >= n) return;
        if (c + 1 < n && a[c + 1] > a[c]) c++;
        if (a[root] >= a[c]) return;
        const int t = a[root];
        a[root] = a[c];
        a[c] = t;
        root = c;
    }
}

__global__ __launch_bounds__(256) void k_sort(Chunk C) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (cell >= C.np) return;
    const int *start = C.start + (size_t)f * (C.np + 1);
    int *a = C.list + (size_t)f * C.npd + start[cell];
    const int n = start[cell + 1] - start[cell];
    if (n < 2) return;
    if (n <= 32) {
        for (int i = 1; i < n; i++) {
            const int v = a[i];
            int j = i - 1;
            while (j >= 0 && a[j] > v) {
                a[j + 1] = a[j];
                j--;
            }
            a[j + 1] = v;
        }
        return;
    }
    for (int r = n / 2 - 1; r >= 0; r--) sift_down(a, r, n);
    for (int e = n - 1; e > 0; e--) {
        const int t = a[0];
        a[0] = a[e];
        a[e] = t;
        sift_down(a, 0, e);
    }
}

// ---------------------------------------------------------------------------
// assemble: per template pixel, the constraint sums over its cells' data
// pixels in ascending index, the regulariser, the resampled value and mask
// ---------------------------------------------------------------------------
template <typename TD>
__global__ __launch_bounds__(256) void k_assemble(Chunk C, const TD *__restrict__ data) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= C.np) return;
    const int th = C.th, tw = C.tw;
    const int r = i % th, c = i / th;
    const double *kx = C.kx + (size_t)f * C.nk, *ky = C.ky + (size_t)f * C.nk;
    const int *start = C.start + (size_t)f * (C.np + 1);
    const int *list = C.list + (size_t)f * C.npd;
    // cells whose corner a is this pixel: (r,c), (r-1,c), (r,c-1), (r-1,c-1)
    int pos[4], end[4];
    const int cellOf[4] = {i, i - 1, i - th, i - th - 1};
    const bool ok[4] = {r <= th - 2 && c <= tw - 2, r >= 1 && c <= tw - 2, r <= th - 2 && c >= 1,
                        r >= 1 && c >= 1};
    for (int a = 0; a < 4; a++) {
        pos[a] = ok[a] ? start[cellOf[a]] : 0;
        end[a] = ok[a] ? start[cellOf[a] + 1] : 0;
    }
    double q[9];
    for (int k = 0; k < 9; k++) q[k] = 0.0;
    double wy = 0.0;
    while (true) {
        int a = -1, best = 0x7fffffff;
        for (int b = 0; b < 4; b++)
            if (pos[b] < end[b] && list[pos[b]] < best) {
                best = list[pos[b]];
                a = b;
            }
        if (a < 0) break;
        pos[a]++;
        const int p = best;
        double x, y;
        position(C.geo[p], kx, ky, &x, &y);
        const double fx = x - floor(x), fy = y - floor(y), cx = 1 - fx, cy = 1 - fy;
        const double w[4] = {cx * cy, cx * fy, fx * cy, fx * fy};
        wy += w[a] * pixel_value(C, data, f, p);
        for (int b = 0; b < 4; b++) {
            const double v = w[a] * w[b];
            if (b == a) q[0] += v;
            else q[1 + kCornerDir[a][b]] += v;
        }
    }
    // sum(ww) > swwthresh (= 0) (steady_scan.m:333-335)
    double sww = 0;
    for (int k = 0; k < 9; k++) sww += q[k];
    const bool swok = sww > 0;
    // + regmat: G'G * regfac with -1 / -vfac^2 off the diagonal and the
    // squared coefficients W E N S summed on it
    double dreg = 0;
    if (c > 0) { q[1 + NB_W] += -C.regfac; dreg += 1; }
    if (c + 1 < tw) { q[1 + NB_E] += -C.regfac; dreg += 1; }
    if (r > 0) { q[1 + NB_N] += -C.regfac * C.v2; dreg += C.v2; }
    if (r + 1 < th) { q[1 + NB_S] += -C.regfac * C.v2; dreg += C.v2; }
    q[0] += C.regfac * dreg;
    // resample_frame: the pushgrid winners in statement order, single precision
    const int *win = C.win + (size_t)f * 4 * C.np;
    float z = 0.f, n = 0.f;
    double ws[4];
    bool has[4];
    for (int s = 0; s < 4; s++) {
        const int p = win[(size_t)s * C.np + i];
        has[s] = p >= 0;
        ws[s] = 0.0;
        if (!has[s]) continue;
        double x, y;
        position(C.geo[p], kx, ky, &x, &y);
        const double xf = x - floor(x), yf = y - floor(y);
        ws[s] = s == 0 ? (1 - xf) * (1 - yf) : s == 1 ? (1 - xf) * yf : s == 2 ? xf * (1 - yf) : xf * yf;
        z = z + (float)(ws[s] * pixel_value(C, data, f, p));
    }
    const int norder[4] = {1, 2, 0, 3};
    for (int o = 0; o < 4; o++)
        if (has[norder[o]]) n = n + (float)ws[norder[o]];
    const float iv = z / n;
    const size_t o = (size_t)f * C.np + i;
    double *qo = C.qraw + o * 9;
    for (int k = 0; k < 9; k++) qo[k] = q[k];
    C.wy[o] = wy;
    C.iv[o] = iv;
    C.mask[o] = (swok && !isnan(iv)) ? 1 : 0;
    if (swok) atomicOr(&C.any[f], 1);
}

// ---------------------------------------------------------------------------
// interact: masked stencil, calc_interactions (steady_scan.m:501-523), the
// right-hand side (steady_scan.m:358-362) and the initial x
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_interact(Chunk C) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= C.np) return;
    const int th = C.th, tw = C.tw;
    const int r = i % th, c = i / th;
    const size_t fo = (size_t)f * C.np;
    const uint8_t *mask = C.mask + fo;
    const bool m = mask[i] != 0;
    const double *qr = C.qraw + (fo + i) * 9;
    double qm[9];
    qm[0] = m ? qr[0] : 0.0;
    for (int d = 0; d < NB_COUNT; d++) {
        const int rr = r + kSlotDr[d], cc = c + kSlotDc[d];
        const bool in = rr >= 0 && rr < th && cc >= 0 && cc < tw;
        qm[1 + d] = (m && in && mask[rr + cc * th]) ? qr[1 + d] : 0.0;
    }
    const float R = (float)qm[0];
    const float Lf = (float)(m ? C.wy[fo + i] : 0.0);
    float L = Lf / R;
    if (isnan(L) || isinf(L)) L = 0.f;
    C.lk[fo + i] = L;
    C.lf[fo + i] = Lf;
    float *rm = C.rm + (fo + i) * 8;
    for (int d = 0; d < NB_COUNT; d++) {
        float v = 0.f;
        if (R != 0.f) {
            bool mm;
            const long j = neighbour(i, d, th, C.np, &mm);
            const int rj = (int)(j % th), cj = (int)(j / th);
            double qij = 0.0;
            if (abs(rj - r) <= 1 && abs(cj - c) <= 1 && j != i) qij = qm[1 + slot_of(rj - r, cj - c)];
            v = ((float)qij / R) * (mm ? 1.0f : 0.0f);
        }
        rm[d] = v;
    }
    double *qs = C.qs + (size_t)f * QS_COUNT * C.np;
    qs[QS_D * (size_t)C.np + i] = qm[0];
    qs[QS_S * (size_t)C.np + i] = qm[1 + NB_S];
    qs[QS_NE * (size_t)C.np + i] = qm[1 + NB_NE];
    qs[QS_E * (size_t)C.np + i] = qm[1 + NB_E];
    qs[QS_SE * (size_t)C.np + i] = qm[1 + NB_SE];
    const float iv = C.iv[fo + i];
    C.x[fo + i] = isnan(iv) ? 0.0f : iv;
}

// ---------------------------------------------------------------------------
// gs: one workgroup per frame
// ---------------------------------------------------------------------------
// The stopping energy single(x'Qx - 2 single(Le x)) over the mask, in the
// fixed order: per row over ascending columns, (Q x)_i over the neighbours
// in ascending index, then the rows in ascending order by thread 0.
__device__ float gs_energy(const Chunk &C, int f, const float *x, const uint8_t *mask,
                           const double *qs, const float *lf, double *rowq, float *rowl) {
#pragma clang fp contract(off)
    const int th = C.th, tw = C.tw;
    const size_t np = C.np;
    const double *qD = qs + QS_D * np, *qS = qs + QS_S * np, *qNE = qs + QS_NE * np,
                 *qE = qs + QS_E * np, *qSE = qs + QS_SE * np;
    for (int r = threadIdx.x; r < th; r += blockDim.x) {
        double qr = 0.0;
        float lr = 0.f;
        for (int c = 0; c < tw; c++) {
            const int i = r + c * th;
            if (!mask[i]) continue;
            double qx = 0.0;
            if (c >= 1) {
                if (r >= 1 && mask[i - th - 1]) qx += (double)x[i - th - 1] * qSE[i - th - 1];  // NW
                if (mask[i - th]) qx += (double)x[i - th] * qE[i - th];                         // W
                if (r + 1 < th && mask[i - th + 1]) qx += (double)x[i - th + 1] * qNE[i - th + 1];  // SW
            }
            if (r >= 1 && mask[i - 1]) qx += (double)x[i - 1] * qS[i - 1];  // N
            qx += (double)x[i] * qD[i];
            if (r + 1 < th && mask[i + 1]) qx += (double)x[i + 1] * qS[i];  // S
            if (c + 1 < tw) {
                if (r >= 1 && mask[i + th - 1]) qx += (double)x[i + th - 1] * qNE[i];  // NE
                if (mask[i + th]) qx += (double)x[i + th] * qE[i];                     // E
                if (r + 1 < th && mask[i + th + 1]) qx += (double)x[i + th + 1] * qSE[i];  // SE
            }
            qr += (double)x[i] * qx;
            lr += lf[i] * x[i];
        }
        rowq[r] = qr;
        rowl[r] = lr;
    }
    __syncthreads();
    float e = 0.f;
    if (threadIdx.x == 0) {
        double quad = 0.0;
        float lin = 0.f;
        for (int r = 0; r < th; r++) {
            quad += rowq[r];
            lin += rowl[r];
        }
        e = (float)(quad - 2.0 * (double)lin);
    }
    return e;  // valid in thread 0
}

// One iteration: the colours A B C D (1-based odd/odd, odd/even, even/odd,
// even/even row/column), each x_i = L_i - sum over N..NW of x(nbr) * Rm.
__device__ void gs_sweep(const Chunk &C, float *x, const float *lk, const float *rm) {
#pragma clang fp contract(off)
    const int th = C.th, tw = C.tw;
    for (int color = 0; color < 4; color++) {
        const int rp = color >> 1, cp = color & 1;
        const int nr = (th - rp + 1) / 2, nc = (tw - cp + 1) / 2;
        const int n = nr * nc;
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const int r = rp + 2 * (k % nr), c = cp + 2 * (k / nr);
            const int i = r + c * th;
            const float4 *rv = (const float4 *)(rm + (size_t)i * 8);
            const float4 a = rv[0], b = rv[1];
            const float w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            float s = 0.f;
            for (int d = 0; d < NB_COUNT; d++) {
                bool mm;
                const long j = neighbour(i, d, th, C.np, &mm);
                s += x[j] * w[d];
            }
            x[i] = lk[i] - s;
        }
        __syncthreads();
    }
}

// The check interval of frame f if already known: frames skipped for an
// empty mask pass their interval on unchanged.
__device__ int known_interval(const Chunk &C, int f) {
    int j = f;
    while (j > 0 && !C.any[j - 1]) j--;
    return __hip_atomic_load(&C.pub[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kGsThreads) void k_gs(Chunk C) {
    const int f = blockIdx.x;
    if (!C.any[f]) return;
    const size_t fo = (size_t)f * C.np;
    float *x = C.x + fo;
    const float *lk = C.lk + fo, *lf = C.lf + fo, *rm = C.rm + fo * 8;
    const uint8_t *mask = C.mask + fo;
    const double *qs = C.qs + (size_t)f * QS_COUNT * C.np;
    double *rowq = C.rowq + (size_t)f * C.th;
    float *rowl = C.rowl + (size_t)f * C.th;
    float *snap = C.snap + (size_t)f * kNSlots * C.np;

    __shared__ float s_eo[kNSched];
    __shared__ int s_iter[kNSched], s_slot[kNSched], s_ref[kNSlots];
    __shared__ unsigned s_active, s_stopped;
    __shared__ int s_known, s_needE, s_copy, s_done;

    float e0 = gs_energy(C, f, x, mask, qs, lf, rowq, rowl);
    if (threadIdx.x == 0) {
        for (int n = 0; n < kNSched; n++) {
            s_eo[n] = e0;
            s_iter[n] = 0;
            s_slot[n] = -1;  // xo starts at zero
        }
        for (int s = 0; s < kNSlots; s++) s_ref[s] = 0;
        s_stopped = 0;
        s_known = known_interval(C, f);
        s_active = s_known ? 1u << (s_known - 1) : (1u << kNSched) - 1;
    }
    __syncthreads();
    for (int k = 1;; k++) {
        gs_sweep(C, x, lk, rm);
        if (threadIdx.x == 0) {
            int need = 0;
            for (int n = 1; n <= kNSched; n++)
                if ((s_active & ~s_stopped & (1u << (n - 1))) && k % n == 0) need = 1;
            s_needE = need;
        }
        __syncthreads();
        float e = 0.f;
        if (s_needE) e = gs_energy(C, f, x, mask, qs, lf, rowq, rowl);
        if (threadIdx.x == 0) {
            unsigned pass = 0;
            if (s_needE) {
                for (int n = 1; n <= kNSched; n++) {
                    const unsigned bit = 1u << (n - 1);
                    if (!(s_active & ~s_stopped & bit) || k % n) continue;
                    if (s_eo[n - 1] <= e) {
                        s_stopped |= bit;
                        s_iter[n - 1] = k;
                    } else {
                        s_eo[n - 1] = e;
                        pass |= bit;
                    }
                }
            }
            int copy = -1;
            if (pass) {
                for (int s = 0; s < kNSlots && copy < 0; s++)
                    if (s_ref[s] == 0) copy = s;
                for (int n = 0; n < kNSched; n++)
                    if (pass & (1u << n)) {
                        if (s_slot[n] >= 0) s_ref[s_slot[n]]--;
                        s_slot[n] = copy;
                        s_ref[copy]++;
                    }
            }
            if (!s_known) {
                const int kn = known_interval(C, f);
                if (kn) {
                    s_known = kn;
                    s_active = 1u << (kn - 1);
                }
            }
            s_copy = copy;
            s_done = (s_active & ~s_stopped) == 0 || k >= C.cap;
        }
        __syncthreads();
        if (s_copy >= 0) {
            float *dst = snap + (size_t)s_copy * C.np;
            for (int i = threadIdx.x; i < C.np; i += blockDim.x) dst[i] = x[i];
        }
        __syncthreads();
        if (s_done) break;
    }
    if (threadIdx.x == 0) {
        int *tab = C.tab + (size_t)f * 3 * kNSched;
        for (int n = 0; n < kNSched; n++) {
            tab[n] = s_iter[n];
            tab[kNSched + n] = s_slot[n];
            tab[2 * kNSched + n] = ((s_stopped >> n) & 1) | ((s_active >> n) & 1) << 1;
        }
        if (s_known && ((s_stopped >> (s_known - 1)) & 1))
            __hip_atomic_store(&C.pub[f + 1], check_interval(s_iter[s_known - 1]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// chosen[f]: -2 skipped frame (NaN, empty mask), -1 zeros, else the snapshot
__global__ __launch_bounds__(256) void k_output(Chunk C, const int *chosen, float *out,
                                                uint8_t *maskOut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= C.np) return;
    const size_t o = (size_t)f * C.np + i;
    const int ch = chosen[f];
    if (ch == -2) {
        out[o] = __builtin_nanf("");
        maskOut[o] = 0;
        return;
    }
    const bool m = C.mask[o] != 0;
    const float v = ch < 0 ? 0.0f : C.snap[((size_t)f * kNSlots + ch) * C.np + i];
    out[o] = m ? v : __builtin_nanf("");
    maskOut[o] = m ? 1 : 0;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" void ss_ibs_settings_default(ss_ibs_settings *s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->regfac = 3e-3;
    s->pixel_aspect = 1.0;
    s->inpaint = 0;
    s->max_iterations = 0;
    s->chunk_frames = 0;
}

extern "C" ss_status ss_invert_bilinear_sampling(ss_handle *h, const ss_resample_in *in,
                                                 const ss_ibs_settings *st, float *out,
                                                 uint8_t *mask_out, int32_t *iterations) {
    using ss::handle_fail;
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    if (!in || !st || !out || !mask_out || !in->geometry || !in->frames || !in->knots_x ||
        !in->knots_y || in->n_frames < 1)
        return handle_fail(h, SS_ERR_INVALID_ARGUMENT, "invert_bilinear_sampling: NULL argument or no frames");
    if (st->inpaint)
        return handle_fail(h, SS_ERR_INVALID_ARGUMENT, "invert_bilinear_sampling: inpainting is not supported");
    if (!(st->pixel_aspect > 0) || !std::isfinite(st->regfac) || st->max_iterations < 0 ||
        st->chunk_frames < 0)
        return handle_fail(h, SS_ERR_INVALID_ARGUMENT, "invert_bilinear_sampling: bad settings");
    const ss_geometry_desc &g = *in->geometry;
    int32_t hc, wc, fh;
    ss_status rc = ss_geometry_size(&g, &hc, &wc, &fh);
    if (rc) return handle_fail(h, rc, "invert_bilinear_sampling: bad geometry");
    const int th = g.template_h, tw = g.template_w, nk = g.nknots;
    if (th < 2 || tw < 2)
        return handle_fail(h, SS_ERR_INVALID_ARGUMENT, "invert_bilinear_sampling: template smaller than 2x2");
    if ((double)th * tw > 2.0e9 || (double)hc * wc > 2.0e9)
        return handle_fail(h, SS_ERR_INVALID_ARGUMENT, "invert_bilinear_sampling: image too large");
    const int npd = hc * wc, np = th * tw;
    std::vector<int32_t> pk(npd);
    std::vector<double> tf(npd), bx(npd), by(npd);
    if ((rc = ss_geometry_init(&g, pk.data(), tf.data(), bx.data(), by.data(), nullptr, nullptr, nullptr,
                               nullptr, nullptr)))
        return handle_fail(h, rc, "invert_bilinear_sampling: bad geometry");
    std::vector<PixGeo> pg(npd);
    for (int i = 0; i < npd; i++) pg[i] = PixGeo{bx[i], by[i], tf[i], pk[i] - 1, 0};
    const size_t esz = in->data_type == SS_F64 ? 8 : in->data_type == SS_F32 ? 4 : 2;
    const size_t raw = (size_t)g.nlines * g.linewidth;

    // device memory per frame of a chunk
    const size_t perFrame =
        align256(sizeof(double) * 2 * nk) + align256(sizeof(int) * (size_t)npd) * 3 +
        align256(sizeof(int) * (size_t)np) + align256(sizeof(int) * ((size_t)np + 1)) +
        align256(sizeof(int) * 4 * (size_t)np) + align256(sizeof(double) * 9 * (size_t)np) +
        align256(sizeof(double) * (size_t)np) + align256(sizeof(float) * (size_t)np) * 5 +
        align256((size_t)np) + align256(sizeof(double) * QS_COUNT * (size_t)np) +
        align256(sizeof(float) * 8 * (size_t)np) + align256(sizeof(float) * kNSlots * (size_t)np) +
        align256(sizeof(double) * th) + align256(sizeof(float) * th) + align256(sizeof(int) * 3 * kNSched) +
        align256(sizeof(int) * 2) + (in->data_on_device ? 0 : align256(esz * raw)) +
        align256(sizeof(float) * (size_t)np) + align256((size_t)np);
    int B = st->chunk_frames > 0 ? st->chunk_frames : (int)std::max<size_t>(1, ((size_t)8 << 30) / perFrame);
    B = std::min(B, in->n_frames);
    B = std::min(B, 65535);
    const int cap = st->max_iterations > 0 ? st->max_iterations : kDefaultCap;

    hipStream_t s = ss::handle_stream(h);
    if (hipSetDevice(ss::handle_device(h)) != hipSuccess)
        return handle_fail(h, SS_ERR_DEVICE, "invert_bilinear_sampling: hipSetDevice failed");
    char *base = nullptr;
    PixGeo *dgeo = nullptr;
    std::string err;
    rc = SS_OK;
#define ICHK(expr)                                                           \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            rc = e_ == hipErrorOutOfMemory ? SS_ERR_OUT_OF_MEMORY : SS_ERR_DEVICE; \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
            goto done;                                                       \
        }                                                                    \
    } while (0)
    {
        const size_t total = perFrame * (size_t)B + 4096;
        ICHK(hipMalloc(&dgeo, sizeof(PixGeo) * npd));
        ICHK(hipMalloc((void **)&base, total));
        ICHK(hipMemcpyAsync(dgeo, pg.data(), sizeof(PixGeo) * npd, hipMemcpyHostToDevice, s));
        size_t off = 0;
        auto take = [&](size_t bytes) {
            char *p = base + off;
            off += align256(bytes);
            return (void *)p;
        };
        Chunk C{};
        C.th = th; C.tw = tw; C.np = np; C.npd = npd; C.hc = hc; C.nl = g.nlines;
        C.cl = g.cropping[0]; C.cb = g.cropping[2]; C.nk = nk; C.cap = cap;
        C.regfac = st->regfac;
        const double vfac = 1.0 / st->pixel_aspect;
        C.v2 = vfac * vfac;
        C.rawStride = raw;
        C.geo = dgeo;
        double *dk = (double *)take(sizeof(double) * 2 * nk * (size_t)B);
        C.kx = dk;
        C.ky = dk + (size_t)nk * B;
        C.cell = (int *)take(sizeof(int) * (size_t)npd * B);
        C.list = (int *)take(sizeof(int) * (size_t)npd * B);
        C.cnt = (int *)take(sizeof(int) * (size_t)np * B);
        C.start = (int *)take(sizeof(int) * ((size_t)np + 1) * B);
        C.win = (int *)take(sizeof(int) * 4 * (size_t)np * B);
        C.qraw = (double *)take(sizeof(double) * 9 * (size_t)np * B);
        C.wy = (double *)take(sizeof(double) * (size_t)np * B);
        C.iv = (float *)take(sizeof(float) * (size_t)np * B);
        C.mask = (uint8_t *)take((size_t)np * B);
        C.any = (int *)take(sizeof(int) * (size_t)B);
        C.qs = (double *)take(sizeof(double) * QS_COUNT * (size_t)np * B);
        C.lf = (float *)take(sizeof(float) * (size_t)np * B);
        C.lk = (float *)take(sizeof(float) * (size_t)np * B);
        C.rm = (float *)take(sizeof(float) * 8 * (size_t)np * B);
        C.x = (float *)take(sizeof(float) * (size_t)np * B);
        C.snap = (float *)take(sizeof(float) * kNSlots * (size_t)np * B);
        C.rowq = (double *)take(sizeof(double) * th * (size_t)B);
        C.rowl = (float *)take(sizeof(float) * th * (size_t)B);
        C.tab = (int *)take(sizeof(int) * 3 * kNSched * (size_t)B);
        C.pub = (int *)take(sizeof(int) * ((size_t)B + 1));
        int *dchosen = (int *)take(sizeof(int) * (size_t)B);
        float *dout = (float *)take(sizeof(float) * (size_t)np * B);
        uint8_t *dmask = (uint8_t *)take((size_t)np * B);
        void *ddata = in->data_on_device ? nullptr : take(esz * raw * B);
        if (off > total) {
            rc = SS_ERR_DEVICE;
            err = "invert_bilinear_sampling: internal sizing error";
            goto done;
        }
        std::vector<int> tab(3 * kNSched * (size_t)B), any(B), chosen(B);
        std::vector<double> kbuf(2 * (size_t)nk * B);
        int carry = 1;  // n_iter_error starts at 1 for the call
        for (int f0 = 0; f0 < in->n_frames; f0 += B) {
            const int nb = std::min(B, in->n_frames - f0);
            C.B = nb;
            const void *fd;
            if (in->data_on_device) {
                fd = (const char *)in->frames + esz * raw * f0;
            } else {
                ICHK(hipMemcpyAsync(ddata, (const char *)in->frames + esz * raw * f0, esz * raw * nb,
                                    hipMemcpyHostToDevice, s));
                fd = ddata;
            }
            for (int f = 0; f < nb; f++)
                for (int k = 0; k < nk; k++) {
                    kbuf[(size_t)f * nk + k] = in->knots_x[(size_t)(f0 + f) * nk + k];
                    kbuf[(size_t)nk * B + (size_t)f * nk + k] = in->knots_y[(size_t)(f0 + f) * nk + k];
                }
            ICHK(hipMemcpyAsync(dk, kbuf.data(), sizeof(double) * 2 * nk * (size_t)B, hipMemcpyHostToDevice, s));
            ICHK(hipMemsetAsync(C.cnt, 0, sizeof(int) * (size_t)np * nb, s));
            ICHK(hipMemsetAsync(C.win, 0xFF, sizeof(int) * 4 * (size_t)np * nb, s));
            ICHK(hipMemsetAsync(C.any, 0, sizeof(int) * (size_t)nb, s));
            ICHK(hipMemsetAsync(C.pub, 0, sizeof(int) * ((size_t)nb + 1), s));
            ICHK(hipMemcpyAsync(C.pub, &carry, sizeof(int), hipMemcpyHostToDevice, s));
            const dim3 gp((npd + 255) / 256, nb), gt((np + 255) / 256, nb);
            hipLaunchKernelGGL(k_points, gp, dim3(256), 0, s, C);
            hipLaunchKernelGGL(k_scan, dim3(nb), dim3(1024), 0, s, C);
            hipLaunchKernelGGL(k_fill, gp, dim3(256), 0, s, C);
            hipLaunchKernelGGL(k_sort, gt, dim3(256), 0, s, C);
            switch (in->data_type) {
            case SS_F64: hipLaunchKernelGGL(k_assemble<double>, gt, dim3(256), 0, s, C, (const double *)fd); break;
            case SS_F32: hipLaunchKernelGGL(k_assemble<float>, gt, dim3(256), 0, s, C, (const float *)fd); break;
            case SS_U16: hipLaunchKernelGGL(k_assemble<uint16_t>, gt, dim3(256), 0, s, C, (const uint16_t *)fd); break;
            default: hipLaunchKernelGGL(k_assemble<int16_t>, gt, dim3(256), 0, s, C, (const int16_t *)fd); break;
            }
            hipLaunchKernelGGL(k_interact, gt, dim3(256), 0, s, C);
            hipLaunchKernelGGL(k_gs, dim3(nb), dim3(kGsThreads), 0, s, C);
            ICHK(hipGetLastError());
            ICHK(hipMemcpyAsync(tab.data(), C.tab, sizeof(int) * 3 * kNSched * nb, hipMemcpyDeviceToHost, s));
            ICHK(hipMemcpyAsync(any.data(), C.any, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
            ICHK(hipStreamSynchronize(s));
            // resolve the check intervals in frame order (steady_scan.m:326-395)
            for (int f = 0; f < nb; f++) {
                if (!any[f]) {
                    chosen[f] = -2;
                    if (iterations) iterations[f0 + f] = 0;
                    continue;
                }
                const int *t = &tab[(size_t)f * 3 * kNSched];
                const int n = carry;
                if (!(t[2 * kNSched + n - 1] & 1)) {
                    rc = SS_ERR_NUMERICAL;
                    err = "invert_bilinear_sampling: frame " + std::to_string(f0 + f) +
                          ": Gauss-Seidel did not stop within " + std::to_string(cap) + " iterations";
                    goto done;
                }
                chosen[f] = t[kNSched + n - 1];
                if (iterations) iterations[f0 + f] = t[n - 1];
                carry = check_interval(t[n - 1]);
            }
            ICHK(hipMemcpyAsync(dchosen, chosen.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_output, gt, dim3(256), 0, s, C, dchosen, dout, dmask);
            ICHK(hipGetLastError());
            ICHK(hipMemcpyAsync(out + (size_t)np * f0, dout, sizeof(float) * (size_t)np * nb,
                                hipMemcpyDeviceToHost, s));
            ICHK(hipMemcpyAsync(mask_out + (size_t)np * f0, dmask, (size_t)np * nb, hipMemcpyDeviceToHost, s));
            ICHK(hipStreamSynchronize(s));
        }
    }
done:
    (void)hipStreamSynchronize(s);
    if (base) (void)hipFree(base);
    if (dgeo) (void)hipFree(dgeo);
#undef ICHK
    if (rc) return handle_fail(h, rc, err);
    return SS_OK;
}
