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
// GPU form, for a chunk of frames at once (by default all frames of the call):
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
//              energy summed in a fixed order (1024 partial sums, one per
//              lane of a row walk, added by a binary tree) and tracks the
//              stopping rule for all ten possible check intervals at once,
//              keeping a snapshot per interval.  Each interval's stop is
//              published as it happens, and each frame publishes its
//              progress every iteration.  A frame drops the intervals its
//              predecessor's progress rules out, and all but one once its
//              own interval follows from the predecessors' stops;
//   resolve    on the host: walks the frames in order, picks each frame's
//              interval from its predecessor, and
//   output     writes the chosen snapshot (NaN outside the mask).
// Everything is deterministic; all floating-point code is compiled without
// contraction so that it rounds as the sequential restatement does.
#include <algorithm>
#include <new>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
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

// The energy's per-pixel record (48 bytes, natural pixel order): the
// symmetric half of the masked stencil -- diagonal, then the couplings to S,
// NE, E, SE -- and a tail holding Le (float bits, low word) and the "all
// eight neighbours masked" flag (bit 32), so that an energy step loads a
// pixel's record with three 16-byte loads.
enum { QR_D, QR_S, QR_NE, QR_E, QR_SE, QR_TAIL, QR_COUNT };

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
    double *qrec;             // B x np x QR_COUNT: energy records
    float *lk;                // B x np, colour-major in sweep order
    float *rm;                // B x np x 8, colour-major in sweep order
    float *x;                 // B x np, per column even rows then odd rows
    float *snap;              // B x kNSlots x np
    int *stop;                // B x kNSched: 1 + the iteration a check interval stopped at, else 0
    int *slot;                // B x kNSched: the snapshot holding that interval's result
    int *diag;                // B x 3: iterations run, iteration the interval became known, energies evaluated
    unsigned long long *prog; // B: progress word (iteration, possible and stopped intervals), 0 before start
    int carry;                // check interval of the chunk's first frame
    int f0;                   // the chunk's first frame within the call
    int chain;                // frames per chain (the interval restarts at 1 on
                              // the call's frames 0, chain, 2 chain, ...)
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

// Exclusive prefix sum of a frame's per-cell counts into start (start[np] =
// total), clearing the counts for k_fill.  Tiles of 4096 counts, 4
// consecutive ones per thread: a wave scan by shuffles, the 16 wave totals
// through LDS, and a carry from tile to tile.
__global__ __launch_bounds__(1024) void k_scan(Chunk C) {
    __shared__ int wsum[16];
    const int f = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    int *cnt = C.cnt + (size_t)f * C.np;
    int *start = C.start + (size_t)f * (C.np + 1);
    const int np = C.np;
    int carry = 0;
    for (int base = 0; base < np; base += 4096) {
        const int i0 = base + 4 * t;
        int v[4];
        for (int q = 0; q < 4; q++) v[q] = i0 + q < np ? cnt[i0 + q] : 0;
        const int own = v[0] + v[1] + v[2] + v[3];
        int incl = own;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        int before = 0, tile = 0;
        for (int k = 0; k < 16; k++) {
            const int ws = wsum[k];
            if (k < w) before += ws;
            tile += ws;
        }
        int acc = carry + before + incl - own;
        for (int q = 0; q < 4; q++)
            if (i0 + q < np) {
                start[i0 + q] = acc;
                acc += v[q];
                cnt[i0 + q] = 0;
            }
        carry += tile;
        __syncthreads();
    }
    if (t == 0) start[np] = carry;
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
    int head[4];  // each list's next data pixel (INT_MAX once it is exhausted)
    for (int a = 0; a < 4; a++) {
        pos[a] = ok[a] ? start[cellOf[a]] : 0;
        end[a] = ok[a] ? start[cellOf[a] + 1] : 0;
        head[a] = pos[a] < end[a] ? list[pos[a]] : INT_MAX;
    }
    double q[9];
    for (int k = 0; k < 9; k++) q[k] = 0.0;
    double wy = 0.0;
    while (true) {
        int a = -1, best = INT_MAX;
        for (int b = 0; b < 4; b++)
            if (head[b] < best) {
                best = head[b];
                a = b;
            }
        if (a < 0) break;
        pos[a]++;
        head[a] = pos[a] < end[a] ? list[pos[a]] : INT_MAX;
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
// Rm and L are stored colour-major, in the order the sweep visits pixels:
// the colours A (even row, even column), B (even, odd), C (odd, even),
// D (odd, odd) in turn, within a colour by column then row.  Rm of a colour
// with n pixels is two planes of n float4: the N NE E SE coefficients, then
// the S SW W NW ones, so a wave's loads are contiguous.
__device__ inline int colour_rows(int th, int rp) { return (th - rp + 1) / 2; }
// x (and its snapshots) keeps each column's even rows first, then its odd
// rows (h0 = (th + 1) / 2 of them even), so the pixels of one colour in a
// column, and each of their neighbours in turn, are contiguous.
__device__ inline int xrow(int r, int h0) { return (r & 1) * h0 + (r >> 1); }

// Whole-wave lane shifts (DPP wave_shr:1 / wave_shl:1): lane l receives lane
// l - 1's (from_below) or lane l + 1's (from_above) value; the end lane
// keeps its own.
__device__ inline int dpp_wave_shr1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ inline int dpp_wave_shl1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
__device__ inline float from_below(float v) { return __int_as_float(dpp_wave_shr1(__float_as_int(v))); }
__device__ inline float from_above(float v) { return __int_as_float(dpp_wave_shl1(__float_as_int(v))); }
__device__ inline double from_below(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_wave_shr1((int)b), hi = dpp_wave_shr1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double from_above(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_wave_shl1((int)b), hi = dpp_wave_shl1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline int colour_count(int th, int tw, int color) {
    return colour_rows(th, color >> 1) * ((tw - (color & 1) + 1) / 2);
}
__device__ inline int colour_base(int th, int tw, int color) {
    int b = 0;
    for (int k = 0; k < color; k++) b += colour_count(th, tw, k);
    return b;
}

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
    const int color = 2 * (r & 1) + (c & 1);
    const int nColour = colour_count(th, tw, color), base = colour_base(th, tw, color);
    const int kk = (r >> 1) + (c >> 1) * colour_rows(th, r & 1);
    C.lk[fo + base + kk] = L;
    float rm[8];
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
    float4 *rm4 = (float4 *)(C.rm + fo * 8) + 2 * (size_t)base;
    rm4[kk] = make_float4(rm[0], rm[1], rm[2], rm[3]);
    rm4[nColour + kk] = make_float4(rm[4], rm[5], rm[6], rm[7]);
    const float iv = C.iv[fo + i];
    C.x[fo + c * th + xrow(r, (th + 1) / 2)] = isnan(iv) ? 0.0f : iv;
    bool full = m && r >= 1 && r + 1 < th && c >= 1 && c + 1 < tw;
    for (int d = 0; d < NB_COUNT && full; d++) full = mask[(r + kSlotDr[d]) + (c + kSlotDc[d]) * th] != 0;
    const unsigned long long tail = (unsigned long long)__float_as_uint(Lf) | ((full ? 1ull : 0ull) << 32);
    double2 *rec = (double2 *)(C.qrec + (fo + i) * QR_COUNT);
    rec[0] = make_double2(qm[0], qm[1 + NB_S]);
    rec[1] = make_double2(qm[1 + NB_NE], qm[1 + NB_E]);
    rec[2] = make_double2(qm[1 + NB_SE], __longlong_as_double((long long)tail));
}

// ---------------------------------------------------------------------------
// gs: one workgroup per frame
// ---------------------------------------------------------------------------
// Coupling `plane` of pixel j's energy record.
__device__ inline double qrec_at(const double *qrec, int j, int plane) { return qrec[(size_t)j * QR_COUNT + plane]; }

// (Q x)_i for pixel i at (r, c): x_j Q_ij over the in-bounds, masked
// neighbours j in ascending index.  Q_ij is read from the symmetric half
// (the pixel with the smaller index holds the coupling).
__device__ inline double qx_generic(const float *x, const uint8_t *mask, const double *qrec, int i, int r, int c,
                                    int th, int tw, int h0) {
#pragma clang fp contract(off)
    const bool up = r >= 1, dn = r + 1 < th, lt = c >= 1, rt = c + 1 < tw;
    double qx = 0.0;
    if (lt) {
        const int w = (c - 1) * th;
        if (up && mask[i - th - 1]) qx += (double)x[w + xrow(r - 1, h0)] * qrec_at(qrec, i - th - 1, QR_SE);  // NW
        if (mask[i - th]) qx += (double)x[w + xrow(r, h0)] * qrec_at(qrec, i - th, QR_E);                     // W
        if (dn && mask[i - th + 1]) qx += (double)x[w + xrow(r + 1, h0)] * qrec_at(qrec, i - th + 1, QR_NE);  // SW
    }
    const int m = c * th;
    if (up && mask[i - 1]) qx += (double)x[m + xrow(r - 1, h0)] * qrec_at(qrec, i - 1, QR_S);  // N
    qx += (double)x[m + xrow(r, h0)] * qrec_at(qrec, i, QR_D);
    if (dn && mask[i + 1]) qx += (double)x[m + xrow(r + 1, h0)] * qrec_at(qrec, i, QR_S);      // S
    if (rt) {
        const int e = (c + 1) * th;
        if (up && mask[i + th - 1]) qx += (double)x[e + xrow(r - 1, h0)] * qrec_at(qrec, i, QR_NE);  // NE
        if (mask[i + th]) qx += (double)x[e + xrow(r, h0)] * qrec_at(qrec, i, QR_E);                  // E
        if (dn && mask[i + th + 1]) qx += (double)x[e + xrow(r + 1, h0)] * qrec_at(qrec, i, QR_SE);   // SE
    }
    return qx;
}

// Units of a wave-per-row-block walk: rows in blocks of `rows`, columns in
// nSeg segments with nRB * nSeg a multiple of the 16 waves (nSeg = 16 /
// gcd(nRB, 16)), so every wave gets the same number of units.
__device__ inline int walk_segments(int nRB) {
    int a = nRB, b = kGsThreads / 64;
    while (b) {
        const int t = a % b;
        a = b;
        b = t;
    }
    return (kGsThreads / 64) / a;
}

// The stopping energy single(x'Qx - 2 single(Le x)) over the mask in the
// fixed order: (Q x)_i over the neighbours in ascending index; 1024 partial
// sums (thread t), added by a binary tree (t += t + 512, then 256, ... 1).
// Rows are cut into blocks of 62 and columns into nSeg segments (segment s:
// columns s tw / nSeg up to (s + 1) tw / nSeg); unit u is row block
// u mod blocks by segment u / blocks and goes to wave u mod 16.  Lane l of
// the wave holds row 62 rb + l - 1: lanes 1..62 sum their row over the
// segment's columns left to right, lanes 0 and 63 only load the rows above
// and below for them.  Walking a row keeps the W column's x and Q from the
// step before and takes the rows above and below from the neighbouring
// lanes, so a step loads one x and the pixel's own five Q values.  Pixels
// whose eight neighbours are all masked take the interior formula; other
// masked pixels test each term.
__device__ float gs_energy(const Chunk &C, const float *x, const uint8_t *mask, const double *qrec, double *sp,
                           float *sl) {
#pragma clang fp contract(off)
    const int th = C.th, tw = C.tw, h0 = (th + 1) / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kWaves = kGsThreads / 64, kRows = 62;
    const int nRB = (th + kRows - 1) / kRows, nSeg = walk_segments(nRB);
    double pq = 0.0;
    float pl = 0.f;
    for (int u = wave; u < nRB * nSeg; u += kWaves) {
        const int rb = u % nRB, sg = u / nRB;
        const int c0 = (int)((long)sg * tw / nSeg), c1 = (int)((long)(sg + 1) * tw / nSeg);
        const int r = rb * kRows + lane - 1;
        const bool inRow = r >= 0 && r < th;
        const bool sums = inRow && lane >= 1 && lane <= kRows;
        const int row0 = xrow(r, h0);
        // x of the columns c - 1 and c at the rows r - 1, r, r + 1, and the W
        // column's couplings to this pixel: qSE(r-1, c-1) qE(r, c-1) qNE(r+1, c-1)
        float xw = 0.f, xc = 0.f;
        double qw = 0.0, wSE = 0.0, wNE = 0.0;
        if (c0 < c1 && inRow) {
            if (c0 >= 1) {
                const int iw = r + (c0 - 1) * th;
                xw = x[(c0 - 1) * th + row0];
                qw = qrec_at(qrec, iw, QR_E);
                wSE = qrec_at(qrec, iw, QR_SE);
                wNE = qrec_at(qrec, iw, QR_NE);
            }
            xc = x[c0 * th + row0];
        }
        float xnw = from_below(xw), xsw = from_above(xw), xn = from_below(xc), xs = from_above(xc);
        double qnw = from_below(wSE), qsw = from_above(wNE);
        for (int c = c0; c < c1; c++) {
            const int i = inRow ? r + c * th : 0;
            const float xe = (inRow && c + 1 < tw) ? x[(c + 1) * th + row0] : 0.f;
            const float xne = from_below(xe), xse = from_above(xe);
            const double2 *rec = (const double2 *)(qrec + (size_t)i * QR_COUNT);
            const double2 r0 = rec[0], r1 = rec[1], r2 = rec[2];
            const double dD = r0.x, dS = r0.y, dNE = r1.x, dE = r1.y, dSE = r2.x;
            const unsigned long long tail = (unsigned long long)__double_as_longlong(r2.y);
            const double dN = from_below(dS);
            const bool fl = sums && ((tail >> 32) & 1);
            const float l = __uint_as_float((unsigned)tail);
            if (fl) {
                double qx = 0.0;
                qx += (double)xnw * qnw;
                qx += (double)xw * qw;
                qx += (double)xsw * qsw;
                qx += (double)xn * dN;
                qx += (double)xc * dD;
                qx += (double)xs * dS;
                qx += (double)xne * dNE;
                qx += (double)xe * dE;
                qx += (double)xse * dSE;
                pq += (double)xc * qx;
                pl += l * xc;
            } else if (sums && mask[i]) {
                const double qx = qx_generic(x, mask, qrec, i, r, c, th, tw, h0);
                pq += (double)xc * qx;
                pl += l * xc;
            }
            xnw = xn;
            xw = xc;
            xsw = xs;
            xn = xne;
            xc = xe;
            xs = xse;
            qnw = from_below(dSE);
            qw = dE;
            qsw = from_above(dNE);
        }
    }
    const int t = threadIdx.x;
    sp[t] = pq;
    sl[t] = pl;
    __syncthreads();
    for (int off = kGsThreads / 2; off >= 1; off >>= 1) {
        if (t < off) {
            sp[t] += sp[t + off];
            sl[t] += sl[t + off];
        }
        __syncthreads();
    }
    const float e = (float)(sp[0] - 2.0 * (double)sl[0]);
    __syncthreads();  // sp/sl are reused by the next call
    return e;
}

// One iteration: the colours A B C D (1-based odd/odd, odd/even, even/odd,
// even/even row/column) in turn.  Pixels of one colour do not read each
// other, so any order within a colour gives the same result.  A wave takes
// 63 consecutive rows of the colour (lanes 1..63; lane 0 holds the row
// above them) and walks a segment of its columns left to right, so that
//   - the W side of a pixel is the E side loaded one step before,
//   - N and NE are the S and SE of the lane above (a lane shift),
// and a step loads 3 neighbours of x instead of 8.  Neighbours outside the
// image read 0: their coefficient is 0 (a rejected or non-adjacent
// neighbour) and x is finite, so the product adds nothing, as when the
// reference's index reads pixel 0 there.  The sum runs in the reference's
// order N NE E SE S SW W NW.
__device__ void gs_sweep(const Chunk &C, float *x, const float *lk, const float *rm) {
#pragma clang fp contract(off)
    const int th = C.th, tw = C.tw, h0 = (th + 1) / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kWaves = kGsThreads / 64, kRows = 63;
    int base = 0;
    for (int color = 0; color < 4; color++) {
        const int rp = color >> 1, cp = color & 1;
        const int nr = (th - rp + 1) / 2, nc = (tw - cp + 1) / 2;
        const int n = nr * nc;
        const float4 *rv = (const float4 *)rm + 2 * (size_t)base;
        const float *lv = lk + base;
        base += n;
        const int nRB = (nr + kRows - 1) / kRows, nSeg = walk_segments(nRB);
        for (int u = wave; u < nRB * nSeg; u += kWaves) {
            const int rb = u % nRB, sg = u / nRB;
            const int ci0 = (int)((long)sg * nc / nSeg), ci1 = (int)((long)(sg + 1) * nc / nSeg);
            const int ri = rb * kRows + lane - 1;
            const int r = rp + 2 * ri;
            const bool upd = lane >= 1 && ri < nr;
            const bool hasS = (lane == 0 || upd) && r + 1 >= 0 && r + 1 < th;
            const int row0 = xrow(r, h0), rowS = xrow(r + 1, h0);
            float xw = 0.f, xsw = 0.f;
            if (ci0 < ci1 && cp + 2 * ci0 >= 1) {
                const int cw = (cp + 2 * ci0 - 1) * th;
                if (upd) xw = x[cw + row0];
                if (hasS) xsw = x[cw + rowS];
            }
            float xnw = from_below(xsw);
            for (int ci = ci0; ci < ci1; ci++) {
                const int c = cp + 2 * ci, cm = c * th, ce = cm + th;
                const bool hasE = c + 1 < tw;
                const int kk = upd ? ri + ci * nr : 0;
                const float4 a = rv[kk], b = rv[n + kk];
                const float l = lv[kk];
                const float xs = hasS ? x[cm + rowS] : 0.f;
                const float xe = (upd && hasE) ? x[ce + row0] : 0.f;
                const float xse = (hasS && hasE) ? x[ce + rowS] : 0.f;
                const float xn = from_below(xs), xne = from_below(xse);
                float sum = 0.f;
                sum += xn * a.x;
                sum += xne * a.y;
                sum += xe * a.z;
                sum += xse * a.w;
                sum += xs * b.x;
                sum += xsw * b.y;
                sum += xw * b.z;
                sum += xnw * b.w;
                if (upd) x[cm + row0] = l - sum;
                xw = xe;
                xsw = xse;
                xnw = xne;
            }
        }
        __syncthreads();
    }
}

// The check interval of frame f, found by walking the frames before it from
// (*wj, *wn): frame j's interval is wn, and its successor's follows from the
// iteration at which j's schedule wn stopped, once j has published it.
// Frames skipped for an empty mask pass the interval on unchanged.  Returns
// 0 while a predecessor's needed schedule is still running; the walk resumes
// from where it stopped at the next call.
__device__ int walk_interval(const Chunk &C, int f, int *wj, int *wn) {
    for (int steps = 0; *wj < f && steps < 64; steps++) {
        const int j = *wj;
        if (C.any[j]) {
            const int v = __hip_atomic_load(&C.stop[(size_t)j * kNSched + *wn - 1], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
            if (v == 0) return 0;
            *wn = check_interval(v - 1);
        }
        (*wj)++;
    }
    return *wj == f ? *wn : 0;
}

// First frame (chunk-local) of frame f's chain when that chain starts inside
// the chunk, else -1 (the chain comes from an earlier chunk, with C.carry).
__device__ inline int chain_start(const Chunk &C, int f) {
    const int g = C.f0 + f, cs = g - g % C.chain;
    return cs >= C.f0 ? cs - C.f0 : -1;
}

// Progress word of a frame: bits 0-31 the last finished iteration k, 32-41
// the intervals still possible for it, 42-51 those of them that stopped.
__device__ inline unsigned long long pack_progress(int k, unsigned active, unsigned stopped) {
    return (unsigned long long)(unsigned)k | ((unsigned long long)(active & 0x3ffu) << 32) |
           ((unsigned long long)(stopped & 0x3ffu) << 42);
}

// The intervals frame f can still need, from its predecessor's progress:
// frame f's interval is check_interval(the predecessor's iteration count),
// which is check_interval(stop) for a possible interval that stopped, and at
// least check_interval(k + 1) for one still running at iteration k.  Frames
// skipped for an empty mask pass the interval on.  Returns all intervals
// when nothing is known yet (or a stop is not visible yet); the set only
// shrinks as the predecessor progresses.
__device__ unsigned possible_intervals(const Chunk &C, int f) {
    const unsigned all = (1u << kNSched) - 1;
    const int cs = chain_start(C, f), lo = cs > 0 ? cs : 0;
    int j = f - 1;
    while (j >= lo && !C.any[j]) j--;
    if (j < lo) return all;
    const unsigned long long w = __hip_atomic_load(&C.prog[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned act = (unsigned)(w >> 32) & 0x3ffu, stp = (unsigned)(w >> 42) & 0x3ffu;
    const int k = (int)(unsigned)(w & 0xffffffffu);
    if (!act) return all;
    unsigned out = 0;
    for (int n = 1; n <= kNSched; n++) {
        const unsigned bit = 1u << (n - 1);
        if (!(act & bit)) continue;
        if (stp & bit) {
            const int v = __hip_atomic_load(&C.stop[(size_t)j * kNSched + n - 1], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
            if (v == 0) return all;
            out |= 1u << (check_interval(v - 1) - 1);
        } else {
            for (int m = check_interval(k + 1); m <= kNSched; m++) out |= 1u << (m - 1);
        }
    }
    return out ? out : all;
}

__global__ __launch_bounds__(kGsThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_gs(Chunk C) {
    const int f = blockIdx.x;
    if (!C.any[f]) return;
    const size_t fo = (size_t)f * C.np;
    float *x = C.x + fo;
    const float *lk = C.lk + fo, *rm = C.rm + fo * 8;
    const uint8_t *mask = C.mask + fo;
    const double *qrec = C.qrec + fo * QR_COUNT;
    float *snap = C.snap + (size_t)f * kNSlots * C.np;

    __shared__ double s_pq[kGsThreads];
    __shared__ float s_pl[kGsThreads];
    __shared__ float s_eo[kNSched];
    __shared__ int s_slot[kNSched], s_ref[kNSlots];
    __shared__ unsigned s_active, s_stopped;
    __shared__ int s_known, s_needE, s_copy, s_done, s_wj, s_wn, s_knownAt, s_k, s_nE;

    const float e0 = gs_energy(C, x, mask, qrec, s_pq, s_pl);
    if (threadIdx.x == 0) {
        for (int n = 0; n < kNSched; n++) {
            s_eo[n] = e0;
            s_slot[n] = -1;  // xo starts at zero
        }
        for (int s = 0; s < kNSlots; s++) s_ref[s] = 0;
        s_stopped = 0;
        s_nE = 0;
        const int cs = chain_start(C, f);
        s_wj = cs >= 0 ? cs : 0;
        s_wn = cs >= 0 ? 1 : C.carry;
        s_known = walk_interval(C, f, &s_wj, &s_wn);
        s_knownAt = s_known ? 0 : -1;
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
        if (s_needE) e = gs_energy(C, x, mask, qrec, s_pq, s_pl);
        if (threadIdx.x == 0) {
            unsigned pass = 0;
            s_nE += s_needE;
            if (s_needE) {
                for (int n = 1; n <= kNSched; n++) {
                    const unsigned bit = 1u << (n - 1);
                    if (!(s_active & ~s_stopped & bit) || k % n) continue;
                    if (s_eo[n - 1] <= e) {
                        s_stopped |= bit;
                        __hip_atomic_store(&C.stop[(size_t)f * kNSched + n - 1], k + 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
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
                int wj = s_wj, wn = s_wn;
                const int kn = walk_interval(C, f, &wj, &wn);
                s_wj = wj;
                s_wn = wn;
                if (kn) {
                    s_knownAt = k;
                    s_known = kn;
                    s_active = 1u << (kn - 1);
                } else {
                    s_active &= possible_intervals(C, f);
                }
            }
            __hip_atomic_store(&C.prog[f], pack_progress(k, s_active, s_stopped), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            s_copy = copy;
            s_k = k;
            s_done = (s_active & ~s_stopped) == 0 || k >= C.cap;
        }
        __syncthreads();
        if (s_copy >= 0) {
            float4 *dst = (float4 *)(snap + (size_t)s_copy * C.np);
            const float4 *src = (const float4 *)x;
            const int n4 = C.np / 4;
            for (int i = threadIdx.x; i < n4; i += kGsThreads) dst[i] = src[i];
            for (int i = 4 * n4 + threadIdx.x; i < C.np; i += kGsThreads) snap[(size_t)s_copy * C.np + i] = x[i];
        }
        __syncthreads();
        if (s_done) break;
    }
    if (threadIdx.x == 0) {
        for (int n = 0; n < kNSched; n++) C.slot[(size_t)f * kNSched + n] = s_slot[n];
        C.diag[3 * f] = s_k;
        C.diag[3 * f + 1] = s_knownAt;
        C.diag[3 * f + 2] = s_nE;
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
    const int th = C.th, c = i / th;
    const float v = ch < 0 ? 0.0f : C.snap[((size_t)f * kNSlots + ch) * C.np + c * th + xrow(i - c * th, (th + 1) / 2)];
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
}

extern "C" ss_status ss_invert_bilinear_sampling(ss_handle *h, const ss_resample_in *in,
                                                 const ss_ibs_settings *st, float *out,
                                                 uint8_t *mask_out, int32_t *iterations) try {
    using ss::handle_fail;
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    if (!in || !st || !out || !mask_out || !in->geometry || !in->frames || !in->knots_x ||
        !in->knots_y || in->n_frames < 1)
        return handle_fail(h, SS_ERR_INVALID_ARGUMENT, "invert_bilinear_sampling: NULL argument or no frames");
    if (st->inpaint)
        return handle_fail(h, SS_ERR_INVALID_ARGUMENT, "invert_bilinear_sampling: inpainting is a separate call (ss_fill_gaps on the result)");
    if (!(st->pixel_aspect > 0) || !std::isfinite(st->regfac) || st->max_iterations < 0 ||
        st->chunk_frames < 0 || st->work_bytes < 0 || st->chain_frames < 0 || st->reserved != 0)
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
    const bool devOut = st->outputs_on_device != 0;

    // device memory: the geometry, then per frame of a chunk
    const size_t geoBytes = align256(sizeof(PixGeo) * (size_t)npd);
    const size_t perFrame =
        align256(sizeof(double) * 2 * nk) + align256(sizeof(int) * (size_t)npd) * 2 +
        align256(sizeof(int) * (size_t)np) + align256(sizeof(int) * ((size_t)np + 1)) +
        align256(sizeof(int) * 4 * (size_t)np) + align256(sizeof(double) * 9 * (size_t)np) +
        align256(sizeof(double) * (size_t)np) + align256(sizeof(float) * (size_t)np) * 3 +
        align256((size_t)np) + align256(sizeof(int)) + align256(sizeof(double) * QR_COUNT * (size_t)np) +
        align256(sizeof(float) * 8 * (size_t)np) +
        align256(sizeof(int) * 2 * kNSched) + align256(sizeof(int)) * 3 + align256(sizeof(unsigned long long)) +
        (in->data_on_device ? 0 : align256(esz * raw)) +
        (devOut ? 0 : align256(sizeof(float) * (size_t)np) + align256((size_t)np));
    size_t budget = (size_t)st->work_bytes;
    if (budget == 0) {
        // half of the device memory that is free or already held by the handle
        size_t freeB = 0, totalB = 0;
        if (hipSetDevice(ss::handle_device(h)) != hipSuccess || hipMemGetInfo(&freeB, &totalB) != hipSuccess)
            return handle_fail(h, SS_ERR_DEVICE, "invert_bilinear_sampling: hipMemGetInfo failed");
        budget = std::max<size_t>((size_t)1 << 30, (freeB + ss::handle_scratch_held(h)) / 2);
    }
    int B = st->chunk_frames > 0 ? st->chunk_frames
                                 : (int)std::min<size_t>(65535, std::max<size_t>(1, budget / perFrame));
    B = std::min(B, in->n_frames);
    B = std::min(B, 65535);
    const int cap = st->max_iterations > 0 ? st->max_iterations : kDefaultCap;
    const int chain = st->chain_frames > 0 ? st->chain_frames : INT_MAX;

    hipStream_t s = ss::handle_stream(h);
    if (hipSetDevice(ss::handle_device(h)) != hipSuccess)
        return handle_fail(h, SS_ERR_DEVICE, "invert_bilinear_sampling: hipSetDevice failed");
    const size_t total = geoBytes + perFrame * (size_t)B;
    char *base = (char *)ss::handle_scratch(h, total, &rc);
    if (!base) return rc;
    std::string err;
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
        size_t off = 0;
        auto take = [&](size_t bytes) {
            char *p = base + off;
            off += align256(bytes);
            return (void *)p;
        };
        PixGeo *dgeo = (PixGeo *)take(sizeof(PixGeo) * (size_t)npd);
        ICHK(hipMemcpyAsync(dgeo, pg.data(), sizeof(PixGeo) * npd, hipMemcpyHostToDevice, s));
        Chunk C{};
        C.th = th; C.tw = tw; C.np = np; C.npd = npd; C.hc = hc; C.nl = g.nlines;
        C.cl = g.cropping[0]; C.cb = g.cropping[2]; C.nk = nk; C.cap = cap;
        C.chain = chain;
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
        C.qrec = (double *)take(sizeof(double) * QR_COUNT * (size_t)np * B);
        C.lk = (float *)take(sizeof(float) * (size_t)np * B);
        C.rm = (float *)take(sizeof(float) * 8 * (size_t)np * B);
        C.x = (float *)take(sizeof(float) * (size_t)np * B);
        // the snapshots reuse the assembly's raw sums, dead once k_interact has run
        static_assert(sizeof(float) * kNSlots <= sizeof(double) * 9, "snapshots must fit the raw sums");
        C.snap = (float *)C.qraw;
        C.stop = (int *)take(sizeof(int) * kNSched * (size_t)B);
        C.slot = (int *)take(sizeof(int) * kNSched * (size_t)B);
        C.prog = (unsigned long long *)take(sizeof(unsigned long long) * (size_t)B);
        C.diag = (int *)take(sizeof(int) * 3 * (size_t)B);
        int *dchosen = (int *)take(sizeof(int) * (size_t)B);
        void *ddata = in->data_on_device ? nullptr : take(esz * raw * B);
        float *dout = devOut ? nullptr : (float *)take(sizeof(float) * (size_t)np * B);
        uint8_t *dmask = devOut ? nullptr : (uint8_t *)take((size_t)np * B);
        if (off > total) {
            rc = SS_ERR_DEVICE;
            err = "invert_bilinear_sampling: internal sizing error";
            goto done;
        }
        std::vector<int> stop(kNSched * (size_t)B), slot(kNSched * (size_t)B), any(B), chosen(B), diag(3 * (size_t)B);
        // SS_IBS_STATS=1 prints, per chunk, the Gauss-Seidel iterations the
        // workgroups ran and when they learned their check interval
        const char *statsEnv = std::getenv("SS_IBS_STATS");
        const bool stats = statsEnv && statsEnv[0] == '1';
        std::vector<double> kbuf(2 * (size_t)nk * B);
        long needE = 0;  // energies the reference's rule evaluates (statistics only)
        int carry = 1;  // n_iter_error starts at 1 for the call (steady_scan.m:326)
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
            ICHK(hipMemsetAsync(C.stop, 0, sizeof(int) * kNSched * (size_t)nb, s));
            ICHK(hipMemsetAsync(C.prog, 0, sizeof(unsigned long long) * (size_t)nb, s));
            C.carry = carry;
            C.f0 = f0;
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
            ICHK(hipMemcpyAsync(stop.data(), C.stop, sizeof(int) * kNSched * nb, hipMemcpyDeviceToHost, s));
            ICHK(hipMemcpyAsync(slot.data(), C.slot, sizeof(int) * kNSched * nb, hipMemcpyDeviceToHost, s));
            if (stats) ICHK(hipMemcpyAsync(diag.data(), C.diag, sizeof(int) * 3 * nb, hipMemcpyDeviceToHost, s));
            ICHK(hipMemcpyAsync(any.data(), C.any, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
            ICHK(hipStreamSynchronize(s));
            // resolve the check intervals in frame order (steady_scan.m:326-395)
            for (int f = 0; f < nb; f++) {
                if (((long)f0 + f) % chain == 0) carry = 1;  // a new chain (a call of the reference)
                if (!any[f]) {
                    chosen[f] = -2;
                    if (iterations) iterations[f0 + f] = 0;
                    continue;
                }
                const int n = carry;
                const int v = stop[(size_t)f * kNSched + n - 1];
                if (v == 0) {
                    rc = SS_ERR_NUMERICAL;
                    err = "invert_bilinear_sampling: frame " + std::to_string(f0 + f) +
                          ": Gauss-Seidel did not stop within " + std::to_string(cap) + " iterations";
                    goto done;
                }
                chosen[f] = slot[(size_t)f * kNSched + n - 1];
                needE += (v - 1) / n;
                if (iterations) iterations[f0 + f] = v - 1;
                carry = check_interval(v - 1);
            }
            if (stats) {
                long runSum = 0, knownSum = 0, needSum = 0, eSum = 0;
                int runMax = 0, knownMax = -1, unknown = 0, nf = 0;
                for (int f = 0; f < nb; f++) {
                    if (!any[f]) continue;
                    nf++;
                    runSum += diag[3 * f];
                    runMax = std::max(runMax, diag[3 * f]);
                    eSum += diag[3 * f + 2];
                    if (diag[3 * f + 1] < 0) unknown++;
                    else { knownSum += diag[3 * f + 1]; knownMax = std::max(knownMax, diag[3 * f + 1]); }
                }
                for (int f0i = 0; f0i < nb; f0i++)
                    if (any[f0i] && iterations) needSum += iterations[f0 + f0i];
                std::fprintf(stderr,
                             "[ibs] chunk of %d frames: iterations run mean %.1f max %d (needed mean %.1f); "
                             "interval known at mean %.1f max %d, never in %d; energies mean %.1f (needed %.1f)\n",
                             nf, nf ? (double)runSum / nf : 0.0, runMax, nf ? (double)needSum / nf : 0.0,
                             nf - unknown ? (double)knownSum / (nf - unknown) : 0.0, knownMax, unknown,
                             nf ? (double)eSum / nf : 0.0, nf ? (double)needE / nf : 0.0);
                needE = 0;
            }
            ICHK(hipMemcpyAsync(dchosen, chosen.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
            float *o = devOut ? out + (size_t)np * f0 : dout;
            uint8_t *m = devOut ? mask_out + (size_t)np * f0 : dmask;
            hipLaunchKernelGGL(k_output, gt, dim3(256), 0, s, C, dchosen, o, m);
            ICHK(hipGetLastError());
            if (!devOut) {
                ICHK(hipMemcpyAsync(out + (size_t)np * f0, dout, sizeof(float) * (size_t)np * nb,
                                    hipMemcpyDeviceToHost, s));
                ICHK(hipMemcpyAsync(mask_out + (size_t)np * f0, dmask, (size_t)np * nb, hipMemcpyDeviceToHost, s));
            }
            ICHK(hipStreamSynchronize(s));
        }
    }
done:
    (void)hipStreamSynchronize(s);
#undef ICHK
    if (rc) return handle_fail(h, rc, err);
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}
