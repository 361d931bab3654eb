// ss_prefilter.hip -- the frame preparation and gap filling of steady_scan.m:
// prefilter_foralignment (steady_scan.m:625-643) and imgapfill_leastsquares
// (steady_scan.m:525-560).
//
// prefilter_foralignment blurs each frame with a Gaussian through 2-D FFTs of
// the frame, zero-padded by L pixels on the far sides, and of a wrapped
// Gaussian on the padded grid: a circular convolution.  The wrapped Gaussian
// is separable, g2d(y, x) = g1(y) g2(x) with each factor normalised, so the
// convolution is two circulant matrix products restricted to the H x W
// output: T = C1 D over the rows (all frames at once) and U = T C2^T over
// the columns, run here as FP64 tiled GEMMs whose circulant operand is
// generated from g1/g2 on the fly.  It agrees with the FFT to FP64 rounding.
// The result takes the input's class (integer frames: rounded half away
// from zero and saturated, NaN to 0).  Median normalisation divides each
// frame by its median, found exactly by an 8-pass radix select on the
// order-preserving 64-bit keys of the values.
//
// imgapfill_leastsquares fills the NaN pixels of each frame with the
// minimiser of the summed squared adjacent differences, the known pixels
// fixed: the Dirichlet Laplacian system  deg_i c_i - sum_{missing j~i} c_j =
// sum_{known j~i} a_j  over the missing pixels.  The reference solves it with
// a sparse direct solve; here one workgroup per frame runs conjugate
// gradients in FP64 to a relative residual of 1e-13, with every reduction in
// a fixed order (deterministic).
#include <algorithm>
#include <new>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include <hip/hip_runtime.h>

#include "ss_host.h"
#include "steadyscan.h"

namespace {

constexpr int TM = 64, TN = 64, TK = 16;
constexpr int kCgThreads = 1024;

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// C(y, x) = sum_k A(y, k) B(k, x), 64 x 64 tile per workgroup, 4 x 4 per thread.
// MODE 0: A(y, k) = g[(y - k) mod S] (generated), B = src (ld H, ncols columns),
//         C = dst (ld H); the M = H, N = ncols, K = H.
// MODE 1: frame f = blockIdx.z: A = src_f (H x W, ld H), B(k, x) = g[(x - k) mod S]
//         (generated), C = dst_f; M = H, N = W, K = W.
template <int MODE>
__global__ __launch_bounds__(256) void k_circ_gemm(const double *__restrict__ src, double *__restrict__ dst,
                                                   const double *__restrict__ g, int S, int H, int W,
                                                   int ncols) {
    __shared__ double As[TK][TM + 1];
    __shared__ double Bs[TK][TN + 1];
    const int M = H;
    const int N = MODE == 0 ? ncols : W;
    const int K = MODE == 0 ? H : W;
    const size_t fo = MODE == 0 ? 0 : (size_t)blockIdx.z * H * W;
    const double *s = src + fo;
    double *d = dst + fo;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += TK) {
        for (int e = 0; e < 4; e++) {
            const int idx = tid + 256 * e;
            const int m = idx % TM, kk = idx / TM;
            const int y = m0 + m, k = k0 + kk;
            double v = 0.0;
            if (y < M && k < K) v = MODE == 0 ? g[(y - k + S) % S] : s[(size_t)k * H + y];
            As[kk][m] = v;
        }
        for (int e = 0; e < 4; e++) {
            const int idx = tid + 256 * e;
            const int kk = idx % TK, n = idx / TK;
            const int k = k0 + kk, x = n0 + n;
            double v = 0.0;
            if (k < K && x < N) v = MODE == 0 ? s[(size_t)x * H + k] : g[(x - k + S) % S];
            Bs[kk][n] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk++) {
            double a[4], b[4];
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty + 16 * i];
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx + 16 * j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const int y = m0 + ty + 16 * i, x = n0 + tx + 16 * j;
            if (y < M && x < N) d[(size_t)x * H + y] = acc[i][j];
        }
}

template <typename T>
__global__ __launch_bounds__(256) void k_to_double(const T *__restrict__ in, double *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

// Assignment of a double into an array of class T (MATLAB's conversion).
template <typename T> __device__ inline T to_class(double v);
template <> __device__ inline double to_class<double>(double v) { return v; }
template <> __device__ inline float to_class<float>(double v) { return (float)v; }
template <> __device__ inline uint16_t to_class<uint16_t>(double v) {
    if (isnan(v)) return 0;
    v = round(v);  // half away from zero
    return (uint16_t)fmin(fmax(v, 0.0), 65535.0);
}
template <> __device__ inline int16_t to_class<int16_t>(double v) {
    if (isnan(v)) return 0;
    v = round(v);
    return (int16_t)fmin(fmax(v, -32768.0), 32767.0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_from_double(const double *__restrict__ in, T *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = to_class<T>(in[i]);
}

__device__ inline unsigned long long order_key(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

__device__ inline double key_value(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    return __longlong_as_double((long long)b);
}

// The rank-th smallest value of the frame (all threads of the workgroup call it).
template <typename T>
__device__ double select_rank(const T *v, int np, int rank, int *hist, unsigned long long *s_prefix,
                              int *s_rank) {
    unsigned long long prefix = 0, pmask = 0;
    for (int pass = 0; pass < 8; pass++) {
        const int shift = 56 - 8 * pass;
        for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < np; i += blockDim.x) {
            const unsigned long long k = order_key((double)v[i]);
            if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255], 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0, d = 0;
            for (d = 0; d < 256; d++) {
                if (acc + hist[d] > rank) break;
                acc += hist[d];
            }
            if (d > 255) d = 255;
            *s_rank = rank - acc;
            *s_prefix = prefix | ((unsigned long long)d << shift);
        }
        __syncthreads();
        rank = *s_rank;
        prefix = *s_prefix;
        pmask |= 255ull << shift;
        __syncthreads();
    }
    return key_value(prefix);
}

// MATLAB median of each frame (NaN if the frame holds a NaN), then the
// division data ./ median in the class's arithmetic.
template <typename T>
__global__ __launch_bounds__(1024) void k_median_divide(T *__restrict__ data, int np) {
    __shared__ int hist[256];
    __shared__ unsigned long long s_prefix;
    __shared__ int s_rank, s_nan;
    T *v = data + (size_t)blockIdx.x * np;
    if (threadIdx.x == 0) s_nan = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += blockDim.x)
        if (isnan((double)v[i])) atomicOr(&s_nan, 1);
    __syncthreads();
    T med;
    if (s_nan) {
        med = (T)__builtin_nan("");
    } else if (np % 2) {
        med = (T)select_rank(v, np, np / 2, hist, &s_prefix, &s_rank);
    } else {
        const T a = (T)select_rank(v, np, np / 2 - 1, hist, &s_prefix, &s_rank);
        const T b = (T)select_rank(v, np, np / 2, hist, &s_prefix, &s_rank);
        med = (T)((a + b) / (T)2);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += blockDim.x) v[i] = v[i] / med;
}

// ---------------------------------------------------------------------------
// gap filling
// ---------------------------------------------------------------------------
struct FillChunk {
    int th, tw, np;
    const double *a;       // B x np input (NaN = missing)
    double *out;           // B x np
    int *list;             // B x np: missing pixels, ascending
    int *nmiss;            // B
    double *x, *r, *ap;    // B x np (first nmiss used)
    double *pfull;         // B x np: search direction on the image grid, 0 at known pixels
                           // (k_fill_cg); compact index of each missing pixel (k_fill_cg_lds)
    double *b;             // B x np: right-hand side
    int *status;           // B: iterations, or -1 if not converged
    int maxit;
};

// The missing (NaN) pixels of a frame in ascending order, and their count:
// tiles of 4 * kCgThreads pixels, 4 consecutive ones per thread, a wave
// scan by shuffles and the wave totals through LDS.
__global__ __launch_bounds__(kCgThreads) void k_fill_list(FillChunk C) {
    constexpr int kW = kCgThreads / 64;
    __shared__ int wsum[kW];
    const int f = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    const double *a = C.a + (size_t)f * C.np;
    int *list = C.list + (size_t)f * C.np;
    const int np = C.np;
    int carry = 0;
    for (int base = 0; base < np; base += 4 * kCgThreads) {
        const int i0 = base + 4 * t;
        bool miss[4];
        int own = 0;
        for (int q = 0; q < 4; q++) {
            miss[q] = i0 + q < np && isnan(a[i0 + q]);
            own += miss[q] ? 1 : 0;
        }
        int incl = own;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        int before = 0, tile = 0;
        for (int k = 0; k < kW; k++) {
            const int ws = wsum[k];
            if (k < w) before += ws;
            tile += ws;
        }
        int pos = carry + before + incl - own;
        for (int q = 0; q < 4; q++)
            if (miss[q]) list[pos++] = i0 + q;
        carry += tile;
        __syncthreads();
    }
    if (t == 0) C.nmiss[f] = carry;
}

// Sum over the workgroup of one value per thread: a butterfly in each wave,
// then the 16 wave sums in order.
__device__ double block_sum(double v, double *sp) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sp[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < kCgThreads / 64; w++) r += sp[w];
    __syncthreads();
    return r;
}

// Conjugate gradients for a frame with at most kFillLds missing pixels, in
// compact form (entry m = the m-th missing pixel): the search direction p
// and the residual r live in LDS, the solution x in global memory, and
// thread t keeps, for its entries t, t + 1024, ..., only the degree and the
// compact indices of the missing W N S E neighbours (14 bits each).  Same
// algebra and dot-product order as k_fill_cg; a neighbour's term enters Ap
// only when it is missing (a known neighbour's direction is zero).
constexpr int kFillK = 8;
constexpr int kFillLds = kFillK * kCgThreads;
constexpr unsigned kNoNbr = 0x3fffu;

__global__ __launch_bounds__(kCgThreads) void k_fill_cg_lds(FillChunk C) {
#pragma clang fp contract(off)
    __shared__ double sp[kCgThreads / 64];
    __shared__ double lp[kFillLds], lr[kFillLds];
    const int f = blockIdx.x, t = threadIdx.x;
    const int th = C.th, tw = C.tw, np = C.np;
    const size_t fo = (size_t)f * np;
    const double *a = C.a + fo;
    double *out = C.out + fo, *x = C.x + fo;
    const int *list = C.list + fo;
    int *map = (int *)(C.pfull + fo);  // compact index of each missing pixel
    const int n = C.nmiss[f];
    if (n > kFillLds) return;  // k_fill_cg takes this frame
    for (int i = t; i < np; i += kCgThreads) out[i] = a[i];
    if (n == 0 || n == np) {  // nothing to fill, or no known pixel (left as NaN)
        if (t == 0) C.status[f] = 0;
        return;
    }
    for (int k = 0; k < kFillK; k++) {
        const int m = t + k * kCgThreads;
        if (m < n) map[list[m]] = m;
    }
    __syncthreads();
    unsigned nb01[kFillK], nb23[kFillK];  // id0 | id1 << 14 | degree << 28, id2 | id3 << 14
    double part = 0.0;
    for (int k = 0; k < kFillK; k++) {
        const int m = t + k * kCgThreads;
        nb01[k] = nb23[k] = kNoNbr | (kNoNbr << 14);
        if (m >= n) continue;
        const int i = list[m], row = i % th, col = i / th;
        // right-hand side: the known neighbours, W N S E (ascending index)
        double s = 0.0;
        unsigned d = 0, id[4] = {kNoNbr, kNoNbr, kNoNbr, kNoNbr};
        const bool in[4] = {col > 0, row > 0, row + 1 < th, col + 1 < tw};
        const int j[4] = {i - th, i - 1, i + 1, i + th};
        for (int q = 0; q < 4; q++) {
            if (!in[q]) continue;
            d++;
            const double v = a[j[q]];
            if (isnan(v)) id[q] = (unsigned)map[j[q]];
            else s += v;
        }
        nb01[k] = id[0] | (id[1] << 14) | (d << 28);
        nb23[k] = id[2] | (id[3] << 14);
        lr[m] = s;
        lp[m] = s;
        x[m] = 0.0;
        part += s * s;
    }
    // (A p)_m from the direction in LDS
    auto apply = [&](int k, int m) {
        const unsigned id[4] = {nb01[k] & kNoNbr, (nb01[k] >> 14) & kNoNbr, nb23[k] & kNoNbr,
                                (nb23[k] >> 14) & kNoNbr};
        double s = 0.0;
        for (int q = 0; q < 4; q++)
            if (id[q] != kNoNbr) s += lp[id[q]];
        return (double)(nb01[k] >> 28) * lp[m] - s;
    };
    double rr_ = block_sum(part, sp);  // also orders the LDS writes before the reads
    const double tol2 = 1e-26 * rr_;
    int it = 0;
    bool ok = rr_ <= tol2;
    while (!ok && it < C.maxit) {
        it++;
        part = 0.0;
        for (int k = 0; k < kFillK; k++) {
            const int m = t + k * kCgThreads;
            if (m >= n) break;
            part += lp[m] * apply(k, m);
        }
        const double pap = block_sum(part, sp);
        const double alpha = rr_ / pap;
        part = 0.0;
        for (int k = 0; k < kFillK; k++) {
            const int m = t + k * kCgThreads;
            if (m >= n) break;
            x[m] += alpha * lp[m];
            const double rm = lr[m] - alpha * apply(k, m);
            lr[m] = rm;
            part += rm * rm;
        }
        const double rn = block_sum(part, sp);
        if (rn <= tol2) {
            ok = true;
            break;
        }
        const double beta = rn / rr_;
        rr_ = rn;
        for (int k = 0; k < kFillK; k++) {
            const int m = t + k * kCgThreads;
            if (m >= n) break;
            lp[m] = lr[m] + beta * lp[m];
        }
        __syncthreads();
    }
    for (int k = 0; k < kFillK; k++) {
        const int m = t + k * kCgThreads;
        if (m < n) out[list[m]] = x[m];
    }
    if (t == 0) C.status[f] = ok ? it : -1;
}

__global__ __launch_bounds__(kCgThreads) void k_fill_cg(FillChunk C) {
#pragma clang fp contract(off)
    __shared__ double sp[kCgThreads / 64];
    const int f = blockIdx.x;
    const int th = C.th, tw = C.tw, np = C.np;
    const size_t fo = (size_t)f * np;
    const double *a = C.a + fo;
    double *out = C.out + fo;
    const int *list = C.list + fo;
    double *__restrict__ x = C.x + fo, *__restrict__ r = C.r + fo, *__restrict__ ap = C.ap + fo,
                         *__restrict__ p = C.pfull + fo, *__restrict__ b = C.b + fo;
    const int n = C.nmiss[f];
    if (n <= kFillLds) return;  // k_fill_cg_lds takes this frame
    for (int i = threadIdx.x; i < np; i += kCgThreads) {
        out[i] = a[i];
        p[i] = 0.0;
    }
    __syncthreads();
    if (n == 0 || n == np) {  // nothing to fill, or no known pixel (left as NaN)
        if (threadIdx.x == 0) C.status[f] = 0;
        return;
    }
    // right-hand side: the known neighbours, W N S E (ascending index)
    double part = 0.0;
    for (int m = threadIdx.x; m < n; m += kCgThreads) {
        const int i = list[m], rr = i % th, cc = i / th;
        double s = 0.0;
        if (cc > 0 && !isnan(a[i - th])) s += a[i - th];
        if (rr > 0 && !isnan(a[i - 1])) s += a[i - 1];
        if (rr + 1 < th && !isnan(a[i + 1])) s += a[i + 1];
        if (cc + 1 < tw && !isnan(a[i + th])) s += a[i + th];
        b[m] = s;
        x[m] = 0.0;
        r[m] = s;
        p[i] = s;
        part += s * s;
    }
    double rr_ = block_sum(part, sp);  // also orders p's writes before the reads
    const double tol2 = 1e-26 * rr_;
    int it = 0;
    bool ok = rr_ <= tol2;
    // Each pass is unrolled by 4 so that a thread has several entries' loads
    // in flight; the partial sums still run over m = t, t + 1024, ... in order.
    while (!ok && it < C.maxit) {
        it++;
        part = 0.0;
#pragma unroll 4
        for (int m = threadIdx.x; m < n; m += kCgThreads) {
            const int i = list[m], row = i % th, col = i / th;
            double deg = 0.0, s = 0.0;
            if (col > 0) { deg += 1.0; s += p[i - th]; }
            if (row > 0) { deg += 1.0; s += p[i - 1]; }
            if (row + 1 < th) { deg += 1.0; s += p[i + 1]; }
            if (col + 1 < tw) { deg += 1.0; s += p[i + th]; }
            const double pi = p[i], api = deg * pi - s;
            ap[m] = api;
            b[m] = pi;
            part += pi * api;
        }
        const double pap = block_sum(part, sp);
        const double alpha = rr_ / pap;
        part = 0.0;
#pragma unroll 4
        for (int m = threadIdx.x; m < n; m += kCgThreads) {
            x[m] += alpha * b[m];
            const double rm = r[m] - alpha * ap[m];
            r[m] = rm;
            part += rm * rm;
        }
        const double rn = block_sum(part, sp);
        if (rn <= tol2) {
            ok = true;
            break;
        }
        const double beta = rn / rr_;
        rr_ = rn;
#pragma unroll 4
        for (int m = threadIdx.x; m < n; m += kCgThreads) p[list[m]] = r[m] + beta * b[m];
        __syncthreads();
    }
    for (int m = threadIdx.x; m < n; m += kCgThreads) out[list[m]] = x[m];
    if (threadIdx.x == 0) C.status[f] = ok ? it : -1;
}

size_t esize(ss_dtype t) { return t == SS_F64 ? 8 : t == SS_F32 ? 4 : 2; }

template <typename F>
void dispatch(ss_dtype t, F &&f) {
    switch (t) {
    case SS_F64: f((double *)nullptr); break;
    case SS_F32: f((float *)nullptr); break;
    case SS_U16: f((uint16_t *)nullptr); break;
    default: f((int16_t *)nullptr); break;
    }
}

}  // namespace

#define PCHK(expr)                                                                            \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ss::handle_fail(h, e_ == hipErrorOutOfMemory ? SS_ERR_OUT_OF_MEMORY : SS_ERR_DEVICE, \
                                   std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

extern "C" ss_status ss_prefilter(ss_handle *h, ss_dtype type, const void *in, void *out, int32_t height,
                                  int32_t width, int32_t n_frames, double pregauss, int32_t mednorm,
                                  int32_t on_device) try {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    if (!in || !out || height < 1 || width < 1 || n_frames < 1 || !(pregauss >= 0) || !std::isfinite(pregauss) ||
        type < SS_F64 || type > SS_I16)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "prefilter: bad arguments");
    if (mednorm && (type == SS_U16 || type == SS_I16))
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT,
                               "prefilter: median normalisation of integer frames is not supported");
    if ((double)height * width > 2.0e9)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "prefilter: frame too large");
    const int H = height, W = width;
    const size_t np = (size_t)H * W, es = esize(type);
    std::vector<double> g1, g2;
    int S1 = H, S2 = W;
    if (pregauss > 0) {
        // L = ceil(sqrt((2 gsd^2) log(10000 / (sqrt(2 pi) gsd)))) (steady_scan.m:628)
        const double gsd = pregauss;
        const int L = (int)std::ceil(std::sqrt((2 * gsd * gsd) * std::log(10000 / (std::sqrt(2 * 3.14159265358979323846) * gsd))));
        S1 = H + L;
        S2 = W + L;
        auto make = [&](int S, std::vector<double> &g) {
            g.resize(S);
            double sum = 0;
            for (int j = 0; j < S; j++) {
                const double d = std::min(j, S - j);
                g[j] = std::exp(-(d * d) / (2 * gsd * gsd));
                sum += g[j];
            }
            for (double &v : g) v /= sum;
        };
        make(S1, g1);
        make(S2, g2);
    }
    const size_t perFrame = align256(np * 8) * 2 + (on_device ? 0 : align256(np * es));
    // work memory: half of what is free on the device or already held by the handle
    PCHK(hipSetDevice(ss::handle_device(h)));
    size_t freeB = 0, totalB = 0;
    PCHK(hipMemGetInfo(&freeB, &totalB));
    const size_t budget = std::max<size_t>((size_t)1 << 30, (freeB + ss::handle_scratch_held(h)) / 2);
    const int B = (int)std::max<size_t>(1, std::min<size_t>({(size_t)n_frames, budget / perFrame, (size_t)65535}));
    ss_status rc;
    const size_t total = align256(8 * (size_t)(S1 + S2)) + perFrame * B;
    char *base = (char *)ss::handle_scratch(h, total, &rc);
    if (!base) return rc;
    hipStream_t s = ss::handle_stream(h);
    PCHK(hipSetDevice(ss::handle_device(h)));
    double *dg = (double *)base;
    double *dD = (double *)(base + align256(8 * (size_t)(S1 + S2)));
    double *dT = (double *)((char *)dD + align256(np * 8 * B));
    char *dio = (char *)dT + align256(np * 8 * B);
    if (pregauss > 0) {
        PCHK(hipMemcpyAsync(dg, g1.data(), 8 * S1, hipMemcpyHostToDevice, s));
        PCHK(hipMemcpyAsync(dg + S1, g2.data(), 8 * S2, hipMemcpyHostToDevice, s));
    }
    for (int f0 = 0; f0 < n_frames; f0 += B) {
        const int nb = std::min(B, n_frames - f0);
        const size_t cnt = np * nb;
        const char *src = (const char *)in + es * np * f0;
        char *dst = (char *)out + es * np * f0;
        char *work = on_device ? dst : dio;
        if (!on_device) PCHK(hipMemcpyAsync(work, src, es * cnt, hipMemcpyHostToDevice, s));
        else if (src != dst) PCHK(hipMemcpyAsync(work, src, es * cnt, hipMemcpyDeviceToDevice, s));
        const dim3 g1d((unsigned)((cnt + 255) / 256));
        if (pregauss > 0) {
            dispatch(type, [&](auto *tp) {
                using T = std::remove_pointer_t<decltype(tp)>;
                hipLaunchKernelGGL(k_to_double<T>, g1d, dim3(256), 0, s, (const T *)work, dD, cnt);
            });
            hipLaunchKernelGGL(k_circ_gemm<0>, dim3((W * nb + TN - 1) / TN, (H + TM - 1) / TM, 1), dim3(256), 0, s,
                               dD, dT, dg, S1, H, W, W * nb);
            hipLaunchKernelGGL(k_circ_gemm<1>, dim3((W + TN - 1) / TN, (H + TM - 1) / TM, nb), dim3(256), 0, s,
                               dT, dD, dg + S1, S2, H, W, 0);
            dispatch(type, [&](auto *tp) {
                using T = std::remove_pointer_t<decltype(tp)>;
                hipLaunchKernelGGL(k_from_double<T>, g1d, dim3(256), 0, s, dD, (T *)work, cnt);
            });
        }
        if (mednorm) {
            if (type == SS_F64) hipLaunchKernelGGL(k_median_divide<double>, dim3(nb), dim3(1024), 0, s, (double *)work, (int)np);
            else hipLaunchKernelGGL(k_median_divide<float>, dim3(nb), dim3(1024), 0, s, (float *)work, (int)np);
        }
        PCHK(hipGetLastError());
        if (!on_device) PCHK(hipMemcpyAsync(dst, work, es * cnt, hipMemcpyDeviceToHost, s));
        PCHK(hipStreamSynchronize(s));
    }
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

extern "C" ss_status ss_fill_gaps(ss_handle *h, ss_dtype type, const void *in, void *out, int32_t height,
                                  int32_t width, int32_t n_frames, int32_t on_device) try {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    if (!in || !out || height < 1 || width < 1 || n_frames < 1 || (type != SS_F64 && type != SS_F32))
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "fill_gaps: bad arguments (float32 or float64 frames)");
    if ((double)height * width > 2.0e9)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "fill_gaps: frame too large");
    const int H = height, W = width;
    const size_t np = (size_t)H * W, es = esize(type);
    const size_t perFrame = align256(np * 8) * 7 + align256(np * 4) + align256(8) * 2 +
                            (on_device ? 0 : align256(np * es));
    // work memory: half of what is free on the device or already held by the handle
    PCHK(hipSetDevice(ss::handle_device(h)));
    size_t freeB = 0, totalB = 0;
    PCHK(hipMemGetInfo(&freeB, &totalB));
    const size_t budget = std::max<size_t>((size_t)1 << 30, (freeB + ss::handle_scratch_held(h)) / 2);
    const int B = (int)std::max<size_t>(1, std::min<size_t>({(size_t)n_frames, budget / perFrame, (size_t)65535}));
    ss_status rc;
    char *base = (char *)ss::handle_scratch(h, perFrame * B, &rc);
    if (!base) return rc;
    hipStream_t s = ss::handle_stream(h);
    PCHK(hipSetDevice(ss::handle_device(h)));
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base + off;
        off += align256(bytes);
        return (void *)p;
    };
    FillChunk C{};
    C.th = H;
    C.tw = W;
    C.np = (int)np;
    C.maxit = 200000;
    double *da = (double *)take(np * 8 * B);
    C.a = da;
    C.out = (double *)take(np * 8 * B);
    C.x = (double *)take(np * 8 * B);
    C.r = (double *)take(np * 8 * B);
    C.ap = (double *)take(np * 8 * B);
    C.pfull = (double *)take(np * 8 * B);
    C.b = (double *)take(np * 8 * B);
    C.list = (int *)take(np * 4 * B);
    C.nmiss = (int *)take(4 * (size_t)B);
    C.status = (int *)take(4 * (size_t)B);
    char *dio = on_device ? nullptr : (char *)take(es * np * B);
    std::vector<int> status(B);
    for (int f0 = 0; f0 < n_frames; f0 += B) {
        const int nb = std::min(B, n_frames - f0);
        const size_t cnt = np * nb;
        const char *src = (const char *)in + es * np * f0;
        char *dst = (char *)out + es * np * f0;
        const char *work = src;
        if (!on_device) {
            PCHK(hipMemcpyAsync(dio, src, es * cnt, hipMemcpyHostToDevice, s));
            work = dio;
        }
        const dim3 g1d((unsigned)((cnt + 255) / 256));
        if (type == SS_F64) hipLaunchKernelGGL(k_to_double<double>, g1d, dim3(256), 0, s, (const double *)work, da, cnt);
        else hipLaunchKernelGGL(k_to_double<float>, g1d, dim3(256), 0, s, (const float *)work, da, cnt);
        hipLaunchKernelGGL(k_fill_list, dim3(nb), dim3(kCgThreads), 0, s, C);
        hipLaunchKernelGGL(k_fill_cg_lds, dim3(nb), dim3(kCgThreads), 0, s, C);
        hipLaunchKernelGGL(k_fill_cg, dim3(nb), dim3(kCgThreads), 0, s, C);
        char *o = on_device ? dst : dio;
        if (type == SS_F64) hipLaunchKernelGGL(k_from_double<double>, g1d, dim3(256), 0, s, C.out, (double *)o, cnt);
        else hipLaunchKernelGGL(k_from_double<float>, g1d, dim3(256), 0, s, C.out, (float *)o, cnt);
        PCHK(hipGetLastError());
        PCHK(hipMemcpyAsync(status.data(), C.status, 4 * nb, hipMemcpyDeviceToHost, s));
        if (!on_device) PCHK(hipMemcpyAsync(dst, dio, es * cnt, hipMemcpyDeviceToHost, s));
        PCHK(hipStreamSynchronize(s));
        if (const char *e = std::getenv("SS_FILL_STATS"); e && e[0] == '1') {
            std::vector<int> nm(nb);
            PCHK(hipMemcpy(nm.data(), C.nmiss, 4 * (size_t)nb, hipMemcpyDeviceToHost));
            long sum = 0, msum = 0;
            int mx = 0, nf = 0, inLds = 0, mmax = 0;
            for (int f = 0; f < nb; f++) {
                msum += nm[f];
                mmax = std::max(mmax, nm[f]);
                inLds += nm[f] <= kFillLds;
                if (status[f] > 0) {
                    sum += status[f];
                    mx = std::max(mx, status[f]);
                    nf++;
                }
            }
            std::fprintf(stderr,
                         "[fill] chunk of %d frames: %d filled, CG iterations mean %.1f max %d; missing pixels "
                         "mean %.0f max %d, %d frames in LDS\n",
                         nb, nf, nf ? (double)sum / nf : 0.0, mx, (double)msum / nb, mmax, inLds);
        }
        for (int f = 0; f < nb; f++)
            if (status[f] < 0)
                return ss::handle_fail(h, SS_ERR_NUMERICAL,
                                       "fill_gaps: frame " + std::to_string(f0 + f) +
                                           ": conjugate gradients did not converge");
    }
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}
