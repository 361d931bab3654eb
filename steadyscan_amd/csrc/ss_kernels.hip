// ss_kernels.hip -- gfx950 kernels of the align_group_to_template path.
//
// A batch holds any number of alignment groups.  Every group runs its own
// copy of the reference's control flow (Task<T>::alignFrames and
// MainIteration<T>::execute) as a small state machine in device memory
// (GroupCtl), so one host submission ("tick") advances every group by one
// stage without host round trips:
//   k_assemble -> k_solve -> (k_ls_round -> k_ls_decide) x3 -> k_ls_finish
//   -> (the same for the second half-iteration) -> k_corr -> k_corr_finish
//   [-> k_final -> k_final_finish]
// Each kernel skips the groups that are not in its phase.  Without the
// final stages, the next tick's k_assemble and k_solve run the final pass
// and end the groups that stopped.
//
// Pixel work: one workgroup (4 wave64) per line-block column (frame f,
// block b = column c = f * nb + b of the block matrix).  The column is cut
// into chunks of 8 rows dealt round-robin to the threads, so a wave reads
// 64 consecutive chunks (fully coalesced 16-byte loads) and, since a block
// lists its pixels column by column, neighbouring lanes sit on neighbouring
// scan-line pixels: a wave's template gathers touch a few cache lines.  Per-column partial sums are reduced in registers, then
// across waves through LDS in FP64.  No floating-point atomics: every result
// is run-to-run deterministic.
#include "ss_device.h"
#include "ss_tuning.h"

#include <type_traits>

namespace ss {

constexpr int kWG = 256;       // pixel kernels
constexpr int kSolveWG = 256;  // per-group kernels
constexpr int kCH = 8;         // pixels per unrolled chunk

// ---------------------------------------------------------------------------
// precision: pixel arithmetic in R = float (precision 'S') or double ('D');
// every reduction is FP64 in both
// ---------------------------------------------------------------------------
template <typename R>
struct Prec;
template <>
struct Prec<float> {
    using V2 = float2;
    using V4 = float4;
    static __device__ __forceinline__ const AffineGeo<float> &aff(const GroupDev &g) { return g.afF; }
    static __device__ __forceinline__ const float4 *geo(const GroupDev &g) { return g.geo; }
    static __device__ __forceinline__ const float *fnext(const GroupDev &g) { return g.fnext; }
    static __device__ __forceinline__ const float2 *tp(const GroupDev &g) { return g.tp; }
    static __device__ __forceinline__ const float2 *gxv(const GroupDev &g) { return g.gxv; }
    static __device__ __forceinline__ const float2 *gyh(const GroupDev &g) { return g.gyh; }
};
template <>
struct Prec<double> {
    using V2 = double2;
    using V4 = double4;
    static __device__ __forceinline__ const AffineGeo<double> &aff(const GroupDev &g) { return g.afD; }
    static __device__ __forceinline__ const double4 *geo(const GroupDev &g) { return g.geoD; }
    static __device__ __forceinline__ const double *fnext(const GroupDev &g) { return g.fnextD; }
    static __device__ __forceinline__ const double2 *tp(const GroupDev &g) { return g.tpD; }
    static __device__ __forceinline__ const double2 *gxv(const GroupDev &g) { return g.gxvD; }
    static __device__ __forceinline__ const double2 *gyh(const GroupDev &g) { return g.gyhD; }
};

__device__ __forceinline__ float mad(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double mad(double a, double b, double c) { return fma(a, b, c); }

template <typename TI>
__host__ __device__ constexpr int img_type() {
    return std::is_same<TI, float>::value      ? IMG_F32
           : std::is_same<TI, uint16_t>::value ? IMG_U16
           : std::is_same<TI, int16_t>::value  ? IMG_I16
                                               : IMG_F64;
}
// true when group g runs in the instantiation <TI, R>
template <typename TI, typename R>
__device__ __forceinline__ bool runs_here(const GroupDev &g) {
    return g.imgType == img_type<TI>() && g.fp64 == (sizeof(R) == 8 ? 1 : 0);
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void load8(const float *p, bool vec, int n, R (&v)[kCH]) {
    if (vec && n >= kCH) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        const float4 b = *reinterpret_cast<const float4 *>(p + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int i = 0; i < kCH; i++) v[i] = i < n ? (R)p[i] : (R)0;
    }
}
template <typename R>
__device__ __forceinline__ void load8(const double *p, bool vec, int n, R (&v)[kCH]) {
    if (vec && n >= kCH) {
#pragma unroll
        for (int i = 0; i < kCH / 2; i++) {
            const double2 a = *reinterpret_cast<const double2 *>(p + 2 * i);
            v[2 * i] = (R)a.x;
            v[2 * i + 1] = (R)a.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kCH; i++) v[i] = i < n ? (R)p[i] : (R)0;
    }
}
template <typename R>
__device__ __forceinline__ void load8(const uint16_t *p, bool vec, int n, R (&v)[kCH]) {
    if (vec && n >= kCH) {
        const uint4 a = *reinterpret_cast<const uint4 *>(p);
        const unsigned w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[2 * i] = (R)(w[i] & 0xffffu);
            v[2 * i + 1] = (R)(w[i] >> 16);
        }
    } else {
#pragma unroll
        for (int i = 0; i < kCH; i++) v[i] = i < n ? (R)p[i] : (R)0;
    }
}
template <typename R>
__device__ __forceinline__ void load8(const int16_t *p, bool vec, int n, R (&v)[kCH]) {
    if (vec && n >= kCH) {
        const uint4 a = *reinterpret_cast<const uint4 *>(p);
        const unsigned w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[2 * i] = (R)(int16_t)(w[i] & 0xffffu);
            v[2 * i + 1] = (R)((int)w[i] >> 16);
        }
    } else {
#pragma unroll
        for (int i = 0; i < kCH; i++) v[i] = i < n ? (R)p[i] : (R)0;
    }
}

// CH consecutive frame values from p (16-byte loads when the chunk is
// aligned, i.e. vec and a full chunk); values past n read as 0.
template <int CH, typename TI>
struct alignas(CH * sizeof(TI) >= 16 ? 16 : CH * sizeof(TI)) Pack {
    TI v[CH];
};
template <int CH, typename TI, typename R>
__device__ __forceinline__ void load_chunk(const TI *p, bool vec, int n, R (&v)[CH]) {
    if constexpr (CH == kCH) {
        load8(p, vec, n, v);
    } else {
        if (vec && n >= CH) {
            constexpr int per = (CH * (int)sizeof(TI) <= 16) ? CH : 16 / (int)sizeof(TI);
#pragma unroll
            for (int j = 0; j < CH / per; j++) {
                const Pack<per, TI> q = *reinterpret_cast<const Pack<per, TI> *>(p + j * per);
#pragma unroll
                for (int i = 0; i < per; i++) v[j * per + i] = (R)q.v[i];
            }
        } else {
#pragma unroll
            for (int i = 0; i < CH; i++) v[i] = i < n ? (R)p[i] : (R)0;
        }
    }
}

// Rows per chunk: enough independent template gathers per wave to hide
// their latency, few enough that the unrolled chunk stays in registers
// (the kernel's *_SAMPLES tuning value).
template <typename R>
__host__ __device__ constexpr int chunk_rows(int samplesPerRow, int budget) {
    const int c = budget / samplesPerRow;
    return c < 1 ? 1 : (c > kCH ? kCH : c);
}

// Sum over the workgroup of NV per-thread doubles; thread i < NV gets sum i
// in out (LDS-backed, returns after a barrier).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *out) {
    __shared__ double sh[kWG / 64][NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double x = v[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[i] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) sh[wave][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kWG / 64; w++) s += sh[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ int block_min(int v) {
    __shared__ int sh[kWG / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    int m = sh[0];
#pragma unroll
    for (int w = 1; w < kWG / 64; w++) m = min(m, sh[w]);
    __syncthreads();
    return m;
}

__device__ __forceinline__ double reduce_pow(double v, double r, int k) {
    for (int i = 0; i < k; i++) v *= r;
    return v;
}

// Geometry of the pixels of one thread segment: affine in the block's
// (column j, line i) coordinates -- the block matrix lists a block's pixels
// column-major, row r = j * L + i -- or read from the per-pixel table.
template <typename R>
struct GeoIter {
    R fj, fi, Lf, byb;
    __device__ __forceinline__ void start(const GroupDev &g, int b, int r0) {
        if (g.affine) {
            const AffineGeo<R> &a = Prec<R>::aff(g);
            const int j = r0 / g.L;
            fj = (R)j;
            fi = (R)(r0 - j * g.L);
            Lf = (R)g.L;
            byb = a.gy0 + a.gyb * (R)b;
        }
    }
    __device__ __forceinline__ void at(const GroupDev &g, const typename Prec<R>::V4 *geo, int r, R &bx,
                                       R &by, R &f, bool &m) const {
        if (g.affine) {
            const AffineGeo<R> &a = Prec<R>::aff(g);
            f = mad(a.fi, fi, mad(a.fj, fj, a.f0));
            bx = mad(a.gxi, fi, mad(a.gxj, fj, a.gx0));
            by = mad(a.gyi, fi, mad(a.gyj, fj, byb));
            m = true;
        } else {
            const typename Prec<R>::V4 q = geo[r];
            bx = q.x;
            by = q.y;
            f = q.z;
            m = q.w != (R)0;
        }
    }
    __device__ __forceinline__ void next() {
        fi += (R)1;
        if (fi == Lf) {
            fi = (R)0;
            fj += (R)1;
        }
    }
};

template <typename R>
__device__ __forceinline__ bool in_bounds(const GroupDev &g, R x, R y) {
    // non-short-circuit: four compares combined without branches
    return (x >= (R)0) & (x < (R)g.xmax) & (y >= (R)0) & (y < (R)g.ymax);
}

// Two BF16 values packed in 32 bits (low half first) as a float pair.
__device__ __forceinline__ float2 bf16_pair(uint32_t u) {
    return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
}

// Bilinear template sample at (x, y) (task_differenceimage.cuh:13-32) from
// the paired layout.  Branch-free: a point outside the mask reads cell 0 (the
// caller discards the value), so a wave can issue all gathers of a chunk
// before waiting on any.
template <typename R>
struct Cell {
    R fx, fy;
    typename Prec<R>::V2 top, bot;
    int t;
};
template <typename R>
__device__ __forceinline__ R sample(const GroupDev &g, R x, R y, bool valid, Cell<R> &cl) {
    const R xs = valid ? x : (R)0, ys = valid ? y : (R)0;
    const int ix = (int)xs, iy = (int)ys;
    cl.fx = xs - (R)ix;
    cl.fy = ys - (R)iy;
    cl.t = iy * g.tw + ix;
    const typename Prec<R>::V2 *tp = Prec<R>::tp(g);
    cl.top = tp[cl.t];
    cl.bot = tp[cl.t + g.tw];
    const R a = mad(cl.fx, cl.top.y - cl.top.x, cl.top.x);
    const R b = mad(cl.fx, cl.bot.y - cl.bot.x, cl.bot.x);
    return mad(cl.fy, b - a, a);
}

// Template gradients at the sample of cl (task_updategradients.cu): from
// the BF16 gradient pairs in mixed precision, else derived from the cell
// corners when they are forward differences, else from the gradient pairs.
template <typename R>
__device__ __forceinline__ void gradients(const GroupDev &g, const Cell<R> &cl, R &wx, R &wy) {
    if constexpr (sizeof(R) == 4) {
        if (g.gradBF16) {  // pairs (G[t], G[t + step]) as two BF16 values
            const float2 gx = bf16_pair(g.gxvB[cl.t]), gy = bf16_pair(g.gyhB[cl.t]);
            wx = mad(cl.fy, gx.y - gx.x, gx.x);
            wy = mad(cl.fx, gy.y - gy.x, gy.x);
            return;
        }
    }
    if (g.gradFD) {
        wx = mad(cl.fy, (cl.bot.y - cl.bot.x) - (cl.top.y - cl.top.x), cl.top.y - cl.top.x);
        wy = mad(cl.fx, (cl.bot.y - cl.top.y) - (cl.bot.x - cl.top.x), cl.bot.x - cl.top.x);
        return;
    }
    const typename Prec<R>::V2 gx = Prec<R>::gxv(g)[cl.t], gy = Prec<R>::gyh(g)[cl.t];
    wx = mad(cl.fy, gx.y - gx.x, gx.x);
    wy = mad(cl.fx, gy.y - gy.x, gy.x);
}

// Knots of a column for candidate k of the line search (or the half-
// iteration start state for k < 0), as x = base_x + a + d * frac.
template <typename R>
struct ColKnots {
    R ax, dx, ay, dy;
};
template <typename R>
__device__ __forceinline__ ColKnots<R> make_knots(double lx, double rx, double ly, double ry) {
    return ColKnots<R>{(R)lx, (R)(rx - lx), (R)ly, (R)(ry - ly)};
}
template <typename R>
__device__ __forceinline__ ColKnots<R> state_knots(const GroupDev &g, int c, int k) {
    const int N = g.N;
    double lx, ly, rx, ry;
    if (k < 0) {  // half-iteration start: p0 with the real right knot
        lx = g.p0[c];
        ly = g.p0[c + N];
        rx = g.p0[c + 1];
        ry = g.p0[c + 1 + N];
    } else {
        const double r = g.rmult;
        lx = g.p0[c] + reduce_pow(g.d0[c], r, k);
        ly = g.p0[c + N] + reduce_pow(g.d0[c + N], r, k);
        if (c + 1 == g.fframeW) {  // pnew of the group's last knot is never written
            rx = 1e6;
            ry = 1e6;
        } else {
            rx = g.p0[c + 1] + reduce_pow(g.d0[c + 1], r, k);
            ry = g.p0[c + 1 + N] + reduce_pow(g.d0[c + 1 + N], r, k);
        }
    }
    return make_knots<R>(lx, rx, ly, ry);
}

// Knots of column c from the current parameter vector p (applyParameters on p).
template <typename R>
__device__ __forceinline__ ColKnots<R> current_knots(const GroupDev &g, int c) {
    const int N = g.N;
    return make_knots<R>(g.p[c], g.p[c + 1], g.p[c + N], g.p[c + 1 + N]);
}

template <typename TI>
__device__ __forceinline__ const TI *column_ptr(const GroupDev &g, int c) {
    return reinterpret_cast<const TI *>(g.img) + (size_t)c * g.frameH;
}

// ---------------------------------------------------------------------------
// assembly pass: warp + residual + errval terms + J^T J / J^T d partials
// (d_applyParameters, d_createDifferenceImage, d_calcBaseErrval,
//  d_updateGradients, d_fillDiagonals, d_fillJTr in one pass)
// ---------------------------------------------------------------------------
template <typename TI, typename R, int V>
__device__ void assemble_column(const GroupDev &g, int c) {
    const int b = c % g.nb;
    const ColKnots<R> kn = current_knots<R>(g, c);
    const TI *img = column_ptr<TI>(g, c);
    const typename Prec<R>::V4 *geo = Prec<R>::geo(g) ? Prec<R>::geo(g) + (size_t)b * g.frameH : nullptr;
    const R *fnx = g.fracUniform ? nullptr : Prec<R>::fnext(g) + (size_t)b * g.frameH;
    // h[1] = sum I, h[i - 1] = column sum i for i >= CS_AXX; the errval term
    // is accumulated in FP64 per pixel: it decides line-search acceptance
    // against near-equal candidates
    double d2 = 0.0;
    R h[CS_N - 1];
#pragma unroll
    for (int i = 0; i < CS_N - 1; i++) h[i] = (R)0;
    int cnt = 0;
    constexpr int CH = chunk_rows<R>(2, tune_variant<TK_ASSEMBLE, R, V>().samples);
    for (int run = threadIdx.x; run * g.runLen < g.frameH; run += kWG) {
        const int rend = min(run * g.runLen + g.runLen, g.frameH);
        GeoIter<R> it;
        it.start(g, b, run * g.runLen);
        for (int r0 = run * g.runLen; r0 < rend; r0 += CH) {
            const int n = min(CH, rend - r0);
            R I[CH];
            load_chunk<CH>(img + r0, g.vec, n, I);
#pragma unroll
            for (int k = 0; k < CH; k++) {
                const int rr = min(r0 + k, g.frameH - 1);
                R bx, by, f;
                bool m;
                it.at(g, geo, rr, bx, by, f, m);
                it.next();
                h[CS_SUMI] += I[k];
                const R x = mad(kn.dx, f, bx + kn.ax);
                const R y = mad(kn.dy, f, by + kn.ay);
                const bool ok = k < n && m && in_bounds(g, x, y);
                Cell<R> cl;
                const R s = sample(g, x, y, ok, cl);
                R wx, wy;
                gradients(g, cl, wx, wy);
                // outside the mask every term below is zero
                const R d = ok ? I[k] - s : (R)0;
                wx = ok ? wx : (R)0;
                wy = ok ? wy : (R)0;
                const R cf = (R)1 - f;
                const R xx = wx * wx, yy = wy * wy, xy = wx * wy;
                const R c2 = cf * cf, f2 = f * f, fc = f - f2;
                cnt += ok ? 1 : 0;
                d2 = fma((double)d, (double)d, d2);
                h[CS_AXX - 1] = mad(xx, c2, h[CS_AXX - 1]);
                h[CS_AYY - 1] = mad(yy, c2, h[CS_AYY - 1]);
                h[CS_AXY - 1] = mad(xy, c2, h[CS_AXY - 1]);
                h[CS_BXX - 1] = mad(xx, f2, h[CS_BXX - 1]);
                h[CS_BYY - 1] = mad(yy, f2, h[CS_BYY - 1]);
                h[CS_BXY - 1] = mad(xy, f2, h[CS_BXY - 1]);
                h[CS_CXX - 1] = mad(xx, fc, h[CS_CXX - 1]);
                h[CS_CYY - 1] = mad(yy, fc, h[CS_CYY - 1]);
                h[CS_CXY - 1] = mad(xy, fc, h[CS_CXY - 1]);
                const R dc = d * cf, df = d * f;
                h[CS_JXA - 1] = mad(dc, wx, h[CS_JXA - 1]);
                h[CS_JYA - 1] = mad(dc, wy, h[CS_JYA - 1]);
                h[CS_JXB - 1] = mad(df, wx, h[CS_JXB - 1]);
                h[CS_JYB - 1] = mad(df, wy, h[CS_JYB - 1]);
                if (fnx) {
                    const R fn = fnx[rr];
                    const R n2 = fn * fn;
                    h[CS_QXX - 1] = mad(xx, n2, h[CS_QXX - 1]);
                    h[CS_QYY - 1] = mad(yy, n2, h[CS_QYY - 1]);
                    h[CS_QXY - 1] = mad(xy, n2, h[CS_QXY - 1]);
                }
            }
        }
    }
    double v[CS_N];
    v[CS_D2] = d2;
    v[CS_SUMI] = (double)h[CS_SUMI];
    v[CS_CNT] = (double)cnt;
#pragma unroll
    for (int i = CS_AXX; i < CS_N; i++) v[i] = (double)h[i - 1];
    __shared__ double out[CS_N];
    block_sum<CS_N>(v, out);
    if (threadIdx.x < CS_N) g.colsum[(size_t)c * CS_N + threadIdx.x] = out[threadIdx.x];
}

template <typename TI, typename R>
__device__ __forceinline__ void final_column(const GroupDev &g, int c);

// The assembly launch also runs the final pass of the groups that stopped
// in the previous tick (phase PH_FINAL), for a tick without final stages.
template <typename TI, typename R, int V>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(tune_variant<TK_ASSEMBLE, R, V>().waves))) void k_assemble(Batch B) {
    for (int item = blockIdx.x; item < B.nItems; item += gridDim.x) {
        const int gi = B.itemGroup[item];
        const int phase = B.ctl[gi].phase;
        if (phase != PH_ASSEMBLE && phase != PH_FINAL) continue;
        const GroupDev &g = B.groups[gi];
        if (!runs_here<TI, R>(g) || g.var[TK_ASSEMBLE] != V) continue;
        if (phase == PH_FINAL) {
            final_column<TI, R>(g, item - g.itemStart);
            continue;
        }
        if ((g.iterTimes || g.funcTimes) && item == g.itemStart && threadIdx.x == 0) {
            GroupCtl &c = B.ctl[gi];  // the first iteration starts (outputs 9 and 10)
            if (c.tIter == 0 || c.tStage == 0) {
                const unsigned long long now = wall_clock64();
                if (c.tIter == 0) c.tIter = now;
                if (c.tStage == 0) c.tStage = now;
            }
        }
        assemble_column<TI, R, V>(g, item - g.itemStart);
    }
}

// ---------------------------------------------------------------------------
// solve: errvals, knot usage, block-tridiagonal normal equations, exact FP64
// solve (partitioned block Thomas + cyclic reduction of the interface
// system), step filter, red/green expansion and line-search set-up.  One
// workgroup per group.
// ---------------------------------------------------------------------------
// level-0 equation of knot i:  O_{i-1} x_{i-1} + D_i x_i + O_i x_{i+1} = g_i,
// with D_i, O_i symmetric 2x2 ([xx xy; xy yy]).
struct Eq0 {
    double b0, b1, b2;  // D
    double c0, c1, c2;  // O (coupling to knot i+1)
    double d0, d1;      // rhs
};

struct M2 {  // general 2x2, row-major
    double a, b, c, d;
};
__device__ __forceinline__ M2 sym(double x0, double x1, double x2) { return M2{x0, x1, x1, x2}; }
__device__ __forceinline__ M2 inv(const M2 &m) {
    const double id = 1.0 / (m.a * m.d - m.b * m.c);
    return M2{m.d * id, -m.b * id, -m.c * id, m.a * id};
}
__device__ __forceinline__ M2 mul(const M2 &x, const M2 &y) {
    return M2{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c,
              x.c * y.b + x.d * y.d};
}
__device__ __forceinline__ M2 add(const M2 &x, const M2 &y) { return M2{x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
__device__ __forceinline__ M2 sub(const M2 &x, const M2 &y) { return M2{x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d}; }
__device__ __forceinline__ M2 neg(const M2 &x) { return M2{-x.a, -x.b, -x.c, -x.d}; }
__device__ __forceinline__ M2 zero2() { return M2{0, 0, 0, 0}; }
__device__ __forceinline__ M2 eye2() { return M2{1, 0, 0, 1}; }
__device__ __forceinline__ double2 mulv(const M2 &m, double2 v) {
    return make_double2(m.a * v.x + m.b * v.y, m.c * v.x + m.d * v.y);
}

// x as an affine function of the neighbouring interface unknowns:
// x = y + V z_prev + W z_own
struct Aff {
    double2 y;
    M2 V, W;
};
__device__ __forceinline__ Aff aff_apply(const M2 &m, const Aff &a) {
    return Aff{mulv(m, a.y), mul(m, a.V), mul(m, a.W)};
}
__device__ __forceinline__ Aff aff_sub(const Aff &a, const Aff &b) {
    return Aff{make_double2(a.y.x - b.y.x, a.y.y - b.y.y), sub(a.V, b.V), sub(a.W, b.W)};
}

struct RedEq {  // reduced (interface) equation, general 2x2 blocks
    M2 A, B, C;
    double2 d, x;
};

struct InteriorRow {  // forward-eliminated interior row: Bh^{-1} and rhs
    M2 Bi;
    Aff R;
};

// Solver id 0 (SolverCuSparse::solve, cusparsesolver.cu:41-174): incomplete
// LU with zero fill of the normal equations ordered [x_0..x_{n-1},
// y_0..y_{n-1}] (the compacted COO order; unused knots are decoupled
// identity rows and change nothing), then the unit-lower and upper
// triangular solves.  With this ordering every row of the pattern is known:
// x_i: {x_{i-1}, x_i, x_{i+1}, y_{i-1}, y_i, y_{i+1}}, y_i likewise, so the
// elimination reduces to the recurrences below, applied in column order
// exactly as a row-wise ILU(0) does (separately rounded products).
struct IluRow {
    double lx;                        // x row: l(x_i, x_{i-1})
    double uxx, uxn, uxyp, uxy, uxyn; // u(x_i, x_i), u(x_i, x_{i+1}), u(x_i, y_{i-1}), u(x_i, y_i), u(x_i, y_{i+1})
    double l1, l2, l3, l4;            // y row: l(y_i, x_{i-1}), l(y_i, x_i), l(y_i, x_{i+1}), l(y_i, y_{i-1})
    double uyy, uyn;                  // u(y_i, y_i), u(y_i, y_{i+1})
};
static_assert(sizeof(IluRow) <= sizeof(InteriorRow), "ILU factors reuse the interior-row workspace");

__device__ void ilu0_solve(const Eq0 *E, IluRow *F, double2 *X, int n) {
#pragma clang fp contract(off)
    // x rows: the only lower entry is x_{i-1}
    for (int i = 0; i < n; i++) {
        const Eq0 q = E[i];
        IluRow r;
        r.uxx = q.b0;
        r.uxn = q.c0;
        r.uxy = q.b1;
        r.uxyn = q.c1;
        r.uxyp = 0.0;
        r.lx = 0.0;
        if (i > 0) {
            const Eq0 qp = E[i - 1];
            const IluRow &p = F[i - 1];
            r.uxyp = qp.c1;
            r.lx = qp.c0 / p.uxx;
            r.uxx = r.uxx - r.lx * p.uxn;
            r.uxyp = r.uxyp - r.lx * p.uxy;
            r.uxy = r.uxy - r.lx * p.uxyn;
        }
        F[i] = r;
    }
    // y rows: lower entries x_{i-1}, x_i, x_{i+1}, y_{i-1}
    for (int i = 0; i < n; i++) {
        const Eq0 q = E[i];
        IluRow &r = F[i];
        double ayxp = 0.0, ayx = q.b1, ayxn = q.c1, ayyp = 0.0, ayy = q.b2, ayyn = q.c2;
        r.l1 = r.l3 = r.l4 = 0.0;
        if (i > 0) {
            const IluRow &p = F[i - 1];
            ayxp = E[i - 1].c1;
            ayyp = E[i - 1].c2;
            r.l1 = ayxp / p.uxx;
            ayx = ayx - r.l1 * p.uxn;
            ayyp = ayyp - r.l1 * p.uxy;
            ayy = ayy - r.l1 * p.uxyn;
        }
        r.l2 = ayx / r.uxx;
        ayxn = ayxn - r.l2 * r.uxn;
        ayyp = ayyp - r.l2 * r.uxyp;
        ayy = ayy - r.l2 * r.uxy;
        ayyn = ayyn - r.l2 * r.uxyn;
        if (i + 1 < n) {
            const IluRow &nx = F[i + 1];
            r.l3 = ayxn / nx.uxx;
            ayy = ayy - r.l3 * nx.uxyp;
            ayyn = ayyn - r.l3 * nx.uxy;
        }
        if (i > 0) {
            const IluRow &p = F[i - 1];
            r.l4 = ayyp / p.uyy;
            ayy = ayy - r.l4 * p.uyn;
        }
        r.uyy = ayy;
        r.uyn = ayyn;
    }
    // forward: z = L^{-1} g (z kept in X)
    for (int i = 0; i < n; i++) {
        double zx = E[i].d0;
        if (i > 0) zx = zx - F[i].lx * X[i - 1].x;
        X[i].x = zx;
    }
    for (int i = 0; i < n; i++) {
        const IluRow &r = F[i];
        double zy = E[i].d1;
        if (i > 0) zy = zy - r.l1 * X[i - 1].x;
        zy = zy - r.l2 * X[i].x;
        if (i + 1 < n) zy = zy - r.l3 * X[i + 1].x;
        if (i > 0) zy = zy - r.l4 * X[i - 1].y;
        X[i].y = zy;
    }
    // backward: y rows from the last, then x rows
    for (int i = n - 1; i >= 0; i--) {
        double s = X[i].y;
        if (i + 1 < n) s = s - F[i].uyn * X[i + 1].y;
        X[i].y = s / F[i].uyy;
    }
    for (int i = n - 1; i >= 0; i--) {
        const IluRow &r = F[i];
        double s = X[i].x;
        if (i + 1 < n) s = s - r.uxn * X[i + 1].x;
        if (i > 0) s = s - r.uxyp * X[i - 1].y;
        s = s - r.uxy * X[i].y;
        if (i + 1 < n) s = s - r.uxyn * X[i + 1].y;
        X[i].x = s / r.uxx;
    }
}

// Knot usage (d_createParameterMask): knot k < n is used when block column
// k or k - 1 reaches its block's minimum point count; knots from n on never
// are.
__device__ __forceinline__ bool knot_used(const double *__restrict__ S, const int *__restrict__ minPts, int k,
                                          int n, int nb) {
    if (k < 0 || k >= n) return false;
    const bool pc = S[(size_t)k * CS_N + CS_CNT] >= (double)minPts[k % nb];
    const bool pp = k > 0 && S[(size_t)(k - 1) * CS_N + CS_CNT] >= (double)minPts[(k - 1) % nb];
    return pc || pp;
}

// Level-0 equations of knots tid, tid + kSolveWG, ... (d_fillDiagonals /
// d_fillJTr entries, lambda extension), and the usage flags of the same
// knots into `used`.  Each knot takes its own and its neighbours' usage from
// the column counts, so that no pass over the flags has to come first.  The
// pointers do not alias, so the next knot's loads may be issued before this
// knot's equation is stored.  E may point to LDS.
__device__ __forceinline__ void level0_equations(const double *__restrict__ S, const int *__restrict__ minPts,
                                                 uint8_t *__restrict__ used, Eq0 *__restrict__ E, int n, int nb,
                                                 double lambda, int fracUniform, int tid) {
#pragma unroll 4
    for (int k = tid; k < n; k += kSolveWG) {
        Eq0 e;
        const double *s = S + (size_t)k * CS_N;
        const bool u = knot_used(S, minPts, k, n, nb);
        const bool up = k > 0 && knot_used(S, minPts, k - 1, n, nb);
        const bool un = (k + 1 < n) && knot_used(S, minPts, k + 1, n, nb);
        used[k] = u ? 1 : 0;
        if (!u) {
            e = Eq0{1, 0, 1, 0, 0, 0, 0, 0};
        } else {
            double dxx = s[CS_AXX], dyy = s[CS_AYY], dxy = s[CS_AXY];
            double jx = s[CS_JXA], jy = s[CS_JYA];
            if (k > 0) {
                const double *sp = s - CS_N;
                const int q = fracUniform ? CS_BXX : CS_QXX;
                dxx += sp[q];
                dyy += sp[q + 1];
                dxy += sp[q + 2];
                jx += sp[CS_JXB];
                jy += sp[CS_JYB];
            }
            const double nl = (double)(up ? 1 : 0) + (double)(un ? 1 : 0);
            e.b0 = dxx + lambda * nl;
            e.b1 = dxy;
            e.b2 = dyy + lambda * nl;
            if (un) {
                e.c0 = s[CS_CXX] - lambda;
                e.c1 = s[CS_CXY];
                e.c2 = s[CS_CYY] - lambda;
            } else {
                e.c0 = e.c1 = e.c2 = 0;
            }
            e.d0 = jx;
            e.d1 = jy;
        }
        E[k] = e;
    }
}

// Interface system of the partitioned solve: thread t < P contributes the
// equation of its partition's last knot z_t in terms of z_{t-1} and z_{t+1}
// (its last interior row as lastA, the next partition's first interior row
// as firstAff[t + 1]); cyclic reduction solves it.  Called by all threads;
// returns z_t and z_{t-1} (0 for t = 0).
__device__ void solve_interfaces(bool interior, const Aff &first, const Aff &lastA, const Eq0 &q, const M2 &A,
                                 Aff *firstAff, int P, int tid, double2 &zo, double2 &zp) {
    __shared__ RedEq red[kSolveWG];
    __shared__ int hasInt[kSolveWG];
    if (tid < P) {
        hasInt[tid] = interior ? 1 : 0;
        firstAff[tid] = first;
    }
    __syncthreads();
    if (tid < P) {
        const M2 C = sym(q.c0, q.c1, q.c2);
        // x_{j+1}: first interior of partition t+1, or z_{t+1}
        Aff F{make_double2(0, 0), zero2(), eye2()};
        if (tid + 1 < P && hasInt[tid + 1]) F = firstAff[tid + 1];
        RedEq r;
        r.A = mul(A, lastA.V);
        r.B = add(add(sym(q.b0, q.b1, q.b2), mul(A, lastA.W)), mul(C, F.V));
        r.C = mul(C, F.W);
        const double2 ay = mulv(A, lastA.y), cy = mulv(C, F.y);
        r.d = make_double2(q.d0 - ay.x - cy.x, q.d1 - ay.y - cy.y);
        r.x = make_double2(0, 0);
        if (tid == 0) r.A = zero2();
        red[tid] = r;
    }
    __syncthreads();
    // cyclic reduction, in place: level-l equation k lives at k*2^l + 2^l - 1
    int nlev = 0;
    for (int mm = P; mm > 1; mm >>= 1) nlev++;
    for (int l = 0; l < nlev; l++) {
        const int st = 1 << l, mm = P >> l;
        for (int jj = tid; jj < (mm >> 1); jj += kSolveWG) {
            const int i = (2 * jj + 1) * st + st - 1;
            const int im = i - st, ip = i + st;
            RedEq e = red[i];
            const RedEq em = red[im];
            const M2 al = neg(mul(e.A, inv(em.B)));
            RedEq o;
            o.A = mul(al, em.A);
            o.B = add(e.B, mul(al, em.C));
            double2 t = mulv(al, em.d);
            o.d = make_double2(e.d.x + t.x, e.d.y + t.y);
            if (2 * jj + 2 < mm) {
                const RedEq ep = red[ip];
                const M2 ga = neg(mul(e.C, inv(ep.B)));
                o.B = add(o.B, mul(ga, ep.A));
                o.C = mul(ga, ep.C);
                t = mulv(ga, ep.d);
                o.d = make_double2(o.d.x + t.x, o.d.y + t.y);
            } else {
                o.C = zero2();
            }
            o.x = make_double2(0, 0);
            red[i] = o;
        }
        __syncthreads();
    }
    if (tid == 0 && P > 0) {
        const int i = (1 << nlev) - 1;
        red[i].x = mulv(inv(red[i].B), red[i].d);
    }
    __syncthreads();
    for (int l = nlev - 1; l >= 0; l--) {
        const int st = 1 << l, mm = P >> l;
        for (int kk = tid; kk < (mm + 1) / 2; kk += kSolveWG) {
            const int k = 2 * kk;  // even equations of level l
            const int i = k * st + st - 1;
            const RedEq e = red[i];
            double2 r = e.d;
            if (k > 0) {
                const double2 t = mulv(e.A, red[i - st].x);
                r.x -= t.x;
                r.y -= t.y;
            }
            if (k + 1 < mm) {
                const double2 t = mulv(e.C, red[i + st].x);
                r.x -= t.x;
                r.y -= t.y;
            }
            red[i].x = mulv(inv(e.B), r);
        }
        __syncthreads();
    }
    zo = make_double2(0, 0);
    zp = make_double2(0, 0);
    if (tid < P) {
        zo = red[tid].x;
        if (tid > 0) zp = red[tid - 1].x;
    }
}

// Exact FP64 solve of the level-0 block-tridiagonal system E (knots 0..n-1)
// into X: partitioned block Thomas + cyclic reduction of the interface
// system.  Thread t owns knots [s, e); its interior rows s..e-2 are
// eliminated in terms of the interfaces z_{t-1} = x_{s-1} and z_t = x_{e-1}.
// Partitions of at most M knots keep their rows and eliminated rows in
// registers (one round of loads); longer ones (M = 0) go through IR in
// memory.  Both do the same arithmetic in the same order.  Called by all
// threads of the workgroup; IR and firstAff are workspace.
template <int M>
__device__ void exact_solve_t(const Eq0 *__restrict__ E, InteriorRow *__restrict__ IR, double2 *__restrict__ X,
                              Aff *firstAff, int n, int tid) {
    const int m = (n + kSolveWG - 1) / kSolveWG;
    const int P = (n + m - 1) / m;
    const int s0 = tid * m, e0 = min(n, s0 + m);
    const int len = e0 - s0;
    constexpr int MR = M > 0 ? M : 1;
    Eq0 q[MR];
    InteriorRow ir[MR];
    Eq0 qm{};  // row s0 - 1
    Aff lastA, first;
    Eq0 qj{};  // row e0 - 1
    M2 A = zero2();
    const bool interior = len >= 2;
    if (tid < P) {
        if constexpr (M > 0) {
#pragma unroll
            for (int r = 0; r < M; r++)
                if (r < len) q[r] = E[s0 + r];
            if (s0 > 0) qm = E[s0 - 1];
        }
        if (interior) {
            // forward elimination
            M2 BiPrev = zero2();
            Aff Rprev;
            if constexpr (M > 0) {
#pragma unroll
                for (int r = 0; r < M - 1; r++) {
                    if (r > len - 2) continue;
                    const Eq0 &qq = q[r];
                    M2 Bh = sym(qq.b0, qq.b1, qq.b2);
                    Aff R{make_double2(qq.d0, qq.d1), zero2(), zero2()};
                    if (r == 0) {
                        if (tid > 0) R.V = neg(sym(qm.c0, qm.c1, qm.c2));
                    } else {
                        const Eq0 &qp = q[r - 1];
                        const M2 Ap = sym(qp.c0, qp.c1, qp.c2);  // A_i = O_{i-1}
                        const M2 Lm = mul(Ap, BiPrev);
                        Bh = sub(Bh, mul(Lm, Ap));
                        R = aff_sub(R, aff_apply(Lm, Rprev));
                    }
                    if (r == len - 2) R.W = neg(sym(qq.c0, qq.c1, qq.c2));
                    const M2 Bi = inv(Bh);
                    ir[r] = InteriorRow{Bi, R};
                    BiPrev = Bi;
                    Rprev = R;
                }
            } else {
                for (int i = s0; i <= e0 - 2; i++) {
                    const Eq0 qq = E[i];
                    M2 Bh = sym(qq.b0, qq.b1, qq.b2);
                    Aff R{make_double2(qq.d0, qq.d1), zero2(), zero2()};
                    if (i == s0) {
                        if (tid > 0) {
                            const Eq0 qp = E[i - 1];
                            R.V = neg(sym(qp.c0, qp.c1, qp.c2));
                        }
                    } else {
                        const Eq0 qp = E[i - 1];
                        const M2 Ap = sym(qp.c0, qp.c1, qp.c2);  // A_i = O_{i-1}
                        const M2 Lm = mul(Ap, BiPrev);
                        Bh = sub(Bh, mul(Lm, Ap));
                        R = aff_sub(R, aff_apply(Lm, Rprev));
                    }
                    if (i == e0 - 2) R.W = neg(sym(qq.c0, qq.c1, qq.c2));
                    const M2 Bi = inv(Bh);
                    IR[i] = InteriorRow{Bi, R};
                    BiPrev = Bi;
                    Rprev = R;
                }
            }
            // back substitution of the affine forms
            Aff Xa = aff_apply(BiPrev, Rprev);
            lastA = Xa;
            if constexpr (M > 0) {
#pragma unroll
                for (int r = M - 3; r >= 0; r--) {
                    if (r > len - 3) continue;
                    const Eq0 &qq = q[r];
                    const Aff t = aff_sub(ir[r].R, aff_apply(sym(qq.c0, qq.c1, qq.c2), Xa));
                    Xa = aff_apply(ir[r].Bi, t);
                }
            } else {
                for (int i = e0 - 3; i >= s0; i--) {
                    const InteriorRow iri = IR[i];
                    const Eq0 qq = E[i];
                    const Aff t = aff_sub(iri.R, aff_apply(sym(qq.c0, qq.c1, qq.c2), Xa));
                    Xa = aff_apply(iri.Bi, t);
                }
            }
            first = Xa;
        } else {
            lastA = Aff{make_double2(0, 0), eye2(), zero2()};      // x_{e-2} := z_prev
            first = Aff{make_double2(0, 0), zero2(), eye2()};      // unused
        }
        // the interface row j = e0 - 1 and its coupling A = O_{j-1}
        const int j = e0 - 1;
        if constexpr (M > 0) {
#pragma unroll
            for (int r = 0; r < M; r++)
                if (r == len - 1) qj = q[r];
            if (len >= 2) {
#pragma unroll
                for (int r = 0; r < M - 1; r++)
                    if (r == len - 2) A = sym(q[r].c0, q[r].c1, q[r].c2);
            } else if (j > 0) {
                A = sym(qm.c0, qm.c1, qm.c2);
            }
        } else {
            qj = E[j];
            if (j > 0) {
                const Eq0 qp = E[j - 1];
                A = sym(qp.c0, qp.c1, qp.c2);
            }
        }
    }
    double2 zo, zp;
    solve_interfaces(interior, first, lastA, qj, A, firstAff, P, tid, zo, zp);
    // interior reconstruction: x_i = Bi_i (R_i(z) - O_i x_{i+1})
    if (tid < P) {
        X[e0 - 1] = zo;
        double2 xn = zo;
        if constexpr (M > 0) {
#pragma unroll
            for (int r = M - 2; r >= 0; r--) {
                if (r > len - 2) continue;
                const InteriorRow &iri = ir[r];
                const double2 v = mulv(iri.R.V, zp), w = mulv(iri.R.W, zo);
                double2 rr = make_double2(iri.R.y.x + v.x + w.x, iri.R.y.y + v.y + w.y);
                if (r < len - 2) {
                    const double2 t = mulv(sym(q[r].c0, q[r].c1, q[r].c2), xn);
                    rr.x -= t.x;
                    rr.y -= t.y;
                }
                xn = mulv(iri.Bi, rr);
                X[s0 + r] = xn;
            }
        } else {
            for (int i = e0 - 2; i >= s0; i--) {
                const InteriorRow iri = IR[i];
                const double2 v = mulv(iri.R.V, zp), w = mulv(iri.R.W, zo);
                double2 rr = make_double2(iri.R.y.x + v.x + w.x, iri.R.y.y + v.y + w.y);
                if (i < e0 - 2) {
                    const Eq0 qq = E[i];
                    const double2 t = mulv(sym(qq.c0, qq.c1, qq.c2), xn);
                    rr.x -= t.x;
                    rr.y -= t.y;
                }
                xn = mulv(iri.Bi, rr);
                X[i] = xn;
            }
        }
    }
    __threadfence_block();
    __syncthreads();
}

// knots per partition that stay in registers
constexpr int kSolveRegRows = 8;
// knots whose level-0 equations the solve keeps in LDS (112 KiB; with the
// kernel's other arrays within the 160 KiB of a workgroup)
constexpr int kSolveLdsRows = 1792;

// Inlined into the solve kernel: as a called function it would save and
// restore the callee-saved registers it uses (over 200 per call).
__device__ __forceinline__ void exact_solve(const Eq0 *E, InteriorRow *IR, double2 *X, Aff *firstAff, int n, int tid) {
    if ((n + kSolveWG - 1) / kSolveWG <= kSolveRegRows)
        exact_solve_t<kSolveRegRows>(E, IR, X, firstAff, n, tid);
    else
        exact_solve_t<0>(E, IR, X, firstAff, n, tid);
}

// Output 10: the device time since the group's last stage end goes to slot
// `slot`.  Called by one thread of a per-group kernel.
__device__ __forceinline__ void stage_time(const GroupDev &g, GroupCtl &ctl, int slot) {
    if (!g.funcTimes) return;
    const unsigned long long now = wall_clock64();
    if (ctl.tStage) g.funcTimes[slot] += (double)(now - ctl.tStage) / g.clockHz;
    ctl.tStage = now;
}

__device__ void final_finish(const GroupDev &g, GroupCtl &ctl);

// The solve launch also ends the groups whose final pass ran in the
// assembly launch before it (phase PH_FINAL).
__global__ __launch_bounds__(kSolveWG) void k_solve(Batch B) {
    const int gi = blockIdx.x;
    GroupCtl &ctl = B.ctl[gi];
    if (ctl.phase == PH_FINAL) {
        final_finish(B.groups[gi], ctl);
        return;
    }
    if (ctl.phase != PH_ASSEMBLE) return;
    const GroupDev &g = B.groups[gi];
    const int tid = threadIdx.x;
    const int nF = g.nF, nb = g.nb, N = g.N, n = g.fframeW;
    const int shift = ctl.shift;
    // filter deactivation after max_filter_iterations (mainiteration.cpp:80-90)
    __shared__ int s_filter;
    if (tid == 0) {
        if (ctl.nFilterIterations != 0 && ctl.iteration > ctl.nFilterIterations) {
            ctl.solverFilter = 0;
            ctl.filterID = 0;
            ctl.nFilterIterations = 0;
        }
        s_filter = ctl.solverFilter;
        stage_time(g, ctl, FT_ASSEMBLY);
    }
    const double *S = g.colsum;
    const int *minPts = g.minPts;
    uint8_t *used = g.used;
    // The loops below load a batch of entries before they use any of it:
    // one thread per frame or knot would otherwise wait out a memory latency
    // per entry.  The additions keep their order.
    constexpr int kBatch = 8;
    // 1. per-frame errval of this shift (d_calcAllGroupErrval): a frame's
    //    columns in batches of 32 (one round of loads for 512x512 frames)
    constexpr int kEvBatch = 32;
    for (int f = tid; f < nF; f += kSolveWG) {
        const double *col = S + (size_t)f * nb * CS_N;
        double s = 0.0, c = 0.0;
        for (int b0 = 0; b0 < nb; b0 += kEvBatch) {
            double vs[kEvBatch], vc[kEvBatch];
#pragma unroll
            for (int u = 0; u < kEvBatch; u++) {
                const int bb = min(b0 + u, nb - 1);
                vs[u] = col[(size_t)bb * CS_N + CS_D2];
                vc[u] = col[(size_t)bb * CS_N + CS_CNT];
            }
#pragma unroll
            for (int u = 0; u < kEvBatch; u++) {
                if (b0 + u < nb) {
                    s += vs[u];
                    c += vc[u];
                }
            }
        }
        g.errval[shift * nF + f] = s / c;
    }
    // 2. knot usage (d_createParameterMask) and level-0 equations
    //    (d_fillDiagonals / d_fillJTr entries, lambda extension); the knots
    //    from n on are never used.  The equations stay in LDS when they fit.
    __shared__ Eq0 sE[kSolveLdsRows];
    Eq0 *E = n <= kSolveLdsRows ? sE : reinterpret_cast<Eq0 *>(g.ws);
    level0_equations(S, minPts, used, E, n, nb, g.lambda, g.fracUniform, tid);
    for (int k = n + tid; k < N; k += kSolveWG) used[k] = 0;
    __threadfence_block();
    __syncthreads();
    InteriorRow *IR = reinterpret_cast<InteriorRow *>(g.ws + (size_t)n * 8);
    double2 *X = reinterpret_cast<double2 *>(IR + n);
    Aff *firstAff = reinterpret_cast<Aff *>(X + n);
    // The exact solve without refinement, its partitions in registers, has
    // read every equation before its first barrier and writes the solution
    // after the last: the solution then takes the equations' LDS.
    if (E == sE && g.solverId != 0 && !g.refine && (n + kSolveWG - 1) / kSolveWG <= kSolveRegRows)
        X = reinterpret_cast<double2 *>(sE);
    if (g.solverId == 0) {
        // 4a. solver id 0: ILU(0) in x-then-y order (one thread, the
        //     reference's recurrence order)
        if (tid == 0) ilu0_solve(E, reinterpret_cast<IluRow *>(IR), X, n);
        __threadfence_block();
        __syncthreads();
    } else {
        exact_solve(E, IR, X, firstAff, n, tid);
        if (g.refine) {
            // one step of iterative refinement in FP64: r = g - H x, H e = r,
            // x += e (the matrix entries stay; each knot's right-hand side is
            // replaced by its residual)
            double2 *Xs = reinterpret_cast<double2 *>(firstAff + kSolveWG);
            for (int k = tid; k < n; k += kSolveWG) {
                const Eq0 q = E[k];
                const double2 x = X[k];
                double r0 = q.d0 - (q.b0 * x.x + q.b1 * x.y);
                double r1 = q.d1 - (q.b1 * x.x + q.b2 * x.y);
                if (k > 0) {
                    const Eq0 qp = E[k - 1];
                    const double2 xp = X[k - 1];
                    r0 -= qp.c0 * xp.x + qp.c1 * xp.y;
                    r1 -= qp.c1 * xp.x + qp.c2 * xp.y;
                }
                if (k + 1 < n) {
                    const double2 xn = X[k + 1];
                    r0 -= q.c0 * xn.x + q.c1 * xn.y;
                    r1 -= q.c1 * xn.x + q.c2 * xn.y;
                }
                Xs[k] = x;
                E[k].d0 = r0;
                E[k].d1 = r1;
            }
            __threadfence_block();
            __syncthreads();
            exact_solve(E, IR, X, firstAff, n, tid);
            for (int k = tid; k < n; k += kSolveWG) {
                const double2 e = X[k], x = Xs[k];
                X[k] = make_double2(x.x + e.x, x.y + e.y);
            }
            __threadfence_block();
            __syncthreads();
        }
    }
    // 5. step filter over the used knots, x sequence then y sequence
    //    (SmearingFilter::apply / RangeFilter::apply on the compacted halves)
    const int filterKind = s_filter;
    const double strength = g.strength;
    if (filterKind == 1) {
        __shared__ double sa[kSolveWG][4];
        const int chunk = (n + kSolveWG - 1) / kSolveWG;
        const int c0 = min(n, tid * chunk), c1 = min(n, c0 + chunk);
        double ax = 1, bx = 0, ay = 1, by = 0;
        for (int k = c0; k < c1; k++) {
            if (!g.used[k]) continue;
            ax *= strength;
            bx = strength * bx + (1 - strength) * X[k].x;
            ay *= strength;
            by = strength * by + (1 - strength) * X[k].y;
        }
        sa[tid][0] = ax;
        sa[tid][1] = bx;
        sa[tid][2] = ay;
        sa[tid][3] = by;
        __syncthreads();
        for (int st = 1; st < kSolveWG; st <<= 1) {
            double v0 = 0, v1 = 0, v2 = 0, v3 = 0;
            const bool has = tid >= st;
            if (has) {
                v0 = sa[tid - st][0];
                v1 = sa[tid - st][1];
                v2 = sa[tid - st][2];
                v3 = sa[tid - st][3];
            }
            __syncthreads();
            if (has) {
                const double a0 = sa[tid][0], b0 = sa[tid][1], a1 = sa[tid][2], b1 = sa[tid][3];
                sa[tid][0] = a0 * v0;
                sa[tid][1] = a0 * v1 + b0;
                sa[tid][2] = a1 * v2;
                sa[tid][3] = a1 * v3 + b1;
            }
            __syncthreads();
        }
        double yx = tid > 0 ? sa[tid - 1][1] : 0.0;
        double yy = tid > 0 ? sa[tid - 1][3] : 0.0;
        for (int k = c0; k < c1; k++) {
            if (!g.used[k]) continue;
            yx = yx * strength + X[k].x * (1 - strength);
            yy = yy * strength + X[k].y * (1 - strength);
            X[k] = make_double2(yx, yy);
        }
        __threadfence_block();
        __syncthreads();
    } else if (filterKind == 2) {
        if (tid == 0) {
            for (int axis = 0; axis < 2; axis++) {
                int pv = -1, cu = -1;
                for (int k = 0; k < n; k++) {
                    if (!g.used[k]) continue;
                    if (pv >= 0 && cu >= 0) {
                        const double vm = axis ? X[pv].y : X[pv].x;
                        const double vi = axis ? X[cu].y : X[cu].x;
                        const double vp = axis ? X[k].y : X[k].x;
                        const double tc = (vm + vp) / 2;
                        const double diff = (tc - vm) * strength;
                        bool rep;
                        if (vm < vp) rep = vi > vp + diff || vi < vm - diff;
                        else rep = vi < vp + diff || vi > vm - diff;
                        if (rep) {
                            if (axis) X[cu].y = tc;
                            else X[cu].x = tc;
                        }
                    }
                    pv = cu;
                    cu = k;
                }
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    // 6. d_expandResults (task_expandresults.cu:10-41) and line-search set-up
    const int G = g.groupSize;
    double *deltaP = g.deltaP, *d0 = g.d0, *p0 = g.p0;
    const double *p = g.p;
    for (int i0 = tid * kBatch; i0 < 2 * N; i0 += kSolveWG * kBatch) {
        double dv[kBatch], pv[kBatch], xv[kBatch];
        bool keep[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int idx = min(i0 + u, 2 * N - 1);
            const int j = idx >= N ? idx - N : idx;
            const bool fixed = ((((idx + 1) - (shift * (G >> 1))) % G) == 0);
            dv[u] = deltaP[idx];
            pv[u] = p[idx];
            const bool uj = j < n && used[j];
            xv[u] = uj ? (idx >= N ? X[j].y : X[j].x) : 0.0;
            keep[u] = fixed;
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int idx = i0 + u;
            if (idx < 2 * N) {
                const double v = keep[u] ? dv[u] : xv[u];
                deltaP[idx] = v;
                d0[idx] = v;
                p0[idx] = pv[u];
            }
        }
    }
    for (int f = tid; f < nF; f += kSolveWG) {
        g.lsActive[f] = 1;
        g.lsNextK[f] = 0;
        g.lsMk[f] = -1;
        g.lsDk[f] = -1;
    }
    if (tid == 0) {
        ctl.slotRounds = 0;
        ctl.roundsThisLS = 0;
        ctl.K = g.K1;
        ctl.phase = PH_LS;
        stage_time(g, ctl, FT_SOLVE);
    }
}

// ---------------------------------------------------------------------------
// line search: K fused candidates p0 + r^k * delta of every active frame
// (the tries of d_linesearch, task_linesearch.cu:35-106, in one pass), then
// one wave per frame resolves the tries sequentially.  A tick runs kLsSlots
// such rounds per half-iteration; after the last one, one wave per group
// does the rounds' bookkeeping.  Separate kernels, so that no device-scope
// fence is needed: on gfx950 each one writes back and invalidates the XCD's
// L2.  Round `slot` of a tick's half-iteration slots runs K = ctl.K
// candidates in slot 0 (K1 after the solve, Kn when the line search goes on
// from the previous slots) and Kn after it; no round runs past maxRounds.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool ls_round_runs(const GroupDev &g, const GroupCtl &ctl, int slot) {
    return ctl.roundsThisLS + slot < g.maxRounds;
}
__device__ __forceinline__ int ls_round_k(const GroupDev &g, const GroupCtl &ctl, int slot) {
    return slot == 0 ? ctl.K : g.Kn;
}

// Candidates k0 .. k0 + K - 1 of column c; their sums go to the column's
// candidate slots slot0 .. slot0 + K - 1.
template <typename TI, typename R, int V, int K>
__device__ void ls_column(const GroupDev &g, int c, int b, int k0, int slot0) {
    ColKnots<R> kn[K];
    {
        const int N = g.N;
        const double r = g.rmult;
        double dlx = reduce_pow(g.d0[c], r, k0), dly = reduce_pow(g.d0[c + N], r, k0);
        double drx = reduce_pow(g.d0[c + 1], r, k0), dry = reduce_pow(g.d0[c + 1 + N], r, k0);
        const bool lastKnot = (c + 1 == g.fframeW);
        const double plx = g.p0[c], ply = g.p0[c + N], prx = g.p0[c + 1], pry = g.p0[c + 1 + N];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const double lx = plx + dlx, ly = ply + dly;
            const double rx = lastKnot ? 1e6 : prx + drx, ry = lastKnot ? 1e6 : pry + dry;
            kn[k] = make_knots<R>(lx, rx, ly, ry);
            dlx *= r;
            dly *= r;
            drx *= r;
            dry *= r;
        }
    }
    const TI *img = column_ptr<TI>(g, c);
    const typename Prec<R>::V4 *geo = Prec<R>::geo(g) ? Prec<R>::geo(g) + (size_t)b * g.frameH : nullptr;
    double e2[K];  // FP64: compared against the half-iteration's errval
    int cnt[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        e2[k] = 0.0;
        cnt[k] = 0;
    }
    constexpr int CH = chunk_rows<R>(K, tune_variant<TK_LS, R, V>().samples);
    for (int run = threadIdx.x; run * g.runLen < g.frameH; run += kWG) {
        const int rend = min(run * g.runLen + g.runLen, g.frameH);
        GeoIter<R> it;
        it.start(g, b, run * g.runLen);
        for (int r0 = run * g.runLen; r0 < rend; r0 += CH) {
            const int n = min(CH, rend - r0);
            R I[CH];
            load_chunk<CH>(img + r0, g.vec, n, I);
#pragma unroll
            for (int q = 0; q < CH; q++) {
                R bx, by, f;
                bool m;
                it.at(g, geo, min(r0 + q, g.frameH - 1), bx, by, f, m);
                it.next();
                m = m && q < n;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const R x = mad(kn[k].dx, f, bx + kn[k].ax);
                    const R y = mad(kn[k].dy, f, by + kn[k].ay);
                    const bool ok = m && in_bounds(g, x, y);
                    Cell<R> cl;
                    const R s = sample(g, x, y, ok, cl);
                    const R d = ok ? I[q] - s : (R)0;
                    e2[k] = fma((double)d, (double)d, e2[k]);
                    cnt[k] += ok ? 1 : 0;
                }
            }
        }
    }
    double v[2 * K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        v[2 * k] = e2[k];
        v[2 * k + 1] = (double)cnt[k];
    }
    __shared__ double out[2 * kMaxCand];
    block_sum<2 * K>(v, out);
    if (threadIdx.x < 2 * K) g.lssum[(size_t)c * 2 * kMaxCand + 2 * slot0 + threadIdx.x] = out[threadIdx.x];
}

// split = 2: work item w covers column w % nItems and, of a round of four
// candidates, the two of half w / nItems (each candidate's sums are the
// same whichever workgroup computes them); other rounds run in half 0.
template <typename TI, typename R, int V>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(tune_variant<TK_LS, R, V>().waves))) void k_ls_round(Batch B, int slot, int split) {
    const int nWork = B.nItems * split;
    for (int w = blockIdx.x; w < nWork; w += gridDim.x) {
        const int half = w >= B.nItems ? 1 : 0, item = w - half * B.nItems;
        const int gi = B.itemGroup[item];
        const GroupCtl &ctl = B.ctl[gi];
        if (ctl.phase != PH_LS) continue;
        const GroupDev &g = B.groups[gi];
        if (!runs_here<TI, R>(g) || g.var[TK_LS] != V || !ls_round_runs(g, ctl, slot)) continue;
        const int c = item - g.itemStart, f = c / g.nb, b = c % g.nb;
        if (!g.lsActive[f]) continue;
        int K = ls_round_k(g, ctl, slot), k0 = g.lsNextK[f], slot0 = 0;
        if (split == 2 && K == 4) {
            K = 2;
            k0 += 2 * half;
            slot0 = 2 * half;
        } else if (half) {
            continue;
        }
        switch (K) {
        case 1: ls_column<TI, R, V, 1>(g, c, b, k0, slot0); break;
        case 2: ls_column<TI, R, V, 2>(g, c, b, k0, slot0); break;
        default: ls_column<TI, R, V, 4>(g, c, b, k0, slot0); break;
        }
    }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    return x;
}
__device__ __forceinline__ int wave_or(int x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x |= __shfl_xor(x, off, 64);
    return x;
}

// Sequential resolution of the candidates of frame f (reproduces the
// reference's try / reduce / movement-threshold order) and commit.  One
// wave per frame.
__device__ void ls_decide_frame(const GroupDev &g, GroupCtl &ctl, int f, int K, int slot) {
    const int nb = g.nb, N = g.N, nF = g.nF;
    const int lane = threadIdx.x;
    double se[kMaxCand], sn[kMaxCand];
#pragma unroll
    for (int k = 0; k < kMaxCand; k++) se[k] = sn[k] = 0.0;
    int pres = 0;  // bit k: some column of candidate k reaches its block's min points
    double dmax = 0.0;
    for (int bb = lane; bb < nb; bb += 64) {
        const double *ls = g.lssum + (size_t)(f * nb + bb) * 2 * kMaxCand;
#pragma unroll
        for (int k = 0; k < kMaxCand; k++) {
            if (k < K) {
                se[k] += ls[2 * k];
                sn[k] += ls[2 * k + 1];
                if (ls[2 * k + 1] >= (double)g.minPts[bb]) pres |= 1 << k;
            }
        }
        dmax = fmax(dmax, fmax(fabs(g.d0[f * nb + bb]), fabs(g.d0[f * nb + bb + N])));
    }
#pragma unroll
    for (int k = 0; k < kMaxCand; k++) {
        if (k < K) {
            se[k] = wave_sum(se[k]);
            sn[k] = wave_sum(sn[k]);
        }
    }
    pres = wave_or(pres);
    dmax = wave_max(dmax);
    int fin = -1, acc = -1;
    if (lane == 0) {
        const int shift = ctl.shift;
        const double r = g.rmult, thr = g.thr;
        const int k0 = g.lsNextK[f];
        double dm = reduce_pow(dmax, r, k0);
        int mk = -1, tries = 0;
        for (int kk = 0; kk < K && fin < 0; kk++) {
            const int k = k0 + kk;
            tries++;
            mk = k;
            if (!((pres >> kk) & 1)) {
                const double dm1 = dm * r;
                if (!(dm1 > thr)) fin = k + 1;
                dm = dm1;
            } else {
                g.lsDk[f] = k;
                double e = 0.0, c = 0.0;
#pragma unroll
                for (int q = 0; q < kMaxCand; q++)
                    if (q == kk) {
                        e = se[q];
                        c = sn[q];
                    }
                e /= c;
                if (e < g.errval[shift * nF + f]) {
                    acc = k;
                    fin = k;
                    if (f == 0) g.errval[shift * nF] = e;
                } else if (!(dm > thr)) {
                    fin = k;
                } else {
                    dm *= r;
                }
            }
        }
        g.lsMk[f] = mk;
        atomicAdd(&ctl.tries, tries);
        atomicAdd(&ctl.lsFramePasses, 1);
        atomicMax(&ctl.slotRounds, slot + 1);
        if (fin < 0) g.lsNextK[f] = k0 + K;
        else g.lsActive[f] = 0;
    }
    fin = __shfl(fin, 0, 64);
    acc = __shfl(acc, 0, 64);
    if (fin >= 0) {
        const double r = g.rmult;
        for (int t = lane; t < 2 * nb; t += 64) {
            const int i = f * nb + (t % nb) + (t >= nb ? N : 0);
            const double d0 = g.d0[i];
            g.deltaP[i] = reduce_pow(d0, r, fin);
            if (acc >= 0) g.p[i] = g.p0[i] + reduce_pow(d0, r, acc);
        }
    }
}

__global__ __launch_bounds__(64) void k_ls_decide(Batch B, int slot) {
    for (int fi = blockIdx.x; fi < B.nFrames; fi += gridDim.x) {
        const int gi = B.frameGroup[fi];
        GroupCtl &ctl = B.ctl[gi];
        if (ctl.phase != PH_LS) continue;
        const GroupDev &g = B.groups[gi];
        if (!ls_round_runs(g, ctl, slot)) continue;
        const int f = fi - g.frameStart;
        if (!g.lsActive[f]) continue;
        ls_decide_frame(g, ctl, f, ls_round_k(g, ctl, slot), slot);
    }
}

// After a tick's last line-search slot: the rounds' bookkeeping, and either
// the next slots' frames or the end of the half-iteration (`improved`,
// task_linesearch.cu:117-125) and of MainIteration::execute
// (mainiteration.cpp:103-118).  A half-iteration that ends in an earlier
// slot leaves the later slots with no active frame.
__global__ __launch_bounds__(64) void k_ls_finish(Batch B) {
    const int gi = blockIdx.x;
    GroupCtl &ctl = B.ctl[gi];
    if (ctl.phase != PH_LS) return;
    const GroupDev &g = B.groups[gi];
    const int nF = g.nF;
    const int shift0 = ctl.shift, it0 = ctl.iteration;
    int active = 0;
    for (int f = threadIdx.x; f < nF; f += 64) active += g.lsActive[f];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) active += __shfl_xor(active, off, 64);
    if (active == 0 && g.rgParams && it0 <= g.maxIter) {  // output 14 (timedmainiteration.cpp:124)
        const int P = 2 * g.N;
        double *dst = g.rgParams + (size_t)P * (size_t)(shift0 + 2 * (it0 - 1));
        for (int i = threadIdx.x; i < P; i += 64) dst[i] = g.p[i];
    }
    if (threadIdx.x != 0) return;
    stage_time(g, ctl, FT_LINESEARCH);
    ctl.rounds += ctl.slotRounds;
    ctl.roundsThisLS += ctl.slotRounds;
    ctl.slotRounds = 0;
    if (active > 0) {
        ctl.K = g.Kn;
        if (ctl.roundsThisLS >= g.maxRounds) {
            ctl.error = 4;  // SS_ERR_NUMERICAL
            ctl.phase = PH_ERROR;
        }
        return;
    }
    // half-iteration complete
    double e0 = 0.0, e1 = 0.0;
    for (int ff = 0; ff < nF; ff++) {
        e0 += g.errval[ff];
        e1 += g.errval[ff + nF];
    }
    const int shift = ctl.shift;
    const int imp = (shift && e0 > e1) || (!shift && e0 < e1);
    ctl.improved &= imp;
    ctl.halfIterations++;
    if (shift == 0) {
        ctl.shift = 1;
        ctl.phase = PH_ASSEMBLE;
        return;
    }
    if (!ctl.improved && ctl.nFilterIterations != 0) {
        ctl.solverFilter = 0;
        ctl.nFilterIterations = 0;
        ctl.improved = 1;  // run this iteration again without the filter
        ctl.shift = 0;
        ctl.phase = PH_ASSEMBLE;
        return;
    }
    ctl.iterImproved = ctl.improved;
    ctl.phase = PH_CORR;
}

// ---------------------------------------------------------------------------
// global correlation statistic (hd_globalCorrelation,
// task_correlation.cu:63-99) of every frame's last-tried state; the
// statistic of a column covers its leading run of masked rows only.
// ---------------------------------------------------------------------------
template <typename TI, typename R, int V>
__device__ void corr_column(const GroupDev &g, int c) {
    const int f = c / g.nb, b = c % g.nb;
    const int mk = g.lsMk[f], dk = g.lsDk[f];
    const bool same = (mk == dk);
    const ColKnots<R> km = state_knots<R>(g, c, mk);
    const ColKnots<R> kd = state_knots<R>(g, c, dk);
    const TI *img = column_ptr<TI>(g, c);
    const typename Prec<R>::V4 *geo = Prec<R>::geo(g) ? Prec<R>::geo(g) + (size_t)b * g.frameH : nullptr;
    // first row outside the mask of the last-tried state: the moments cover
    // the rows before it
    int firstOut = g.frameH;
    constexpr int CH = chunk_rows<R>(1, tune_variant<TK_CORR, R, V>().samples);
    for (int run = threadIdx.x; run * g.runLen < g.frameH; run += kWG) {
        const int rend = min(run * g.runLen + g.runLen, g.frameH);
        GeoIter<R> it;
        it.start(g, b, run * g.runLen);
        for (int r0 = run * g.runLen; r0 < rend && r0 < firstOut; r0 += CH) {
            const int n = min(CH, rend - r0);
#pragma unroll
            for (int q = 0; q < CH; q++) {
                R bx, by, f1;
                bool m;
                it.at(g, geo, min(r0 + q, g.frameH - 1), bx, by, f1, m);
                it.next();
                const R x = mad(km.dx, f1, bx + km.ax);
                const R y = mad(km.dy, f1, by + km.ay);
                if (q < n && !(m && in_bounds(g, x, y))) firstOut = min(firstOut, r0 + q);
            }
        }
    }
    firstOut = block_min(firstOut);
    double v[CR_N];
#pragma unroll
    for (int i = 0; i < CR_N; i++) v[i] = 0.0;
    for (int run = threadIdx.x; run * g.runLen < g.frameH; run += kWG) {
        const int rend = min(run * g.runLen + g.runLen, g.frameH);
        GeoIter<R> it;
        it.start(g, b, run * g.runLen);
        for (int r0 = run * g.runLen; r0 < rend; r0 += CH) {
            const int n = min(CH, rend - r0);
            R I[CH];
            load_chunk<CH>(img + r0, g.vec, n, I);
#pragma unroll
            for (int q = 0; q < CH; q++) {
                R bx, by, f1;
                bool m;
                it.at(g, geo, min(r0 + q, g.frameH - 1), bx, by, f1, m);
                it.next();
                m = m && q < n;
                const R x = mad(km.dx, f1, bx + km.ax);
                const R y = mad(km.dy, f1, by + km.ay);
                const bool mm = m && in_bounds(g, x, y);
                // the difference image holds the last-decided state's residual
                R xd = x, yd = y;
                bool okd = mm;
                if (!same) {
                    xd = mad(kd.dx, f1, bx + kd.ax);
                    yd = mad(kd.dy, f1, by + kd.ay);
                    okd = m && in_bounds(g, xd, yd);
                }
                Cell<R> cl;
                const R s = sample(g, xd, yd, okd, cl);
                const R dbuf = okd ? I[q] - s : (R)0;
                const double Iq = q < n ? (double)I[q] : 0.0;
                v[CR_SUMD] += mm ? (double)(I[q] - dbuf) : (double)dbuf;
                v[CR_NMASK] += mm ? 1.0 : 0.0;
                const bool run = r0 + q < firstOut;  // rows before it are all inside the mask
                const double a = run ? (double)(I[q] - dbuf) : 0.0, bb = run ? Iq : 0.0;
                v[CR_SA] += a;
                v[CR_SAA] += a * a;
                v[CR_SB] += bb;
                v[CR_SBB] += bb * bb;
                v[CR_SAB] += a * bb;
                v[CR_NS] += run ? 1.0 : 0.0;
            }
        }
    }
    __shared__ double out[CR_N];
    block_sum<CR_N>(v, out);
    if (threadIdx.x < CR_N) g.corrsum[(size_t)c * CR_N + threadIdx.x] = out[threadIdx.x];
}

template <typename TI, typename R, int V>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(tune_variant<TK_CORR, R, V>().waves))) void k_corr(Batch B) {
    for (int item = blockIdx.x; item < B.nItems; item += gridDim.x) {
        const int gi = B.itemGroup[item];
        if (B.ctl[gi].phase != PH_CORR) continue;
        const GroupDev &g = B.groups[gi];
        if (!runs_here<TI, R>(g) || g.var[TK_CORR] != V) continue;
        corr_column<TI, R, V>(g, item - g.itemStart);
    }
}

// Flag 'c': the frame correlation taken right after the global correlation
// (task.cpp:286-293).  The global correlation swaps every masked pixel to
// image - difference but swaps back only the row-major pixels below its
// grid, fframeW rounded up to 256 (task_correlation.cu:97); the frame
// correlation then swaps every masked pixel once more
// (task_correlation.cu:176-178).  Sums over all of the column's rows, in
// finalsum's slots (free while the group is in this phase).  A kernel of its
// own, launched only for batches that ask for it, so that the correlation
// kernel's register allocation does not carry it.
template <typename TI, typename R>
__device__ void frame_corr_column(const GroupDev &g, int c) {
    const int f = c / g.nb, b = c % g.nb;
    const int mk = g.lsMk[f], dk = g.lsDk[f];
    const bool same = (mk == dk);
    const ColKnots<R> km = state_knots<R>(g, c, mk);
    const ColKnots<R> kd = state_knots<R>(g, c, dk);
    const TI *img = column_ptr<TI>(g, c);
    const typename Prec<R>::V4 *geo = Prec<R>::geo(g) ? Prec<R>::geo(g) + (size_t)b * g.frameH : nullptr;
    constexpr int CH = chunk_rows<R>(1, sizeof(R) == 8 ? kFinalSamplesF64 : kFinalSamplesF32);
    const long long swapBack = min((long long)((g.fframeW + 255) / 256) * 256, (long long)g.frameH * g.fframeW);
    double w[FC_N];
#pragma unroll
    for (int i = 0; i < FC_N; i++) w[i] = 0.0;
    for (int run = threadIdx.x; run * g.runLen < g.frameH; run += kWG) {
        const int rend = min(run * g.runLen + g.runLen, g.frameH);
        GeoIter<R> it;
        it.start(g, b, run * g.runLen);
        for (int r0 = run * g.runLen; r0 < rend; r0 += CH) {
            const int n = min(CH, rend - r0);
            R I[CH];
            load_chunk<CH>(img + r0, g.vec, n, I);
#pragma unroll
            for (int q = 0; q < CH; q++) {
                R bx, by, f1;
                bool m;
                it.at(g, geo, min(r0 + q, g.frameH - 1), bx, by, f1, m);
                it.next();
                m = m && q < n;
                const R x = mad(km.dx, f1, bx + km.ax);
                const R y = mad(km.dy, f1, by + km.ay);
                const bool mm = m && in_bounds(g, x, y);
                R xd = x, yd = y;
                bool okd = mm;
                if (!same) {
                    xd = mad(kd.dx, f1, bx + kd.ax);
                    yd = mad(kd.dy, f1, by + kd.ay);
                    okd = m && in_bounds(g, xd, yd);
                }
                Cell<R> cl;
                const R smp = sample(g, xd, yd, okd, cl);
                const R Ir = q < n ? I[q] : (R)0;
                const R dbuf = okd ? Ir - smp : (R)0;
                const bool pre = (long long)(r0 + q) * g.fframeW + c < swapBack;
                const R d1 = Ir - dbuf;           // the global correlation's swap
                const R d2 = pre ? Ir - d1 : d1;  // its partial swap-back
                const R d3 = mm ? Ir - d2 : dbuf; // the frame correlation's swap
                const double a = mm ? (double)d3 : 0.0, bb = mm ? (double)Ir : 0.0;
                // an unmasked pixel enters the reference's second pass times 0
                // (task_correlation.cu:150-156): NaN or Inf there gives NaN
                const double za = q < n ? (double)d3 * 0.0 : 0.0, zb = (double)Ir * 0.0;
                w[FC_CNT] += mm ? 1.0 : 0.0;
                w[FC_SA] += a;
                w[FC_SAA] += mm ? a * a : za * za;
                w[FC_SB] += bb;
                w[FC_SBB] += mm ? bb * bb : zb * zb;
                w[FC_SAB] += a * bb;
            }
        }
    }
    __shared__ double outF[FC_N];
    block_sum<FC_N>(w, outF);
    if (threadIdx.x < FC_N) g.finalsum[(size_t)c * FS_N + threadIdx.x] = outF[threadIdx.x];
}

template <typename TI, typename R>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(2))) void k_corr_frames(Batch B) {
    for (int item = blockIdx.x; item < B.nItems; item += gridDim.x) {
        const int gi = B.itemGroup[item];
        if (B.ctl[gi].phase != PH_CORR) continue;
        const GroupDev &g = B.groups[gi];
        if (!g.fcIter || !runs_here<TI, R>(g)) continue;
        frame_corr_column<TI, R>(g, item - g.itemStart);
    }
}

// Global correlation of the group and the stop tests of Task::alignFrames
// (task.cpp:281-342).
__global__ __launch_bounds__(kSolveWG) void k_corr_finish(Batch B) {
    const int gi = blockIdx.x;
    GroupCtl &ctl = B.ctl[gi];
    if (ctl.phase != PH_CORR) return;
    const GroupDev &g = B.groups[gi];
    // read before the barriers below: thread 0 advances it afterwards
    const int it = ctl.iteration;
    double v[CR_N + 1];
#pragma unroll
    for (int i = 0; i <= CR_N; i++) v[i] = 0.0;
    for (int c = threadIdx.x; c < g.fframeW; c += kSolveWG) {
        for (int i = 0; i < CR_N; i++) v[i] += g.corrsum[(size_t)c * CR_N + i];
        v[CR_N] += g.colsum[(size_t)c * CS_N + CS_SUMI];  // image over all pixels
    }
    __shared__ double tot[CR_N + 1];
    block_sum<CR_N + 1>(v, tot);
    const int P = 2 * g.N;
    if (g.fcIter && it <= g.maxIter)  // flag 'c' (task.cpp:304-306)
        for (int f = threadIdx.x; f < g.nF; f += kSolveWG) {
            double sf[FC_N];
            for (int i = 0; i < FC_N; i++) sf[i] = 0.0;
            for (int bb = 0; bb < g.nb; bb++)
                for (int i = 0; i < FC_N; i++) sf[i] += g.finalsum[(size_t)(f * g.nb + bb) * FS_N + i];
            const double nn = sf[FC_CNT], mA = sf[FC_SA] / nn, mB = sf[FC_SB] / nn;
            const double assq = sf[FC_SAA] - nn * mA * mA;
            const double bssq = sf[FC_SBB] - nn * mB * mB;
            const double q = sf[FC_SAB] - nn * mA * mB;
            g.fcIter[(size_t)(it - 1) * g.nF + f] = q / sqrt(assq * bssq);
        }
    if (g.allParams && it <= g.maxIter)
        for (int i = threadIdx.x; i < P; i += kSolveWG) g.allParams[(size_t)(it - 1) * P + i] = g.p[i];
    __shared__ int s_copy, s_save;
    if (threadIdx.x == 0) {
        const double nm = tot[CR_NMASK];
        const double mA = tot[CR_N] / nm, mB = tot[CR_SUMD] / nm, ns = tot[CR_NS];
        const double assq = tot[CR_SAA] - 2 * mA * tot[CR_SA] + ns * mA * mA;
        const double bssq = tot[CR_SBB] - 2 * mB * tot[CR_SB] + ns * mB * mB;
        const double q = tot[CR_SAB] - mB * tot[CR_SA] - mA * tot[CR_SB] + ns * mA * mB;
        const double corr = q / sqrt(assq * bssq);
        ctl.corr = corr;
        stage_time(g, ctl, FT_CORRELATION);
        if (g.iterCorr && it <= g.maxIter) g.iterCorr[it - 1] = corr;
        if (g.iterTimes) {  // output 9: this iteration's time
            const unsigned long long now = wall_clock64();
            if (it <= g.maxIter) g.iterTimes[it - 1] = (double)(now - ctl.tIter) / g.clockHz;
            ctl.tIter = now;
        }
        const double correlation = ctl.correlation;
        int copy = 0, save = 0;
        if (corr - correlation < g.cit) {
            ctl.retv = 0;  // CorrelationIncreaseThreshold
            ctl.phase = PH_FINAL;
        } else if (corr - correlation < g.fct && ctl.filterID != 0) {
            ctl.solverFilter = 0;
            ctl.nFilterIterations = 0;
            ctl.filterID = 0;
            if (it >= g.maxIter) {  // never run past max_iterations
                ctl.retv = 3;
                ctl.phase = PH_FINAL;
            } else {
                ctl.iteration = it + 1;
                ctl.improved = 1;
                ctl.shift = 0;
                ctl.phase = PH_ASSEMBLE;
            }
        } else if (!ctl.iterImproved) {
            ctl.retv = 2;  // MovementThreshold
            ctl.phase = PH_FINAL;
        } else if (g.maxIter == it) {
            ctl.retv = 3;  // IterationLimit
            ctl.phase = PH_FINAL;
        } else if (correlation >= g.halt) {
            ctl.retv = 1;  // HaltCorrelationReached
            ctl.phase = PH_FINAL;
        } else {
            ctl.correlation = corr;
            copy = 1;
            save = 1;
            ctl.iteration = it + 1;
            ctl.improved = 1;
            ctl.shift = 0;
            ctl.phase = PH_ASSEMBLE;
        }
        if (ctl.phase == PH_FINAL && it == 1) copy = 1;  // task.cpp:381-383
        s_copy = copy;
        s_save = save;
    }
    __syncthreads();
    if (s_copy)
        for (int i = threadIdx.x; i < P; i += kSolveWG) g.pLast[i] = g.p[i];
    if (s_save && g.iterErr)  // flag 'e': an iteration that continues (task.cpp:362-370)
        for (int f = threadIdx.x; f < g.nF; f += kSolveWG) {
            const size_t o = (size_t)f * 2 * g.maxIter + 2 * (size_t)(it - 1);
            g.iterErr[o] = g.errval[f];
            g.iterErr[o + 1] = g.errval[g.nF + f];
        }
}

// ---------------------------------------------------------------------------
// final state: errvals, frame correlations and mask of the last-tried state
// (hd_calculateGroupErrvals(0) + hd_frameCorrelation, task.cpp:389-405)
// ---------------------------------------------------------------------------
template <typename TI, typename R>
__device__ __forceinline__ void final_column(const GroupDev &g, int c) {
    const int f = c / g.nb, b = c % g.nb;
    const ColKnots<R> kn = state_knots<R>(g, c, g.lsMk[f]);
    const TI *img = column_ptr<TI>(g, c);
    const typename Prec<R>::V4 *geo = Prec<R>::geo(g) ? Prec<R>::geo(g) + (size_t)b * g.frameH : nullptr;
    uint8_t *mo = g.maskOut ? g.maskOut + (size_t)c * g.frameH : nullptr;
    double v[FS_N];
#pragma unroll
    for (int i = 0; i < FS_N; i++) v[i] = 0.0;
    constexpr int CH = chunk_rows<R>(1, sizeof(R) == 8 ? kFinalSamplesF64 : kFinalSamplesF32);
    for (int run = threadIdx.x; run * g.runLen < g.frameH; run += kWG) {
        const int rend = min(run * g.runLen + g.runLen, g.frameH);
        GeoIter<R> it;
        it.start(g, b, run * g.runLen);
        for (int r0 = run * g.runLen; r0 < rend; r0 += CH) {
            const int n = min(CH, rend - r0);
            R I[CH];
            load_chunk<CH>(img + r0, g.vec, n, I);
#pragma unroll
            for (int q = 0; q < CH; q++) {
                R bx, by, f1;
                bool m;
                it.at(g, geo, min(r0 + q, g.frameH - 1), bx, by, f1, m);
                it.next();
                const R x = mad(kn.dx, f1, bx + kn.ax);
                const R y = mad(kn.dy, f1, by + kn.ay);
                const bool mm = q < n && m && in_bounds(g, x, y);
                if (mo && q < n) mo[r0 + q] = mm ? 1 : 0;
                Cell<R> cl;
                const R s = sample(g, x, y, mm, cl);
                const double wI = mm ? (double)s : 0.0, Iq = mm ? (double)I[q] : 0.0;
                const double d = Iq - wI;
                v[FS_D2] += d * d;
                v[FS_CNT] += mm ? 1.0 : 0.0;
                v[FS_SA] += wI;
                v[FS_SAA] += wI * wI;
                v[FS_SB] += Iq;
                v[FS_SBB] += Iq * Iq;
                v[FS_SAB] += wI * Iq;
            }
        }
    }
    __shared__ double out[FS_N];
    block_sum<FS_N>(v, out);
    if (threadIdx.x < FS_N) g.finalsum[(size_t)c * FS_N + threadIdx.x] = out[threadIdx.x];
}

template <typename TI, typename R>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 8 ? kFinalWavesF64 : kFinalWavesF32))) void k_final(Batch B) {
    for (int item = blockIdx.x; item < B.nItems; item += gridDim.x) {
        const int gi = B.itemGroup[item];
        if (B.ctl[gi].phase != PH_FINAL) continue;
        const GroupDev &g = B.groups[gi];
        if (!runs_here<TI, R>(g)) continue;
        final_column<TI, R>(g, item - g.itemStart);
    }
}

// Residual RMS and frame correlations of the final state, and the end of
// the group.  Called by every thread of a kSolveWG workgroup.
__device__ void final_finish(const GroupDev &g, GroupCtl &ctl) {
    for (int f = threadIdx.x; f < g.nF; f += kSolveWG) {
        double s[FS_N];
        for (int i = 0; i < FS_N; i++) s[i] = 0.0;
        for (int bb = 0; bb < g.nb; bb++)
            for (int i = 0; i < FS_N; i++) s[i] += g.finalsum[(size_t)(f * g.nb + bb) * FS_N + i];
        const double nn = s[FS_CNT];
        g.resid[f] = sqrt(s[FS_D2] / nn);
        const double mA = s[FS_SA] / nn, mB = s[FS_SB] / nn;
        const double assq = s[FS_SAA] - nn * mA * mA;
        const double bssq = s[FS_SBB] - nn * mB * mB;
        const double q = s[FS_SAB] - nn * mA * mB;
        // flag 'c': the last iteration's correlations stand (task.cpp:392-394)
        g.fcorr[f] = g.fcIter ? g.fcIter[(size_t)(ctl.iteration - 1) * g.nF + f] : q / sqrt(assq * bssq);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        stage_time(g, ctl, FT_CORRELATION);
        ctl.phase = PH_DONE;
    }
}

__global__ __launch_bounds__(kSolveWG) void k_final_finish(Batch B) {
    const int gi = blockIdx.x;
    GroupCtl &ctl = B.ctl[gi];
    if (ctl.phase != PH_FINAL) return;
    final_finish(B.groups[gi], ctl);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// one launch per (frame type, precision) present in the batch
#define SS_LAUNCH_ONE(kernel, TI, R, grid, L, s)                                                  \
    do {                                                                                          \
        if ((L).hasType[img_type<TI>() + IMG_NTYPES * (sizeof(R) == 8 ? 1 : 0)])                  \
            hipLaunchKernelGGL((kernel<TI, R>), grid, dim3(kWG), 0, s, (L).b);                   \
    } while (0)
// kernels built in kVariants variants: one launch per variant some group of
// the type and precision uses
#define SS_LAUNCH_VAR_ONE(kernel, TK, TI, R, grid, L, s, ...)                                                 \
    do {                                                                                                      \
        const int m_ = (L).hasVar[TK][img_type<TI>() + IMG_NTYPES * (sizeof(R) == 8 ? 1 : 0)];                \
        if (m_ & 1) hipLaunchKernelGGL((kernel<TI, R, 0>), grid, dim3(kWG), 0, s, (L).b, ##__VA_ARGS__);      \
        if (m_ & 2) hipLaunchKernelGGL((kernel<TI, R, 1>), grid, dim3(kWG), 0, s, (L).b, ##__VA_ARGS__);      \
        if (m_ & 4) hipLaunchKernelGGL((kernel<TI, R, 2>), grid, dim3(kWG), 0, s, (L).b, ##__VA_ARGS__);      \
    } while (0)
#define SS_LAUNCH_VAR(kernel, TK, grid, L, s, ...)                                   \
    do {                                                                             \
        SS_LAUNCH_VAR_ONE(kernel, TK, float, float, grid, L, s, ##__VA_ARGS__);      \
        SS_LAUNCH_VAR_ONE(kernel, TK, uint16_t, float, grid, L, s, ##__VA_ARGS__);   \
        SS_LAUNCH_VAR_ONE(kernel, TK, int16_t, float, grid, L, s, ##__VA_ARGS__);    \
        SS_LAUNCH_VAR_ONE(kernel, TK, double, float, grid, L, s, ##__VA_ARGS__);     \
        SS_LAUNCH_VAR_ONE(kernel, TK, float, double, grid, L, s, ##__VA_ARGS__);     \
        SS_LAUNCH_VAR_ONE(kernel, TK, uint16_t, double, grid, L, s, ##__VA_ARGS__);  \
        SS_LAUNCH_VAR_ONE(kernel, TK, int16_t, double, grid, L, s, ##__VA_ARGS__);   \
        SS_LAUNCH_VAR_ONE(kernel, TK, double, double, grid, L, s, ##__VA_ARGS__);    \
    } while (0)
#define SS_LAUNCH_TYPED(kernel, grid, L, s)                  \
    do {                                                     \
        SS_LAUNCH_ONE(kernel, float, float, grid, L, s);     \
        SS_LAUNCH_ONE(kernel, uint16_t, float, grid, L, s);  \
        SS_LAUNCH_ONE(kernel, int16_t, float, grid, L, s);   \
        SS_LAUNCH_ONE(kernel, double, float, grid, L, s);    \
        SS_LAUNCH_ONE(kernel, float, double, grid, L, s);    \
        SS_LAUNCH_ONE(kernel, uint16_t, double, grid, L, s); \
        SS_LAUNCH_ONE(kernel, int16_t, double, grid, L, s);  \
        SS_LAUNCH_ONE(kernel, double, double, grid, L, s);   \
    } while (0)

hipError_t launch_stage(const LaunchBatch &L, int entry, hipStream_t s) {
    const dim3 items(L.itemGrid), groups(L.b.nGroups);
    const int slot = entry_slot(entry);
    switch (entry_stage(entry)) {
    case ST_ASSEMBLE: SS_LAUNCH_VAR(k_assemble, TK_ASSEMBLE, items, L, s); break;
    case ST_SOLVE: hipLaunchKernelGGL(k_solve, groups, dim3(kSolveWG), 0, s, L.b); break;
    case ST_LS:
        if (slot == 0) SS_LAUNCH_VAR(k_ls_round, TK_LS, items, L, s, slot, 1);
        else SS_LAUNCH_VAR(k_ls_round, TK_LS, dim3(L.lsSplitGrid), L, s, slot, L.lsSplit);
        break;
    case ST_LS_DECIDE:
        hipLaunchKernelGGL(k_ls_decide, dim3(L.frameGrid), dim3(64), 0, s, L.b, slot);
        if (slot == kLsSlots - 1) hipLaunchKernelGGL(k_ls_finish, groups, dim3(64), 0, s, L.b);
        break;
    case ST_CORR:
        SS_LAUNCH_VAR(k_corr, TK_CORR, items, L, s);
        if (L.hasFrameCorr) SS_LAUNCH_TYPED(k_corr_frames, items, L, s);
        break;
    case ST_CORR_FINISH: hipLaunchKernelGGL(k_corr_finish, groups, dim3(kSolveWG), 0, s, L.b); break;
    case ST_FINAL: SS_LAUNCH_TYPED(k_final, items, L, s); break;
    case ST_FINAL_FINISH: hipLaunchKernelGGL(k_final_finish, groups, dim3(kSolveWG), 0, s, L.b); break;
    }
    return hipGetLastError();
}

size_t solve_workspace_bytes(int n) {
    // Eq0 + InteriorRow + solution per knot, first-interior forms per
    // partition, the refinement's saved solution per knot
    return (size_t)n * (sizeof(Eq0) + sizeof(InteriorRow) + 2 * sizeof(double2)) + kSolveWG * sizeof(Aff) + 256;
}

}  // namespace ss
