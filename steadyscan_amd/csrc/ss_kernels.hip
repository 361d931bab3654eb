// ss_kernels.hip -- gfx950 kernels of the align_group_to_template path.
//
// Work decomposition: one workgroup (4 wave64) per line-block column of a
// frame (frame f, block b = column c = f * nb + b of the reference's block
// matrix).  A column is contiguous in HBM (frame_h pixels), so a wave reads
// 64 consecutive pixels per load; per-column partial sums are reduced in
// registers, then across the 4 waves through LDS, and written once as FP64.
// Nothing is accumulated with atomics, so every result is run-to-run
// deterministic.
#include "ss_device.h"

#include <float.h>

namespace ss {

constexpr int kWG = 256;
constexpr int kSolveWG = 256;

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
template <typename TI>
__device__ __forceinline__ float load_px(const TI *p, int r) { return (float)p[r]; }

// Reduces NV per-thread doubles over the workgroup; thread i < NV writes out[i].
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *out) {
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
}

__device__ __forceinline__ int block_min_int(int v) {
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

// Bilinear sample of the row-major template at (x, y), x/y inside the mask
// bounds (task_differenceimage.cuh:13-32).  Also returns the cell.
__device__ __forceinline__ float bilinear(const float *__restrict__ T, int tw, float x, float y,
                                          int &t, float &fx, float &fy) {
    const int ix = (int)x, iy = (int)y;
    fx = x - (float)ix;
    fy = y - (float)iy;
    t = iy * tw + ix;
    const float t00 = T[t], t01 = T[t + 1], t10 = T[t + tw], t11 = T[t + tw + 1];
    return (1.f - fx) * ((1.f - fy) * t00 + fy * t10) + fx * ((1.f - fy) * t01 + fy * t11);
}

__device__ __forceinline__ bool in_mask(const GroupDev &g, float m, float x, float y) {
    return m != 0.f && x >= 0.f && x < g.xmax && y >= 0.f && y < g.ymax;
}

__device__ __forceinline__ double reduce_pow(double v, double r, int k) {
    for (int i = 0; i < k; i++) v *= r;
    return v;
}

// ---------------------------------------------------------------------------
// assembly: warp + residual + errval terms + J^T J / J^T d partials
// (d_applyParameters, d_createDifferenceImage, d_calcBaseErrval,
//  d_updateGradients, d_fillDiagonals, d_fillJTr fused into one pass)
// ---------------------------------------------------------------------------
template <typename TI>
__global__ __launch_bounds__(kWG) void k_assemble(GroupDev g) {
    const int c = blockIdx.x;
    const int b = c % g.nb;
    const int N = g.N;
    const float lx = (float)g.p[c], rx = (float)g.p[c + 1];
    const float ly = (float)g.p[c + N], ry = (float)g.p[c + 1 + N];
    const TI *img = (const TI *)g.img + (size_t)c * g.frameH;
    const float4 *geo = g.geo + (size_t)b * g.frameH;
    const float *fn = g.fracUniform ? nullptr : g.fnext + (size_t)b * g.frameH;
    const int tw = g.tw;
    double d2 = 0.0;
    int cnt = 0;
    float h[CS_N - CS_AXX];
#pragma unroll
    for (int i = 0; i < CS_N - CS_AXX; i++) h[i] = 0.f;
    for (int r = threadIdx.x; r < g.frameH; r += kWG) {
        const float4 q = geo[r];
        const float f = q.z, cf = 1.f - f;
        const float x = q.x + lx * cf + rx * f;
        const float y = q.y + ly * cf + ry * f;
        if (in_mask(g, q.w, x, y)) {
            int t;
            float fx, fy;
            const float wI = bilinear(g.tmpl, tw, x, y, t, fx, fy);
            const float d = load_px(img, r) - wI;
            const float wx = (1.f - fy) * g.xg[t] + fy * g.xg[t + tw];
            const float wy = (1.f - fx) * g.yg[t] + fx * g.yg[t + 1];
            d2 += (double)d * (double)d;
            cnt++;
            const float xx = wx * wx, yy = wy * wy, xy = wx * wy;
            const float c2 = cf * cf, f2 = f * f, fc = f - f2;
            h[CS_AXX - CS_AXX] += xx * c2;
            h[CS_AYY - CS_AXX] += yy * c2;
            h[CS_AXY - CS_AXX] += xy * c2;
            h[CS_BXX - CS_AXX] += xx * f2;
            h[CS_BYY - CS_AXX] += yy * f2;
            h[CS_BXY - CS_AXX] += xy * f2;
            h[CS_CXX - CS_AXX] += xx * fc;
            h[CS_CYY - CS_AXX] += yy * fc;
            h[CS_CXY - CS_AXX] += xy * fc;
            const float dc = d * cf, df = d * f;
            h[CS_JXA - CS_AXX] += dc * wx;
            h[CS_JYA - CS_AXX] += dc * wy;
            h[CS_JXB - CS_AXX] += df * wx;
            h[CS_JYB - CS_AXX] += df * wy;
            if (fn) {
                const float n2 = fn[r] * fn[r];
                h[CS_QXX - CS_AXX] += xx * n2;
                h[CS_QYY - CS_AXX] += yy * n2;
                h[CS_QXY - CS_AXX] += xy * n2;
            }
        }
    }
    double v[CS_N];
    v[CS_D2] = d2;
    v[CS_CNT] = (double)cnt;
#pragma unroll
    for (int i = CS_AXX; i < CS_N; i++) v[i] = (double)h[i - CS_AXX];
    block_reduce_store<CS_N>(v, g.colsum + (size_t)c * CS_N);
}

// ---------------------------------------------------------------------------
// solve: errvals, knot usage, block-tridiagonal normal equations, exact FP64
// solve by block cyclic reduction, step filter, red/green expansion.
// One workgroup per group.
// ---------------------------------------------------------------------------
struct Eq {  // A x_{i-1} + B x_i + C x_{i+1} = d (2x2 row-major blocks)
    double A[4], B[4], C[4], d[2], x[2];
};

__device__ __forceinline__ void inv2(const double *M, double *R) {
    const double det = M[0] * M[3] - M[1] * M[2];
    const double id = 1.0 / det;
    R[0] = M[3] * id;
    R[1] = -M[1] * id;
    R[2] = -M[2] * id;
    R[3] = M[0] * id;
}
__device__ __forceinline__ void mul2(const double *A, const double *B, double *R) {
    R[0] = A[0] * B[0] + A[1] * B[2];
    R[1] = A[0] * B[1] + A[1] * B[3];
    R[2] = A[2] * B[0] + A[3] * B[2];
    R[3] = A[2] * B[1] + A[3] * B[3];
}
__device__ __forceinline__ void mulv2(const double *A, const double *v, double *r) {
    r[0] = A[0] * v[0] + A[1] * v[1];
    r[1] = A[2] * v[0] + A[3] * v[1];
}

__device__ __forceinline__ int level_size(int n, int lvl) {
    int m = n;
    for (int i = 0; i < lvl; i++) m >>= 1;
    return m;
}

__global__ __launch_bounds__(kSolveWG) void k_solve(GroupDev g, int shift, int filterKind,
                                                    double strength, double lambda) {
    const int tid = threadIdx.x;
    const int nF = g.nF, nb = g.nb, N = g.N, n = g.fframeW;
    const double *S = g.colsum;
    // 1. per-frame errval of this shift (d_calcAllGroupErrval)
    for (int f = tid; f < nF; f += kSolveWG) {
        double s = 0.0, c = 0.0;
        for (int bb = 0; bb < nb; bb++) {
            s += S[(size_t)(f * nb + bb) * CS_N + CS_D2];
            c += S[(size_t)(f * nb + bb) * CS_N + CS_CNT];
        }
        g.errval[shift * nF + f] = s / c;
    }
    // 2. knot usage (d_createParameterMask); the last knot is never used
    for (int k = tid; k < N; k += kSolveWG) {
        uint8_t u = 0;
        if (k < n) {
            const bool pc = S[(size_t)k * CS_N + CS_CNT] >= (double)g.minPts[k % nb];
            const bool pp = k > 0 && S[(size_t)(k - 1) * CS_N + CS_CNT] >= (double)g.minPts[(k - 1) % nb];
            u = (pc || pp) ? 1 : 0;
        }
        g.used[k] = u;
    }
    __syncthreads();
    // 3. level-0 equations (d_fillDiagonals / d_fillJTr entries)
    Eq *L0 = reinterpret_cast<Eq *>(g.ws);
    for (int k = tid; k < n; k += kSolveWG) {
        Eq e;
        const double *s = S + (size_t)k * CS_N;
        const bool u = g.used[k];
        const bool up = k > 0 && g.used[k - 1];
        const bool un = (k + 1 < n) && g.used[k + 1];
        if (!u) {
            for (int i = 0; i < 4; i++) { e.A[i] = 0; e.C[i] = 0; e.B[i] = 0; }
            e.B[0] = 1;
            e.B[3] = 1;
            e.d[0] = e.d[1] = 0;
        } else {
            double dxx = s[CS_AXX], dyy = s[CS_AYY], dxy = s[CS_AXY];
            double jx = s[CS_JXA], jy = s[CS_JYA];
            if (k > 0) {
                const double *sp = s - CS_N;
                const int qx = g.fracUniform ? CS_BXX : CS_QXX;
                dxx += sp[qx];
                dyy += sp[qx + 1];
                dxy += sp[qx + 2];
                jx += sp[CS_JXB];
                jy += sp[CS_JYB];
            }
            const double nl = (double)(up ? 1 : 0) + (double)(un ? 1 : 0);
            e.B[0] = dxx + lambda * nl;
            e.B[1] = dxy;
            e.B[2] = dxy;
            e.B[3] = dyy + lambda * nl;
            if (un) {
                e.C[0] = s[CS_CXX] - lambda;
                e.C[1] = s[CS_CXY];
                e.C[2] = s[CS_CXY];
                e.C[3] = s[CS_CYY] - lambda;
            } else {
                e.C[0] = e.C[1] = e.C[2] = e.C[3] = 0;
            }
            if (up) {
                const double *sp = s - CS_N;
                e.A[0] = sp[CS_CXX] - lambda;
                e.A[1] = sp[CS_CXY];
                e.A[2] = sp[CS_CXY];
                e.A[3] = sp[CS_CYY] - lambda;
            } else {
                e.A[0] = e.A[1] = e.A[2] = e.A[3] = 0;
            }
            e.d[0] = jx;
            e.d[1] = jy;
        }
        e.x[0] = e.x[1] = 0;
        L0[k] = e;
    }
    __threadfence_block();
    __syncthreads();
    // 4. block cyclic reduction, forward
    int nlev = 0;
    {
        int m = n;
        while (m > 1) { m >>= 1; nlev++; }
    }
    int off = 0;
    for (int lvl = 0; lvl < nlev; lvl++) {
        const int m = level_size(n, lvl), mn = m >> 1;
        Eq *Lc = reinterpret_cast<Eq *>(g.ws) + off;
        Eq *Ln = Lc + m;
        for (int j = tid; j < mn; j += kSolveWG) {
            const int i = 2 * j + 1;
            const Eq &e = Lc[i], &em = Lc[i - 1];
            double iv[4], al[4], t[4];
            Eq o;
            inv2(em.B, iv);
            mul2(e.A, iv, al);
            for (int q = 0; q < 4; q++) al[q] = -al[q];
            mul2(al, em.A, o.A);
            mul2(al, em.C, t);
            for (int q = 0; q < 4; q++) o.B[q] = e.B[q] + t[q];
            double tv[2];
            mulv2(al, em.d, tv);
            o.d[0] = e.d[0] + tv[0];
            o.d[1] = e.d[1] + tv[1];
            if (i + 1 < m) {
                const Eq &ep = Lc[i + 1];
                double ga[4];
                inv2(ep.B, iv);
                mul2(e.C, iv, ga);
                for (int q = 0; q < 4; q++) ga[q] = -ga[q];
                mul2(ga, ep.A, t);
                for (int q = 0; q < 4; q++) o.B[q] += t[q];
                mul2(ga, ep.C, o.C);
                mulv2(ga, ep.d, tv);
                o.d[0] += tv[0];
                o.d[1] += tv[1];
            } else {
                o.C[0] = o.C[1] = o.C[2] = o.C[3] = 0;
            }
            o.x[0] = o.x[1] = 0;
            Ln[j] = o;
        }
        off += m;
        __threadfence_block();
        __syncthreads();
    }
    // top level has one equation
    if (tid == 0 && n > 0) {
        Eq *Lt = reinterpret_cast<Eq *>(g.ws) + off;
        double iv[4];
        inv2(Lt[0].B, iv);
        mulv2(iv, Lt[0].d, Lt[0].x);
    }
    __threadfence_block();
    __syncthreads();
    // 5. back substitution
    for (int lvl = nlev - 1; lvl >= 0; lvl--) {
        const int m = level_size(n, lvl);
        off -= m;
        Eq *Lc = reinterpret_cast<Eq *>(g.ws) + off;
        Eq *Ln = Lc + m;
        for (int i = tid; i < m; i += kSolveWG) {
            if (i & 1) {
                Lc[i].x[0] = Ln[i >> 1].x[0];
                Lc[i].x[1] = Ln[i >> 1].x[1];
            }
        }
        __threadfence_block();
        __syncthreads();
        for (int i = tid; i < m; i += kSolveWG) {
            if (!(i & 1)) {
                const Eq &e = Lc[i];
                double r[2] = {e.d[0], e.d[1]}, tv[2];
                if (i > 0) {
                    mulv2(e.A, Lc[i - 1].x, tv);
                    r[0] -= tv[0];
                    r[1] -= tv[1];
                }
                if (i + 1 < m) {
                    mulv2(e.C, Lc[i + 1].x, tv);
                    r[0] -= tv[0];
                    r[1] -= tv[1];
                }
                double iv[4];
                inv2(e.B, iv);
                mulv2(iv, r, Lc[i].x);
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    // 6. step filter over the used knots, x sequence then y sequence
    //    (SmearingFilter::apply / RangeFilter::apply on the compacted halves)
    if (filterKind == 1) {
        // y_i = s * y_{i-1} + (1 - s) * v_i over used knots: affine scan
        __shared__ double sa[kSolveWG][4];
        const int chunk = (n + kSolveWG - 1) / kSolveWG;
        const int c0 = min(n, tid * chunk), c1 = min(n, c0 + chunk);
        double ax = 1, bx = 0, ay = 1, by = 0;
        for (int k = c0; k < c1; k++) {
            if (!g.used[k]) continue;
            ax *= strength;
            bx = strength * bx + (1 - strength) * L0[k].x[0];
            ay *= strength;
            by = strength * by + (1 - strength) * L0[k].x[1];
        }
        sa[tid][0] = ax; sa[tid][1] = bx; sa[tid][2] = ay; sa[tid][3] = by;
        __syncthreads();
        // inclusive Hillis-Steele scan of affine maps
        for (int st = 1; st < kSolveWG; st <<= 1) {
            double v0 = 0, v1 = 0, v2 = 0, v3 = 0;
            const bool has = tid >= st;
            if (has) { v0 = sa[tid - st][0]; v1 = sa[tid - st][1]; v2 = sa[tid - st][2]; v3 = sa[tid - st][3]; }
            __syncthreads();
            if (has) {
                // compose: earlier (v) then current (sa[tid])
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
            yx = yx * strength + L0[k].x[0] * (1 - strength);
            yy = yy * strength + L0[k].x[1] * (1 - strength);
            L0[k].x[0] = yx;
            L0[k].x[1] = yy;
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
                        const double vm = L0[pv].x[axis], vi = L0[cu].x[axis], vp = L0[k].x[axis];
                        const double tc = (vm + vp) / 2;
                        const double diff = (tc - vm) * strength;
                        bool rep;
                        if (vm < vp) rep = vi > vp + diff || vi < vm - diff;
                        else rep = vi < vp + diff || vi > vm - diff;
                        if (rep) L0[cu].x[axis] = tc;
                    }
                    pv = cu;
                    cu = k;
                }
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    // 7. d_expandResults (task_expandresults.cu:10-41) and line-search set-up
    const int G = g.groupSize;
    for (int idx = tid; idx < 2 * N; idx += kSolveWG) {
        const int j = idx >= N ? idx - N : idx;
        const bool fixed = ((((idx + 1) - (shift * (G >> 1))) % G) == 0);
        double dv = g.deltaP[idx];
        if (!fixed) {
            if (j < n && g.used[j]) dv = L0[j].x[idx >= N ? 1 : 0];
            else dv = 0.0;
        }
        g.deltaP[idx] = dv;
        g.d0[idx] = dv;
        g.p0[idx] = g.p[idx];
    }
    for (int f = tid; f < nF; f += kSolveWG) {
        g.lsActive[f] = 1;
        g.lsNextK[f] = 0;
        g.lsMk[f] = -1;
        g.lsDk[f] = -1;
    }
}

// ---------------------------------------------------------------------------
// line search: K fused candidates p0 + r^k * delta for every active frame
// (d_linesearch's tries, task_linesearch.cu:35-106, evaluated in one pass)
// ---------------------------------------------------------------------------
template <typename TI, int K>
__global__ __launch_bounds__(kWG) void k_ls_round(GroupDev g) {
    const int c = blockIdx.x, f = c / g.nb, b = c % g.nb;
    if (!g.lsActive[f]) return;
    const int k0 = g.lsNextK[f];
    const int N = g.N;
    const double r = g.rmult;
    double dlx = reduce_pow(g.d0[c], r, k0), dly = reduce_pow(g.d0[c + N], r, k0);
    double drx = reduce_pow(g.d0[c + 1], r, k0), dry = reduce_pow(g.d0[c + 1 + N], r, k0);
    const bool lastKnot = (c + 1 == g.fframeW);
    const double plx = g.p0[c], ply = g.p0[c + N], prx = g.p0[c + 1], pry = g.p0[c + 1 + N];
    float klx[K], kly[K], krx[K], kry[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        klx[k] = (float)(plx + dlx);
        kly[k] = (float)(ply + dly);
        krx[k] = lastKnot ? 1e6f : (float)(prx + drx);
        kry[k] = lastKnot ? 1e6f : (float)(pry + dry);
        dlx *= r; dly *= r; drx *= r; dry *= r;
    }
    const TI *img = (const TI *)g.img + (size_t)c * g.frameH;
    const float4 *geo = g.geo + (size_t)b * g.frameH;
    double e2[K];
    int cnt[K];
#pragma unroll
    for (int k = 0; k < K; k++) { e2[k] = 0.0; cnt[k] = 0; }
    for (int rr = threadIdx.x; rr < g.frameH; rr += kWG) {
        const float4 q = geo[rr];
        const float f1 = q.z, cf = 1.f - f1;
        const float I = load_px(img, rr);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const float x = q.x + klx[k] * cf + krx[k] * f1;
            const float y = q.y + kly[k] * cf + kry[k] * f1;
            if (in_mask(g, q.w, x, y)) {
                int t;
                float fx, fy;
                const float d = I - bilinear(g.tmpl, g.tw, x, y, t, fx, fy);
                e2[k] += (double)d * (double)d;
                cnt[k]++;
            }
        }
    }
    double v[2 * K];
#pragma unroll
    for (int k = 0; k < K; k++) { v[2 * k] = e2[k]; v[2 * k + 1] = (double)cnt[k]; }
    block_reduce_store<2 * K>(v, g.lssum + (size_t)c * 2 * kMaxCand);
}

// Sequential resolution of the fused candidates per frame: reproduces the
// reference's try / reduce / movement-threshold order exactly.
__global__ __launch_bounds__(kSolveWG) void k_ls_decide(GroupDev g, int shift, int K,
                                                        double thr) {
    const int nF = g.nF, nb = g.nb, N = g.N;
    const double r = g.rmult;
    int tries = 0, pendingCount = 0;
    bool pending = false;
    for (int f = threadIdx.x; f < nF; f += kSolveWG) {
        if (!g.lsActive[f]) continue;
        const int k0 = g.lsNextK[f];
        double dmax = 0.0;
        for (int t = 0; t < nb; t++) {
            const double a = fabs(g.d0[f * nb + t]), bq = fabs(g.d0[f * nb + t + N]);
            if (a > dmax) dmax = a;
            if (bq > dmax) dmax = bq;
        }
        double dm = reduce_pow(dmax, r, k0);
        int fin = -1;      // exponent of the final delta
        int acc = -1;      // accepted candidate
        int mk = -1;
        for (int kk = 0; kk < K && fin < 0; kk++) {
            const int k = k0 + kk;
            tries++;
            mk = k;
            bool present = false;
            double s = 0.0, cn = 0.0;
            for (int bb = 0; bb < nb; bb++) {
                const double *ls = g.lssum + (size_t)(f * nb + bb) * 2 * kMaxCand + 2 * kk;
                if (ls[1] >= (double)g.minPts[bb]) present = true;
                s += ls[0];
                cn += ls[1];
            }
            if (!present) {
                const double dm1 = dm * r;
                if (!(dm1 > thr)) fin = k + 1;
                dm = dm1;
            } else {
                g.lsDk[f] = k;
                const double e = s / cn;
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
        if (fin >= 0) {
            for (int t = 0; t < nb; t++) {
                for (int ax = 0; ax < 2; ax++) {
                    const int i = f * nb + t + ax * N;
                    const double d0 = g.d0[i];
                    g.deltaP[i] = reduce_pow(d0, r, fin);
                    if (acc >= 0) g.p[i] = g.p0[i] + reduce_pow(d0, r, acc);
                }
            }
            g.lsActive[f] = 0;
        } else {
            g.lsNextK[f] = k0 + K;
            pending = true;
            pendingCount++;
        }
    }
    // workgroup-wide: anything pending?  then the `improved` flag
    __shared__ int shp, sht;
    if (threadIdx.x == 0) { shp = 0; sht = 0; }
    __syncthreads();
    if (pending) atomicAdd(&shp, pendingCount);
    atomicAdd(&sht, tries);
    __syncthreads();
    if (threadIdx.x == 0) {
        g.status->need_more = shp ? 1 : 0;
        g.status->active = shp;
        g.status->tries += sht;
        if (!shp) {
            double e0 = 0.0, e1 = 0.0;
            for (int f = 0; f < nF; f++) {
                e0 += g.errval[f];
                e1 += g.errval[f + nF];
            }
            g.status->improved = (shift && e0 > e1) || (!shift && e0 < e1);
        }
    }
}

// ---------------------------------------------------------------------------
// global correlation statistic (hd_globalCorrelation, task_correlation.cu:63-99)
// of the last-tried state of every frame.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void state_knots(const GroupDev &g, int c, int k, float &lx, float &rx,
                                            float &ly, float &ry) {
    const int N = g.N;
    if (k < 0) {  // state of the half-iteration start: p0 with the real right knot
        lx = (float)g.p0[c];
        ly = (float)g.p0[c + N];
        rx = (float)g.p0[c + 1];
        ry = (float)g.p0[c + 1 + N];
        return;
    }
    const double r = g.rmult;
    lx = (float)(g.p0[c] + reduce_pow(g.d0[c], r, k));
    ly = (float)(g.p0[c + N] + reduce_pow(g.d0[c + N], r, k));
    if (c + 1 == g.fframeW) {
        rx = 1e6f;
        ry = 1e6f;
    } else {
        rx = (float)(g.p0[c + 1] + reduce_pow(g.d0[c + 1], r, k));
        ry = (float)(g.p0[c + 1 + N] + reduce_pow(g.d0[c + 1 + N], r, k));
    }
}

template <typename TI>
__global__ __launch_bounds__(kWG) void k_corr(GroupDev g) {
    const int c = blockIdx.x, f = c / g.nb, b = c % g.nb;
    const int mk = g.lsMk[f], dk = g.lsDk[f];
    const bool same = (mk == dk);
    float mlx, mrx, mly, mry, dlx, drx, dly, dry;
    state_knots(g, c, mk, mlx, mrx, mly, mry);
    state_knots(g, c, dk, dlx, drx, dly, dry);
    const TI *img = (const TI *)g.img + (size_t)c * g.frameH;
    const float4 *geo = g.geo + (size_t)b * g.frameH;
    // pass 1: first unmasked row of the column (the reference's walk stops there)
    int first = g.frameH;
    for (int rr = threadIdx.x; rr < g.frameH; rr += kWG) {
        const float4 q = geo[rr];
        const float x = q.x + mlx * (1.f - q.z) + mrx * q.z;
        const float y = q.y + mly * (1.f - q.z) + mry * q.z;
        if (!in_mask(g, q.w, x, y)) { first = rr; break; }
    }
    const int r0 = block_min_int(first);
    double v[CR_N];
#pragma unroll
    for (int i = 0; i < CR_N; i++) v[i] = 0.0;
    for (int rr = threadIdx.x; rr < g.frameH; rr += kWG) {
        const float4 q = geo[rr];
        const float cf = 1.f - q.z;
        const float I = load_px(img, rr);
        const float x = q.x + mlx * cf + mrx * q.z;
        const float y = q.y + mly * cf + mry * q.z;
        const bool mm = in_mask(g, q.w, x, y);
        int t;
        float fx, fy;
        float dbuf = 0.f;
        if (same) {
            if (mm) dbuf = I - bilinear(g.tmpl, g.tw, x, y, t, fx, fy);
        } else {
            const float xd = q.x + dlx * cf + drx * q.z;
            const float yd = q.y + dly * cf + dry * q.z;
            if (in_mask(g, q.w, xd, yd)) dbuf = I - bilinear(g.tmpl, g.tw, xd, yd, t, fx, fy);
        }
        v[CR_SUMI] += (double)I;
        v[CR_SUMD] += mm ? (double)(I - dbuf) : (double)dbuf;
        v[CR_NMASK] += mm ? 1.0 : 0.0;
        if (rr < r0) {
            const double a = (double)(I - dbuf), bb = (double)I;
            v[CR_SA] += a;
            v[CR_SAA] += a * a;
            v[CR_SB] += bb;
            v[CR_SBB] += bb * bb;
            v[CR_SAB] += a * bb;
            v[CR_NS] += 1.0;
        }
    }
    block_reduce_store<CR_N>(v, g.corrsum + (size_t)c * CR_N);
}

__global__ __launch_bounds__(kSolveWG) void k_corr_reduce(GroupDev g) {
    __shared__ double sh[CR_N];
    if (threadIdx.x < CR_N) {
        double s = 0.0;
        for (int c = 0; c < g.fframeW; c++) s += g.corrsum[(size_t)c * CR_N + threadIdx.x];
        sh[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double n = sh[CR_NMASK];
        const double mA = sh[CR_SUMI] / n, mB = sh[CR_SUMD] / n, ns = sh[CR_NS];
        const double assq = sh[CR_SAA] - 2 * mA * sh[CR_SA] + ns * mA * mA;
        const double bssq = sh[CR_SBB] - 2 * mB * sh[CR_SB] + ns * mB * mB;
        const double q = sh[CR_SAB] - mB * sh[CR_SA] - mA * sh[CR_SB] + ns * mA * mB;
        g.status->corr = q / sqrt(assq * bssq);
    }
}

// ---------------------------------------------------------------------------
// final state: errvals, frame correlations and mask of the last-tried state
// (hd_calculateGroupErrvals(0) + hd_frameCorrelation, task.cpp:389-405)
// ---------------------------------------------------------------------------
template <typename TI>
__global__ __launch_bounds__(kWG) void k_final(GroupDev g) {
    const int c = blockIdx.x, f = c / g.nb, b = c % g.nb;
    float lx, rx, ly, ry;
    state_knots(g, c, g.lsMk[f], lx, rx, ly, ry);
    const TI *img = (const TI *)g.img + (size_t)c * g.frameH;
    const float4 *geo = g.geo + (size_t)b * g.frameH;
    uint8_t *mo = g.maskOut ? g.maskOut + (size_t)c * g.frameH : nullptr;
    double v[FS_N];
#pragma unroll
    for (int i = 0; i < FS_N; i++) v[i] = 0.0;
    for (int rr = threadIdx.x; rr < g.frameH; rr += kWG) {
        const float4 q = geo[rr];
        const float cf = 1.f - q.z;
        const float x = q.x + lx * cf + rx * q.z;
        const float y = q.y + ly * cf + ry * q.z;
        const bool mm = in_mask(g, q.w, x, y);
        if (mo) mo[rr] = mm ? 1 : 0;
        if (mm) {
            int t;
            float fx, fy;
            const float I = load_px(img, rr);
            const float wI = bilinear(g.tmpl, g.tw, x, y, t, fx, fy);
            const double d = (double)(I - wI);
            v[FS_D2] += d * d;
            v[FS_CNT] += 1.0;
            v[FS_SA] += (double)wI;
            v[FS_SAA] += (double)wI * (double)wI;
            v[FS_SB] += (double)I;
            v[FS_SBB] += (double)I * (double)I;
            v[FS_SAB] += (double)wI * (double)I;
        }
    }
    block_reduce_store<FS_N>(v, g.finalsum + (size_t)c * FS_N);
}

__global__ __launch_bounds__(kSolveWG) void k_final_reduce(GroupDev g) {
    for (int f = threadIdx.x; f < g.nF; f += kSolveWG) {
        double s[FS_N];
        for (int i = 0; i < FS_N; i++) s[i] = 0.0;
        for (int bb = 0; bb < g.nb; bb++)
            for (int i = 0; i < FS_N; i++) s[i] += g.finalsum[(size_t)(f * g.nb + bb) * FS_N + i];
        const double n = s[FS_CNT];
        g.resid[f] = sqrt(s[FS_D2] / n);
        const double mA = s[FS_SA] / n, mB = s[FS_SB] / n;
        const double assq = s[FS_SAA] - n * mA * mA;
        const double bssq = s[FS_SBB] - n * mB * mB;
        const double q = s[FS_SAB] - n * mA * mB;
        g.fcorr[f] = q / sqrt(assq * bssq);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename TI>
static hipError_t launch_assemble_t(const GroupDev &g, hipStream_t s) {
    hipLaunchKernelGGL(k_assemble<TI>, dim3(g.fframeW), dim3(kWG), 0, s, g);
    return hipGetLastError();
}
hipError_t launch_assemble(const GroupDev &g, hipStream_t s) {
    switch (g.imgType) {
    case IMG_U16: return launch_assemble_t<uint16_t>(g, s);
    case IMG_I16: return launch_assemble_t<int16_t>(g, s);
    default: return launch_assemble_t<float>(g, s);
    }
}

hipError_t launch_solve(const GroupDev &g, int shift, int filterKind, double strength,
                        double lambda, hipStream_t s) {
    hipLaunchKernelGGL(k_solve, dim3(1), dim3(kSolveWG), 0, s, g, shift, filterKind, strength, lambda);
    return hipGetLastError();
}

template <typename TI>
static hipError_t launch_ls_round_t(const GroupDev &g, int K, hipStream_t s) {
    switch (K) {
    case 1: hipLaunchKernelGGL((k_ls_round<TI, 1>), dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    case 2: hipLaunchKernelGGL((k_ls_round<TI, 2>), dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    case 4: hipLaunchKernelGGL((k_ls_round<TI, 4>), dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    default: hipLaunchKernelGGL((k_ls_round<TI, 8>), dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    }
    return hipGetLastError();
}
hipError_t launch_ls_round(const GroupDev &g, int K, hipStream_t s) {
    switch (g.imgType) {
    case IMG_U16: return launch_ls_round_t<uint16_t>(g, K, s);
    case IMG_I16: return launch_ls_round_t<int16_t>(g, K, s);
    default: return launch_ls_round_t<float>(g, K, s);
    }
}

hipError_t launch_ls_decide(const GroupDev &g, int shift, int K, double thr, hipStream_t s) {
    hipLaunchKernelGGL(k_ls_decide, dim3(1), dim3(kSolveWG), 0, s, g, shift, K, thr);
    return hipGetLastError();
}

hipError_t launch_corr(const GroupDev &g, hipStream_t s) {
    switch (g.imgType) {
    case IMG_U16: hipLaunchKernelGGL(k_corr<uint16_t>, dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    case IMG_I16: hipLaunchKernelGGL(k_corr<int16_t>, dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    default: hipLaunchKernelGGL(k_corr<float>, dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    }
    hipLaunchKernelGGL(k_corr_reduce, dim3(1), dim3(kSolveWG), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_final(const GroupDev &g, hipStream_t s) {
    switch (g.imgType) {
    case IMG_U16: hipLaunchKernelGGL(k_final<uint16_t>, dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    case IMG_I16: hipLaunchKernelGGL(k_final<int16_t>, dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    default: hipLaunchKernelGGL(k_final<float>, dim3(g.fframeW), dim3(kWG), 0, s, g); break;
    }
    hipLaunchKernelGGL(k_final_reduce, dim3(1), dim3(kSolveWG), 0, s, g);
    return hipGetLastError();
}

size_t solve_workspace_bytes(int n) {
    size_t eqs = 0;
    int m = n;
    for (;;) {
        eqs += (size_t)(m > 0 ? m : 1);
        if (m <= 1) break;
        m >>= 1;
    }
    return eqs * sizeof(Eq);
}

}  // namespace ss
