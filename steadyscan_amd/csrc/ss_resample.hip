// ss_resample.hip -- corrected frames: resample_frame + pushgrid
// (steady_scan.m:562-599) on the GPU.
//
// pushgrid splats every in-bounds data pixel onto the 4 template pixels of
// its cell with bilinear weights, through MATLAB indexed assignments
// z(ind) = z(ind) + v.  With repeated indices such an assignment keeps the
// value of the LAST occurrence, so for each template pixel and each of the
// four corner offsets exactly one data pixel contributes: the one with the
// largest (column-major) pixel index.  The GPU form is therefore
//   pass 0: the winner table of the launch's frames is set to -1 (a fill
//           kernel, so that the pass also runs under stream capture);
//   pass 1: per frame and corner offset, atomicMax of the pixel index into
//           the target (an integer maximum: order-independent, deterministic);
//   pass 2: per frame and template pixel, gather the <= 4 winners and
//           accumulate in single precision in the statement order of pushgrid.
// Positions are computed in FP64 without contraction, as MATLAB does.
//
// All frames of a launch share the pixel geometry; grid.y is the frame.  The
// work memory (geometry, knots, winner tables and the staging of host frames
// and host output) is the handle's, kept across calls, and everything runs
// on the handle's stream with one synchronisation at the end of the call.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include <hip/hip_runtime.h>

#include "ss_host.h"
#include "steadyscan.h"

namespace {

struct PixGeo {
    double bx, by, tf;
    int k;  // prevknotind - 1
    int pad;
};

__global__ __launch_bounds__(256) void k_fill_minus1(int *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = -1;
}

// Displaced position of pixel q for the knots of one frame (1-based template
// coordinates, steady_scan.m:565-566).
__device__ __forceinline__ void position(const PixGeo &q, const double *kx, const double *ky, double &x,
                                         double &y) {
#pragma clang fp contract(off)
    const double tc = 1.0 - q.tf;
    x = q.bx + kx[q.k] * tc + kx[q.k + 1] * q.tf;
    y = q.by + ky[q.k] * tc + ky[q.k + 1] * q.tf;
}

__global__ __launch_bounds__(256) void k_splat_winners(const PixGeo *__restrict__ geo, int npix,
                                                       const double *__restrict__ knots, int nk, int th,
                                                       int tw, int *__restrict__ win) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const int f = blockIdx.y;
    const double *kx = knots + (size_t)f * 2 * nk, *ky = kx + nk;
    double x, y;
    position(geo[i], kx, ky, x, y);
    if (!(x >= 1.0 && x < (double)tw && y >= 1.0 && y < (double)th)) return;
    const int xi = (int)floor(x), yi = (int)floor(y);
    const int m0 = (yi - 1) + th * (xi - 1);
    const size_t tp = (size_t)th * tw;
    int *w = win + (size_t)f * 4 * tp;
    atomicMax(&w[m0], i);
    atomicMax(&w[tp + m0 + 1], i);
    atomicMax(&w[2 * tp + m0 + th], i);
    atomicMax(&w[3 * tp + m0 + th + 1], i);
}

template <typename TD>
__global__ __launch_bounds__(256) void k_splat_gather(const PixGeo *__restrict__ geo, const TD *__restrict__ data,
                                                      size_t framePix, int hc, int nl, int cl, int cb,
                                                      const double *__restrict__ knots, int nk, int th, int tw,
                                                      const int *__restrict__ win, float *__restrict__ out) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tp = (size_t)th * tw;
    if (t >= (int)tp) return;
    const int f = blockIdx.y;
    const double *kx = knots + (size_t)f * 2 * nk, *ky = kx + nk;
    const int *wf = win + (size_t)f * 4 * tp;
    const TD *df = data + framePix * f;
    float z = 0.f, n = 0.f;
    double w[4];
    double d[4];
    bool has[4];
    for (int s = 0; s < 4; s++) {
        const int i = wf[s * tp + t];
        has[s] = i >= 0;
        w[s] = 0.0;
        d[s] = 0.0;
        if (!has[s]) continue;
        double x, y;
        position(geo[i], kx, ky, x, y);
        const double xf = x - floor(x), yf = y - floor(y);
        // corner offsets 0, 1, th, th+1 <-> weights (steady_scan.m:589-592)
        w[s] = s == 0 ? (1 - xf) * (1 - yf) : s == 1 ? (1 - xf) * yf : s == 2 ? xf * (1 - yf) : xf * yf;
        const int ic = i % hc, jc = i / hc;  // cropped (row, column)
        d[s] = (double)df[(size_t)(jc + cl) * nl + (ic + cb)];
    }
    // z statements in order: offsets 0, 1, th, th+1 (steady_scan.m:589-592)
    for (int s = 0; s < 4; s++)
        if (has[s]) z = z + (float)(w[s] * d[s]);
    // n statements in order: offsets 1, th, 0, th+1 (steady_scan.m:595-598)
    const int norder[4] = {1, 2, 0, 3};
    for (int o = 0; o < 4; o++) {
        const int s = norder[o];
        if (has[s]) n = n + (float)w[s];
    }
    out[(size_t)f * tp + t] = z / n;
}

template <typename TD>
hipError_t launch_frames(hipStream_t s, const PixGeo *geo, int npix, const void *data, size_t framePix,
                         const ss_geometry_desc &g, int hc, const double *knots, int nk, int nf, int *win,
                         float *out) {
    const int th = g.template_h, tw = g.template_w;
    const size_t tp = (size_t)th * tw, nwin = (size_t)nf * 4 * tp;
    const int fillGrid = (int)std::min<size_t>((nwin + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_minus1, dim3(fillGrid), dim3(256), 0, s, win, nwin);
    hipLaunchKernelGGL(k_splat_winners, dim3((npix + 255) / 256, nf), dim3(256), 0, s, geo, npix, knots, nk, th, tw,
                       win);
    hipLaunchKernelGGL(k_splat_gather<TD>, dim3((unsigned)((tp + 255) / 256), nf), dim3(256), 0, s, geo,
                       (const TD *)data, framePix, hc, g.nlines, g.cropping[0], g.cropping[2], knots, nk, th, tw,
                       win, out);
    return hipGetLastError();
}

}  // namespace

extern "C" ss_status ss_resample(ss_handle *h, const ss_resample_in *in, float *out) try {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    if (!in || !out || !in->geometry || !in->frames || !in->knots_x || !in->knots_y || in->n_frames < 1)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample: NULL argument or no frames");
    if (in->data_type < SS_F64 || in->data_type > SS_I16)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample: invalid data_type");
    const ss_geometry_desc &g = *in->geometry;
    int32_t hc, wc, fh;
    ss_status st = ss_geometry_size(&g, &hc, &wc, &fh);
    if (st) return ss::handle_fail(h, st, "resample: invalid geometry");
    const int th = g.template_h, tw = g.template_w, nk = g.nknots, nF = in->n_frames;
    if (th < 2 || tw < 2) return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample: template smaller than 2x2");
    if ((long long)hc * wc >= (1ll << 31) || (long long)th * tw >= (1ll << 31))
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample: frame or template too large");
    if (hipSetDevice(ss::handle_device(h)) != hipSuccess) return ss::handle_fail(h, SS_ERR_DEVICE, "hipSetDevice");
    hipStream_t s = ss::handle_stream(h);
    const int npix = hc * wc;
    std::vector<int32_t> pk(npix);
    std::vector<double> tf(npix), bx(npix), by(npix);
    if ((st = ss_geometry_init(&g, pk.data(), tf.data(), bx.data(), by.data(), nullptr, nullptr, nullptr, nullptr,
                               nullptr)))
        return ss::handle_fail(h, st, "resample: invalid geometry");
    std::vector<PixGeo> pg(npix);
    for (int i = 0; i < npix; i++) pg[i] = PixGeo{bx[i], by[i], tf[i], pk[i] - 1, 0};
    // knots of all frames, per frame [x(0..nk-1), y(0..nk-1)]
    std::vector<double> kn((size_t)2 * nk * nF);
    for (int f = 0; f < nF; f++) {
        std::memcpy(kn.data() + (size_t)f * 2 * nk, in->knots_x + (size_t)f * nk, sizeof(double) * nk);
        std::memcpy(kn.data() + (size_t)f * 2 * nk + nk, in->knots_y + (size_t)f * nk, sizeof(double) * nk);
    }
    const size_t tp = (size_t)th * tw;
    const size_t esz = in->data_type == SS_F64 ? 8 : in->data_type == SS_F32 ? 4 : 2;
    const size_t fpix = (size_t)g.nlines * g.linewidth;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // work memory: geometry + knots, then per frame of a launch the winner
    // table and the staging of host frames / host output
    const size_t fixed = al(sizeof(PixGeo) * npix) + al(kn.size() * sizeof(double));
    const size_t perFrame = 4 * tp * sizeof(int) + (in->data_on_device ? 0 : esz * fpix) +
                            (in->out_on_device ? 0 : tp * sizeof(float));
    const size_t budget = std::max<size_t>(ss::handle_scratch_held(h), (size_t)256 << 20);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(
                                                   {(size_t)nF, (budget > fixed ? budget - fixed : 0) / perFrame, 65535}));
    char *mem = (char *)ss::handle_scratch(h, fixed + al(perFrame * chunk) + 4096, &st);
    if (!mem) return st;
    PixGeo *dgeo = (PixGeo *)mem;
    double *dk = (double *)(mem + al(sizeof(PixGeo) * npix));
    int *dwin = (int *)(mem + fixed);
    char *dstage = (char *)dwin + al(4 * tp * sizeof(int) * chunk);
    char *ddata = in->data_on_device ? nullptr : dstage;
    float *dout = in->out_on_device ? nullptr : (float *)(dstage + (in->data_on_device ? 0 : al(esz * fpix * chunk)));
#define RCHK(expr)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return ss::handle_fail(h, SS_ERR_DEVICE, std::string("resample: ") + #expr + \
                                                                           ": " + hipGetErrorString(e_)); \
    } while (0)
    RCHK(hipMemcpyAsync(dgeo, pg.data(), sizeof(PixGeo) * npix, hipMemcpyHostToDevice, s));
    RCHK(hipMemcpyAsync(dk, kn.data(), kn.size() * sizeof(double), hipMemcpyHostToDevice, s));
    for (int f0 = 0; f0 < nF; f0 += chunk) {
        const int nf = std::min(chunk, nF - f0);
        const void *src = (const char *)in->frames + esz * fpix * f0;
        if (ddata) {
            RCHK(hipMemcpyAsync(ddata, src, esz * fpix * nf, hipMemcpyHostToDevice, s));
            src = ddata;
        }
        float *o = dout ? dout : out + tp * f0;
        const double *kf = dk + (size_t)2 * nk * f0;
        switch (in->data_type) {
        case SS_F64: RCHK((launch_frames<double>(s, dgeo, npix, src, fpix, g, hc, kf, nk, nf, dwin, o))); break;
        case SS_F32: RCHK((launch_frames<float>(s, dgeo, npix, src, fpix, g, hc, kf, nk, nf, dwin, o))); break;
        case SS_U16: RCHK((launch_frames<uint16_t>(s, dgeo, npix, src, fpix, g, hc, kf, nk, nf, dwin, o))); break;
        default: RCHK((launch_frames<int16_t>(s, dgeo, npix, src, fpix, g, hc, kf, nk, nf, dwin, o))); break;
        }
        if (dout) RCHK(hipMemcpyAsync(out + tp * f0, dout, sizeof(float) * tp * nf, hipMemcpyDeviceToHost, s));
    }
    RCHK(hipStreamSynchronize(s));
#undef RCHK
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}
