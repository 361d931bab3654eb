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
//
// ss_resample_noalloc runs the same two passes with no allocation at all:
// the kernels compute each pixel's geometry from the scan formulas
// (ss_scan.h) instead of reading a table, and the knots and winner tables
// live in a fixed static device array (kNoAllocBytes per device).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include <hip/hip_runtime.h>

#include "ss_host.h"
#include "ss_scan.h"
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

// Pixel geometry from the table that ss_resample uploads.
struct TableGeo {
    const PixGeo *g;
    __device__ __forceinline__ PixGeo at(int i) const { return g[i]; }
};

// Pixel geometry computed from the scan formulas (the same functions that
// build the table on the host, so the same doubles); kt: the knot times.
struct ScanGeo {
    ss::ScanParams s;
    const double *kt;
    __device__ __forceinline__ PixGeo at(int i) const {
        const int ic = i % s.Hc, jc = i / s.Hc;  // cropped (row, column), column-major
        const double t = ss::scan_cropped_time(s, ic, jc);
        // t lies in (0, 1) and the knot times run from 0 to 1, so the bin is
        // 1 .. nknots - 1; the clamp only guards the knot reads
        const int k = min(max(ss::scan_bin(kt, s.nknots, t), 1), s.nknots - 1);
        return PixGeo{ss::scan_base_x(s, jc), ss::scan_base_y(s, ic), ss::scan_tfrac(kt, k, t), k - 1, 0};
    }
};

// Displaced position of pixel q for the knots of one frame (1-based template
// coordinates, steady_scan.m:565-566).
__device__ __forceinline__ void position(const PixGeo &q, const double *kx, const double *ky, double &x,
                                         double &y) {
#pragma clang fp contract(off)
    const double tc = 1.0 - q.tf;
    x = q.bx + kx[q.k] * tc + kx[q.k + 1] * q.tf;
    y = q.by + ky[q.k] * tc + ky[q.k + 1] * q.tf;
}

// Knots of the launch's frames: frame f's x knots at kx + f * stride, its y
// knots at ky + f * stride.
struct FrameKnots {
    const double *kx, *ky;
    size_t stride;
};

template <typename Geo>
__global__ __launch_bounds__(256) void k_splat_winners(Geo geo, int npix, FrameKnots kn, int th, int tw,
                                                       int *__restrict__ win) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const int f = blockIdx.y;
    const double *kx = kn.kx + (size_t)f * kn.stride, *ky = kn.ky + (size_t)f * kn.stride;
    double x, y;
    position(geo.at(i), kx, ky, x, y);
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

template <typename TD, typename Geo>
__global__ __launch_bounds__(256) void k_splat_gather(Geo geo, const TD *__restrict__ data, size_t framePix, int hc,
                                                      int nl, int cl, int cb, FrameKnots kn, int th, int tw,
                                                      const int *__restrict__ win, float *__restrict__ out) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tp = (size_t)th * tw;
    if (t >= (int)tp) return;
    const int f = blockIdx.y;
    const double *kx = kn.kx + (size_t)f * kn.stride, *ky = kn.ky + (size_t)f * kn.stride;
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
        position(geo.at(i), kx, ky, x, y);
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

template <typename TD, typename Geo>
hipError_t launch_frames(hipStream_t s, const Geo &geo, int npix, const void *data, size_t framePix,
                         const ss_geometry_desc &g, int hc, const FrameKnots &kn, int nf, int *win, float *out) {
    const int th = g.template_h, tw = g.template_w;
    const size_t tp = (size_t)th * tw, nwin = (size_t)nf * 4 * tp;
    const int fillGrid = (int)std::min<size_t>((nwin + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_minus1, dim3(fillGrid), dim3(256), 0, s, win, nwin);
    hipLaunchKernelGGL((k_splat_winners<Geo>), dim3((npix + 255) / 256, nf), dim3(256), 0, s, geo, npix, kn, th, tw,
                       win);
    hipLaunchKernelGGL((k_splat_gather<TD, Geo>), dim3((unsigned)((tp + 255) / 256), nf), dim3(256), 0, s, geo,
                       (const TD *)data, framePix, hc, g.nlines, g.cropping[0], g.cropping[2], kn, th, tw, win, out);
    return hipGetLastError();
}

template <typename Geo>
hipError_t launch_typed(ss_dtype type, hipStream_t s, const Geo &geo, int npix, const void *data, size_t framePix,
                        const ss_geometry_desc &g, int hc, const FrameKnots &kn, int nf, int *win, float *out) {
    switch (type) {
    case SS_F64: return launch_frames<double>(s, geo, npix, data, framePix, g, hc, kn, nf, win, out);
    case SS_F32: return launch_frames<float>(s, geo, npix, data, framePix, g, hc, kn, nf, win, out);
    case SS_U16: return launch_frames<uint16_t>(s, geo, npix, data, framePix, g, hc, kn, nf, win, out);
    default: return launch_frames<int16_t>(s, geo, npix, data, framePix, g, hc, kn, nf, win, out);
    }
}

// ---------------------------------------------------------------------------
// the allocation-free variant's work area: one static array per device
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
constexpr int kNoAllocMaxKnots = 4096;
constexpr int kMaxDevices = 64;

// In host code `scratch` names a host-side stand-in, not device memory: only
// the address that hipGetSymbolAddress returns for a device may be used.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratchLock;
void *g_scratchPtr[kMaxDevices] = {};
hipError_t g_scratchErr[kMaxDevices] = {};
bool g_scratchDone[kMaxDevices] = {};

}  // namespace

namespace ss {

// Looks up this device's copy of the static work area, once per device, with
// the device current (called from ss_create).
hipError_t resample_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratchLock);
    if (!g_scratchDone[device]) {
        void *ptr = nullptr;
        g_scratchErr[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratchPtr[device] = g_scratchErr[device] == hipSuccess ? ptr : nullptr;
        g_scratchDone[device] = true;
    }
    return g_scratchErr[device];
}

}  // namespace ss

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
        const FrameKnots fk{kf, kf + nk, (size_t)2 * nk};
        RCHK((launch_typed(in->data_type, s, TableGeo{dgeo}, npix, src, fpix, g, hc, fk, nf, dwin, o)));
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

extern "C" ss_status ss_resample_noalloc(ss_handle *h, const ss_resample_in *in, float *out) try {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    if (!in || !out || !in->geometry || !in->frames || !in->knots_x || !in->knots_y || in->n_frames < 1)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample_noalloc: NULL argument or no frames");
    if (in->data_type < SS_F64 || in->data_type > SS_I16)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample_noalloc: invalid data_type");
    if (!in->data_on_device || !in->out_on_device)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT,
                               "resample_noalloc: frames and output must be device memory (data_on_device and "
                               "out_on_device set)");
    const ss_geometry_desc &g = *in->geometry;
    if (g.nknots < 2 || g.nknots > kNoAllocMaxKnots)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample_noalloc: nknots must be 2 .. 4096");
    ss::ScanParams sp;
    double kt[kNoAllocMaxKnots];
    // the same validity rules as ss_geometry_size: vertical cropping mis-sizes
    // the default basey (steady_scan.m:799, 822-824)
    if (!ss::scan_setup(g, sp, kt) || sp.cb + sp.ct != 0)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample_noalloc: invalid geometry");
    const int th = g.template_h, tw = g.template_w, nk = g.nknots, nF = in->n_frames;
    if (th < 2 || tw < 2)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample_noalloc: template smaller than 2x2");
    if ((long long)sp.Hc * sp.Wc >= (1ll << 31) || (long long)th * tw >= (1ll << 31))
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT, "resample_noalloc: frame or template too large");
    const int dev = ss::handle_device(h);
    if (hipSetDevice(dev) != hipSuccess) return ss::handle_fail(h, SS_ERR_DEVICE, "hipSetDevice");
    const hipError_t se = ss::resample_scratch_init(dev);  // done by ss_create; cached
    char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratchLock);
        base = se == hipSuccess ? (char *)g_scratchPtr[dev] : nullptr;
    }
    if (!base)
        return ss::handle_fail(h, SS_ERR_DEVICE, std::string("resample_noalloc: static work area unavailable: ") +
                                                     hipGetErrorString(se));
    // work area: [knot times][x knots of a launch][y knots][winner tables]
    const size_t tp = (size_t)th * tw;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t fixed = al(sizeof(double) * nk);
    auto need = [&](size_t nf) { return fixed + 2 * al(sizeof(double) * nk * nf) + 4 * tp * sizeof(int) * nf; };
    if (need(1) > kNoAllocBytes)
        return ss::handle_fail(h, SS_ERR_INVALID_ARGUMENT,
                               "resample_noalloc: one frame needs " + std::to_string(need(1)) +
                                   " bytes of work memory; the static work area holds " +
                                   std::to_string(kNoAllocBytes) + " bytes (template_h x template_w up to about 4.1 "
                                                                   "million pixels)");
    size_t chunk = std::min<size_t>({(size_t)nF, (kNoAllocBytes - fixed) / (4 * tp * sizeof(int) + 2 * sizeof(double) * nk),
                                     65535});
    while (chunk > 1 && need(chunk) > kNoAllocBytes) chunk--;
    double *dkt = (double *)base;
    double *dkx = (double *)(base + fixed);
    double *dky = (double *)(base + fixed + al(sizeof(double) * nk * chunk));
    int *dwin = (int *)(base + fixed + 2 * al(sizeof(double) * nk * chunk));
    hipStream_t s = ss::handle_stream(h);
    const size_t esz = in->data_type == SS_F64 ? 8 : in->data_type == SS_F32 ? 4 : 2;
    const size_t fpix = (size_t)g.nlines * g.linewidth;
    const int npix = sp.Hc * sp.Wc;
#define NCHK(expr)                                                                                              \
    do {                                                                                                        \
        hipError_t e_ = (expr);                                                                                 \
        if (e_ != hipSuccess) {                                                                                 \
            (void)hipStreamSynchronize(s);                                                                      \
            return ss::handle_fail(h, SS_ERR_DEVICE,                                                            \
                                   std::string("resample_noalloc: ") + #expr + ": " + hipGetErrorString(e_)); \
        }                                                                                                       \
    } while (0)
    NCHK(hipMemcpyAsync(dkt, kt, sizeof(double) * nk, hipMemcpyHostToDevice, s));
    const ScanGeo geo{sp, dkt};
    for (int f0 = 0; f0 < nF; f0 += (int)chunk) {
        const int nf = std::min((int)chunk, nF - f0);
        NCHK(hipMemcpyAsync(dkx, in->knots_x + (size_t)nk * f0, sizeof(double) * nk * nf, hipMemcpyHostToDevice, s));
        NCHK(hipMemcpyAsync(dky, in->knots_y + (size_t)nk * f0, sizeof(double) * nk * nf, hipMemcpyHostToDevice, s));
        const void *src = (const char *)in->frames + esz * fpix * f0;
        const FrameKnots fk{dkx, dky, (size_t)nk};
        NCHK((launch_typed(in->data_type, s, geo, npix, src, fpix, g, sp.Hc, fk, nf, dwin, out + tp * f0)));
    }
    NCHK(hipStreamSynchronize(s));
#undef NCHK
    return SS_OK;
} catch (...) {
    return SS_ERR_DEVICE;
}
