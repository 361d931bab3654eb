// ss_resample.hip -- corrected frames: resample_frame + pushgrid
// (steady_scan.m:562-599) on the GPU.
//
// pushgrid splats every in-bounds data pixel onto the 4 template pixels of
// its cell with bilinear weights, through MATLAB indexed assignments
// z(ind) = z(ind) + v.  With repeated indices such an assignment keeps the
// value of the LAST occurrence, so for each template pixel and each of the
// four corner offsets exactly one data pixel contributes: the one with the
// largest (column-major) pixel index.  The GPU form is therefore
//   pass 1: per corner offset, atomicMax of the pixel index into the target
//           (an integer maximum: order-independent and deterministic);
//   pass 2: per template pixel, gather the <= 4 winners and accumulate in
//           single precision in the statement order of pushgrid.
// Positions are computed in FP64 without contraction, as MATLAB does.
#include <cmath>
#include <new>
#include <cstring>
#include <vector>

#include <hip/hip_runtime.h>

#include "steadyscan.h"

namespace {

struct PixGeo {
    double bx, by, tf;
    int k;  // prevknotind - 1
    int pad;
};

template <typename TD>
__global__ __launch_bounds__(256) void k_splat_winners(const PixGeo *__restrict__ geo,
                                                       int npix, const double *__restrict__ kx,
                                                       const double *__restrict__ ky, int th,
                                                       int tw, int *__restrict__ win) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const PixGeo q = geo[i];
    const double tc = 1.0 - q.tf;
    const double x = q.bx + kx[q.k] * tc + kx[q.k + 1] * q.tf;
    const double y = q.by + ky[q.k] * tc + ky[q.k + 1] * q.tf;
    if (!(x >= 1.0 && x < (double)tw && y >= 1.0 && y < (double)th)) return;
    const int xi = (int)floor(x), yi = (int)floor(y);
    const int m0 = (yi - 1) + th * (xi - 1);
    const size_t tp = (size_t)th * tw;
    atomicMax(&win[m0], i);
    atomicMax(&win[tp + m0 + 1], i);
    atomicMax(&win[2 * tp + m0 + th], i);
    atomicMax(&win[3 * tp + m0 + th + 1], i);
}

template <typename TD>
__global__ __launch_bounds__(256) void k_splat_gather(const PixGeo *__restrict__ geo,
                                                      const TD *__restrict__ data, int hc, int nl,
                                                      int cl, int cb, const double *__restrict__ kx,
                                                      const double *__restrict__ ky, int th, int tw,
                                                      const int *__restrict__ win,
                                                      float *__restrict__ out) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tp = (size_t)th * tw;
    if (t >= (int)tp) return;
    float z = 0.f, n = 0.f;
    double w[4];
    double d[4];
    bool has[4];
    for (int s = 0; s < 4; s++) {
        const int i = win[s * tp + t];
        has[s] = i >= 0;
        w[s] = 0.0;
        d[s] = 0.0;
        if (!has[s]) continue;
        const PixGeo q = geo[i];
        const double tc = 1.0 - q.tf;
        const double x = q.bx + kx[q.k] * tc + kx[q.k + 1] * q.tf;
        const double y = q.by + ky[q.k] * tc + ky[q.k + 1] * q.tf;
        const double xf = x - floor(x), yf = y - floor(y);
        // corner offsets 0, 1, th, th+1 <-> weights (steady_scan.m:589-592)
        w[s] = s == 0 ? (1 - xf) * (1 - yf) : s == 1 ? (1 - xf) * yf : s == 2 ? xf * (1 - yf) : xf * yf;
        const int ic = i % hc, jc = i / hc;  // cropped (row, column)
        d[s] = (double)data[(size_t)(jc + cl) * nl + (ic + cb)];
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
    out[t] = z / n;
}

}  // namespace

extern "C" ss_status ss_resample(ss_handle *h, const ss_resample_in *in, float *out) try {
    (void)h;
    if (!in || !out || !in->geometry || !in->frames || !in->knots_x || !in->knots_y || in->n_frames < 1)
        return SS_ERR_INVALID_ARGUMENT;
    const ss_geometry_desc &g = *in->geometry;
    int32_t hc, wc, fh;
    ss_status st = ss_geometry_size(&g, &hc, &wc, &fh);
    if (st) return st;
    if (in->data_on_device && in->data_type == SS_F64) return SS_ERR_INVALID_ARGUMENT;
    const int th = g.template_h, tw = g.template_w, nk = g.nknots;
    if (th < 2 || tw < 2) return SS_ERR_INVALID_ARGUMENT;
    const int npix = hc * wc;
    std::vector<int32_t> pk(npix);
    std::vector<double> tf(npix), bx(npix), by(npix);
    if ((st = ss_geometry_init(&g, pk.data(), tf.data(), bx.data(), by.data(), nullptr, nullptr,
                               nullptr, nullptr, nullptr)))
        return st;
    std::vector<PixGeo> pg(npix);
    for (int i = 0; i < npix; i++) pg[i] = PixGeo{bx[i], by[i], tf[i], pk[i] - 1, 0};
    const size_t tp = (size_t)th * tw;
    const size_t esz = in->data_type == SS_F64 ? 8 : in->data_type == SS_F32 ? 4 : 2;
    const size_t fpix = (size_t)g.nlines * g.linewidth;
    PixGeo *dgeo = nullptr;
    double *dk = nullptr;
    int *dwin = nullptr;
    float *dout = nullptr;
    void *ddata = nullptr;
    hipStream_t s = nullptr;
    ss_status rc = SS_OK;
#define RCHK(expr)                                  \
    do {                                            \
        if ((expr) != hipSuccess) {                 \
            rc = SS_ERR_DEVICE;                     \
            goto done;                              \
        }                                           \
    } while (0)
    RCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    RCHK(hipMalloc(&dgeo, sizeof(PixGeo) * npix));
    RCHK(hipMalloc(&dk, sizeof(double) * 2 * nk));
    RCHK(hipMalloc(&dwin, sizeof(int) * 4 * tp));
    RCHK(hipMalloc(&dout, sizeof(float) * tp));
    if (!in->data_on_device) RCHK(hipMalloc(&ddata, esz * fpix));
    RCHK(hipMemcpyAsync(dgeo, pg.data(), sizeof(PixGeo) * npix, hipMemcpyHostToDevice, s));
    for (int f = 0; f < in->n_frames; f++) {
        const char *src = (const char *)in->frames + esz * fpix * f;
        const void *fd = src;
        if (!in->data_on_device) {
            RCHK(hipMemcpyAsync(ddata, src, esz * fpix, hipMemcpyHostToDevice, s));
            fd = ddata;
        }
        RCHK(hipMemcpyAsync(dk, in->knots_x + (size_t)f * nk, sizeof(double) * nk, hipMemcpyHostToDevice, s));
        RCHK(hipMemcpyAsync(dk + nk, in->knots_y + (size_t)f * nk, sizeof(double) * nk, hipMemcpyHostToDevice, s));
        RCHK(hipMemsetAsync(dwin, 0xFF, sizeof(int) * 4 * tp, s));
        const int gb = (npix + 255) / 256, gt = (int)((tp + 255) / 256);
        switch (in->data_type) {
        case SS_F64:
            hipLaunchKernelGGL(k_splat_winners<double>, dim3(gb), dim3(256), 0, s, dgeo, npix, dk, dk + nk, th, tw, dwin);
            hipLaunchKernelGGL(k_splat_gather<double>, dim3(gt), dim3(256), 0, s, dgeo, (const double *)fd, hc, g.nlines, g.cropping[0], g.cropping[2], dk, dk + nk, th, tw, dwin, dout);
            break;
        case SS_F32:
            hipLaunchKernelGGL(k_splat_winners<float>, dim3(gb), dim3(256), 0, s, dgeo, npix, dk, dk + nk, th, tw, dwin);
            hipLaunchKernelGGL(k_splat_gather<float>, dim3(gt), dim3(256), 0, s, dgeo, (const float *)fd, hc, g.nlines, g.cropping[0], g.cropping[2], dk, dk + nk, th, tw, dwin, dout);
            break;
        case SS_U16:
            hipLaunchKernelGGL(k_splat_winners<uint16_t>, dim3(gb), dim3(256), 0, s, dgeo, npix, dk, dk + nk, th, tw, dwin);
            hipLaunchKernelGGL(k_splat_gather<uint16_t>, dim3(gt), dim3(256), 0, s, dgeo, (const uint16_t *)fd, hc, g.nlines, g.cropping[0], g.cropping[2], dk, dk + nk, th, tw, dwin, dout);
            break;
        default:
            hipLaunchKernelGGL(k_splat_winners<int16_t>, dim3(gb), dim3(256), 0, s, dgeo, npix, dk, dk + nk, th, tw, dwin);
            hipLaunchKernelGGL(k_splat_gather<int16_t>, dim3(gt), dim3(256), 0, s, dgeo, (const int16_t *)fd, hc, g.nlines, g.cropping[0], g.cropping[2], dk, dk + nk, th, tw, dwin, dout);
            break;
        }
        RCHK(hipGetLastError());
        RCHK(hipMemcpyAsync(out + tp * f, dout, sizeof(float) * tp, hipMemcpyDeviceToHost, s));
        RCHK(hipStreamSynchronize(s));
    }
done:
    if (s) (void)hipStreamSynchronize(s);
    if (dgeo) (void)hipFree(dgeo);
    if (dk) (void)hipFree(dk);
    if (dwin) (void)hipFree(dwin);
    if (dout) (void)hipFree(dout);
    if (ddata) (void)hipFree(ddata);
    if (s) (void)hipStreamDestroy(s);
#undef RCHK
    return rc;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}
