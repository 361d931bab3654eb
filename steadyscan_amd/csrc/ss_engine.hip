// ss_engine.hip -- the C ABI of include/steadyscan.h: handle, resident
// inputs, and the host loop of one alignment group (Task<T>::alignFrames,
// task/task.cpp:259-412, with MainIteration<T>::execute,
// task/mainiteration/mainiteration.cpp:61-119).
//
// Per half-iteration the device runs: assembly (1 kernel), solve + filter +
// expand (1 kernel), then fused line-search rounds (round + decide kernels)
// until every frame has resolved its sequential backtracking.  The host only
// reads a 24-byte status block per line-search round and per iteration.
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ss_device.h"
#include "steadyscan.h"

namespace ss {
hipError_t launch_assemble(const GroupDev &g, hipStream_t s);
hipError_t launch_solve(const GroupDev &g, int shift, int filterKind, double strength,
                        double lambda, hipStream_t s);
hipError_t launch_ls_round(const GroupDev &g, int K, hipStream_t s);
hipError_t launch_ls_decide(const GroupDev &g, int shift, int K, double thr, hipStream_t s);
hipError_t launch_corr(const GroupDev &g, hipStream_t s);
hipError_t launch_final(const GroupDev &g, hipStream_t s);
size_t solve_workspace_bytes(int n);
}  // namespace ss

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

struct ResidentTemplate {
    int th = 0, tw = 0;
    float *tmpl = nullptr, *xg = nullptr, *yg = nullptr;
};

struct ResidentGeometry {
    int frameH = 0, nb = 0, fracUniform = 1;
    float4 *geo = nullptr;
    float *fnext = nullptr;
};

constexpr int kMaxRounds = 512;  // 4096 candidates at K = 8

}  // namespace

struct Profile {
    bool on = false;
    std::vector<hipEvent_t> pool;
    struct Pending {
        int cls;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Pending> pending;
    int64_t launches[SS_K_COUNT] = {};
    double ms[SS_K_COUNT] = {};
    double bytes[SS_K_COUNT] = {};
};

struct ss_handle {
    int device = 0;
    Profile prof;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, DevBuf> bufs;
    std::map<int, ResidentTemplate> templates;
    std::map<int, ResidentGeometry> geometries;
    int next_id = 1;
    ss::Status *hstatus = nullptr;  // pinned
    ss::Status *dstatus = nullptr;
};

namespace {

ss_status fail(ss_handle *h, ss_status st, const std::string &msg) {
    if (h) h->err = msg;
    return st;
}

hipEvent_t prof_event(ss_handle *h) {
    if (!h->prof.pool.empty()) {
        hipEvent_t e = h->prof.pool.back();
        h->prof.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Folds completed event pairs into the counters (synchronises the stream).
void prof_collect(ss_handle *h) {
    if (h->prof.pending.empty()) return;
    hipError_t e = hipStreamSynchronize(h->stream);
    (void)e;
    for (auto &p : h->prof.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            h->prof.ms[p.cls] += ms;
            h->prof.launches[p.cls] += 1;
            h->prof.bytes[p.cls] += p.bytes;
        }
        h->prof.pool.push_back(p.a);
        h->prof.pool.push_back(p.b);
    }
    h->prof.pending.clear();
}

// Runs one launcher, bracketed by events when profiling is on.
template <typename F>
hipError_t timed(ss_handle *h, int cls, double bytes, F &&launch) {
    if (!h->prof.on) return launch();
    hipEvent_t a = prof_event(h), b = prof_event(h);
    if (!a || !b) return launch();
    hipError_t e = hipEventRecord(a, h->stream);
    if (e != hipSuccess) return e;
    e = launch();
    if (e != hipSuccess) return e;
    e = hipEventRecord(b, h->stream);
    h->prof.pending.push_back({cls, a, b, bytes});
    return e;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(h, e_ == hipErrorOutOfMemory ? SS_ERR_OUT_OF_MEMORY : SS_ERR_DEVICE, \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

// Grow-only named device buffer owned by the handle.
ss_status buffer(ss_handle *h, const char *name, size_t bytes, void **out) {
    DevBuf &b = h->bufs[name];
    if (b.n < bytes) {
        if (b.p) {
            hipError_t e = hipFree(b.p);
            (void)e;
            b.p = nullptr;
            b.n = 0;
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e != hipSuccess)
            return fail(h, SS_ERR_OUT_OF_MEMORY, std::string("hipMalloc(") + name + ") failed: " + hipGetErrorString(e));
        b.p = p;
        b.n = bytes;
    }
    *out = b.p;
    return SS_OK;
}

template <typename T>
ss_status buffer_t(ss_handle *h, const char *name, size_t count, T **out) {
    void *p = nullptr;
    ss_status st = buffer(h, name, count * sizeof(T), &p);
    *out = (T *)p;
    return st;
}

// Template and gradients to row-major float (mexconverter.cpp:223-285 makes
// the same transposition for RowMajor targets).
ss_status upload_template(ss_handle *h, int th, int tw, const double *t, const double *xg,
                          const double *yg, ResidentTemplate &rt, bool resident) {
    if (th < 2 || tw < 2 || !t || !xg || !yg) return fail(h, SS_ERR_INVALID_ARGUMENT, "template: bad size or NULL array");
    const size_t n = (size_t)th * tw;
    std::vector<float> a(n), bx(n), by(n);
    for (int iy = 0; iy < th; iy++)
        for (int ix = 0; ix < tw; ix++) {
            const size_t cm = (size_t)ix * th + iy, rm = (size_t)iy * tw + ix;
            a[rm] = (float)t[cm];
            bx[rm] = (float)xg[cm];
            by[rm] = (float)yg[cm];
        }
    rt.th = th;
    rt.tw = tw;
    if (resident) {
        HIPCHK(hipMalloc(&rt.tmpl, n * sizeof(float)));
        HIPCHK(hipMalloc(&rt.xg, n * sizeof(float)));
        HIPCHK(hipMalloc(&rt.yg, n * sizeof(float)));
    } else {
        ss_status st;
        if ((st = buffer_t(h, "tmpl", n, &rt.tmpl)) || (st = buffer_t(h, "xg", n, &rt.xg)) ||
            (st = buffer_t(h, "yg", n, &rt.yg)))
            return st;
    }
    HIPCHK(hipMemcpyAsync(rt.tmpl, a.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(rt.xg, bx.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(rt.yg, by.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SS_OK;
}

// Geometry block matrices to {base_x, base_y, frac, base_mask} float4 per
// pixel, column-contiguous; frac of the next block column only where the
// frac columns differ (the diagonal term of task_filldiagonals.cu:66-103
// reads column idx % frameW).
ss_status upload_geometry(ss_handle *h, int fh, int nb, const double *frac, const double *bx,
                          const double *by, const uint8_t *bm, ResidentGeometry &rg,
                          bool resident) {
    if (fh < 1 || nb < 1 || !frac || !bx || !by || !bm)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "geometry: bad size or NULL array");
    const size_t n = (size_t)fh * nb;
    std::vector<float4> g(n);
    bool uniform = true;
    for (int b = 0; b < nb; b++)
        for (int r = 0; r < fh; r++) {
            const size_t i = (size_t)b * fh + r;
            g[i] = make_float4((float)bx[i], (float)by[i], (float)frac[i], bm[i] ? 1.f : 0.f);
            if (frac[i] != frac[r]) uniform = false;
        }
    rg.frameH = fh;
    rg.nb = nb;
    rg.fracUniform = uniform ? 1 : 0;
    std::vector<float> fn;
    if (!uniform) {
        fn.resize(n);
        for (int b = 0; b < nb; b++)
            for (int r = 0; r < fh; r++) fn[(size_t)b * fh + r] = (float)frac[(size_t)((b + 1) % nb) * fh + r];
    }
    if (resident) {
        HIPCHK(hipMalloc(&rg.geo, n * sizeof(float4)));
        if (!uniform) HIPCHK(hipMalloc(&rg.fnext, n * sizeof(float)));
    } else {
        ss_status st;
        if ((st = buffer_t(h, "geo", n, &rg.geo))) return st;
        if (!uniform && (st = buffer_t(h, "fnext", n, &rg.fnext))) return st;
    }
    HIPCHK(hipMemcpyAsync(rg.geo, g.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    if (!uniform)
        HIPCHK(hipMemcpyAsync(rg.fnext, fn.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SS_OK;
}

// ParameterOrganizer::organize2D1 (parameterorganizer.cpp:30-55): shared
// knots of consecutive frames are averaged.
void knots_in(const double *px, const double *py, int ld, int nF, int nb, double *p) {
    const int N = nF * nb + 1;
    for (int f = 0; f < nF; f++) {
        const double *fx = px + (size_t)f * ld, *fy = py + (size_t)f * ld;
        for (int i = (f != 0); i < nb; i++) {
            p[f * nb + i] = fx[i];
            p[f * nb + i + N] = fy[i];
        }
        const int j = f * nb + nb;
        if (f + 1 < nF) {
            p[j] = (fx[nb] + px[(size_t)(f + 1) * ld]) / 2;
            p[j + N] = (fy[nb] + py[(size_t)(f + 1) * ld]) / 2;
        } else {
            p[j] = fx[nb];
            p[j + N] = fy[nb];
        }
    }
}

// ParameterOrganizer::organize1D2 (parameterorganizer.cpp:79-99): a frame's
// last output knot repeats its second-to-last knot.
void knots_out(const double *p, int nF, int nb, double *px, double *py) {
    const int N = nF * nb + 1;
    for (int f = 0; f < nF; f++) {
        for (int i = 0; i < nb; i++) {
            px[f * (nb + 1) + i] = p[f * nb + i];
            py[f * (nb + 1) + i] = p[f * nb + i + N];
        }
        px[f * (nb + 1) + nb] = p[f * nb + nb - 1];
        py[f * (nb + 1) + nb] = p[f * nb + nb - 1 + N];
    }
}

ss_status validate(ss_handle *h, const ss_group_in *in, const ss_settings *s) {
    if (!in || !s) return fail(h, SS_ERR_INVALID_ARGUMENT, "not enough input arguments");
    if (in->frame_h < 1 || in->n_blocks < 1 || in->n_frames < 1)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "frame_h, n_blocks and n_frames must be positive");
    if (in->group_size > 1024)  // mexgateway.cpp:186-190
        return fail(h, SS_ERR_INVALID_ARGUMENT, "groupsize is too big\nmax groupsize: 1024");
    if (in->group_size < 1) return fail(h, SS_ERR_INVALID_ARGUMENT, "group_size must be positive");
    if (!in->data || !in->init_x || !in->init_y || !in->min_points)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "NULL input array");
    if (in->data_type < SS_F64 || in->data_type > SS_I16)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "invalid data_type");
    if (in->data_on_device && in->data_type == SS_F64)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "device-resident data must be f32, u16 or i16");
    if (s->max_iterations < 1) return fail(h, SS_ERR_INVALID_ARGUMENT, "max_iterations must be >= 1");
    if (s->solver_id < 0 || s->solver_id > 2)  // task.cpp:74-88
        return fail(h, SS_ERR_INVALID_ARGUMENT, "invalid solver id\n0:CuSparse solver\n1:CuSolver solver\n2:Eigen solver");
    const int fsel = s->max_filter_iterations != 0 ? s->filter_id : 0;  // task.cpp:54
    if (fsel < 0 || fsel > 2)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "invalid filter ID\n0: disable filter\n1: smearingfiter\n2:rangefilter");
    if (s->precision != 'S' && s->precision != 'D')
        return fail(h, SS_ERR_INVALID_ARGUMENT, "invalid precision type! use 'S' or 'D'");
    const long long npix = (long long)in->frame_h * in->n_blocks * in->n_frames;
    if (npix > (1ll << 31)) return fail(h, SS_ERR_INVALID_ARGUMENT, "group too large");
    return SS_OK;
}

struct Loop {
    ss_handle *h;
    ss::GroupDev g;
    const ss_settings *s;
    int K;
    double frameBytes = 0;  // algorithmic bytes of one frame (pixels x element size)
    // filter lifecycle (task.cpp:54-70, mainiteration.cpp:80-115, task.cpp:313-325)
    int solverFilter = 0, filterID = 0, nFilterIterations = 0;
    int halfIterations = 0, rounds = 0;

    ss_status read_status() {
        HIPCHK(hipMemcpyAsync(h->hstatus, h->dstatus, sizeof(ss::Status), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return SS_OK;
    }

    ss_status linesearch(int shift, int *improved) {
        int active = g.nF;
        for (int round = 0;; round++) {
            if (round >= kMaxRounds)
                return fail(h, SS_ERR_NUMERICAL, "line search did not reach the movement threshold (non-finite step?)");
            HIPCHK(timed(h, SS_K_LS_ROUND, (double)active * frameBytes, [&] { return ss::launch_ls_round(g, K, h->stream); }));
            HIPCHK(timed(h, SS_K_LS_DECIDE, 0.0, [&] { return ss::launch_ls_decide(g, shift, K, s->move_threshold, h->stream); }));
            rounds++;
            ss_status st = read_status();
            if (st) return st;
            active = h->hstatus->active;
            if (!h->hstatus->need_more) break;
        }
        *improved = h->hstatus->improved;
        return SS_OK;
    }

    // MainIteration::execute (mainiteration.cpp:61-119)
    ss_status main_iteration(int iteration, int *improved_out) {
        int improved = 1;
        for (int shift = 0; shift < 2; shift++) {
            HIPCHK(timed(h, SS_K_ASSEMBLE, frameBytes * g.nF, [&] { return ss::launch_assemble(g, h->stream); }));
            if (nFilterIterations != 0 && iteration > nFilterIterations) {
                solverFilter = 0;
                filterID = 0;
                nFilterIterations = 0;
            }
            HIPCHK(timed(h, SS_K_SOLVE, 0.0, [&] { return ss::launch_solve(g, shift, solverFilter, s->filter_strength, s->lambda, h->stream); }));
            int imp = 0;
            ss_status st = linesearch(shift, &imp);
            if (st) return st;
            improved &= imp;
            halfIterations++;
        }
        if (!improved) {
            if (nFilterIterations != 0) {
                solverFilter = 0;
                nFilterIterations = 0;
                return main_iteration(iteration, improved_out);
            }
            *improved_out = 0;
            return SS_OK;
        }
        *improved_out = 1;
        return SS_OK;
    }
};

ss_status run_group(ss_handle *h, const ss_group_in *in, const ss_settings *s, ss_group_out *out) {
    ss_status st = validate(h, in, s);
    if (st) return st;
    if (s->solver_id == 0)
        return fail(h, SS_ERR_NOT_AVAILABLE, "solver_id 0 (ILU(0)) is not available in this build");
    const auto t0 = std::chrono::steady_clock::now();
    const int fh = in->frame_h, nb = in->n_blocks, nF = in->n_frames;
    const int fframeW = nb * nF, N = fframeW + 1;
    // inputs
    ResidentTemplate rt;
    if (in->template_id > 0) {
        auto it = h->templates.find(in->template_id);
        if (it == h->templates.end()) return fail(h, SS_ERR_INVALID_ARGUMENT, "unknown template_id");
        rt = it->second;
    } else {
        if ((st = upload_template(h, in->template_h, in->template_w, in->template_image, in->xgrad,
                                  in->ygrad, rt, false)))
            return st;
    }
    ResidentGeometry rg;
    if (in->geometry_id > 0) {
        auto it = h->geometries.find(in->geometry_id);
        if (it == h->geometries.end()) return fail(h, SS_ERR_INVALID_ARGUMENT, "unknown geometry_id");
        rg = it->second;
        if (rg.frameH != fh || rg.nb != nb)
            return fail(h, SS_ERR_INVALID_ARGUMENT, "resident geometry does not match frame_h x n_blocks");
    } else {
        if ((st = upload_geometry(h, fh, nb, in->frac, in->base_x, in->base_y, in->base_mask, rg, false)))
            return st;
    }
    const size_t npix = (size_t)fh * fframeW;
    const void *dimg = nullptr;
    int imgType = ss::IMG_F32;
    if (in->data_type == SS_U16) imgType = ss::IMG_U16;
    else if (in->data_type == SS_I16) imgType = ss::IMG_I16;
    if (in->data_on_device) {
        dimg = in->data;
    } else {
        void *dp = nullptr;
        const size_t esz = (imgType == ss::IMG_F32) ? 4 : 2;
        if ((st = buffer(h, "img", npix * esz, &dp))) return st;
        if (in->data_type == SS_F64) {
            std::vector<float> tmp(npix);
            const double *src = (const double *)in->data;
            for (size_t i = 0; i < npix; i++) tmp[i] = (float)src[i];
            HIPCHK(hipMemcpyAsync(dp, tmp.data(), npix * 4, hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        } else {
            HIPCHK(hipMemcpyAsync(dp, in->data, npix * esz, hipMemcpyHostToDevice, h->stream));
        }
        dimg = dp;
    }
    int *minPts = nullptr;
    std::vector<int> mp(in->min_points, in->min_points + nb);
    if ((st = buffer_t(h, "minpts", (size_t)nb, &minPts))) return st;
    HIPCHK(hipMemcpyAsync(minPts, mp.data(), nb * sizeof(int), hipMemcpyHostToDevice, h->stream));
    // state (task_allocate.cu:124-184)
    ss::GroupDev g;
    std::memset(&g, 0, sizeof(g));
    g.frameH = fh;
    g.nb = nb;
    g.nF = nF;
    g.fframeW = fframeW;
    g.N = N;
    g.th = rt.th;
    g.tw = rt.tw;
    g.groupSize = in->group_size;
    g.fracUniform = rg.fracUniform;
    const double eps = s->precision == 'S' ? (double)FLT_EPSILON : DBL_EPSILON;
    g.xmax = (float)((double)(rt.tw - 1) - eps);
    g.ymax = (float)((double)(rt.th - 1) - eps);
    g.rmult = s->linesearch_reduction_multiplier;
    g.img = dimg;
    g.imgType = imgType;
    g.geo = rg.geo;
    g.fnext = rg.fnext;
    g.tmpl = rt.tmpl;
    g.xg = rt.xg;
    g.yg = rt.yg;
    g.minPts = minPts;
    const size_t P = 2 * (size_t)N;
    if ((st = buffer_t(h, "p", P, &g.p)) || (st = buffer_t(h, "pLast", P, &g.pLast)) ||
        (st = buffer_t(h, "p0", P, &g.p0)) || (st = buffer_t(h, "d0", P, &g.d0)) ||
        (st = buffer_t(h, "deltaP", P, &g.deltaP)) ||
        (st = buffer_t(h, "colsum", (size_t)fframeW * ss::CS_N, &g.colsum)) ||
        (st = buffer_t(h, "lssum", (size_t)fframeW * 2 * ss::kMaxCand, &g.lssum)) ||
        (st = buffer_t(h, "errval", 2 * (size_t)nF, &g.errval)) ||
        (st = buffer_t(h, "corrsum", (size_t)fframeW * ss::CR_N, &g.corrsum)) ||
        (st = buffer_t(h, "finalsum", (size_t)fframeW * ss::FS_N, &g.finalsum)) ||
        (st = buffer(h, "ws", ss::solve_workspace_bytes(fframeW), (void **)&g.ws)) ||
        (st = buffer_t(h, "used", (size_t)N, &g.used)) ||
        (st = buffer_t(h, "lsActive", (size_t)nF, &g.lsActive)) ||
        (st = buffer_t(h, "lsNextK", (size_t)nF, &g.lsNextK)) ||
        (st = buffer_t(h, "lsMk", (size_t)nF, &g.lsMk)) ||
        (st = buffer_t(h, "lsDk", (size_t)nF, &g.lsDk)) ||
        (st = buffer_t(h, "resid", (size_t)nF, &g.resid)) ||
        (st = buffer_t(h, "fcorr", (size_t)nF, &g.fcorr)))
        return st;
    if (out->mask && (st = buffer_t(h, "maskOut", npix, &g.maskOut))) return st;
    g.status = h->dstatus;
    std::vector<double> hp(P, 0.0), ev(2 * (size_t)nF, 1e6);
    knots_in(in->init_x, in->init_y, nb + 1, nF, nb, hp.data());
    HIPCHK(hipMemcpyAsync(g.p, hp.data(), P * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemsetAsync(g.deltaP, 0, P * sizeof(double), h->stream));
    HIPCHK(hipMemcpyAsync(g.errval, ev.data(), ev.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemsetAsync(h->dstatus, 0, sizeof(ss::Status), h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    out->device_init_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    Loop L;
    L.h = h;
    L.g = g;
    L.s = s;
    L.K = s->linesearch_candidates <= 0 ? 4 : s->linesearch_candidates;
    L.frameBytes = (double)fh * nb * (imgType == ss::IMG_F32 ? 4.0 : 2.0);
    if (L.K != 1 && L.K != 2 && L.K != 4 && L.K != 8)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "linesearch_candidates must be 1, 2, 4 or 8");
    L.filterID = s->filter_id;
    L.nFilterIterations = s->max_filter_iterations;
    switch (s->max_filter_iterations != 0 ? s->filter_id : 0) {
    case 0: L.solverFilter = 0; L.nFilterIterations = 0; break;
    case 1: L.solverFilter = 1; break;
    default: L.solverFilter = 2; break;
    }
    // Task::alignFrames (task.cpp:259-412)
    double correlation = 0, corr = 0;
    int iteration = 0, retv = SS_STOP_ERROR;
    for (;;) {
        iteration++;
        int improved = 0;
        if ((st = L.main_iteration(iteration, &improved))) return st;
        if (out->all_params && iteration <= s->max_iterations)
            HIPCHK(hipMemcpyAsync(out->all_params + P * (iteration - 1), g.p, P * sizeof(double),
                                  hipMemcpyDeviceToHost, h->stream));
        HIPCHK(timed(h, SS_K_CORR, L.frameBytes * nF, [&] { return ss::launch_corr(g, h->stream); }));
        if ((st = L.read_status())) return st;
        corr = h->hstatus->corr;
        if (out->iteration_corr && iteration <= s->max_iterations) out->iteration_corr[iteration - 1] = corr;
        if (corr - correlation < s->correlation_increase_threshold) {
            retv = SS_STOP_CORRELATION_INCREASE;
            break;
        } else if (corr - correlation < s->filter_correlation_threshold) {
            if (L.filterID != 0) {
                L.solverFilter = 0;
                L.nFilterIterations = 0;
                L.filterID = 0;
                if (iteration >= s->max_iterations) {  // never run past max_iterations
                    retv = SS_STOP_ITERATION_LIMIT;
                    break;
                }
                continue;
            }
        }
        if (!improved) {
            retv = SS_STOP_MOVEMENT;
            break;
        } else if (s->max_iterations == iteration) {
            retv = SS_STOP_ITERATION_LIMIT;
            break;
        } else if (correlation >= s->halt_correlation) {
            retv = SS_STOP_HALT_CORRELATION;
            break;
        }
        correlation = corr;
        HIPCHK(hipMemcpyAsync(g.pLast, g.p, P * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    }
    if (iteration == 1)
        HIPCHK(hipMemcpyAsync(g.pLast, g.p, P * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(timed(h, SS_K_FINAL, L.frameBytes * nF, [&] { return ss::launch_final(g, h->stream); }));
    std::vector<double> pl(P);
    HIPCHK(hipMemcpyAsync(pl.data(), g.pLast, P * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (out->residrms)
        HIPCHK(hipMemcpyAsync(out->residrms, g.resid, nF * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (out->frame_corr)
        HIPCHK(hipMemcpyAsync(out->frame_corr, g.fcorr, nF * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (out->mask) HIPCHK(hipMemcpyAsync(out->mask, g.maskOut, npix, hipMemcpyDeviceToHost, h->stream));
    if ((st = L.read_status())) return st;
    if (out->knots_x && out->knots_y) knots_out(pl.data(), nF, nb, out->knots_x, out->knots_y);
    out->iterations = iteration;
    out->stop_criterion = retv;
    out->half_iterations = L.halfIterations;
    out->linesearch_rounds = L.rounds;
    out->linesearch_tries = h->hstatus->tries;
    if (h->prof.on) prof_collect(h);
    return SS_OK;
}

}  // namespace

extern "C" {

ss_status ss_create(int32_t device_id, ss_handle **out) {
    if (!out) return SS_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SS_ERR_NOT_AVAILABLE;
    if (device_id < 0 || device_id >= n) return SS_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device_id) != hipSuccess) return SS_ERR_DEVICE;
    ss_handle *h = new ss_handle();
    h->device = device_id;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&h->hstatus, sizeof(ss::Status), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&h->dstatus, sizeof(ss::Status)) != hipSuccess) {
        ss_destroy(h);
        return SS_ERR_DEVICE;
    }
    *out = h;
    return SS_OK;
}

void ss_destroy(ss_handle *h) {
    if (!h) return;
    hipError_t e = hipSetDevice(h->device);
    if (h->stream) e = hipStreamSynchronize(h->stream);
    for (auto &kv : h->bufs)
        if (kv.second.p) e = hipFree(kv.second.p);
    for (auto &kv : h->templates) {
        e = hipFree(kv.second.tmpl);
        e = hipFree(kv.second.xg);
        e = hipFree(kv.second.yg);
    }
    for (auto &kv : h->geometries) {
        e = hipFree(kv.second.geo);
        if (kv.second.fnext) e = hipFree(kv.second.fnext);
    }
    for (auto &p : h->prof.pending) {
        e = hipEventDestroy(p.a);
        e = hipEventDestroy(p.b);
    }
    for (auto ev : h->prof.pool) e = hipEventDestroy(ev);
    if (h->dstatus) e = hipFree(h->dstatus);
    if (h->hstatus) e = hipHostFree(h->hstatus);
    if (h->stream) e = hipStreamDestroy(h->stream);
    (void)e;
    delete h;
}

const char *ss_last_error(const ss_handle *h) { return h ? h->err.c_str() : "no handle"; }

ss_status ss_upload_template(ss_handle *h, int32_t th, int32_t tw, const double *t,
                             const double *xg, const double *yg, int32_t *id) {
    if (!h || !id) return SS_ERR_INVALID_ARGUMENT;
    HIPCHK(hipSetDevice(h->device));
    ResidentTemplate rt;
    ss_status st = upload_template(h, th, tw, t, xg, yg, rt, true);
    if (st) return st;
    *id = h->next_id++;
    h->templates[*id] = rt;
    return SS_OK;
}

ss_status ss_upload_geometry(ss_handle *h, int32_t fh, int32_t nb, const double *frac,
                             const double *bx, const double *by, const uint8_t *bm, int32_t *id) {
    if (!h || !id) return SS_ERR_INVALID_ARGUMENT;
    HIPCHK(hipSetDevice(h->device));
    ResidentGeometry rg;
    ss_status st = upload_geometry(h, fh, nb, frac, bx, by, bm, rg, true);
    if (st) return st;
    *id = h->next_id++;
    h->geometries[*id] = rg;
    return SS_OK;
}

ss_status ss_release(ss_handle *h, int32_t id) {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    auto t = h->templates.find(id);
    if (t != h->templates.end()) {
        HIPCHK(hipFree(t->second.tmpl));
        HIPCHK(hipFree(t->second.xg));
        HIPCHK(hipFree(t->second.yg));
        h->templates.erase(t);
        return SS_OK;
    }
    auto g = h->geometries.find(id);
    if (g != h->geometries.end()) {
        HIPCHK(hipFree(g->second.geo));
        if (g->second.fnext) HIPCHK(hipFree(g->second.fnext));
        h->geometries.erase(g);
        return SS_OK;
    }
    return fail(h, SS_ERR_INVALID_ARGUMENT, "unknown resident id");
}

ss_status ss_align_group(ss_handle *h, const ss_group_in *in, const ss_settings *s,
                         ss_group_out *out) {
    if (!h || !out) return SS_ERR_INVALID_ARGUMENT;
    if (s && s->device_id != h->device)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "settings.device_id differs from the handle's device");
    HIPCHK(hipSetDevice(h->device));
    h->err.clear();
    return run_group(h, in, s, out);
}

ss_status ss_profile_enable(ss_handle *h, int32_t on) {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    HIPCHK(hipSetDevice(h->device));
    prof_collect(h);
    h->prof.on = on != 0;
    for (int i = 0; i < SS_K_COUNT; i++) {
        h->prof.launches[i] = 0;
        h->prof.ms[i] = 0;
        h->prof.bytes[i] = 0;
    }
    return SS_OK;
}

ss_status ss_profile_read(ss_handle *h, int64_t *launches, double *ms, double *frame_bytes) {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    HIPCHK(hipSetDevice(h->device));
    prof_collect(h);
    for (int i = 0; i < SS_K_COUNT; i++) {
        if (launches) launches[i] = h->prof.launches[i];
        if (ms) ms[i] = h->prof.ms[i];
        if (frame_bytes) frame_bytes[i] = h->prof.bytes[i];
    }
    return SS_OK;
}

ss_status ss_align_batch(ss_handle *h, int32_t n, const ss_group_in *in, const ss_settings *s,
                         ss_group_out *out) {
    if (!h || n < 0 || (n > 0 && (!in || !s || !out))) return SS_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < n; i++) {
        ss_status st = ss_align_group(h, in + i, s + i, out + i);
        if (st) return st;
    }
    return SS_OK;
}

}  // extern "C"
