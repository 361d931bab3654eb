// ss_engine.hip -- the C ABI of include/steadyscan.h: handle, resident
// inputs, batch set-up and the submission loop.
//
// A call (one group, or a batch of groups) is laid out in one device arena.
// Each group carries its own control state machine on the device (the
// reference's Task<T>::alignFrames / MainIteration<T>::execute, task.cpp:
// 259-412, mainiteration.cpp:61-119), so the host only submits "ticks" --
// a fixed sequence of stage kernels -- and reads the groups' phases once per
// tick, until every group is done.
#include <algorithm>
#include <new>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "ss_device.h"
#include "ss_host.h"
#include "ss_pinned.h"
#include "steadyscan.h"

namespace {

// Device copies of a template in the pixel precisions it serves (resident
// templates hold both, per-call templates the one of the call).
struct DevTemplate {
    int th = 0, tw = 0, gradFD = 1;
    float2 *tp = nullptr, *gxv = nullptr, *gyh = nullptr;
    double2 *tpD = nullptr, *gxvD = nullptr, *gyhD = nullptr;
    uint32_t *gxvB = nullptr, *gyhB = nullptr;
    void *mem = nullptr;  // owning allocation (resident only)
};

struct DevGeometry {
    int frameH = 0, nb = 0, fracUniform = 1, affine = 0, L = 0, runLen = 8;
    ss::AffineGeo<float> afF{};
    ss::AffineGeo<double> afD{};
    float4 *geo = nullptr;
    double4 *geoD = nullptr;
    float *fnext = nullptr;
    double *fnextD = nullptr;
    void *mem = nullptr;
};

// Host-side converted arrays, per precision.
struct TemplateHost {
    std::vector<float2> tp, gxv, gyh;
    std::vector<double2> tpD, gxvD, gyhD;
    std::vector<uint32_t> gxvB, gyhB;
};

// Round to nearest even BF16 (through FP32), quiet NaN kept.
uint32_t to_bf16(double v) {
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (u >> 16) | 0x40u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

// Gradient pairs as packed BF16 (the Jacobian factors of mixed precision):
// gxv low half xgrad[iy][ix], high half xgrad[iy+1][ix]; gyh low half
// ygrad[iy][ix], high half ygrad[iy][ix+1].  Always built, also when the
// gradients are the template's forward differences.
void gradient_pairs_bf16(int th, int tw, const double *xg, const double *yg, std::vector<uint32_t> &gxv,
                         std::vector<uint32_t> &gyh) {
    const size_t n = (size_t)th * tw;
    auto at = [&](const double *a, int iy, int ix) { return a[(size_t)ix * th + iy]; };
    auto pack = [](double lo, double hi) { return to_bf16(lo) | (to_bf16(hi) << 16); };
    gxv.assign(n, 0u);
    gyh.assign(n, 0u);
    for (int iy = 0; iy < th; iy++)
        for (int ix = 0; ix < tw; ix++) {
            const size_t o = (size_t)iy * tw + ix;
            gxv[o] = pack(at(xg, iy, ix), iy + 1 < th ? at(xg, iy + 1, ix) : 0.0);
            gyh[o] = pack(at(yg, iy, ix), ix + 1 < tw ? at(yg, iy, ix + 1) : 0.0);
        }
}
struct GeometryHost {
    std::vector<float4> tab;
    std::vector<double4> tabD;
    std::vector<float> fn;
    std::vector<double> fnD;
};

struct Profile {
    bool on = false;
    std::vector<hipEvent_t> pool;
    struct Pending {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    int64_t launches[SS_K_COUNT] = {};
    double ms[SS_K_COUNT] = {};
    double bytes[SS_K_COUNT] = {};
};

constexpr int kTickCap = 200000;
// ss_tuning_mode: the default, launch variants by the size of each batch
constexpr int kTuneBySize = -2;

// The group shape the autotuner keys its choices on: what sets the cost of a
// column (precision, frame element type, mixed precision, gradients read or
// derived, geometry form, column height, lines per run) and the number of
// columns (blocks, frame count class).
struct TuneKey {
    int fp64, bf16, gradFD, imgType, affine, frameH, runLen, nb, framesClass;
    bool operator<(const TuneKey &o) const {
        const int a[] = {fp64, bf16, gradFD, imgType, affine, frameH, runLen, nb, framesClass};
        const int b[] = {o.fp64, o.bf16, o.gradFD, o.imgType, o.affine, o.frameH, o.runLen, o.nb, o.framesClass};
        return std::lexicographical_compare(a, a + 9, b, b + 9);
    }
};

}  // namespace

struct ss_handle {
    int device = 0;
    int numCU = 256;
    double clockHz = 1e8;  // device wall clock (wall_clock64) ticks per second
    hipStream_t stream = nullptr;
    std::string err;
    std::map<int, DevTemplate> templates;
    std::map<int, DevGeometry> geometries;
    int next_id = 1;
    void *arena = nullptr;
    size_t arenaBytes = 0;
    void *harena = nullptr;  // pinned host staging
    size_t harenaBytes = 0;
    void *scratch = nullptr;  // work memory of the reconstruction
    size_t scratchBytes = 0;
    Profile prof;
    // the tick's launches, captured once per group table and replayed
    bool useGraph = true;
    hipGraphExec_t tickExec = nullptr;
    ss::LaunchBatch tickFor{};
    bool tickWithFinal = false;
    hipEvent_t tickDone[2] = {nullptr, nullptr};  // control blocks of a tick read back
    // autotuning: -1 tunes each new group shape, 0..kVariants-1 forces that variant
    int tuneMode = kTuneBySize;  // variants by batch size; -1 tunes each new group shape
    std::map<TuneKey, ss_tuning_entry> tuned;
    void *trialTable = nullptr;  // group table of the trial ticks
    size_t trialTableBytes = 0;
    void *trialStage = nullptr;  // its pinned staging
    size_t trialStageBytes = 0;
};

namespace {

ss_status fail(ss_handle *h, ss_status st, const std::string &msg) {
    if (h) h->err = msg;
    return st;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(h, e_ == hipErrorOutOfMemory ? SS_ERR_OUT_OF_MEMORY : SS_ERR_DEVICE, \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

// ---------------------------------------------------------------------------
// profiling: HIP events around stage launches on the handle's stream
// ---------------------------------------------------------------------------
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

void prof_collect(ss_handle *h) {
    if (h->prof.pending.empty()) return;
    hipError_t e = hipStreamSynchronize(h->stream);
    (void)e;
    for (auto &p : h->prof.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            h->prof.ms[p.cls] += ms;
            h->prof.launches[p.cls] += 1;
        }
        h->prof.pool.push_back(p.a);
        h->prof.pool.push_back(p.b);
    }
    h->prof.pending.clear();
}

int stage_class(int entry) {
    switch (ss::entry_stage(entry)) {
    case ss::ST_ASSEMBLE: return SS_K_ASSEMBLE;
    case ss::ST_SOLVE: return SS_K_SOLVE;
    case ss::ST_LS: return SS_K_LS_ROUND;
    case ss::ST_LS_DECIDE: return SS_K_LS_DECIDE;
    case ss::ST_CORR:
    case ss::ST_CORR_FINISH: return SS_K_CORR;
    default: return SS_K_FINAL;
    }
}

hipError_t run_stage(ss_handle *h, const ss::LaunchBatch &L, int entry) {
    if (!h->prof.on) return ss::launch_stage(L, entry, h->stream);
    hipEvent_t a = prof_event(h), b = prof_event(h);
    if (!a || !b) return ss::launch_stage(L, entry, h->stream);
    hipError_t e = hipEventRecord(a, h->stream);
    if (e != hipSuccess) return e;
    e = ss::launch_stage(L, entry, h->stream);
    if (e != hipSuccess) return e;
    e = hipEventRecord(b, h->stream);
    h->prof.pending.push_back({stage_class(entry), a, b});
    return e;
}

// ---------------------------------------------------------------------------
// device arena
// ---------------------------------------------------------------------------
struct Arena {
    size_t off = 0;
    template <typename T>
    size_t take(size_t count) {
        const size_t o = off;
        off += (count * sizeof(T) + 255) & ~(size_t)255;
        return o;
    }
};

ss_status ensure_arena(ss_handle *h, size_t bytes, size_t hostBytes) {
    if (h->arenaBytes < bytes) {
        if (h->arena) {
            HIPCHK(hipStreamSynchronize(h->stream));
            HIPCHK(hipFree(h->arena));
            h->arena = nullptr;
            h->arenaBytes = 0;
        }
        hipError_t e = hipMalloc(&h->arena, bytes);
        if (e != hipSuccess) return fail(h, SS_ERR_OUT_OF_MEMORY, "unable to allocate device memory");
        h->arenaBytes = bytes;
    }
    if (h->harenaBytes < hostBytes) {
        if (h->harena) {
            HIPCHK(hipStreamSynchronize(h->stream));
            HIPCHK(hipHostFree(h->harena));
            h->harena = nullptr;
            h->harenaBytes = 0;
        }
        hipError_t e = hipHostMalloc(&h->harena, hostBytes, hipHostMallocDefault);
        if (e != hipSuccess) return fail(h, SS_ERR_OUT_OF_MEMORY, "unable to allocate pinned host memory");
        h->harenaBytes = hostBytes;
    }
    return SS_OK;
}

// ---------------------------------------------------------------------------
// host-side conversions
// ---------------------------------------------------------------------------
// Template (column-major doubles) -> row-major pairs (T[t], T[t+1]) in the
// requested precisions; gradient pairs only when the gradients are not the
// template's forward differences (steady_scan.m:178-179 makes them so).
template <typename P, typename S>
void template_pairs(int th, int tw, const double *t, const double *xg, const double *yg, bool fd,
                    std::vector<P> &tp, std::vector<P> &gxv, std::vector<P> &gyh) {
    const size_t n = (size_t)th * tw;
    auto at = [&](const double *a, int iy, int ix) { return (S)a[(size_t)ix * th + iy]; };
    tp.assign(n, P{});
    for (int iy = 0; iy < th; iy++)
        for (int ix = 0; ix < tw; ix++) {
            P &q = tp[(size_t)iy * tw + ix];
            q.x = at(t, iy, ix);
            q.y = ix + 1 < tw ? at(t, iy, ix + 1) : (S)0;
        }
    gxv.clear();
    gyh.clear();
    if (fd) return;
    gxv.assign(n, P{});
    gyh.assign(n, P{});
    for (int iy = 0; iy < th; iy++)
        for (int ix = 0; ix < tw; ix++) {
            const size_t o = (size_t)iy * tw + ix;
            gxv[o].x = at(xg, iy, ix);
            gxv[o].y = iy + 1 < th ? at(xg, iy + 1, ix) : (S)0;
            gyh[o].x = at(yg, iy, ix);
            gyh[o].y = ix + 1 < tw ? at(yg, iy, ix + 1) : (S)0;
        }
}

void convert_template(int th, int tw, const double *t, const double *xg, const double *yg, bool wantF,
                      bool wantD, bool wantB, TemplateHost &out, int &gradFD) {
    auto T = [&](int iy, int ix) { return t[(size_t)ix * th + iy]; };
    auto GX = [&](int iy, int ix) { return xg[(size_t)ix * th + iy]; };
    auto GY = [&](int iy, int ix) { return yg[(size_t)ix * th + iy]; };
    double tmax = 0;
    for (size_t i = 0; i < (size_t)th * tw; i++) tmax = std::max(tmax, std::fabs(t[i]));
    // every gradient value the warp can read: xgrad(:, 1:end-1), ygrad(1:end-1, :)
    bool fd = true;
    const double tol = 1e-12 * std::max(1.0, tmax);
    for (int iy = 0; iy < th && fd; iy++)
        for (int ix = 0; ix + 1 < tw; ix++)
            if (!(std::fabs(GX(iy, ix) - (T(iy, ix + 1) - T(iy, ix))) <= tol)) { fd = false; break; }
    for (int iy = 0; iy + 1 < th && fd; iy++)
        for (int ix = 0; ix < tw; ix++)
            if (!(std::fabs(GY(iy, ix) - (T(iy + 1, ix) - T(iy, ix))) <= tol)) { fd = false; break; }
    gradFD = fd ? 1 : 0;
    if (wantF) template_pairs<float2, float>(th, tw, t, xg, yg, fd, out.tp, out.gxv, out.gyh);
    if (wantD) template_pairs<double2, double>(th, tw, t, xg, yg, fd, out.tpD, out.gxvD, out.gyhD);
    if (wantB) gradient_pairs_bf16(th, tw, xg, yg, out.gxvB, out.gyhB);
}

// Geometry block matrices -> affine description (when the base coordinates
// and time fractions are affine in the block's (column, line) coordinates,
// the case of every unidirectional integer-spacing scan) or the per-pixel
// {base_x, base_y, frac, base_mask} table.
void convert_geometry(int fh, int nb, const double *frac, const double *bx, const double *by,
                      const uint8_t *bm, bool wantF, bool wantD, DevGeometry &g, GeometryHost &out) {
    const size_t n = (size_t)fh * nb;
    g.frameH = fh;
    g.nb = nb;
    bool uniform = true, allMask = true;
    for (int b = 0; b < nb && uniform; b++)
        for (int r = 0; r < fh; r++)
            if (frac[(size_t)b * fh + r] != frac[r]) { uniform = false; break; }
    for (size_t i = 0; i < n; i++)
        if (!bm[i]) { allMask = false; break; }
    g.fracUniform = uniform ? 1 : 0;
    // affine detection: L = first row where base_x changes within block 0
    g.affine = 0;
    if (uniform && allMask && fh >= 1) {
        int L = fh;
        for (int r = 1; r < fh; r++)
            if (bx[r] != bx[0]) { L = r; break; }
        if (fh % L == 0) {
            const int J = fh / L;
            auto at = [&](const double *a, int b, int r) { return a[(size_t)b * fh + r]; };
            const double gx0 = at(bx, 0, 0), gxi = L > 1 ? at(bx, 0, 1) - gx0 : 0.0,
                         gxj = J > 1 ? at(bx, 0, L) - gx0 : 0.0;
            const double gy0 = at(by, 0, 0), gyi = L > 1 ? at(by, 0, 1) - gy0 : 0.0,
                         gyj = J > 1 ? at(by, 0, L) - gy0 : 0.0, gyb = nb > 1 ? at(by, 1, 0) - gy0 : 0.0;
            const double f0 = frac[0], fi = L > 1 ? frac[1] - f0 : 0.0, fj = J > 1 ? frac[L] - f0 : 0.0;
            bool ok = true;
            const double tol = 1e-12;
            for (int b = 0; b < nb && ok; b++)
                for (int r = 0; r < fh; r++) {
                    const int j = r / L, i = r % L;
                    if (std::fabs(at(bx, b, r) - (gx0 + gxj * j + gxi * i)) > tol * std::max(1.0, std::fabs(gx0) + std::fabs(gxj) * J) ||
                        std::fabs(at(by, b, r) - (gy0 + gyb * b + gyj * j + gyi * i)) > tol * std::max(1.0, std::fabs(gy0) + std::fabs(gyb) * nb) ||
                        std::fabs(at(frac, b, r) - (f0 + fj * j + fi * i)) > tol) {
                        ok = false;
                        break;
                    }
                }
            if (ok) {
                g.affine = 1;
                g.L = L;
                g.afD = ss::AffineGeo<double>{gx0, gxj, gxi, gy0, gyb, gyj, gyi, f0, fj, fi};
                g.afF = ss::AffineGeo<float>{(float)gx0, (float)gxj, (float)gxi, (float)gy0, (float)gyb,
                                             (float)gyj, (float)gyi, (float)f0, (float)fj, (float)fi};
            }
        }
    }
    // rows a thread walks in order: the lines of one image column of a block
    // (image2blockmatrix lists a block column by column), so that a wave's
    // lanes sit on neighbouring pixels of a scan line
    g.runLen = 8;
    {
        int L = fh;
        for (int r = 1; r < fh; r++)
            if (bx[r] != bx[0]) { L = r; break; }
        if (g.affine) L = g.L;
        if (L >= 1 && L <= 64 && fh % L == 0) g.runLen = L;
    }
    out = GeometryHost{};
    if (!g.affine) {
        if (wantF) {
            out.tab.resize(n);
            for (size_t i = 0; i < n; i++)
                out.tab[i] = make_float4((float)bx[i], (float)by[i], (float)frac[i], bm[i] ? 1.f : 0.f);
        }
        if (wantD) {
            out.tabD.resize(n);
            for (size_t i = 0; i < n; i++) out.tabD[i] = make_double4(bx[i], by[i], frac[i], bm[i] ? 1.0 : 0.0);
        }
    }
    if (!uniform) {
        if (wantF) out.fn.resize(n);
        if (wantD) out.fnD.resize(n);
        for (int b = 0; b < nb; b++)
            for (int r = 0; r < fh; r++) {
                const double v = frac[(size_t)((b + 1) % nb) * fh + r];
                if (wantF) out.fn[(size_t)b * fh + r] = (float)v;
                if (wantD) out.fnD[(size_t)b * fh + r] = v;
            }
    }
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

size_t img_size(int imgType) {
    return imgType == ss::IMG_F64 ? 8 : imgType == ss::IMG_F32 ? 4 : 2;
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
    if (in->template_id <= 0 && (!in->template_image || !in->xgrad || !in->ygrad || in->template_h < 2 || in->template_w < 2))
        return fail(h, SS_ERR_INVALID_ARGUMENT, "template: bad size or NULL array");
    if (in->geometry_id <= 0 && (!in->frac || !in->base_x || !in->base_y || !in->base_mask))
        return fail(h, SS_ERR_INVALID_ARGUMENT, "geometry: NULL array");
    if (s->max_iterations < 1) return fail(h, SS_ERR_INVALID_ARGUMENT, "max_iterations must be >= 1");
    if (s->solver_id < 0 || s->solver_id > 2)  // task.cpp:74-88
        return fail(h, SS_ERR_INVALID_ARGUMENT, "invalid solver id\n0:CuSparse solver\n1:CuSolver solver\n2:Eigen solver");
    const int fsel = s->max_filter_iterations != 0 ? s->filter_id : 0;  // task.cpp:54
    if (fsel < 0 || fsel > 2)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "invalid filter ID\n0: disable filter\n1: smearingfiter\n2:rangefilter");
    if (s->precision != 'S' && s->precision != 'D')
        return fail(h, SS_ERR_INVALID_ARGUMENT, "invalid precision type! use 'S' or 'D'");
    if (s->template_bf16 != 0 && s->template_bf16 != 1)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "template_bf16 must be 0 or 1");
    if (s->template_bf16 && s->precision != 'S')
        return fail(h, SS_ERR_INVALID_ARGUMENT, "a BF16 template needs precision 'S'");
    const int K = s->linesearch_candidates <= 0 ? 2 : s->linesearch_candidates;
    if (K != 1 && K != 2 && K != 4)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "linesearch_candidates must be 1, 2 or 4");
    const long long npix = (long long)in->frame_h * in->n_blocks * in->n_frames;
    if (npix >= (1ll << 31)) return fail(h, SS_ERR_INVALID_ARGUMENT, "group too large");
    return SS_OK;
}

// Per-group host-side plan: the group's device block and host staging, as
// offsets from the group's own base.
struct Plan {
    const ss_group_in *in;
    const ss_settings *s;
    ss_group_out *out;
    DevTemplate tmpl;
    DevGeometry geom;
    TemplateHost th;
    GeometryHost gh;
    int fp64 = 0, bf16 = 0;
    size_t oTp = 0, oGxv = 0, oGyh = 0, oGeo = 0, oFn = 0, oImg = 0, oMin = 0;
    size_t oP = 0, oPL = 0, oP0 = 0, oD0 = 0, oDP = 0, oErr = 0, oCol = 0, oLs = 0, oCorr = 0,
           oFin = 0, oWs = 0, oUsed = 0, oAct = 0, oNext = 0, oMk = 0, oDk = 0, oMask = 0,
           oRes = 0, oFc = 0, oAll = 0, oIc = 0, oFci = 0, oIe = 0, oRg = 0, oIt = 0, oFt = 0;
    size_t hImg = 0, hP = 0;  // host staging offsets
    size_t dBytes = 0, hBytes = 0;
    int imgType = 0;
    size_t imgBytes = 0;
    bool staged = false;  // host frames already in the staging (stage_host_frames)
    int var[ss::TK_COUNT] = {0, 0, 0};  // launch variants of the pixel kernels
    bool varBySize = false;             // chosen per batch by its size instead
};

// Validates one call and lays out its device block and host staging.
ss_status plan_group(ss_handle *h, Plan &P) {
    ss_status st = validate(h, P.in, P.s);
    if (st) return st;
    Arena A, H;
    const ss_group_in &in = *P.in;
    const int fh = in.frame_h, nb = in.n_blocks, nF = in.n_frames, fw = nb * nF, N = fw + 1;
    P.fp64 = P.s->precision == 'D' ? 1 : 0;
    P.bf16 = P.s->template_bf16 ? 1 : 0;
    const bool wF = !P.fp64, wD = P.fp64;
    if (in.template_id > 0) {
        auto it = h->templates.find(in.template_id);
        if (it == h->templates.end()) return fail(h, SS_ERR_INVALID_ARGUMENT, "unknown template_id");
        P.tmpl = it->second;
    } else {
        P.tmpl.th = in.template_h;
        P.tmpl.tw = in.template_w;
        convert_template(in.template_h, in.template_w, in.template_image, in.xgrad, in.ygrad, wF, wD, P.bf16,
                         P.th, P.tmpl.gradFD);
        const size_t tn = (size_t)in.template_h * in.template_w;
        const size_t pb = P.fp64 ? sizeof(double2) : sizeof(float2);
        P.oTp = A.take<char>(tn * pb);
        if (P.bf16) {  // BF16 gradient pairs, whether or not they are forward differences
            P.oGxv = A.take<uint32_t>(tn);
            P.oGyh = A.take<uint32_t>(tn);
        } else if (!P.tmpl.gradFD) {
            P.oGxv = A.take<char>(tn * pb);
            P.oGyh = A.take<char>(tn * pb);
        }
    }
    if (in.geometry_id > 0) {
        auto it = h->geometries.find(in.geometry_id);
        if (it == h->geometries.end()) return fail(h, SS_ERR_INVALID_ARGUMENT, "unknown geometry_id");
        P.geom = it->second;
        if (P.geom.frameH != fh || P.geom.nb != nb)
            return fail(h, SS_ERR_INVALID_ARGUMENT, "resident geometry does not match frame_h x n_blocks");
    } else {
        convert_geometry(fh, nb, in.frac, in.base_x, in.base_y, in.base_mask, wF, wD, P.geom, P.gh);
        if (!P.gh.tab.empty()) P.oGeo = A.take<float4>(P.gh.tab.size());
        if (!P.gh.tabD.empty()) P.oGeo = A.take<double4>(P.gh.tabD.size());
        if (!P.gh.fn.empty()) P.oFn = A.take<float>(P.gh.fn.size());
        if (!P.gh.fnD.empty()) P.oFn = A.take<double>(P.gh.fnD.size());
    }
    // f64 host frames are converted to f32 for precision 'S' (as the mex converts them)
    P.imgType = in.data_type == SS_U16   ? ss::IMG_U16
                : in.data_type == SS_I16 ? ss::IMG_I16
                : in.data_type == SS_F32 ? ss::IMG_F32
                : (in.data_on_device || P.fp64) ? ss::IMG_F64
                                                : ss::IMG_F32;
    P.imgBytes = (size_t)fh * fw * img_size(P.imgType);
    if (!in.data_on_device) {
        P.oImg = A.take<char>(P.imgBytes);
        P.hImg = H.take<char>(P.imgBytes);
    }
    P.oMin = A.take<int>(nb);
    const size_t Pn = 2 * (size_t)N;
    P.oP = A.take<double>(Pn);
    P.oPL = A.take<double>(Pn);
    P.oP0 = A.take<double>(Pn);
    P.oD0 = A.take<double>(Pn);
    P.oDP = A.take<double>(Pn);
    P.oErr = A.take<double>(2 * (size_t)nF);
    P.oCol = A.take<double>((size_t)fw * ss::CS_N);
    P.oLs = A.take<double>((size_t)fw * 2 * ss::kMaxCand);
    P.oCorr = A.take<double>((size_t)fw * ss::CR_N);
    P.oFin = A.take<double>((size_t)fw * ss::FS_N);
    P.oWs = A.take<char>(ss::solve_workspace_bytes(fw));
    P.oUsed = A.take<uint8_t>(N);
    P.oAct = A.take<int>(nF);
    P.oNext = A.take<int>(nF);
    P.oMk = A.take<int>(nF);
    P.oDk = A.take<int>(nF);
    if (P.out->mask && !P.out->mask_on_device) P.oMask = A.take<uint8_t>((size_t)fh * fw);
    P.oRes = A.take<double>(nF);
    P.oFc = A.take<double>(nF);
    if (P.out->all_params) P.oAll = A.take<double>(Pn * P.s->max_iterations);
    if (P.out->iteration_corr) P.oIc = A.take<double>(P.s->max_iterations);
    if (P.s->save_frame_correlations) P.oFci = A.take<double>((size_t)nF * P.s->max_iterations);
    if (P.out->iteration_errvals) P.oIe = A.take<double>(2 * (size_t)nF * P.s->max_iterations);
    if (P.out->rg_params) P.oRg = A.take<double>(Pn * 2 * P.s->max_iterations);
    if (P.out->iteration_times) P.oIt = A.take<double>(P.s->max_iterations);
    if (P.out->function_times) P.oFt = A.take<double>(ss::kFuncSlots);
    P.hP = H.take<double>(Pn + 2 * (size_t)nF);
    P.dBytes = A.off;
    P.hBytes = H.off;
    return SS_OK;
}

// Fills the group's device description and initial control state and
// enqueues its uploads on the handle's stream (D: the group's device block,
// HS: its pinned staging, which must stay untouched until the stream syncs).
// Pixels [b, e) of a group's host frames into its pinned staging HS, as they
// are uploaded: copied, or converted for f64 frames in precision 'S'.
void stage_frames(const Plan &P, char *HS, size_t b, size_t e) {
    char *stage = HS + P.hImg;
    if (P.in->data_type == SS_F64 && P.imgType == ss::IMG_F32) {
        const double *src = (const double *)P.in->data;
        float *dst = (float *)stage;
        for (size_t i = b; i < e; i++) dst[i] = (float)src[i];
    } else {
        const size_t es = img_size(P.imgType);
        std::memcpy(stage + b * es, (const char *)P.in->data + b * es, (e - b) * es);
    }
}

// The host frames of all groups of a batch into their staging (group gi at
// HG[gi]) on up to 16 host threads, in pieces of about 8 MB: one thread's
// copy into pinned memory is several times slower than the upload itself.
// Pieces whose thread cannot be started run on the calling thread.
void stage_host_frames(std::vector<Plan> &plans, const std::vector<char *> &HG) {
    struct Piece {
        int g;
        size_t b, e;
    };
    std::vector<Piece> pieces;
    for (size_t gi = 0; gi < plans.size(); gi++) {
        const Plan &P = plans[gi];
        if (P.in->data_on_device) continue;
        const size_t npix = (size_t)P.in->frame_h * P.in->n_blocks * P.in->n_frames;
        const size_t step = std::max<size_t>(1, ((size_t)8 << 20) / img_size(P.imgType));
        for (size_t b = 0; b < npix; b += step) pieces.push_back({(int)gi, b, std::min(npix, b + step)});
    }
    if (pieces.empty()) return;
    const size_t nt = std::min<size_t>(pieces.size(), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    auto run = [&](size_t t) {
        for (size_t k = t; k < pieces.size(); k += nt)
            stage_frames(plans[pieces[k].g], HG[pieces[k].g], pieces[k].b, pieces[k].e);
    };
    std::vector<std::thread> pool;
    size_t started = 1;
    try {
        for (; started < nt; started++) pool.emplace_back(run, started);
    } catch (...) {  // no thread available: the rest runs here
    }
    run(0);
    for (size_t t = started; t < nt; t++) run(t);
    for (auto &th : pool) th.join();
    for (auto &P : plans)
        if (!P.in->data_on_device) P.staged = true;
}

ss_status setup_group(ss_handle *h, Plan &P, char *D, char *HS, ss::GroupDev &g, ss::GroupCtl &c) {
    auto dptr = [&](size_t o) { return (void *)(D + o); };
    const ss_group_in &in = *P.in;
    const ss_settings &s = *P.s;
    const int fh = in.frame_h, nb = in.n_blocks, nF = in.n_frames, fw = nb * nF, N = fw + 1;
    const size_t Pn = 2 * (size_t)N;
    std::memset(&g, 0, sizeof(g));
    g.frameH = fh;
    g.nb = nb;
    g.nF = nF;
    g.fframeW = fw;
    g.N = N;
    g.th = P.tmpl.th;
    g.tw = P.tmpl.tw;
    g.groupSize = in.group_size;
    for (int k = 0; k < ss::TK_COUNT; k++) g.var[k] = P.var[k];
    g.varBySize = P.varBySize ? 1 : 0;
    const double eps = s.precision == 'S' ? (double)FLT_EPSILON : DBL_EPSILON;
    g.fp64 = P.fp64;
    g.xmax = (double)(g.tw - 1) - eps;
    g.ymax = (double)(g.th - 1) - eps;
    g.rmult = s.linesearch_reduction_multiplier;
    g.thr = s.move_threshold;
    g.halt = s.halt_correlation;
    g.cit = s.correlation_increase_threshold;
    g.fct = s.filter_correlation_threshold;
    g.strength = s.filter_strength;
    g.lambda = s.lambda;
    g.solverId = s.solver_id;
    g.maxIter = s.max_iterations;
    g.K1 = s.linesearch_candidates <= 0 ? 2 : s.linesearch_candidates;
    g.Kn = std::max(g.K1, 4);
    g.maxRounds = 4096 / g.Kn + 1;
    // template
    auto up = [&](size_t o, const void *src, size_t bytes) {
        return hipMemcpyAsync(dptr(o), src, bytes, hipMemcpyHostToDevice, h->stream);
    };
    g.gradBF16 = P.bf16;
    g.refine = P.bf16;  // mixed precision: one FP64 refinement step of each knot update
    if (in.template_id > 0) {
        g.tp = P.tmpl.tp;
        g.gxv = P.tmpl.gxv;
        g.gyh = P.tmpl.gyh;
        g.tpD = P.tmpl.tpD;
        g.gxvD = P.tmpl.gxvD;
        g.gyhD = P.tmpl.gyhD;
        g.gxvB = P.tmpl.gxvB;
        g.gyhB = P.tmpl.gyhB;
    } else {
        const TemplateHost &T = P.th;
        if (P.bf16) {
            g.gxvB = (const uint32_t *)dptr(P.oGxv);
            g.gyhB = (const uint32_t *)dptr(P.oGyh);
            HIPCHK(up(P.oGxv, T.gxvB.data(), T.gxvB.size() * sizeof(uint32_t)));
            HIPCHK(up(P.oGyh, T.gyhB.data(), T.gyhB.size() * sizeof(uint32_t)));
        }
        if (P.fp64) {
            g.tpD = (const double2 *)dptr(P.oTp);
            HIPCHK(up(P.oTp, T.tpD.data(), T.tpD.size() * sizeof(double2)));
            if (!P.tmpl.gradFD) {
                g.gxvD = (const double2 *)dptr(P.oGxv);
                g.gyhD = (const double2 *)dptr(P.oGyh);
                HIPCHK(up(P.oGxv, T.gxvD.data(), T.gxvD.size() * sizeof(double2)));
                HIPCHK(up(P.oGyh, T.gyhD.data(), T.gyhD.size() * sizeof(double2)));
            }
        } else {
            g.tp = (const float2 *)dptr(P.oTp);
            HIPCHK(up(P.oTp, T.tp.data(), T.tp.size() * sizeof(float2)));
            if (!P.tmpl.gradFD && !P.bf16) {  // mixed precision reads the BF16 pairs instead
                g.gxv = (const float2 *)dptr(P.oGxv);
                g.gyh = (const float2 *)dptr(P.oGyh);
                HIPCHK(up(P.oGxv, T.gxv.data(), T.gxv.size() * sizeof(float2)));
                HIPCHK(up(P.oGyh, T.gyh.data(), T.gyh.size() * sizeof(float2)));
            }
        }
    }
    g.gradFD = P.tmpl.gradFD;
    // geometry
    const DevGeometry &G = P.geom;
    g.affine = G.affine;
    g.L = G.L;
    g.runLen = G.runLen;
    g.afF = G.afF;
    g.afD = G.afD;
    g.fracUniform = G.fracUniform;
    if (in.geometry_id > 0) {
        g.geo = G.geo;
        g.geoD = G.geoD;
        g.fnext = G.fnext;
        g.fnextD = G.fnextD;
    } else {
        const GeometryHost &Q = P.gh;
        if (!Q.tab.empty()) {
            g.geo = (const float4 *)dptr(P.oGeo);
            HIPCHK(up(P.oGeo, Q.tab.data(), Q.tab.size() * sizeof(float4)));
        }
        if (!Q.tabD.empty()) {
            g.geoD = (const double4 *)dptr(P.oGeo);
            HIPCHK(up(P.oGeo, Q.tabD.data(), Q.tabD.size() * sizeof(double4)));
        }
        if (!Q.fn.empty()) {
            g.fnext = (const float *)dptr(P.oFn);
            HIPCHK(up(P.oFn, Q.fn.data(), Q.fn.size() * sizeof(float)));
        }
        if (!Q.fnD.empty()) {
            g.fnextD = (const double *)dptr(P.oFn);
            HIPCHK(up(P.oFn, Q.fnD.data(), Q.fnD.size() * sizeof(double)));
        }
    }
    // frame data
    g.imgType = P.imgType;
    if (in.data_on_device) {
        g.img = in.data;
    } else {
        void *stage = HS + P.hImg;
        if (!P.staged) stage_frames(P, HS, 0, (size_t)fh * fw);
        HIPCHK(hipMemcpyAsync(dptr(P.oImg), stage, P.imgBytes, hipMemcpyHostToDevice, h->stream));
        g.img = dptr(P.oImg);
    }
    // chunks of rows start 16-byte aligned when the columns do
    g.vec = ((fh % 8) == 0 && (g.runLen % 8) == 0 && ((uintptr_t)g.img % 16) == 0) ? 1 : 0;
    // min points (first n_blocks entries are read: minPointsPerBlock[c % frameW])
    HIPCHK(up(P.oMin, in.min_points, nb * sizeof(int)));
    g.minPts = (const int *)dptr(P.oMin);
    g.p = (double *)dptr(P.oP);
    g.pLast = (double *)dptr(P.oPL);
    g.p0 = (double *)dptr(P.oP0);
    g.d0 = (double *)dptr(P.oD0);
    g.deltaP = (double *)dptr(P.oDP);
    g.errval = (double *)dptr(P.oErr);
    g.colsum = (double *)dptr(P.oCol);
    g.lssum = (double *)dptr(P.oLs);
    g.corrsum = (double *)dptr(P.oCorr);
    g.finalsum = (double *)dptr(P.oFin);
    g.ws = (double *)dptr(P.oWs);
    g.used = (uint8_t *)dptr(P.oUsed);
    g.lsActive = (int *)dptr(P.oAct);
    g.lsNextK = (int *)dptr(P.oNext);
    g.lsMk = (int *)dptr(P.oMk);
    g.lsDk = (int *)dptr(P.oDk);
    // the mask goes straight to a device-resident output, else to the group's
    // block and is copied to the host when the group finishes
    g.maskOut = !P.out->mask ? nullptr : P.out->mask_on_device ? P.out->mask : (uint8_t *)dptr(P.oMask);
    g.resid = (double *)dptr(P.oRes);
    g.fcorr = (double *)dptr(P.oFc);
    g.allParams = P.out->all_params ? (double *)dptr(P.oAll) : nullptr;
    g.iterCorr = P.out->iteration_corr ? (double *)dptr(P.oIc) : nullptr;
    // flags 'c' and 'e': entries of iterations that do not run stay 0
    const size_t mi = (size_t)s.max_iterations;
    g.fcIter = s.save_frame_correlations ? (double *)dptr(P.oFci) : nullptr;
    g.iterErr = P.out->iteration_errvals ? (double *)dptr(P.oIe) : nullptr;
    if (g.fcIter) HIPCHK(hipMemsetAsync(g.fcIter, 0, (size_t)nF * mi * sizeof(double), h->stream));
    if (g.iterErr) HIPCHK(hipMemsetAsync(g.iterErr, 0, 2 * (size_t)nF * mi * sizeof(double), h->stream));
    // outputs 14 and 9
    g.rgParams = P.out->rg_params ? (double *)dptr(P.oRg) : nullptr;
    g.iterTimes = P.out->iteration_times ? (double *)dptr(P.oIt) : nullptr;
    g.clockHz = h->clockHz;
    if (g.rgParams) HIPCHK(hipMemsetAsync(g.rgParams, 0, Pn * 2 * mi * sizeof(double), h->stream));
    if (g.iterTimes) HIPCHK(hipMemsetAsync(g.iterTimes, 0, mi * sizeof(double), h->stream));
    g.funcTimes = P.out->function_times ? (double *)dptr(P.oFt) : nullptr;  // output 10
    if (g.funcTimes) HIPCHK(hipMemsetAsync(g.funcTimes, 0, ss::kFuncSlots * sizeof(double), h->stream));
    // initial knots (task_allocate.cu:130-134), deltaP = 0, errval = 1e6 (task_allocate.cu:150-184)
    double *hp = (double *)(HS + P.hP);
    std::memset(hp, 0, sizeof(double) * Pn);
    knots_in(in.init_x, in.init_y, nb + 1, nF, nb, hp);
    for (int i = 0; i < 2 * nF; i++) hp[Pn + i] = 1e6;
    HIPCHK(hipMemcpyAsync(g.p, hp, Pn * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(g.errval, hp + Pn, 2 * nF * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemsetAsync(g.deltaP, 0, Pn * sizeof(double), h->stream));
    // control state (task.cpp:54-70 filter selection)
    std::memset(&c, 0, sizeof(c));
    c.phase = ss::PH_ASSEMBLE;
    c.shift = 0;
    c.iteration = 1;
    c.improved = 1;
    c.filterID = s.filter_id;
    c.nFilterIterations = s.max_filter_iterations;
    switch (s.max_filter_iterations != 0 ? s.filter_id : 0) {
    case 0: c.solverFilter = 0; c.nFilterIterations = 0; break;
    case 1: c.solverFilter = 1; break;
    default: c.solverFilter = 2; break;
    }
    c.retv = SS_STOP_ERROR;
    return SS_OK;
}

// Copies a finished group's outputs to the caller's buffers.  finalInAssembly:
// the group's final pass ran in an assembly launch (a tick without final
// stages).
ss_status finish_group(ss_handle *h, const Plan &P, const ss::GroupDev &g, const ss::GroupCtl &c,
                       double initTime, bool finalInAssembly) {
    const ss_group_in &in = *P.in;
    ss_group_out &o = *P.out;
    const int nb = in.n_blocks, nF = in.n_frames, fw = nb * nF;
    const size_t Pn = 2 * (size_t)(fw + 1);
    if (c.phase == ss::PH_ERROR)
        return fail(h, (ss_status)c.error, "line search did not reach the movement threshold (non-finite step?)");
    std::vector<double> pl(Pn);
    HIPCHK(hipMemcpyAsync(pl.data(), g.pLast, Pn * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (o.residrms) HIPCHK(hipMemcpyAsync(o.residrms, g.resid, nF * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (o.frame_corr) HIPCHK(hipMemcpyAsync(o.frame_corr, g.fcorr, nF * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (o.mask && !o.mask_on_device)
        HIPCHK(hipMemcpyAsync(o.mask, g.maskOut, (size_t)in.frame_h * fw, hipMemcpyDeviceToHost, h->stream));
    const int nrec = std::min(c.iteration, P.s->max_iterations);
    if (o.all_params && nrec > 0)
        HIPCHK(hipMemcpyAsync(o.all_params, g.allParams, Pn * nrec * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (o.iteration_corr && nrec > 0)
        HIPCHK(hipMemcpyAsync(o.iteration_corr, g.iterCorr, nrec * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    const size_t mi = (size_t)P.s->max_iterations;
    if (o.frame_corr_iter && g.fcIter)
        HIPCHK(hipMemcpyAsync(o.frame_corr_iter, g.fcIter, nF * mi * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (o.iteration_errvals)
        HIPCHK(hipMemcpyAsync(o.iteration_errvals, g.iterErr, 2 * nF * mi * sizeof(double), hipMemcpyDeviceToHost,
                              h->stream));
    if (o.rg_params)
        HIPCHK(hipMemcpyAsync(o.rg_params, g.rgParams, Pn * 2 * mi * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (o.iteration_times)
        HIPCHK(hipMemcpyAsync(o.iteration_times, g.iterTimes, mi * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (o.function_times)
        HIPCHK(hipMemcpyAsync(o.function_times, g.funcTimes, ss::kFuncSlots * sizeof(double), hipMemcpyDeviceToHost,
                              h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (o.knots_x && o.knots_y) knots_out(pl.data(), nF, nb, o.knots_x, o.knots_y);
    if (o.raw_params) std::memcpy(o.raw_params, pl.data(), Pn * sizeof(double));  // output 13
    o.iterations = c.iteration;
    o.stop_criterion = c.retv;
    o.half_iterations = c.halfIterations;
    o.linesearch_rounds = c.rounds;
    o.linesearch_tries = c.tries;
    o.device_init_time = initTime;
    // algorithmic frame bytes per stage class for the roofline
    if (h->prof.on) {
        const double fb = (double)in.frame_h * nb * (double)img_size(P.imgType);
        h->prof.bytes[SS_K_ASSEMBLE] += fb * nF * c.halfIterations;
        h->prof.bytes[SS_K_LS_ROUND] += fb * c.lsFramePasses;
        h->prof.bytes[SS_K_CORR] += fb * nF * c.iteration;
        h->prof.bytes[finalInAssembly ? SS_K_ASSEMBLE : SS_K_FINAL] += fb * nF;
    }
    return SS_OK;
}

// Device table of the groups in flight (GroupDev, GroupCtl, work-item maps).
struct Table {
    void *mem = nullptr;
    size_t bytes = 0;
};

// Uploads the group table (rows in gd / gc order) and describes the launches.
ss_status upload_table(ss_handle *h, Table &T, std::vector<ss::GroupDev> &gd, const ss::GroupCtl *gc,
                       void *hstage, size_t hstageBytes, ss::LaunchBatch &L) {
    const int n = (int)gd.size();
    int nItems = 0, nFrames = 0;
    for (const auto &g : gd) {
        nItems += g.fframeW;
        nFrames += g.nF;
    }
    Arena A;
    const size_t oG = A.take<ss::GroupDev>(n), oC = A.take<ss::GroupCtl>(n), oI = A.take<int>(nItems),
                 oF = A.take<int>(nFrames);
    if (A.off > hstageBytes) return fail(h, SS_ERR_OUT_OF_MEMORY, "group table staging too small");
    if (T.bytes < A.off) {
        const size_t want = std::max(A.off, T.bytes * 2);
        if (T.mem) {
            HIPCHK(hipStreamSynchronize(h->stream));
            HIPCHK(hipFree(T.mem));
            T.mem = nullptr;
            T.bytes = 0;
        }
        if (hipMalloc(&T.mem, want) != hipSuccess) return fail(h, SS_ERR_OUT_OF_MEMORY, "unable to allocate device memory");
        T.bytes = want;
    }
    char *HS = (char *)hstage;
    ss::GroupDev *hg = (ss::GroupDev *)(HS + oG);
    int *hi = (int *)(HS + oI), *hf = (int *)(HS + oF);
    int item = 0, frame = 0;
    for (int i = 0; i < 2 * ss::IMG_NTYPES; i++) {
        L.hasType[i] = 0;
        for (int k = 0; k < ss::TK_COUNT; k++) L.hasVar[k][i] = 0;
    }
    L.hasFrameCorr = 0;
    const bool small = nItems <= 8 * h->numCU;  // see ss_tuning.h
    for (int gi = 0; gi < n; gi++) {
        ss::GroupDev &g = gd[gi];
        if (g.varBySize)
            for (int k = 0; k < ss::TK_COUNT; k++) g.var[k] = small ? ss::kSmallBatchVariant[k][g.fp64 ? 1 : 0] : 0;
        g.itemStart = item;
        g.frameStart = frame;
        hg[gi] = g;
        for (int i = 0; i < g.fframeW; i++) hi[item + i] = gi;
        for (int i = 0; i < g.nF; i++) hf[frame + i] = gi;
        item += g.fframeW;
        frame += g.nF;
        L.hasType[g.imgType + ss::IMG_NTYPES * g.fp64] = 1;
        for (int k = 0; k < ss::TK_COUNT; k++) L.hasVar[k][g.imgType + ss::IMG_NTYPES * g.fp64] |= 1 << g.var[k];
        if (g.fcIter) L.hasFrameCorr = 1;
    }
    std::memcpy(HS + oC, gc, n * sizeof(ss::GroupCtl));
    char *D = (char *)T.mem;
    HIPCHK(hipMemcpyAsync(D, HS, A.off, hipMemcpyHostToDevice, h->stream));
    L.b.nGroups = n;
    L.b.nItems = nItems;
    L.b.nFrames = nFrames;
    L.b.groups = (ss::GroupDev *)(D + oG);
    L.b.ctl = (ss::GroupCtl *)(D + oC);
    L.b.itemGroup = (int *)(D + oI);
    L.b.frameGroup = (int *)(D + oF);
    L.itemGrid = std::max(1, std::min(nItems, 8 * h->numCU));
    L.frameGrid = std::max(1, std::min(nFrames, 16 * h->numCU));
    // A batch of at most 8 columns per CU leaves most of the chip idle in
    // the rounds after a half-iteration's first (few frames remain): there a
    // column's two halves of the candidates go to two workgroups, which
    // halves the chain of gathers a workgroup waits on.
    L.lsSplit = small ? 2 : 1;
    L.lsSplitGrid = L.lsSplit == 2 ? 2 * nItems : L.itemGrid;
    return SS_OK;
}

size_t table_bytes(int nGroups, int nItems, int nFrames) {
    Arena A;
    A.take<ss::GroupDev>(nGroups);
    A.take<ss::GroupCtl>(nGroups);
    A.take<int>(nItems);
    A.take<int>(nFrames);
    return A.off;
}

// One tick: every group advances through the stage sequence (a kernel skips
// groups in other phases); the control blocks come back to hc.
// kLsSlots = 3 line-search rounds per half-iteration: most half-iterations
// finish within them, so a tick is usually one Gauss-Newton iteration.
// kTick ends a group that stops in it with the final stages; its first
// kTickNoFinalLen entries leave the group in the final phase, and the
// assembly and solve launches of the next tick run its final pass and end
// it (run_batch keeps that tick in flight, so the groups end no later).
using ss::tick_entry;
constexpr int kTick[] = {
    tick_entry(ss::ST_ASSEMBLE),    tick_entry(ss::ST_SOLVE),
    tick_entry(ss::ST_LS, 0),       tick_entry(ss::ST_LS_DECIDE, 0),
    tick_entry(ss::ST_LS, 1),       tick_entry(ss::ST_LS_DECIDE, 1),
    tick_entry(ss::ST_LS, 2),       tick_entry(ss::ST_LS_DECIDE, 2),
    tick_entry(ss::ST_ASSEMBLE),    tick_entry(ss::ST_SOLVE),
    tick_entry(ss::ST_LS, 0),       tick_entry(ss::ST_LS_DECIDE, 0),
    tick_entry(ss::ST_LS, 1),       tick_entry(ss::ST_LS_DECIDE, 1),
    tick_entry(ss::ST_LS, 2),       tick_entry(ss::ST_LS_DECIDE, 2),
    tick_entry(ss::ST_CORR),        tick_entry(ss::ST_CORR_FINISH),
    tick_entry(ss::ST_FINAL),       tick_entry(ss::ST_FINAL_FINISH)};
static_assert(ss::kLsSlots == 3, "kTick holds three line-search slots per half-iteration");
constexpr int kTickNoFinalLen = sizeof(kTick) / sizeof(kTick[0]) - 2;
static_assert(ss::entry_stage(kTick[kTickNoFinalLen]) == ss::ST_FINAL, "the final stages end kTick");

// Enqueues one tick on the handle's stream: kTick, or without its final
// stages.
ss_status enqueue_tick(ss_handle *h, const ss::LaunchBatch &L, bool withFinal) {
    const int len = withFinal ? (int)(sizeof(kTick) / sizeof(kTick[0])) : kTickNoFinalLen;
    if (h->prof.on || !h->useGraph) {  // per-stage events need plain launches
        for (int i = 0; i < len; i++) HIPCHK(run_stage(h, L, kTick[i]));
    } else {
        // the tick holds kernel launches only (no memset nodes)
        if (!h->tickExec || h->tickWithFinal != withFinal || std::memcmp(&h->tickFor, &L, sizeof(L)) != 0) {
            if (h->tickExec) {
                HIPCHK(hipGraphExecDestroy(h->tickExec));
                h->tickExec = nullptr;
            }
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < len; i++) {
                const hipError_t e = ss::launch_stage(L, kTick[i], h->stream);
                if (e != hipSuccess) {
                    (void)hipStreamEndCapture(h->stream, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    return fail(h, SS_ERR_DEVICE, std::string("tick capture: ") + hipGetErrorString(e));
                }
            }
            HIPCHK(hipStreamEndCapture(h->stream, &graph));
            const hipError_t e = hipGraphInstantiate(&h->tickExec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                h->tickExec = nullptr;
                return fail(h, SS_ERR_DEVICE, std::string("tick instantiate: ") + hipGetErrorString(e));
            }
            h->tickFor = L;
            h->tickWithFinal = withFinal;
        }
        HIPCHK(hipGraphLaunch(h->tickExec, h->stream));
    }
    return SS_OK;
}

// One tick with the final stages; the control blocks come back to hc.
ss_status run_tick(ss_handle *h, const ss::LaunchBatch &L, ss::GroupCtl *hc) {
    ss_status st = enqueue_tick(h, L, true);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(hc, L.b.ctl, L.b.nGroups * sizeof(ss::GroupCtl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SS_OK;
}

bool finished(const ss::GroupCtl &c) { return c.phase == ss::PH_DONE || c.phase == ss::PH_ERROR; }

// ---------------------------------------------------------------------------
// autotuning of the pixel kernels' launch variants
// ---------------------------------------------------------------------------
TuneKey tune_key(const Plan &P) {
    int fc = 0;
    for (int n = P.in->n_frames; n > 1; n >>= 1) fc++;  // floor(log2(frames))
    return TuneKey{P.fp64, P.bf16, P.tmpl.gradFD, P.imgType, P.geom.affine, P.in->frame_h, P.geom.runLen,
                   P.in->n_blocks, fc};
}

// One tick of the batch L with plain launches, HIP events around the
// assembly, line-search and correlation launches; ms receives each kernel's
// summed time.
ss_status trial_tick(ss_handle *h, const ss::LaunchBatch &L, double ms[ss::TK_COUNT]) {
    struct Ev {
        int k;
        hipEvent_t a, b;
    };
    std::vector<Ev> ev;
    ss_status st = SS_OK;
    for (int entry : kTick) {
        const int stg = ss::entry_stage(entry);
        const int k = stg == ss::ST_ASSEMBLE ? ss::TK_ASSEMBLE : stg == ss::ST_LS ? ss::TK_LS
                      : stg == ss::ST_CORR   ? ss::TK_CORR
                                             : -1;
        hipEvent_t a = k >= 0 ? prof_event(h) : nullptr, b = k >= 0 ? prof_event(h) : nullptr;
        if (a && b && hipEventRecord(a, h->stream) != hipSuccess) st = fail(h, SS_ERR_DEVICE, "trial tick event");
        if (!st && ss::launch_stage(L, entry, h->stream) != hipSuccess) st = fail(h, SS_ERR_DEVICE, "trial tick launch");
        if (a && b && !st && hipEventRecord(b, h->stream) != hipSuccess) st = fail(h, SS_ERR_DEVICE, "trial tick event");
        if (a && b) ev.push_back({k, a, b});
        if (st) break;
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess && !st) st = fail(h, SS_ERR_DEVICE, "trial tick");
    for (auto &e : ev) {
        float t = 0.f;
        if (!st && hipEventElapsedTime(&t, e.a, e.b) == hipSuccess) ms[e.k] += t;
        h->prof.pool.push_back(e.a);
        h->prof.pool.push_back(e.b);
    }
    return st;
}

// Picks the launch variants of group P (device block D, pinned staging HS):
// the forced variant, the handle's choice for the group's shape, or, for a
// shape seen for the first time, the fastest variant per kernel measured
// on the group's own first tick, run twice per variant from the initial
// state (the real set-up follows and starts again from it).
ss_status tune_group(ss_handle *h, Plan &P, char *D, char *HS) {
    P.varBySize = h->tuneMode == kTuneBySize;
    if (h->tuneMode >= 0 || P.varBySize) {
        for (int k = 0; k < ss::TK_COUNT; k++) P.var[k] = P.varBySize ? 0 : h->tuneMode;
        return SS_OK;
    }
    const TuneKey key = tune_key(P);
    auto it = h->tuned.find(key);
    if (it != h->tuned.end()) {
        for (int k = 0; k < ss::TK_COUNT; k++) P.var[k] = it->second.chosen[k];
        return SS_OK;
    }
    ss_tuning_entry e;
    std::memset(&e, 0, sizeof(e));
    e.precision = P.fp64 ? 'D' : 'S';
    e.mixed = P.bf16;
    e.data_type = P.in->data_type;
    e.frame_h = P.in->frame_h;
    e.n_blocks = P.in->n_blocks;
    e.frames = P.in->n_frames;
    for (int k = 0; k < ss::TK_COUNT; k++)
        for (int v = 0; v < ss::kVariants; v++) {
            e.samples[k][v] = ss::kTuneTable[k][P.fp64][v].samples;
            e.waves[k][v] = ss::kTuneTable[k][P.fp64][v].waves;
        }
    const int fw = P.in->n_blocks * P.in->n_frames;
    const size_t tb = table_bytes(1, fw, P.in->n_frames);
    if (h->trialStageBytes < tb) {
        if (h->trialStage) HIPCHK(hipHostFree(h->trialStage));
        h->trialStage = nullptr;
        h->trialStageBytes = 0;
        if (hipHostMalloc(&h->trialStage, tb, hipHostMallocDefault) != hipSuccess)
            return fail(h, SS_ERR_OUT_OF_MEMORY, "unable to allocate pinned host memory");
        h->trialStageBytes = tb;
    }
    // two passes over the variants: the first only loads and warms each
    // variant's code, the second is timed
    for (int pass = 0; pass < 2; pass++)
    for (int v = 0; v < ss::kVariants; v++) {
        for (int k = 0; k < ss::TK_COUNT; k++) P.var[k] = v;
        std::vector<ss::GroupDev> gd(1);
        ss::GroupCtl c;
        ss_status st = setup_group(h, P, D, HS, gd[0], c);
        if (st) return st;
        Table T;
        T.mem = h->trialTable;
        T.bytes = h->trialTableBytes;
        ss::LaunchBatch L{};
        st = upload_table(h, T, gd, &c, h->trialStage, h->trialStageBytes, L);
        h->trialTable = T.mem;
        h->trialTableBytes = T.bytes;
        if (st) return st;
        double ms[ss::TK_COUNT] = {0, 0, 0};
        st = trial_tick(h, L, ms);  // ends with a synchronisation: HS may be written again
        if (st) return st;
        if (pass == 1)
            for (int k = 0; k < ss::TK_COUNT; k++) e.trial_us[k][v] = ms[k] * 1e3;
    }
    for (int k = 0; k < ss::TK_COUNT; k++) {
        int best = 0;
        for (int v = 1; v < ss::kVariants; v++)
            if (e.trial_us[k][v] < e.trial_us[k][best]) best = v;
        e.chosen[k] = best;
        P.var[k] = best;
    }
    h->tuned[key] = e;
    return SS_OK;
}

ss_status run_batch(ss_handle *h, int n, const ss_group_in *ins, const ss_settings *ss_,
                    ss_group_out *outs) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<Plan> plans(n);
    int nItems = 0, nFrames = 0;
    size_t dTotal = 0, hTotal = 0;
    for (int gi = 0; gi < n; gi++) {
        Plan &P = plans[gi];
        P.in = ins + gi;
        P.s = ss_ + gi;
        P.out = outs + gi;
        ss_status st = plan_group(h, P);
        if (st) return st;
        dTotal += (P.dBytes + 255) & ~(size_t)255;
        hTotal += (P.hBytes + 255) & ~(size_t)255;
        nItems += P.in->n_blocks * P.in->n_frames;
        nFrames += P.in->n_frames;
    }
    const size_t tb = (table_bytes(n, nItems, nFrames) + 255) & ~(size_t)255;
    const size_t cb = (n * sizeof(ss::GroupCtl) + 255) & ~(size_t)255;
    ss_status st = ensure_arena(h, dTotal + tb, hTotal + tb + 2 * cb);
    if (st) return st;
    // [table][group 0][group 1]... in the arena; the same order in staging,
    // after two read-back buffers of the control blocks
    char *D = (char *)h->arena + tb;
    char *HS = (char *)h->harena;
    ss::GroupCtl *hcb[2] = {(ss::GroupCtl *)(HS + tb), (ss::GroupCtl *)(HS + tb + cb)};
    ss::GroupCtl *hc = hcb[0];
    std::vector<char *> HG(n);
    HG[0] = HS + tb + 2 * cb;
    for (int gi = 1; gi < n; gi++) HG[gi] = HG[gi - 1] + ((plans[gi - 1].hBytes + 255) & ~(size_t)255);
    stage_host_frames(plans, HG);
    std::vector<ss::GroupDev> gd(n);
    for (int gi = 0; gi < n; gi++) {
        Plan &P = plans[gi];
        st = tune_group(h, P, D, HG[gi]);
        if (st) return st;
        st = setup_group(h, P, D, HG[gi], gd[gi], hc[gi]);
        if (st) return st;
        D += (P.dBytes + 255) & ~(size_t)255;
    }
    Table T;
    T.mem = h->arena;
    T.bytes = tb;
    ss::LaunchBatch L{};
    st = upload_table(h, T, gd, hc, HS, tb, L);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(h->stream));
    const double initTime = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // One tick stays in flight: tick t is enqueued with the read-back of its
    // control blocks before the host waits for tick t-1's, so the device
    // does not idle while the host reads them.  Kernels skip finished
    // groups.  The ticks have no final stages: a group that stops in tick
    // t-1 ends in the assembly and solve launches of tick t, so the loop
    // stops once every group has finished or stopped by tick t-1.
    for (int i = 0; i < 2; i++)
        if (!h->tickDone[i]) HIPCHK(hipEventCreateWithFlags(&h->tickDone[i], hipEventDisableTiming));
    const size_t ctlBytes = n * sizeof(ss::GroupCtl);
    auto all_finished = [&](const ss::GroupCtl *c) {
        for (int gi = 0; gi < n; gi++)
            if (!finished(c[gi])) return false;
        return true;
    };
    auto all_stopped = [&](const ss::GroupCtl *c) {
        for (int gi = 0; gi < n; gi++)
            if (!finished(c[gi]) && c[gi].phase != ss::PH_FINAL) return false;
        return true;
    };
    // a failure leaves no read-back in flight into the staging, which the
    // next call writes from the host
    auto stop = [&](ss_status e) {
        (void)hipStreamSynchronize(h->stream);
        return e;
    };
    int t = 0;
    for (; t < kTickCap; t++) {
        st = enqueue_tick(h, L, false);
        if (st) return stop(st);
        hipError_t e = hipMemcpyAsync(hcb[t & 1], L.b.ctl, ctlBytes, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipEventRecord(h->tickDone[t & 1], h->stream);
        if (e == hipSuccess && t > 0) e = hipEventSynchronize(h->tickDone[(t - 1) & 1]);
        if (e != hipSuccess) return stop(fail(h, SS_ERR_DEVICE, std::string("tick read-back: ") + hipGetErrorString(e)));
        if (t > 0 && all_stopped(hcb[(t - 1) & 1])) break;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    if (t == kTickCap) t--;
    hc = hcb[t & 1];  // the last tick's control blocks
    if (!all_finished(hc)) return fail(h, SS_ERR_NUMERICAL, "alignment did not terminate");
    ss_status result = SS_OK;
    std::string firstErr;
    for (int gi = 0; gi < n; gi++) {
        st = finish_group(h, plans[gi], gd[gi], hc[gi], initTime, true);
        if (st && result == SS_OK) {
            result = st;
            firstErr = h->err + " (group " + std::to_string(gi) + ")";
        }
    }
    if (result) h->err = firstErr;
    if (h->prof.on) prof_collect(h);
    return result;
}

// ---------------------------------------------------------------------------
// serving: continuous admission of groups into the running batch
// ---------------------------------------------------------------------------
// First-fit sub-allocator over device chunks, bounded by a byte budget.
struct DevicePool {
    struct Chunk {
        char *base = nullptr;
        size_t size = 0;
        std::map<size_t, size_t> free;  // offset -> length
    };
    std::vector<Chunk> chunks;
    size_t budget = 0, reserved = 0, inUse = 0;
    static constexpr size_t kAlign = 256, kChunk = (size_t)1 << 30;

    char *alloc(size_t n) {
        n = (n + kAlign - 1) & ~(kAlign - 1);
        for (auto &c : chunks)
            for (auto it = c.free.begin(); it != c.free.end(); ++it)
                if (it->second >= n) {
                    const size_t off = it->first, len = it->second;
                    c.free.erase(it);
                    if (len > n) c.free[off + n] = len - n;
                    inUse += n;
                    return c.base + off;
                }
        const size_t want = std::max(std::min(kChunk, budget - reserved), n);
        if (reserved + want > budget) return nullptr;
        Chunk c;
        if (hipMalloc(&c.base, want) != hipSuccess) return nullptr;
        c.size = want;
        if (want > n) c.free[n] = want - n;
        reserved += want;
        inUse += n;
        chunks.push_back(std::move(c));
        return chunks.back().base;
    }
    void release(char *p, size_t n) {
        n = (n + kAlign - 1) & ~(kAlign - 1);
        for (auto &c : chunks) {
            if (p < c.base || p >= c.base + c.size) continue;
            size_t off = (size_t)(p - c.base);
            inUse -= n;
            auto next = c.free.lower_bound(off);
            if (next != c.free.end() && next->first == off + n) {  // merge with the following block
                n += next->second;
                next = c.free.erase(next);
            }
            if (next != c.free.begin()) {  // and with the preceding one
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    prev->second += n;
                    return;
                }
            }
            c.free[off] = n;
            return;
        }
    }
    void destroy() {
        for (auto &c : chunks) (void)hipFree(c.base);
        chunks.clear();
        reserved = inUse = 0;
    }
};

}  // namespace

struct ss_server {
    ss_handle *h = nullptr;
    int maxGroups = 256;
    DevicePool pool;
    struct Request {
        int64_t ticket;
        ss_group_in in;
        ss_settings s;
        ss_group_out *out;
    };
    struct Slot {
        Request req;
        Plan P;
        char *mem = nullptr;
        ss::GroupDev g;
        ss::GroupCtl c;
        double admitTime = 0;
    };
    struct Done {
        int64_t ticket;
        ss_status status;
        std::string message;  // empty on success
    };
    std::deque<Request> queue;
    std::vector<Slot> resident;
    std::vector<Done> completed;
    std::vector<Done> reported;  // the last ss_server_completed's entries (their messages)
    Table table;
    void *stage = nullptr;  // pinned: group staging + table
    size_t stageBytes = 0;
    size_t stageLimit = 0;  // 0: no limit
    int64_t nextTicket = 1;
    bool dirty = true;      // membership changed since the table was uploaded
    ss::LaunchBatch L{};
    std::vector<ss::GroupCtl> hc;
};

namespace {

// Grows the page-locked staging to at least `bytes`.  The limit applies to
// the staging of admitted groups (`groups`), not to the group table's.
ss_status server_stage(ss_server *S, size_t bytes, bool groups) {
    ss_handle *h = S->h;
    if (S->stageBytes >= bytes) return SS_OK;
    if (groups && S->stageLimit && bytes > S->stageLimit)
        return fail(h, SS_ERR_OUT_OF_MEMORY, "the group's page-locked staging would exceed the server's limit");
    size_t want = std::max(bytes, S->stageBytes * 2);
    if (groups && S->stageLimit) want = std::min(want, S->stageLimit);
    if (S->stage) {
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipHostFree(S->stage));
        S->stage = nullptr;
        S->stageBytes = 0;
    }
    if (hipHostMalloc(&S->stage, want, hipHostMallocDefault) != hipSuccess)
        return fail(h, SS_ERR_OUT_OF_MEMORY, "unable to allocate pinned host memory");
    S->stageBytes = want;
    return SS_OK;
}

// Completes a ticket that never ran (a bad request, no memory, or a failed
// set-up) with its status and message.
void complete_failed(ss_server *S, int64_t ticket, ss_status st, const std::string &msg) {
    S->completed.push_back({ticket, st, msg.empty() ? std::string("request failed") : msg});
}

// Admits queued requests in order while slots and the memory budget allow.
// A request joins the resident set only once its device block is set up; a
// request that cannot be set up completes with its error and returns its
// memory, so the group table never holds a half-admitted group.  The
// return value is an error only when the device itself failed.
ss_status server_admit(ss_server *S) {
    ss_handle *h = S->h;
    std::vector<ss_server::Slot> fresh;
    size_t hNeed = 0;
    while (!S->queue.empty() && (int)(S->resident.size() + fresh.size()) < S->maxGroups) {
        ss_server::Slot slot;
        slot.req = S->queue.front();
        slot.P.in = &slot.req.in;
        slot.P.s = &slot.req.s;
        slot.P.out = slot.req.out;
        ss_status st = plan_group(h, slot.P);
        if (st) {  // a bad request completes with its error; the others go on
            complete_failed(S, slot.req.ticket, st, h->err);
            S->queue.pop_front();
            continue;
        }
        const size_t hb = (slot.P.hBytes + 255) & ~(size_t)255;
        if (S->stageLimit && hNeed + hb > S->stageLimit) {
            if (fresh.empty()) {  // its staging alone exceeds the limit
                complete_failed(S, slot.req.ticket, SS_ERR_OUT_OF_MEMORY,
                                "the group's page-locked staging would exceed the server's limit");
                S->queue.pop_front();
                continue;
            }
            break;  // admitted at a later step
        }
        slot.mem = S->pool.alloc(slot.P.dBytes);
        if (!slot.mem) {
            if (S->resident.empty() && fresh.empty()) {  // can never fit
                complete_failed(S, slot.req.ticket, SS_ERR_OUT_OF_MEMORY,
                                "the group does not fit in the server's HBM budget");
                S->queue.pop_front();
                continue;
            }
            break;  // wait for memory to free up
        }
        S->queue.pop_front();
        hNeed += hb;
        fresh.push_back(std::move(slot));
    }
    if (fresh.empty()) return SS_OK;
    auto drop = [&](ss_server::Slot &slot, ss_status st) {
        S->pool.release(slot.mem, slot.P.dBytes);
        slot.mem = nullptr;
        complete_failed(S, slot.req.ticket, st, h->err);
    };
    ss_status st = server_stage(S, hNeed, true);
    if (st) {
        for (auto &slot : fresh) drop(slot, st);
        return SS_OK;
    }
    char *HS = (char *)S->stage;
    std::vector<bool> ok(fresh.size(), false);
    for (size_t k = 0; k < fresh.size(); k++) {
        auto &slot = fresh[k];
        slot.P.in = &slot.req.in;  // the Plan points at the slot's own request copy
        slot.P.s = &slot.req.s;
        st = tune_group(h, slot.P, slot.mem, HS);
        if (!st) st = setup_group(h, slot.P, slot.mem, HS, slot.g, slot.c);
        if (st) {
            drop(slot, st);
            continue;
        }
        ok[k] = true;
        HS += (slot.P.hBytes + 255) & ~(size_t)255;
    }
    // staged uploads must land before the staging is reused for the table
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < fresh.size(); k++) {
        if (!ok[k]) continue;
        S->resident.push_back(std::move(fresh[k]));
        S->dirty = true;
    }
    for (auto &slot : S->resident) {  // plans point into the moved requests
        slot.P.in = &slot.req.in;
        slot.P.s = &slot.req.s;
    }
    return SS_OK;
}

}  // namespace

extern "C" {

ss_status ss_server_create(ss_handle *h, const ss_server_config *cfg, ss_server **out) try {
    if (!h || !out) return SS_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    ss_server *S = new ss_server();
    S->h = h;
    const int64_t budget = (cfg && cfg->hbm_budget_bytes > 0) ? cfg->hbm_budget_bytes : ((int64_t)64 << 30);
    S->pool.budget = (size_t)budget;
    S->maxGroups = (cfg && cfg->max_groups > 0) ? cfg->max_groups : 256;
    S->stageLimit = (cfg && cfg->staging_limit_bytes > 0) ? (size_t)cfg->staging_limit_bytes : 0;
    *out = S;
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

void ss_server_destroy(ss_server *S) {
    if (!S) return;
    ss_handle *h = S->h;
    hipError_t e = hipSetDevice(h->device);
    e = hipStreamSynchronize(h->stream);
    S->pool.destroy();
    if (S->table.mem) e = hipFree(S->table.mem);
    if (S->stage) e = hipHostFree(S->stage);
    (void)e;
    delete S;
}

ss_status ss_server_submit(ss_server *S, const ss_group_in *in, const ss_settings *s, ss_group_out *out,
                           int64_t *ticket) try {
    if (!S || !in || !s || !out || !ticket) return SS_ERR_INVALID_ARGUMENT;
    ss_handle *h = S->h;
    if (s->device_id != h->device)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "settings.device_id differs from the handle's device");
    ss_server::Request r;
    r.ticket = S->nextTicket++;
    r.in = *in;
    r.s = *s;
    r.out = out;
    S->queue.push_back(r);
    *ticket = r.ticket;
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

ss_status ss_server_step(ss_server *S, int32_t *n_resident, int32_t *n_queued) try {
    if (!S) return SS_ERR_INVALID_ARGUMENT;
    ss_handle *h = S->h;
    HIPCHK(hipSetDevice(h->device));
    h->err.clear();
    ss_status st = server_admit(S);
    if (st) return st;
    if (!S->resident.empty()) {
        const int n = (int)S->resident.size();
        S->hc.resize(n);
        if (S->dirty) {
            std::vector<ss::GroupDev> gd(n);
            for (int i = 0; i < n; i++) {
                gd[i] = S->resident[i].g;
                S->hc[i] = S->resident[i].c;
            }
            int nItems = 0, nFrames = 0;
            for (auto &g : gd) {
                nItems += g.fframeW;
                nFrames += g.nF;
            }
            const size_t tb = table_bytes(n, nItems, nFrames);
            st = server_stage(S, tb, false);
            if (st) return st;
            st = upload_table(h, S->table, gd, S->hc.data(), S->stage, S->stageBytes, S->L);
            if (st) return st;
            for (int i = 0; i < n; i++) S->resident[i].g = gd[i];
            S->dirty = false;
        }
        st = run_tick(h, S->L, S->hc.data());
        if (st) return st;
        // retire finished groups
        std::vector<ss_server::Slot> keep;
        keep.reserve(n);
        for (int i = 0; i < n; i++) {
            auto &slot = S->resident[i];
            slot.c = S->hc[i];
            if (finished(slot.c)) {
                slot.P.in = &slot.req.in;
                slot.P.s = &slot.req.s;
                const ss_status fs = finish_group(h, slot.P, slot.g, slot.c, 0.0, false);
                S->pool.release(slot.mem, slot.P.dBytes);
                S->completed.push_back({slot.req.ticket, fs, fs ? h->err : std::string()});
                S->dirty = true;
            } else {
                keep.push_back(std::move(slot));
            }
        }
        S->resident.swap(keep);
        for (auto &slot : S->resident) {  // plans point into the moved requests
            slot.P.in = &slot.req.in;
            slot.P.s = &slot.req.s;
        }
    }
    if (n_resident) *n_resident = (int32_t)S->resident.size();
    if (n_queued) *n_queued = (int32_t)S->queue.size();
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

int32_t ss_server_completed(ss_server *S, int64_t *tickets, ss_status *status, const char **messages,
                            int32_t cap) {
    if (!S || cap <= 0) return 0;
    const int32_t n = std::min<int32_t>(cap, (int32_t)S->completed.size());
    S->reported.assign(S->completed.begin(), S->completed.begin() + n);
    S->completed.erase(S->completed.begin(), S->completed.begin() + n);
    for (int32_t i = 0; i < n; i++) {
        if (tickets) tickets[i] = S->reported[i].ticket;
        if (status) status[i] = S->reported[i].status;
        if (messages) messages[i] = S->reported[i].message.c_str();
    }
    return n;
}

ss_status ss_server_memory(const ss_server *S, int64_t *reserved, int64_t *in_use) {
    if (!S) return SS_ERR_INVALID_ARGUMENT;
    if (reserved) *reserved = (int64_t)S->pool.reserved;
    if (in_use) *in_use = (int64_t)S->pool.inUse;
    return SS_OK;
}

}  // extern "C"

namespace {

ss_status upload_owned(ss_handle *h, const void *src, size_t bytes, void *base, size_t off) {
    if (!bytes) return SS_OK;
    HIPCHK(hipMemcpy((char *)base + off, src, bytes, hipMemcpyHostToDevice));
    return SS_OK;
}

}  // namespace

namespace ss {
ss_status handle_fail(ss_handle *h, ss_status st, const std::string &msg) { return fail(h, st, msg); }
int handle_device(const ss_handle *h) { return h->device; }
hipStream_t handle_stream(const ss_handle *h) { return h->stream; }
void *handle_scratch(ss_handle *h, size_t bytes, ss_status *st) {
    *st = SS_OK;
    if (h->scratchBytes >= bytes) return h->scratch;
    if (h->scratch) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipFree(h->scratch);
        h->scratch = nullptr;
        h->scratchBytes = 0;
    }
    hipError_t e = hipMalloc(&h->scratch, bytes);
    if (e != hipSuccess) {
        h->scratch = nullptr;
        *st = fail(h, e == hipErrorOutOfMemory ? SS_ERR_OUT_OF_MEMORY : SS_ERR_DEVICE,
                   std::string("reconstruction work memory: ") + hipGetErrorString(e));
        return nullptr;
    }
    h->scratchBytes = bytes;
    return h->scratch;
}

size_t handle_scratch_held(const ss_handle *h) { return h->scratchBytes; }

void *pinned_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocPortable) != hipSuccess) {
        (void)hipGetLastError();  // the caller falls back to pageable memory
        return nullptr;
    }
    return p;
}

void pinned_free(void *p) {
    if (p) (void)hipHostFree(p);
}
}  // namespace ss

extern "C" {

ss_status ss_create(int32_t device_id, ss_handle **out) try {
    if (!out) return SS_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SS_ERR_NOT_AVAILABLE;
    if (device_id < 0 || device_id >= n) return SS_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device_id) != hipSuccess) return SS_ERR_DEVICE;
    // the allocation-free resampling's static work area on this device (its
    // entry point reports a failure here)
    (void)ss::resample_scratch_init(device_id);
    ss_handle *h = new ss_handle();
    h->device = device_id;
    if (const char *g = std::getenv("SS_TICK_GRAPH")) h->useGraph = std::atoi(g) != 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_id) == hipSuccess && prop.multiProcessorCount > 0)
        h->numCU = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_id) == hipSuccess && khz > 0)
        h->clockHz = 1e3 * khz;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        ss_destroy(h);
        return SS_ERR_DEVICE;
    }
    *out = h;
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

void ss_destroy(ss_handle *h) {
    if (!h) return;
    hipError_t e = hipSetDevice(h->device);
    if (h->stream) e = hipStreamSynchronize(h->stream);
    for (auto &kv : h->templates)
        if (kv.second.mem) e = hipFree(kv.second.mem);
    for (auto &kv : h->geometries)
        if (kv.second.mem) e = hipFree(kv.second.mem);
    for (auto &p : h->prof.pending) {
        e = hipEventDestroy(p.a);
        e = hipEventDestroy(p.b);
    }
    for (auto ev : h->prof.pool) e = hipEventDestroy(ev);
    if (h->tickExec) e = hipGraphExecDestroy(h->tickExec);
    for (auto ev : h->tickDone)
        if (ev) e = hipEventDestroy(ev);
    if (h->arena) e = hipFree(h->arena);
    if (h->harena) e = hipHostFree(h->harena);
    if (h->scratch) e = hipFree(h->scratch);
    if (h->trialTable) e = hipFree(h->trialTable);
    if (h->trialStage) e = hipHostFree(h->trialStage);
    if (h->stream) e = hipStreamDestroy(h->stream);
    (void)e;
    delete h;
}

const char *ss_last_error(const ss_handle *h) { return h ? h->err.c_str() : "no handle"; }

ss_status ss_tuning_mode(ss_handle *h, int32_t mode) {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    if (mode < kTuneBySize || mode >= ss::kVariants)
        return fail(h, SS_ERR_INVALID_ARGUMENT, "tuning mode must be -2 .. 2");
    h->tuneMode = mode;
    return SS_OK;
}

int32_t ss_tuning_entries(const ss_handle *h, ss_tuning_entry *out, int32_t cap) {
    if (!h) return 0;
    int32_t i = 0;
    for (const auto &kv : h->tuned) {
        if (out && i < cap) out[i] = kv.second;
        i++;
    }
    return i;
}

ss_status ss_upload_template(ss_handle *h, int32_t th, int32_t tw, const double *t,
                             const double *xg, const double *yg, int32_t *id) try {
    if (!h || !id) return SS_ERR_INVALID_ARGUMENT;
    if (th < 2 || tw < 2 || !t || !xg || !yg) return fail(h, SS_ERR_INVALID_ARGUMENT, "template: bad size or NULL array");
    HIPCHK(hipSetDevice(h->device));
    DevTemplate dt;
    TemplateHost T;
    dt.th = th;
    dt.tw = tw;
    convert_template(th, tw, t, xg, yg, true, true, true, T, dt.gradFD);  // every width
    const size_t tn = (size_t)th * tw;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t fb = al(tn * sizeof(float2)), db = al(tn * sizeof(double2)), bb = al(tn * sizeof(uint32_t));
    const int planes = dt.gradFD ? 1 : 3;
    HIPCHK(hipMalloc(&dt.mem, planes * (fb + db) + 2 * bb));
    char *m = (char *)dt.mem;
    // [planes x float2][planes x double2][2 x bf16 gradient pairs]; plane 0
    // the template, planes 1, 2 the gradient pairs
    const size_t oF = 0, oD = planes * fb, oB = planes * (fb + db);
    ss_status st = SS_OK;
    dt.tp = (float2 *)(m + oF);
    dt.tpD = (double2 *)(m + oD);
    dt.gxvB = (uint32_t *)(m + oB);
    dt.gyhB = (uint32_t *)(m + oB + bb);
    st = upload_owned(h, T.tp.data(), tn * sizeof(float2), m, oF);
    if (!st) st = upload_owned(h, T.tpD.data(), tn * sizeof(double2), m, oD);
    if (!st) st = upload_owned(h, T.gxvB.data(), tn * sizeof(uint32_t), m, oB);
    if (!st) st = upload_owned(h, T.gyhB.data(), tn * sizeof(uint32_t), m, oB + bb);
    if (!st && !dt.gradFD) {
        dt.gxv = (float2 *)(m + oF + fb);
        dt.gyh = (float2 *)(m + oF + 2 * fb);
        dt.gxvD = (double2 *)(m + oD + db);
        dt.gyhD = (double2 *)(m + oD + 2 * db);
        st = upload_owned(h, T.gxv.data(), tn * sizeof(float2), m, oF + fb);
        if (!st) st = upload_owned(h, T.gyh.data(), tn * sizeof(float2), m, oF + 2 * fb);
        if (!st) st = upload_owned(h, T.gxvD.data(), tn * sizeof(double2), m, oD + db);
        if (!st) st = upload_owned(h, T.gyhD.data(), tn * sizeof(double2), m, oD + 2 * db);
    }
    if (st) {
        (void)hipFree(dt.mem);
        return st;
    }
    *id = h->next_id++;
    h->templates[*id] = dt;
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

ss_status ss_upload_geometry(ss_handle *h, int32_t fh, int32_t nb, const double *frac,
                             const double *bx, const double *by, const uint8_t *bm, int32_t *id) try {
    if (!h || !id) return SS_ERR_INVALID_ARGUMENT;
    if (fh < 1 || nb < 1 || !frac || !bx || !by || !bm) return fail(h, SS_ERR_INVALID_ARGUMENT, "geometry: bad size or NULL array");
    HIPCHK(hipSetDevice(h->device));
    DevGeometry g;
    GeometryHost Q;
    convert_geometry(fh, nb, frac, bx, by, bm, true, true, g, Q);  // both precisions
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o1 = al(Q.tab.size() * sizeof(float4));
    const size_t o2 = o1 + al(Q.tabD.size() * sizeof(double4));
    const size_t o3 = o2 + al(Q.fn.size() * sizeof(float));
    const size_t total = o3 + al(Q.fnD.size() * sizeof(double));
    HIPCHK(hipMalloc(&g.mem, total + 256));
    char *m = (char *)g.mem;
    ss_status st = SS_OK;
    if (!Q.tab.empty()) {
        g.geo = (float4 *)m;
        g.geoD = (double4 *)(m + o1);
        st = upload_owned(h, Q.tab.data(), Q.tab.size() * sizeof(float4), m, 0);
        if (!st) st = upload_owned(h, Q.tabD.data(), Q.tabD.size() * sizeof(double4), m, o1);
    }
    if (!st && !Q.fn.empty()) {
        g.fnext = (float *)(m + o2);
        g.fnextD = (double *)(m + o3);
        st = upload_owned(h, Q.fn.data(), Q.fn.size() * sizeof(float), m, o2);
        if (!st) st = upload_owned(h, Q.fnD.data(), Q.fnD.size() * sizeof(double), m, o3);
    }
    if (st) {
        (void)hipFree(g.mem);
        return st;
    }
    *id = h->next_id++;
    h->geometries[*id] = g;
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

ss_status ss_release(ss_handle *h, int32_t id) {
    if (!h) return SS_ERR_INVALID_ARGUMENT;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    auto t = h->templates.find(id);
    if (t != h->templates.end()) {
        HIPCHK(hipFree(t->second.mem));
        h->templates.erase(t);
        return SS_OK;
    }
    auto g = h->geometries.find(id);
    if (g != h->geometries.end()) {
        HIPCHK(hipFree(g->second.mem));
        h->geometries.erase(g);
        return SS_OK;
    }
    return fail(h, SS_ERR_INVALID_ARGUMENT, "unknown resident id");
}

ss_status ss_align_group(ss_handle *h, const ss_group_in *in, const ss_settings *s,
                         ss_group_out *out) try {
    return ss_align_batch(h, 1, in, s, out);
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

ss_status ss_align_batch(ss_handle *h, int32_t n, const ss_group_in *in, const ss_settings *s,
                         ss_group_out *out) try {
    if (!h || n < 0 || (n > 0 && (!in || !s || !out))) return SS_ERR_INVALID_ARGUMENT;
    if (n == 0) return SS_OK;
    for (int i = 0; i < n; i++)
        if (s[i].device_id != h->device)
            return fail(h, SS_ERR_INVALID_ARGUMENT, "settings.device_id differs from the handle's device");
    HIPCHK(hipSetDevice(h->device));
    h->err.clear();
    return run_batch(h, n, in, s, out);
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
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

}  // extern "C"
