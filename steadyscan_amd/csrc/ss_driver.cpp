// ss_driver.cpp -- the steady_scan driver (steady_scan.m:93-288): a whole
// frame stack aligned set by set (estimate_displacements) and reconstructed
// set by set (the loop of steady_scan.m:124-154) through the C ABI.
//
// The reference aligns its sets one after another; here every set of the
// stack is one group of a single batched submission (ss_align_batch), with
// the template and geometry resident.  The set partition, the +1-frame
// overlap, the prefiltering, the unpacking and the per-frame success tests
// follow the reference line by line.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "steadyscan.h"
#include "ss_pinned.h"

namespace {

// dst[k] = src[k] (float to double) for k < n and, when mask_src is given,
// mask_dst[k] = mask_src[k], on a few host threads; the parts whose thread
// could not be started are done on the calling thread.
void widen(const float *src, double *dst, size_t n, const uint8_t *mask_src = nullptr,
           uint8_t *mask_dst = nullptr) {
    // the destination is often fresh memory: its page faults, more than the
    // conversion, set the time, and they too spread over the threads
    const size_t parts = n < ((size_t)1 << 20) ? 1 : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto part = [=](size_t p) {
        const size_t b = n * p / parts, e = n * (p + 1) / parts;
        for (size_t k = b; k < e; k++) dst[k] = src[k];
        if (mask_src) std::memcpy(mask_dst + b, mask_src + b, e - b);
    };
    std::vector<std::thread> pool;
    size_t started = 1;
    try {
        for (; started < parts; started++) pool.emplace_back(part, started);
    } catch (...) {  // no thread available: the rest runs here
    }
    part(0);
    for (size_t p = started; p < parts; p++) part(p);
    for (auto &t : pool) t.join();
}

// set_start (steady_scan.m:97-99), 0-based.  The merge of a short last set
// (steady_scan.m:101-106) only rewrites set_size, and estimate_displacements
// takes each set's frames from set_start (steady_scan.m:216), so it does not
// change the sets that are aligned.
void set_starts(int nframes, int per_set, std::vector<int> &start) {
    start.clear();
    for (int s = 0; s < nframes; s += per_set) start.push_back(s);
}

// extend_firstandlast (steady_scan.m:615-623): knots before the first used
// block take the first used block's left knot, knots after the last used
// block take the last used block's left knot (mask has nknots-1 entries).
void extend_first_last(double *x, int nknots, const std::vector<uint8_t> &used) {
    int first = -1, last = -1;
    for (int b = 0; b < (int)used.size(); b++)
        if (used[b]) {
            if (first < 0) first = b;
            last = b;
        }
    if (first < 0) return;
    const double xf = x[first], xl = x[last];
    for (int k = 0; k < first; k++) x[k] = xf;
    for (int k = last + 1; k < nknots; k++) x[k] = xl;
}

size_t elem_size(ss_dtype t) { return t == SS_F64 ? 8 : t == SS_F32 ? 4 : 2; }

}  // namespace

extern "C" void ss_driver_settings_default(ss_driver_settings *d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->frames_per_set = 500;
    d->min_frames_per_set = 100;
    d->corr_thresh = 0.75;
    d->minblockfrac = 0.25;
    d->frame_groupsize = 2;
    d->pregauss = 0;
    d->mednorm = 0;
}

extern "C" ss_status ss_estimate_displacements(ss_handle *h, const ss_geometry_desc *geom,
                                               const ss_driver_settings *ds, const ss_settings *s,
                                               const double *template_image, ss_dtype type, const void *frames,
                                               int32_t n_frames, const double *est_dx, const double *est_dy,
                                               ss_displacements_out *out) try {
    if (!h || !geom || !ds || !s || !template_image || !frames || !out || n_frames < 1)
        return SS_ERR_INVALID_ARGUMENT;
    if (ds->frames_per_set < 1 || ds->frame_groupsize < 1 || !(ds->pregauss >= 0)) return SS_ERR_INVALID_ARGUMENT;
    int32_t Hc, Wc, fh;
    ss_status st = ss_geometry_size(geom, &Hc, &Wc, &fh);
    if (st) return st;
    const int nk = geom->nknots, nb = nk - 1, th = geom->template_h, tw = geom->template_w;
    const size_t nbm = (size_t)fh * nb;
    // geometry block matrices and the template gradients (steady_scan.m:166-179)
    std::vector<double> frac(nbm), bx(nbm), by(nbm);
    std::vector<uint8_t> bm(nbm);
    std::vector<int32_t> minpts(nk);
    st = ss_geometry_init(geom, nullptr, nullptr, nullptr, nullptr, frac.data(), bx.data(), by.data(), bm.data(),
                          minpts.data());
    if (st) return st;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    std::vector<double> xg((size_t)th * tw, nan), yg((size_t)th * tw, nan);
    auto T = [&](int iy, int ix) { return template_image[(size_t)ix * th + iy]; };
    for (int ix = 0; ix + 1 < tw; ix++)
        for (int iy = 0; iy < th; iy++) xg[(size_t)ix * th + iy] = T(iy, ix + 1) - T(iy, ix);
    for (int ix = 0; ix < tw; ix++)
        for (int iy = 0; iy + 1 < th; iy++) yg[(size_t)ix * th + iy] = T(iy + 1, ix) - T(iy, ix);
    int32_t tid = 0, gid = 0;
    st = ss_upload_template(h, th, tw, template_image, xg.data(), yg.data(), &tid);
    if (st) return st;
    st = ss_upload_geometry(h, fh, nb, frac.data(), bx.data(), by.data(), bm.data(), &gid);
    if (st) {
        ss_release(h, tid);
        return st;
    }
    // outputs start as "not estimated" (steady_scan.m:161-163)
    for (size_t i = 0; i < (size_t)nk * n_frames; i++) out->dx_knots[i] = out->dy_knots[i] = nan;
    for (int j = 0; j < n_frames; j++) {
        if (out->corr) out->corr[j] = nan;
        if (out->iter_used) out->iter_used[j] = nan;
        if (out->residrms) out->residrms[j] = nan;
        if (out->failed) out->failed[j] = 0;
    }
    std::vector<int> start;
    set_starts(n_frames, ds->frames_per_set, start);
    const int nsets = (int)start.size();
    // every set is extended by the next set's first frame (steady_scan.m:216)
    std::vector<int> nset(nsets);
    for (int i = 0; i < nsets; i++) {
        const int next = (i + 1 < nsets) ? start[i + 1] : n_frames;
        nset[i] = std::min(next, n_frames - 1) - start[i] + 1;
    }
    // estimates: the whole arrays are passed and each set reads columns
    // 1..nsetframes of them (steady_scan.m:240, parameterorganizer.cpp:38-39)
    std::vector<double> zeros;
    if (!est_dx || !est_dy) {
        zeros.assign((size_t)nk * n_frames, 0.0);
        est_dx = est_dy = zeros.data();
    }
    const size_t esz = elem_size(type), rawFrame = (size_t)geom->nlines * geom->linewidth * esz;
    // prefilter_foralignment of each set's (uncropped) frames (steady_scan.m:228);
    // a frame's result does not depend on its set, so the stack is filtered once
    std::vector<char> filtered;
    if (ds->pregauss > 0 || ds->mednorm) {
        filtered.resize(rawFrame * n_frames);
        st = ss_prefilter(h, type, frames, filtered.data(), geom->nlines, geom->linewidth, n_frames, ds->pregauss,
                          ds->mednorm, 0);
        if (st) {
            ss_release(h, tid);
            ss_release(h, gid);
            return st;
        }
        frames = filtered.data();
    }
    std::vector<std::vector<char>> packed(nsets);
    std::vector<ss_group_in> gin(nsets);
    std::vector<ss_settings> gs(nsets, *s);
    std::vector<ss_group_out> gout(nsets);
    std::vector<std::vector<double>> kx(nsets), ky(nsets), fc(nsets), rr(nsets);
    std::vector<std::vector<uint8_t>> mk(nsets);
    for (int i = 0; i < nsets; i++) {
        const int n = nset[i];
        packed[i].resize(nbm * n * esz);
        st = ss_pack_frames(geom, type, (const char *)frames + (size_t)start[i] * rawFrame, n, packed[i].data());
        if (st) break;
        ss_group_in &g = gin[i];
        std::memset(&g, 0, sizeof(g));
        g.frame_h = fh;
        g.n_blocks = nb;
        g.n_frames = n;
        g.group_size = nb * ds->frame_groupsize;
        g.template_h = th;
        g.template_w = tw;
        g.template_id = tid;
        g.geometry_id = gid;
        g.data_type = type;
        g.data = packed[i].data();
        g.init_x = est_dx;
        g.init_y = est_dy;
        g.min_points = minpts.data();
        kx[i].resize((size_t)nk * n);
        ky[i].resize((size_t)nk * n);
        fc[i].resize(n);
        rr[i].resize(n);
        mk[i].resize(nbm * n);
        ss_group_out &o = gout[i];
        std::memset(&o, 0, sizeof(o));
        o.knots_x = kx[i].data();
        o.knots_y = ky[i].data();
        o.frame_corr = fc[i].data();
        o.residrms = rr[i].data();
        o.mask = mk[i].data();
    }
    if (!st) st = ss_align_batch(h, nsets, gin.data(), gs.data(), gout.data());
    ss_release(h, tid);
    ss_release(h, gid);
    if (st) return st;
    // unpacking and success tests (steady_scan.m:246-284)
    for (int i = 0; i < nsets; i++) {
        for (int sj = 0; sj + 1 < nset[i]; sj++) {
            const int j = start[i] + sj;
            double *dx = out->dx_knots + (size_t)j * nk, *dy = out->dy_knots + (size_t)j * nk;
            std::memcpy(dx, kx[i].data() + (size_t)sj * nk, nk * sizeof(double));
            std::memcpy(dy, ky[i].data() + (size_t)sj * nk, nk * sizeof(double));
            double corr = fc[i][sj];
            if (out->iter_used) out->iter_used[j] = gout[i].iterations;
            if (out->residrms) out->residrms[j] = rr[i][sj];
            // blocks with at least one masked pixel
            std::vector<uint8_t> used(nb, 0);
            bool any = false;
            for (int b = 0; b < nb; b++) {
                const uint8_t *col = mk[i].data() + ((size_t)sj * nb + b) * fh;
                for (int r = 0; r < fh && !used[b]; r++) used[b] = col[r] && bm[(size_t)b * fh + r];
                any |= used[b] != 0;
            }
            if (!any) {
                if (out->failed) out->failed[j] = 1;
                if (out->corr) out->corr[j] = nan;
                continue;
            }
            if (!used[0] || !used[nb - 1]) {
                extend_first_last(dx, nk, used);
                extend_first_last(dy, nk, used);
            }
            if (out->corr) out->corr[j] = corr;
            // numel(blocksused) / (nknots - 1) is 1 (steady_scan.m:283)
            if (out->failed) out->failed[j] = (!(corr >= ds->corr_thresh) || 1.0 < ds->minblockfrac) ? 1 : 0;
        }
    }
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

extern "C" ss_status ss_reconstruct(ss_handle *h, const ss_geometry_desc *geom, const ss_driver_settings *ds,
                                    const ss_ibs_settings *ibs, ss_dtype type, const void *frames, int32_t n_frames,
                                    const double *dx_knots, const double *dy_knots, double *data_corrected,
                                    uint8_t *mask) try {
    if (!h || !geom || !ds || !ibs || !frames || !dx_knots || !dy_knots || !data_corrected || !mask ||
        n_frames < 1 || ds->frames_per_set < 1 || !(ds->pregauss >= 0))
        return SS_ERR_INVALID_ARGUMENT;
    int32_t Hc, Wc, fh;
    ss_status st = ss_geometry_size(geom, &Hc, &Wc, &fh);
    if (st) return st;
    const int nl = geom->nlines, lw = geom->linewidth, nk = geom->nknots;
    const size_t np = (size_t)geom->template_h * geom->template_w;
    const size_t esz = elem_size(type), rawPix = (size_t)nl * lw;
    const int cl = geom->cropping[0], cb = geom->cropping[2];
    const bool prefilter = ds->pregauss > 0 || ds->mednorm;
    // The sets restart the energy-check interval (each is one call of
    // invert_bilinear_sampling in the reference); here whole sets go to one
    // call as chains of frames_per_set frames, so that a call fills the GPU.
    // A call takes as many whole sets as fit in about 4096 frames, which
    // bounds the host staging buffers.
    ss_ibs_settings is = *ibs;
    is.inpaint = 0;
    is.chain_frames = ds->frames_per_set;
    const int perCall = ds->frames_per_set * std::max(1, 4096 / ds->frames_per_set);
    std::vector<char> cropped, raw;
    std::vector<float> out;
    std::vector<double> out64;
    // The single-precision results and the masks come back through
    // page-locked staging, which the device writes at DMA speed (a copy into
    // pageable memory goes through the runtime's bounce buffers at a fraction
    // of it); without page-locked memory they go straight to pageable memory.
    // SS_DRIVER_STATS=1 prints where the loop's time goes
    const char *statsEnv = std::getenv("SS_DRIVER_STATS");
    const bool stats = statsEnv && statsEnv[0] == '1';
    auto clock = [] { return std::chrono::steady_clock::now(); };
    auto secs = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double>(b - a).count();
    };
    auto t0 = clock();
    struct Stage {
        void *p;
        bool stats;
        ~Stage() {
            const auto a = std::chrono::steady_clock::now();
            ss::pinned_free(p);
            if (stats)
                std::fprintf(stderr, "[driver] staging released in %.4f s\n",
                             std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count());
        }
    } stage{ss::pinned_alloc((sizeof(float) + 1) * np * std::min(perCall, n_frames)), stats};
    if (stats)
        std::fprintf(stderr, "[driver] %s staging of %zu bytes in %.4f s\n", stage.p ? "page-locked" : "no page-locked",
                     (sizeof(float) + 1) * np * std::min(perCall, n_frames), secs(t0, clock()));
    for (int f0 = 0; f0 < n_frames; f0 += perCall) {
        const int n = std::min(perCall, n_frames - f0);
        const char *src = (const char *)frames + rawPix * esz * f0;
        if (prefilter) {
            // prefilter_foralignment(cropimage(frames)) (steady_scan.m:132): crop,
            // filter the cropped frames, and put them back into raw-size frames
            // for the reconstruction, which crops again
            cropped.assign((size_t)Hc * Wc * n * esz, 0);
            raw.assign(rawPix * n * esz, 0);
            for (int f = 0; f < n; f++)
                for (int c = 0; c < Wc; c++)
                    std::memcpy(cropped.data() + (((size_t)f * Wc + c) * Hc) * esz,
                                src + ((size_t)f * rawPix + (size_t)(c + cl) * nl + cb) * esz, (size_t)Hc * esz);
            st = ss_prefilter(h, type, cropped.data(), cropped.data(), Hc, Wc, n, ds->pregauss, ds->mednorm, 0);
            if (st) return st;
            for (int f = 0; f < n; f++)
                for (int c = 0; c < Wc; c++)
                    std::memcpy(raw.data() + ((size_t)f * rawPix + (size_t)(c + cl) * nl + cb) * esz,
                                cropped.data() + (((size_t)f * Wc + c) * Hc) * esz, (size_t)Hc * esz);
            src = raw.data();
        }
        ss_resample_in in;
        std::memset(&in, 0, sizeof(in));
        in.geometry = geom;
        in.n_frames = n;
        in.data_type = type;
        in.frames = src;
        in.knots_x = dx_knots + (size_t)nk * f0;
        in.knots_y = dy_knots + (size_t)nk * f0;
        float *o = (float *)stage.p;
        uint8_t *m = mask + np * f0, *ms = nullptr;  // ms: the staged masks, copied out with the widening
        if (o) {
            ms = (uint8_t *)(o + np * std::min(perCall, n_frames));
        } else {
            out.resize(np * n);
            o = out.data();
        }
        t0 = clock();
        st = ss_invert_bilinear_sampling(h, &in, &is, o, ms ? ms : m, nullptr);
        if (st) return st;
        const auto t1 = clock();
        double *dc = data_corrected + np * f0;
        if (ibs->inpaint && type == SS_F32) {
            // the frames' class is single: filled in double, rounded to single
            st = ss_fill_gaps(h, SS_F32, o, o, geom->template_h, geom->template_w, n, 0);
            if (st) return st;
            widen(o, dc, np * n, ms, m);
        } else if (ibs->inpaint) {
            out64.resize(np * n);
            widen(o, out64.data(), np * n, ms, m);
            st = ss_fill_gaps(h, SS_F64, out64.data(), dc, geom->template_h, geom->template_w, n, 0);
            if (st) return st;
        } else {
            widen(o, dc, np * n, ms, m);
        }
        if (stats)
            std::fprintf(stderr, "[driver] frames %d..%d: reconstruction %.4f s, gap fill + widening %.4f s\n", f0,
                         f0 + n, secs(t0, t1), secs(t1, clock()));
    }
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}
