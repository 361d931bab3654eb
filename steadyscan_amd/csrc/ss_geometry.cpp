// ss_geometry.cpp -- host-side scan geometry: the parts of steady_scan.m that
// build align_group_to_template's geometry arguments (init_settings,
// steady_scan.m:748-828; the block-matrix packing of estimate_displacements,
// steady_scan.m:166-175, 229-234; image2blockmatrix, steady_scan.m:601-613)
// and default_alignment_settings (steady_scan.m:836-867).
#include <cmath>
#include <new>
#include <cstring>
#include <vector>

#include "ss_scan.h"
#include "steadyscan.h"

namespace ss {

bool scan_setup(const ss_geometry_desc &g, ScanParams &s, double *kt) {
    s.H = g.nlines;
    s.W = g.linewidth;
    s.cl = g.cropping[0];
    s.cr = g.cropping[1];
    s.cb = g.cropping[2];
    s.ct = g.cropping[3];
    s.Hc = s.H - s.cb - s.ct;
    s.Wc = s.W - s.cl - s.cr;
    s.nknots = g.nknots;
    s.bidirectional = g.bidirectional != 0;
    if (s.H < 1 || s.W < 1 || s.Hc < 1 || s.Wc < 1 || s.nknots < 2) return false;
    if (s.cl < 0 || s.cr < 0 || s.cb < 0 || s.ct < 0) return false;
    const int nknots = s.nknots;
    for (int k = 0; k < nknots; k++) kt[k] = 0.0;
    const double dwell = 1.0 / ((double)s.H * (double)s.W);
    const double lines_per_knot = (double)s.Hc / (double)(nknots - 1);
    if (lines_per_knot == std::floor(lines_per_knot)) {
        // knots at the first pixel of every lines_per_knot-th raw line
        // (steady_scan.m:801-804)
        const int L = (int)lines_per_knot;
        for (int k = 0; k + 1 < nknots; k++) {
            double mn = scan_pixel_time(s, k * L, 0);
            for (int j = 1; j < s.W; j++) mn = std::fmin(mn, scan_pixel_time(s, k * L, j));
            kt[k] = mn - dwell / 2;
        }
        kt[nknots - 1] = 1;
        kt[0] = 0;
    } else {
        // equal spacing over the cropped scan (steady_scan.m:805-811)
        double lo = scan_cropped_time(s, 0, 0), hi = lo;
        for (int i = 0; i < s.Hc; i++)
            for (int j = 0; j < s.Wc; j++) {
                const double t = scan_cropped_time(s, i, j);
                lo = std::fmin(lo, t);
                hi = std::fmax(hi, t);
            }
        for (int k = 0; k < nknots; k++) kt[k] = matlab_linspace(lo - dwell / 2, hi + dwell / 2, nknots, k);
        kt[0] = 0;
        kt[nknots - 1] = 1;
    }
    // subrec: given, the frame's own extent when it matches the template
    // (steady_scan.m:777-778), else the whole template (steady_scan.m:781)
    if (g.has_subrec) {
        for (int i = 0; i < 4; i++) s.subrec[i] = g.subrec[i];
    } else if (s.Hc == g.template_h && s.Wc == g.template_w) {
        s.subrec[0] = 1 - g.cropping[0];
        s.subrec[1] = g.template_w + g.cropping[1];
        s.subrec[2] = 1 - g.cropping[2];
        s.subrec[3] = g.template_h + g.cropping[3];
    } else {
        s.subrec[0] = 1;
        s.subrec[1] = g.template_w;
        s.subrec[2] = 1;
        s.subrec[3] = g.template_h;
    }
    return true;
}

}  // namespace ss

namespace {

using ss::ScanParams;

struct ScanTiming {
    ScanParams s{};
    std::vector<double> knot_time;

    bool build(const ss_geometry_desc &g) {
        if (g.nknots < 2) return false;
        knot_time.assign(g.nknots, 0.0);
        return ss::scan_setup(g, s, knot_time.data());
    }
    double cropped_time(int i, int j) const { return ss::scan_cropped_time(s, i, j); }
    int bin(double t) const { return ss::scan_bin(knot_time.data(), s.nknots, t); }
};

}  // namespace

extern "C" void ss_settings_default(ss_settings *s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->max_iterations = 120;
    s->move_threshold = 0.010;
    s->halt_correlation = 0.995;
    s->correlation_increase_threshold = -1;
    s->filter_strength = 0.8;
    s->max_filter_iterations = 10;
    s->filter_correlation_threshold = 0.01;
    s->lambda = 0;
    s->linesearch_reduction_multiplier = 0.5;
    s->solver_id = 2;
    s->filter_id = 1;
    s->device_id = 0;
    s->precision = 'D';
    s->linesearch_candidates = 2;
    s->template_bf16 = 0;
}

extern "C" ss_status ss_geometry_size(const ss_geometry_desc *g, int32_t *hc, int32_t *wc,
                                      int32_t *frame_h) {
    if (!g) return SS_ERR_INVALID_ARGUMENT;
    ScanTiming st;
    if (!st.build(*g)) return SS_ERR_INVALID_ARGUMENT;
    // The default basey is built over the cropped line count and cropped
    // again (steady_scan.m:799, 822-824): vertical cropping mis-sizes it.
    if (st.s.cb + st.s.ct != 0) return SS_ERR_INVALID_ARGUMENT;
    std::vector<int> count(st.s.nknots - 1, 0);
    for (int j = 0; j < st.s.Wc; j++)
        for (int i = 0; i < st.s.Hc; i++) {
            const int k = st.bin(st.cropped_time(i, j));
            if (k <= 0 || k >= st.s.nknots) return SS_ERR_INVALID_ARGUMENT;  // steady_scan.m:814
            count[k - 1]++;
        }
    int fh = 0;
    for (int c : count) fh = c > fh ? c : fh;
    if (hc) *hc = st.s.Hc;
    if (wc) *wc = st.s.Wc;
    if (frame_h) *frame_h = fh;
    return SS_OK;
}

extern "C" ss_status ss_geometry_init(const ss_geometry_desc *g, int32_t *prevknotind,
                                      double *tfrac, double *basex, double *basey,
                                      double *frac_bm, double *base_x_bm, double *base_y_bm,
                                      uint8_t *base_mask_bm, int32_t *min_points) try {
    int32_t Hc, Wc, fh;
    ss_status rc = ss_geometry_size(g, &Hc, &Wc, &fh);
    if (rc != SS_OK) return rc;
    ScanTiming st;
    st.build(*g);
    const int nb = st.s.nknots - 1;
    std::vector<int> fill(nb, 0), per_knot(st.s.nknots, 0);
    const size_t bm = (size_t)fh * nb;
    if (frac_bm) std::memset(frac_bm, 0, sizeof(double) * bm);
    if (base_x_bm) std::memset(base_x_bm, 0, sizeof(double) * bm);
    if (base_y_bm) std::memset(base_y_bm, 0, sizeof(double) * bm);
    if (base_mask_bm) std::memset(base_mask_bm, 0, bm);
    // column-major walk of the cropped image: the order of x(blockind == b)
    for (int j = 0; j < Wc; j++) {
        // meshgrid(linspace(0,1,linewidth)*diff(subrec(1:2)) + subrec(1), ...)
        // cropped by columns (steady_scan.m:820-823)
        const double bx = ss::scan_base_x(st.s, j);
        for (int i = 0; i < Hc; i++) {
            const double by = ss::scan_base_y(st.s, i);
            const double t = st.cropped_time(i, j);
            const int k = st.bin(t);
            const double tf = ss::scan_tfrac(st.knot_time.data(), k, t);
            const size_t o = (size_t)j * Hc + i;
            if (prevknotind) prevknotind[o] = k;
            if (tfrac) tfrac[o] = tf;
            if (basex) basex[o] = bx;
            if (basey) basey[o] = by;
            per_knot[k - 1]++;
            const size_t q = (size_t)(k - 1) * fh + fill[k - 1]++;
            if (frac_bm) frac_bm[q] = tf;
            if (base_x_bm) base_x_bm[q] = bx - 1;   // basex - 1 (steady_scan.m:171)
            if (base_y_bm) base_y_bm[q] = by - 1;
            if (base_mask_bm) base_mask_bm[q] = 1;
        }
    }
    if (min_points) {
        // histc(prevknotind, 1:nknots) * ratio, truncated toward zero by the
        // gateway's int conversion (steady_scan.m:174-175, mexconverter.cpp:40-59)
        for (int k = 0; k < st.s.nknots; k++)
            min_points[k] = (int32_t)((double)per_knot[k] * g->min_blockpoints_ratio);
    }
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

extern "C" ss_status ss_pack_frames(const ss_geometry_desc *g, ss_dtype type, const void *frames,
                                    int32_t n, void *out) try {
    int32_t Hc, Wc, fh;
    ss_status rc = ss_geometry_size(g, &Hc, &Wc, &fh);
    if (rc != SS_OK) return rc;
    if (!frames || !out || n < 1) return SS_ERR_INVALID_ARGUMENT;
    ScanTiming st;
    st.build(*g);
    const int nb = st.s.nknots - 1;
    size_t esz = type == SS_F64 ? 8 : type == SS_F32 ? 4 : 2;
    std::vector<int> bin_of((size_t)Hc * Wc);
    for (int j = 0; j < Wc; j++)
        for (int i = 0; i < Hc; i++) bin_of[(size_t)j * Hc + i] = st.bin(st.cropped_time(i, j)) - 1;
    const size_t per_frame_out = (size_t)fh * nb;
    const char *src = (const char *)frames;
    char *dst = (char *)out;
    std::memset(out, 0, per_frame_out * n * esz);
    for (int f = 0; f < n; f++) {
        std::vector<int> fill(nb, 0);
        const char *fr = src + (size_t)f * g->nlines * g->linewidth * esz;
        char *fo = dst + (size_t)f * per_frame_out * esz;
        for (int j = 0; j < Wc; j++)
            for (int i = 0; i < Hc; i++) {
                const int b = bin_of[(size_t)j * Hc + i];
                // cropimage (steady_scan.m:830-834) then column-major order
                const size_t si = (size_t)(j + g->cropping[0]) * g->nlines + (size_t)(i + g->cropping[2]);
                const size_t di = (size_t)b * fh + fill[b]++;
                std::memcpy(fo + di * esz, fr + si * esz, esz);
            }
    }
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}
