// ss_rom.cpp -- capture and replay of one align_group_to_template call
// (the role of the 'F' flag's ROM file, mex/mexsimulator/mexsimulator.cpp:
// 52-98), in a portable format of its own:
//
//   header   "SSROM001" (8 bytes), u32 byte-order mark 0x01020304,
//            u32 record count
//   record   4-char tag, u32 element code (1 f64, 2 f32, 3 u16, 4 i16,
//            5 u8, 6 i32), u32 rows, u32 cols, rows*cols elements
//            (column-major, little-endian)
//
// Scalars are 1x1 records; the settings are written field by field, never
// as a struct image.
#include <cstdio>
#include <new>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "steadyscan.h"

struct ss_rom {
    std::map<std::string, std::vector<char>> rec;
    std::map<std::string, uint32_t> code, rows, cols;
};

namespace {

const char kMagic[8] = {'S', 'S', 'R', 'O', 'M', '0', '0', '1'};
enum { E_F64 = 1, E_F32 = 2, E_U16 = 3, E_I16 = 4, E_U8 = 5, E_I32 = 6 };

size_t esize(uint32_t c) {
    switch (c) {
    case E_F64: return 8;
    case E_F32: case E_I32: return 4;
    case E_U16: case E_I16: return 2;
    default: return 1;
    }
}

bool little_endian() {
    const uint32_t v = 1;
    uint8_t b;
    std::memcpy(&b, &v, 1);
    return b == 1;
}

struct Writer {
    std::vector<char> buf;
    uint32_t count = 0;
    void put(const void *p, size_t n) { buf.insert(buf.end(), (const char *)p, (const char *)p + n); }
    void u32(uint32_t v) { put(&v, 4); }
    void rec(const char *tag, uint32_t code, uint32_t rows, uint32_t cols, const void *data) {
        put(tag, 4);
        u32(code);
        u32(rows);
        u32(cols);
        put(data, esize(code) * (size_t)rows * cols);
        count++;
    }
    void f64(const char *tag, double v) { rec(tag, E_F64, 1, 1, &v); }
    void i32(const char *tag, int32_t v) { rec(tag, E_I32, 1, 1, &v); }
};

}  // namespace

extern "C" ss_status ss_rom_write(const char *path, const ss_group_in *in, const ss_settings *s) try {
    if (!path || !in || !s) return SS_ERR_INVALID_ARGUMENT;
    if (!little_endian()) return SS_ERR_NOT_AVAILABLE;
    if (in->template_id > 0 || in->geometry_id > 0 || in->data_on_device)
        return SS_ERR_INVALID_ARGUMENT;  // a ROM holds the arrays themselves
    if (!in->template_image || !in->xgrad || !in->ygrad || !in->frac || !in->base_x || !in->base_y ||
        !in->base_mask || !in->data || !in->init_x || !in->init_y || !in->min_points)
        return SS_ERR_INVALID_ARGUMENT;
    Writer w;
    const uint32_t fh = in->frame_h, nb = in->n_blocks, nF = in->n_frames, th = in->template_h, tw = in->template_w;
    const char prec[2] = {s->precision, 0};
    w.rec("PREC", E_U8, 1, 1, prec);
    w.i32("MXIT", s->max_iterations);
    w.f64("MOVT", s->move_threshold);
    w.f64("HALT", s->halt_correlation);
    w.f64("CINC", s->correlation_increase_threshold);
    w.f64("FSTR", s->filter_strength);
    w.i32("FMXI", s->max_filter_iterations);
    w.f64("FCTH", s->filter_correlation_threshold);
    w.f64("LAMB", s->lambda);
    w.f64("LSRM", s->linesearch_reduction_multiplier);
    w.i32("SOLV", s->solver_id);
    w.i32("FILT", s->filter_id);
    w.i32("DEVI", s->device_id);
    w.i32("LSCK", s->linesearch_candidates);
    w.i32("TBF6", s->template_bf16);
    w.i32("SVFC", s->save_frame_correlations);
    w.i32("GRPS", in->group_size);
    w.i32("NFRM", in->n_frames);
    w.rec("TMPL", E_F64, th, tw, in->template_image);
    w.rec("XGRD", E_F64, th, tw, in->xgrad);
    w.rec("YGRD", E_F64, th, tw, in->ygrad);
    w.rec("FRAC", E_F64, fh, nb, in->frac);
    w.rec("BASX", E_F64, fh, nb, in->base_x);
    w.rec("BASY", E_F64, fh, nb, in->base_y);
    w.rec("BMSK", E_U8, fh, nb, in->base_mask);
    const uint32_t dc = in->data_type == SS_F64 ? E_F64 : in->data_type == SS_F32 ? E_F32
                        : in->data_type == SS_U16 ? E_U16 : E_I16;
    w.rec("DATA", dc, fh, nb * nF, in->data);
    w.rec("ESTX", E_F64, nb + 1, nF, in->init_x);
    w.rec("ESTY", E_F64, nb + 1, nF, in->init_y);
    w.rec("MINP", E_I32, nb + 1, 1, in->min_points);
    FILE *f = std::fopen(path, "wb");
    if (!f) return SS_ERR_INVALID_ARGUMENT;
    const uint32_t bom = 0x01020304u;
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&bom, 4, 1, f) == 1 &&
              std::fwrite(&w.count, 4, 1, f) == 1 && std::fwrite(w.buf.data(), 1, w.buf.size(), f) == w.buf.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? SS_OK : SS_ERR_INVALID_ARGUMENT;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

extern "C" ss_status ss_rom_read(const char *path, ss_rom **out, ss_group_in *in, ss_settings *s) try {
    if (!path || !out || !in || !s) return SS_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!little_endian()) return SS_ERR_NOT_AVAILABLE;
    FILE *f = std::fopen(path, "rb");
    if (!f) return SS_ERR_INVALID_ARGUMENT;
    struct Closer {  // closes the file also when a record allocation throws
        FILE *&f;
        ~Closer() {
            if (f) std::fclose(f);
        }
    } closer{f};
    // bytes the records may still claim: a header never makes the reader
    // allocate more than the file holds
    long fileBytes = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) fileBytes = std::ftell(f);
    if (fileBytes < 16 || std::fseek(f, 0, SEEK_SET) != 0) return SS_ERR_INVALID_ARGUMENT;
    size_t left = (size_t)fileBytes - 16;
    char magic[8];
    uint32_t bom = 0, count = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
              std::fread(&bom, 4, 1, f) == 1 && bom == 0x01020304u && std::fread(&count, 4, 1, f) == 1;
    std::unique_ptr<ss_rom> r(new ss_rom());
    for (uint32_t i = 0; ok && i < count; i++) {
        char tag[5] = {0};
        uint32_t hdr[3];
        ok = std::fread(tag, 1, 4, f) == 4 && std::fread(hdr, 4, 3, f) == 3 && hdr[0] >= E_F64 && hdr[0] <= E_I32;
        if (!ok) break;
        const uint64_t cells = (uint64_t)hdr[1] * hdr[2];  // both below 2^32: no wrap
        if (left < 16 || cells > (left - 16) / esize(hdr[0])) { ok = false; break; }
        const size_t n = esize(hdr[0]) * cells;
        left -= 16 + n;
        std::vector<char> d(n);
        ok = n == 0 || std::fread(d.data(), 1, n, f) == n;
        r->rec[tag] = std::move(d);
        r->code[tag] = hdr[0];
        r->rows[tag] = hdr[1];
        r->cols[tag] = hdr[2];
    }
    std::fclose(f);
    f = nullptr;
    auto has = [&](const char *t, uint32_t c) { return r->rec.count(t) && r->code[t] == c; };
    static const char *f64s[] = {"MOVT", "HALT", "CINC", "FSTR", "FCTH", "LAMB", "LSRM", "TMPL", "XGRD",
                                 "YGRD", "FRAC", "BASX", "BASY", "ESTX", "ESTY"};
    static const char *i32s[] = {"MXIT", "FMXI", "SOLV", "FILT", "DEVI", "LSCK", "TBF6", "GRPS", "NFRM", "MINP"};
    for (const char *t : f64s) ok = ok && has(t, E_F64);
    for (const char *t : i32s) ok = ok && has(t, E_I32);
    ok = ok && has("PREC", E_U8) && has("BMSK", E_U8) && r->rec.count("DATA");
    if (!ok) return SS_ERR_INVALID_ARGUMENT;
    // the shapes must agree with each other, since the returned call reads
    // every array at the sizes the template, block and frame counts give
    auto shape = [&](const char *t, uint64_t rows, uint64_t cols) {
        return r->rows[t] == rows && r->cols[t] == cols;
    };
    const uint64_t th = r->rows["TMPL"], tw = r->cols["TMPL"], fh = r->rows["FRAC"], nb = r->cols["FRAC"];
    int32_t nF = 0;
    std::memcpy(&nF, r->rec["NFRM"].data(), r->rec["NFRM"].size() < 4 ? 0 : 4);
    for (const char *t : f64s) ok = ok && r->rec[t].size() >= 8;
    for (const char *t : i32s) ok = ok && r->rec[t].size() >= 4;
    ok = ok && r->rec["PREC"].size() >= 1 && nF > 0 && shape("XGRD", th, tw) && shape("YGRD", th, tw) &&
         shape("BASX", fh, nb) && shape("BASY", fh, nb) && shape("BMSK", fh, nb) &&
         shape("DATA", fh, nb * (uint64_t)nF) && shape("ESTX", nb + 1, (uint64_t)nF) &&
         shape("ESTY", nb + 1, (uint64_t)nF) && shape("MINP", nb + 1, 1);
    const uint32_t dcode = r->code["DATA"];
    ok = ok && (dcode == E_F64 || dcode == E_F32 || dcode == E_U16 || dcode == E_I16);
    if (!ok) return SS_ERR_INVALID_ARGUMENT;
    auto F = [&](const char *t) { double v; std::memcpy(&v, r->rec[t].data(), 8); return v; };
    auto I = [&](const char *t) { int32_t v; std::memcpy(&v, r->rec[t].data(), 4); return v; };
    ss_settings_default(s);
    s->precision = r->rec["PREC"][0];
    s->max_iterations = I("MXIT");
    s->move_threshold = F("MOVT");
    s->halt_correlation = F("HALT");
    s->correlation_increase_threshold = F("CINC");
    s->filter_strength = F("FSTR");
    s->max_filter_iterations = I("FMXI");
    s->filter_correlation_threshold = F("FCTH");
    s->lambda = F("LAMB");
    s->linesearch_reduction_multiplier = F("LSRM");
    s->solver_id = I("SOLV");
    s->filter_id = I("FILT");
    s->device_id = I("DEVI");
    s->linesearch_candidates = I("LSCK");
    s->template_bf16 = I("TBF6");
    if (r->rec.count("SVFC") && r->rec["SVFC"].size() >= 4) s->save_frame_correlations = I("SVFC");
    std::memset(in, 0, sizeof(*in));
    in->frame_h = (int32_t)r->rows["FRAC"];
    in->n_blocks = (int32_t)r->cols["FRAC"];
    in->n_frames = I("NFRM");
    in->group_size = I("GRPS");
    in->template_h = (int32_t)r->rows["TMPL"];
    in->template_w = (int32_t)r->cols["TMPL"];
    auto P = [&](const char *t) { return (const double *)r->rec[t].data(); };
    in->template_image = P("TMPL");
    in->xgrad = P("XGRD");
    in->ygrad = P("YGRD");
    in->frac = P("FRAC");
    in->base_x = P("BASX");
    in->base_y = P("BASY");
    in->base_mask = (const uint8_t *)r->rec["BMSK"].data();
    const uint32_t dc = r->code["DATA"];
    in->data_type = dc == E_F64 ? SS_F64 : dc == E_F32 ? SS_F32 : dc == E_U16 ? SS_U16 : SS_I16;
    in->data = r->rec["DATA"].data();
    in->init_x = P("ESTX");
    in->init_y = P("ESTY");
    in->min_points = (const int32_t *)r->rec["MINP"].data();
    *out = r.release();
    return SS_OK;
} catch (const std::bad_alloc &) {
    return SS_ERR_OUT_OF_MEMORY;  // host memory; C callers never see an exception
} catch (...) {
    return SS_ERR_DEVICE;
}

extern "C" void ss_rom_free(ss_rom *r) { delete r; }
