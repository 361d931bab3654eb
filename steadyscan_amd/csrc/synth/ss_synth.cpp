// ss_synth.cpp -- seeded synthetic acquisitions (see include/steadyscan_synth.h).
#include "steadyscan_synth.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    double normal() {  // Box-Muller, one value per pair
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    long poisson(double lam) {
        if (!(lam > 0)) return 0;
        if (lam < 10) {  // Knuth: multiply uniforms until below exp(-lam)
            double L = std::exp(-lam), p = 1.0;
            long k = 0;
            do { k++; p *= uniform(); } while (p > L);
            return k - 1;
        }
        // PTRS, W. Hormann, "The transformed rejection method for generating
        // Poisson random variables", Insurance: Math. & Econ. 12 (1993).
        double slam = std::sqrt(lam), loglam = std::log(lam);
        double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
        double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
        for (;;) {
            double U = uniform() - 0.5, V = uniform();
            double us = 0.5 - std::fabs(U);
            long k = (long)std::floor((2 * a / us + b) * U + lam + 0.43);
            if (us >= 0.07 && V <= vr) return k;
            if (k < 0 || (us < 0.013 && V > us)) continue;
            if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
                -lam + k * loglam - std::lgamma((double)k + 1))
                return k;
        }
    }
};

double bilinear_clamped(const double *T, int th, int tw, double x, double y) {
    x = std::min(std::max(x, 0.0), (double)(tw - 1));
    y = std::min(std::max(y, 0.0), (double)(th - 1));
    int ix = std::min((int)x, tw - 2), iy = std::min((int)y, th - 2);
    double fx = x - ix, fy = y - iy;
    auto at = [&](int yy, int xx) { return T[(size_t)xx * th + yy]; };
    return (1 - fx) * ((1 - fy) * at(iy, ix) + fy * at(iy + 1, ix)) +
           fx * ((1 - fy) * at(iy, ix + 1) + fy * at(iy + 1, ix + 1));
}

}  // namespace

extern "C" int ss_synth_template(const ss_synth_template_desc *d, double *out) {
    if (!d || !out || d->h < 2 || d->w < 2) return 1;
    Rng rng(d->seed);
    const int h = d->h, w = d->w;
    std::vector<double> img((size_t)h * w, d->offset);
    auto add_blob = [&](double cx, double cy, double sg, double amp) {
        double r = 5 * sg;
        int x0 = std::max(0, (int)std::floor(cx - r)), x1 = std::min(w - 1, (int)std::ceil(cx + r));
        int y0 = std::max(0, (int)std::floor(cy - r)), y1 = std::min(h - 1, (int)std::ceil(cy + r));
        double inv = 1.0 / (2 * sg * sg);
        for (int x = x0; x <= x1; x++)
            for (int y = y0; y <= y1; y++) {
                double dx = x - cx, dy = y - cy;
                img[(size_t)x * h + y] += amp * std::exp(-(dx * dx + dy * dy) * inv);
            }
    };
    for (int b = 0; b < d->n_background; b++) {
        double cx = rng.uniform(0, w), cy = rng.uniform(0, h);
        double amp = rng.uniform(d->background_amp_lo, d->background_amp_hi);
        double sg = d->background_sigma, inv = 1.0 / (2 * sg * sg);
        for (int x = 0; x < w; x++)
            for (int y = 0; y < h; y++) {
                double dx = x - cx, dy = y - cy;
                img[(size_t)x * h + y] += amp * std::exp(-(dx * dx + dy * dy) * inv);
            }
    }
    for (int b = 0; b < d->n_blobs; b++) {
        double cx = rng.uniform(0, w), cy = rng.uniform(0, h);
        double sg = rng.uniform(d->sigma_lo, d->sigma_hi);
        double amp = rng.uniform(d->amp_lo, d->amp_hi);
        add_blob(cx, cy, sg, amp);
    }
    if (d->target_mean > 0) {
        double s = 0;
        for (double v : img) s += v;
        double scale = d->target_mean / (s / (double)img.size());
        for (double &v : img) v *= scale;
    }
    std::memcpy(out, img.data(), sizeof(double) * img.size());
    return 0;
}

extern "C" int ss_synth_motion(const ss_synth_motion_desc *d, double *dx, double *dy) {
    if (!d || !dx || !dy || d->n_frames < 1 || d->nlines < 1) return 1;
    Rng rng(d->seed);
    double wx = 0, wy = 0;
    const double two_pi = 6.283185307179586;
    size_t n = (size_t)d->n_frames * d->nlines;
    for (size_t k = 0; k < n; k++) {
        double t = (double)k / ((double)d->frame_rate * d->nlines);
        wx = wx * (1 - d->walk_revert) + d->walk_sigma * rng.normal();
        wy = wy * (1 - d->walk_revert) + d->walk_sigma * rng.normal();
        if (d->jump_prob > 0) {
            if (rng.uniform() < d->jump_prob) wx += (rng.uniform() < 0.5 ? -1 : 1) * d->jump_size;
            if (rng.uniform() < d->jump_prob) wy += (rng.uniform() < 0.5 ? -1 : 1) * d->jump_size;
        }
        dx[k] = wx + d->sin_amp_x * std::sin(two_pi * d->sin_freq * t + d->sin_phase_x);
        dy[k] = wy + d->sin_amp_y * std::sin(two_pi * d->sin_freq * t + d->sin_phase_y);
    }
    return 0;
}

extern "C" int ss_synth_frames(const double *tmpl, int32_t th, int32_t tw, int32_t h, int32_t w,
                               int32_t n_frames, const double *dx, const double *dy,
                               int32_t poisson, uint64_t seed, int32_t out_type, void *out,
                               int32_t n_threads) {
    if (!tmpl || !dx || !dy || !out || th < 2 || tw < 2 || h < 1 || w < 1 || n_frames < 1) return 1;
    if (out_type < 0 || out_type > 3) return 1;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 1)
    for (int f = 0; f < n_frames; f++) {
        uint64_t s = seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(f + 1));
        Rng rng(splitmix64(s));
        size_t base = (size_t)f * h * w;
        for (int j = 0; j < w; j++)
            for (int i = 0; i < h; i++) {
                double v = bilinear_clamped(tmpl, th, tw, j + dx[(size_t)f * h + i], i + dy[(size_t)f * h + i]);
                if (poisson) v = (double)rng.poisson(v);
                size_t o = base + (size_t)j * h + i;
                switch (out_type) {
                case 0: ((double *)out)[o] = v; break;
                case 1: ((float *)out)[o] = (float)v; break;
                case 2: ((uint16_t *)out)[o] = (uint16_t)std::min(std::max(std::nearbyint(v), 0.0), 65535.0); break;
                case 3: ((int16_t *)out)[o] = (int16_t)std::min(std::max(std::nearbyint(v), -32768.0), 32767.0); break;
                }
            }
    }
    return 0;
}
