// ss_scan.h -- the raster-scan clock of init_settings (steady_scan.m:786-827),
// as functions that the host geometry (ss_geometry.cpp) and the device
// resampling (ss_resample.hip) share, so that both compute the same doubles.
//
// Every function is plain IEEE arithmetic (+, -, *, / on doubles, no
// contraction), which rounds identically on the host and on the GPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "steadyscan.h"

namespace ss {

// MATLAB linspace(a, b, n), element i: a + i*(b-a)/(n-1) with both ends exact.
__host__ __device__ inline double matlab_linspace(double a, double b, int n, int i) {
#pragma clang fp contract(off)
    if (n == 1 || i == n - 1) return b;
    if (i == 0) return a;
    return a + ((double)i * (b - a)) / (double)(n - 1);
}

// Scan parameters of one geometry, everything the per-pixel formulas need.
struct ScanParams {
    int H, W;                           // raw lines, pixels per line
    int Hc, Wc;                         // cropped size
    int cl, cr, cb, ct;                 // cropping: left, right, bottom, top
    int bidirectional;
    int nknots;
    double subrec[4];                   // template coordinates of the frame: x0, x1, y0, y1
};

// Normalised time of raw pixel (line i, column j), 0-based (pixelscantime,
// steady_scan.m:790-795).
__host__ __device__ inline double scan_pixel_time(const ScanParams &s, int i, int j) {
#pragma clang fp contract(off)
    const double npix = (double)s.H * (double)s.W;
    const int jj = (s.bidirectional && (i % 2 == 1)) ? (s.W - 1 - j) : j;
    return (1.0 / npix) / 2 + (double)((long long)i * s.W + jj) / npix;
}

// Time of cropped pixel (i, j).
__host__ __device__ inline double scan_cropped_time(const ScanParams &s, int i, int j) {
    return scan_pixel_time(s, i + s.cb, j + s.cl);
}

// histc bin of t against the n knot times kt (1-based; 0 outside).
__host__ __device__ inline int scan_bin(const double *kt, int n, double t) {
    if (!(t >= kt[0]) || t > kt[n - 1]) return 0;
    if (t == kt[n - 1]) return n;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (kt[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo + 1;
}

// Undisplaced template coordinates (1-based, steady_scan.m:820-823) of
// cropped pixel (i, j).
__host__ __device__ inline double scan_base_x(const ScanParams &s, int j) {
#pragma clang fp contract(off)
    return matlab_linspace(0, 1, s.W, j + s.cl) * (s.subrec[1] - s.subrec[0]) + s.subrec[0];
}
__host__ __device__ inline double scan_base_y(const ScanParams &s, int i) {
#pragma clang fp contract(off)
    return matlab_linspace(0, 1, s.Hc, i) * (s.subrec[3] - s.subrec[2]) + s.subrec[2];
}

// Fraction of time t between knot times kt[k-1] and kt[k] (tfrac).
__host__ __device__ inline double scan_tfrac(const double *kt, int k, double t) {
#pragma clang fp contract(off)
    return (t - kt[k - 1]) / (kt[k] - kt[k - 1]);
}

// Host: fills s and the knot times kt[0 .. nknots) of geometry g without
// allocating (steady_scan.m:786-811, 777-781).  False if g is malformed.
bool scan_setup(const ss_geometry_desc &g, ScanParams &s, double *kt);

}  // namespace ss
