"""Seeded synthetic workloads for tests and benchmarks.

The reference ships no data.  These workloads follow the benchmark
configurations of this project; the choices the configurations leave open
are fixed here:
  - 30 frames/s, displacement defined per scan line, continuous across frames;
  - x and y independent; sinusoid phases 0 (x) and pi/2 (y);
  - 512x512 walks are mean-reverting (reversion 5.6e-4 per line, stationary
    std ~3 px) so long acquisitions stay on the template; 128x128 walks are
    plain random walks;
  - blob amplitudes U[50, 200]; 512x512 background: 4 broad Gaussians
    (sigma 150 px, amplitude U[50, 200]); 512x512 templates scaled to a mean
    of 20 photons; sinusoid amplitude U[2, 8] px per axis; jumps of 5 px with
    random sign at p = 0.002 per line per axis;
  - initial knot estimates are zero.
"""
import ctypes
from ctypes import POINTER, byref, c_double

import numpy as np

from . import _native as N
from .api import image2blockmatrix, init_settings


def make_template(h, w, n_blobs, sigma, amp=(50.0, 200.0), offset=0.0, n_background=0,
                  background_sigma=150.0, background_amp=(50.0, 200.0), target_mean=0.0, seed=1):
    d = N.SynthTemplateDesc(h=h, w=w, n_blobs=n_blobs, sigma_lo=sigma[0], sigma_hi=sigma[1],
                            amp_lo=amp[0], amp_hi=amp[1], offset=offset, n_background=n_background,
                            background_sigma=background_sigma, background_amp_lo=background_amp[0],
                            background_amp_hi=background_amp[1], target_mean=target_mean, seed=seed)
    out = np.zeros((h, w), order="F")
    if N.synth().ss_synth_template(byref(d), out.ctypes.data_as(POINTER(c_double))):
        raise RuntimeError("ss_synth_template failed")
    return out


def make_motion(n_frames, nlines, walk_sigma, walk_revert=0.0, sin_amp=(3.0, 3.0), sin_freq=0.5,
                jump_prob=0.0, jump_size=5.0, frame_rate=30.0, seed=2):
    d = N.SynthMotionDesc(n_frames=n_frames, nlines=nlines, frame_rate=frame_rate,
                          walk_sigma=walk_sigma, walk_revert=walk_revert, sin_amp_x=sin_amp[0],
                          sin_amp_y=sin_amp[1], sin_freq=sin_freq, sin_phase_x=0.0,
                          sin_phase_y=np.pi / 2, jump_prob=jump_prob, jump_size=jump_size, seed=seed)
    dx = np.zeros(n_frames * nlines)
    dy = np.zeros(n_frames * nlines)
    if N.synth().ss_synth_motion(byref(d), dx.ctypes.data_as(POINTER(c_double)),
                                 dy.ctypes.data_as(POINTER(c_double))):
        raise RuntimeError("ss_synth_motion failed")
    return dx.reshape(n_frames, nlines), dy.reshape(n_frames, nlines)


_OUT = {np.float64: 0, np.float32: 1, np.uint16: 2, np.int16: 3}


def make_frames(template, dx, dy, h, w, dtype=np.float32, poisson=True, seed=3, threads=8):
    template = np.asfortranarray(template, dtype=np.float64)
    n = dx.shape[0]
    out = np.zeros((h, w, n), dtype=dtype, order="F")
    dxc = np.ascontiguousarray(dx, dtype=np.float64)
    dyc = np.ascontiguousarray(dy, dtype=np.float64)
    rc = N.synth().ss_synth_frames(template.ctypes.data_as(POINTER(c_double)), template.shape[0],
                                   template.shape[1], h, w, n, dxc.ctypes.data_as(POINTER(c_double)),
                                   dyc.ctypes.data_as(POINTER(c_double)), 1 if poisson else 0, seed,
                                   _OUT[np.dtype(dtype).type], ctypes.c_void_p(out.ctypes.data), threads)
    if rc:
        raise RuntimeError("ss_synth_frames failed")
    return out


def gradients(template):
    """[diff(T,1,2) NaN] and [diff(T,1,1); NaN] (steady_scan.m:178-179)."""
    t = np.asarray(template, dtype=np.float64)
    xg = np.full(t.shape, np.nan)
    yg = np.full(t.shape, np.nan)
    xg[:, :-1] = np.diff(t, axis=1)
    yg[:-1, :] = np.diff(t, axis=0)
    return np.asfortranarray(xg), np.asfortranarray(yg)


def default_args():
    """Scalar arguments of the mex call from default_alignment_settings
    (steady_scan.m:195-208, 836-867), as a dict."""
    return dict(max_iter=120, move_thresh=0.010, halt_correlation=0.995,
                correlation_increase_threshold=-1.0, filter_strength=0.8,
                max_filter_iterations=10, filter_deactivation_threshold=0.01, lambda_=0.0,
                linesearch_reduction_multiplier=0.5, solver_id=2, filter_id=1, gpu_device_id=0)


def group_args(template, frames, geometry, est=None, first=0, count=None, **overrides):
    """The 26 mex arguments for frames[:, :, first:first+count], as a dict
    keyed by the names of steady_scan.m:239-242."""
    n = frames.shape[2] if count is None else count
    nk = geometry["nknots"]
    nb = nk - 1
    xg, yg = gradients(template)
    bm = image2blockmatrix(frames[:, :, first:first + n], geometry)
    ex = np.zeros((nk, n)) if est is None else est[0]
    ey = np.zeros((nk, n)) if est is None else est[1]
    a = dict(flags="D", template_image=np.asfortranarray(template), xgrad=xg, ygrad=yg,
             tfrac=geometry["tfrac_bm"], basex=geometry["basex_bm"], basey=geometry["basey_bm"],
             base_mask=geometry["base_mask_bm"], bmData=bm, est_dx_knots=ex, est_dy_knots=ey,
             group_size=2 * nb, nsetframes=n,
             min_points_per_block=geometry["min_points_per_block"].astype(np.float64))
    a.update(default_args())
    a.update(overrides)
    return a


ARG_ORDER = ["flags", "template_image", "xgrad", "ygrad", "tfrac", "basex", "basey", "base_mask",
             "bmData", "est_dx_knots", "est_dy_knots", "group_size", "nsetframes",
             "min_points_per_block", "max_iter", "move_thresh", "halt_correlation",
             "correlation_increase_threshold", "filter_strength", "max_filter_iterations",
             "filter_deactivation_threshold", "lambda_", "linesearch_reduction_multiplier",
             "solver_id", "filter_id", "gpu_device_id"]


def config0(n_frames=64, size=128, seed=100):
    """1 group of 64 frames, 128x128 float32, knots every 8 lines."""
    T = make_template(size, size, 40, (2.0, 6.0), offset=10.0, seed=seed + 1)
    dx, dy = make_motion(n_frames, size, 0.05, 0.0, (3.0, 3.0), seed=seed + 2)
    fr = make_frames(T, dx, dy, size, size, np.float32, True, seed=seed + 3)
    geom = init_settings(size, size, T.shape, nknots=size // 8 + 1)
    return T, fr, geom, (dx, dy)


def config1(n_frames=1000, size=512, seed=200, dtype=np.uint16):
    """512x512 uint16 stream, knots every 16 lines (groups of 50 frames)."""
    T = make_template(size, size, 400, (1.5, 8.0), n_background=4, background_sigma=150.0,
                      target_mean=20.0, seed=seed + 1)
    rng = np.random.default_rng(seed)
    amp = (float(rng.uniform(2, 8)), float(rng.uniform(2, 8)))
    dx, dy = make_motion(n_frames, size, 0.1, 5.6e-4, amp, jump_prob=0.002, jump_size=5.0,
                         seed=seed + 2)
    fr = make_frames(T, dx, dy, size, size, dtype, True, seed=seed + 3)
    geom = init_settings(size, size, T.shape, nknots=size // 16 + 1)
    return T, fr, geom, (dx, dy)


def config4(duration=4.0, seed=400, n512=16, n1024=2, dtype=np.int16):
    """Acquisition streams for the serving workload: n512 streams of 512x512
    at 30 Hz and n1024 of 1024x1024 at 15 Hz, int16, each with its own
    template and motion (as config1), cut into groups of U{16..64} frames.
    A group arrives when its last frame has been scanned, plus an
    exponential jitter of mean 10 ms.  Returns (streams, groups) where a
    group is a dict(stream, first, count, arrival)."""
    rng = np.random.default_rng(seed)
    streams, groups = [], []
    for k, (size, rate) in enumerate([(512, 30.0)] * n512 + [(1024, 15.0)] * n1024):
        s = seed + 10 * (k + 1)
        n = max(16, int(round(duration * rate)))
        T = make_template(size, size, 400 * (size // 512) ** 2, (1.5, 8.0), n_background=4,
                          background_sigma=150.0 * size / 512, target_mean=20.0, seed=s + 1)
        amp = (float(rng.uniform(2, 8)), float(rng.uniform(2, 8)))
        dx, dy = make_motion(n, size, 0.1, 5.6e-4 * 512 / size, amp, jump_prob=0.002, jump_size=5.0,
                             frame_rate=rate, seed=s + 2)
        fr = make_frames(T, dx, dy, size, size, dtype, True, seed=s + 3)
        streams.append(dict(T=T, frames=fr, size=size, rate=rate))
        first = 0
        while first < n:
            count = min(int(rng.integers(16, 65)), n - first)
            if count < 2:
                break
            arrival = (first + count) / rate + float(rng.exponential(0.010))
            groups.append(dict(stream=k, first=first, count=count, arrival=arrival))
            first += count
    groups.sort(key=lambda g: g["arrival"])
    return streams, groups


def config2_group(size, n_frames, lines_per_knot, seed, dtype=np.uint16):
    """One group of the batched uneven workload: its own template and motion
    (as config1) at size x size, knots every lines_per_knot lines."""
    rng = np.random.default_rng(seed)
    T = make_template(size, size, 400 * size * size // (512 * 512), (1.5, 8.0), n_background=4,
                      background_sigma=150.0 * size / 512, target_mean=20.0, seed=seed + 1)
    amp = (float(rng.uniform(2, 8)), float(rng.uniform(2, 8)))
    dx, dy = make_motion(n_frames, size, 0.1, 5.6e-4 * 512 / size, amp, jump_prob=0.002, jump_size=5.0,
                         seed=seed + 2)
    fr = make_frames(T, dx, dy, size, size, dtype, True, seed=seed + 3)
    return dict(T=T, frames=fr, size=size, lines_per_knot=lines_per_knot)


def config2(n_groups=256, seed=300, dtype=np.uint16):
    """Batched uneven groups: frame counts U{8..96}, half 256x256 and half
    512x512, each group with its own template and motion (as config1) and
    knots every 8 or 16 lines (coin flip).  Returns a list of dicts
    (T, frames, size, lines_per_knot)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_groups):
        size = 256 if k % 2 == 0 else 512
        n = int(rng.integers(8, 97))
        lpk = 8 if rng.random() < 0.5 else 16
        out.append(config2_group(size, n, lpk, seed + 10 * (k + 1), dtype))
    return out


def config2_args(groups, **overrides):
    """The mex arguments of every config2 group (one set of frames each)."""
    calls = []
    for g in groups:
        geom = init_settings(g["size"], g["size"], g["T"].shape, nknots=g["size"] // g["lines_per_knot"] + 1)
        calls.append(group_args(g["T"], g["frames"], geom, **overrides))
    return calls
