"""Writes the golden fixtures of the alignment path: small seeded groups with
all 26 mex arguments and the FP64 oracle's outputs (knots, frame
correlations, residuals, masks, iteration counts, stop criteria, and the
per-iteration diagnostics of flags 'c' and 'e').

    python tests/golden/make_golden.py

The reference ships no test vectors; these come from the oracle, whose
functions are pinned by tests/test_oracle_pins.py and
tests/test_oracle_closed_forms.py.  tests/test_golden.py checks that the
oracle still reproduces them bit for bit; tests/test_gpu_golden.py checks the
GPU path against them without running the oracle."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import steadyscan_amd as ssa  # noqa: E402
from oracle import oracle as O  # noqa: E402
from steadyscan_amd import workloads as W  # noqa: E402

OUTPUTS = ("knots_x", "knots_y", "frame_corr", "residrms", "mask", "iterations", "stop_criterion",
           "frame_corr_iter", "iteration_errvals")


def cases():
    # 48x48 float32, knots every 8 lines, default settings (smearing filter,
    # exact solve), precision D with the per-iteration diagnostics, and S
    T, fr, geom, _ = W.config0(n_frames=6, size=48, seed=701)
    yield "c0_48_f32_D", W.group_args(T, fr, geom, max_iter=15, flags="Dce")
    yield "c0_48_f32_S", W.group_args(T, fr, geom, max_iter=15, flags="S")
    # bidirectional scan (per-pixel geometry table), uint16 frames
    T = W.make_template(48, 48, 16, (2.0, 5.0), offset=10.0, seed=711)
    dx, dy = W.make_motion(5, 48, 0.05, 0.0, (1.5, 1.5), seed=712)
    fr = W.make_frames(T, dx, dy, 48, 48, np.uint16, True, seed=713)
    geom = ssa.init_settings(48, 48, T.shape, nknots=7, isbidirectional=True)
    yield "bidir_48_u16_D", W.group_args(T, fr, geom, max_iter=12, flags="D")
    # 40 lines in 6 knot intervals (padded blocks), warm-start knots,
    # lambda 1e-3, range filter, ILU(0) solve
    T = W.make_template(40, 40, 14, (2.0, 5.0), offset=10.0, seed=721)
    dx, dy = W.make_motion(5, 40, 0.05, 0.0, (1.0, 1.0), seed=722)
    fr = W.make_frames(T, dx, dy, 40, 40, np.float32, True, seed=723)
    geom = ssa.init_settings(40, 40, T.shape, nknots=7)
    nk = geom["nknots"]
    rng = np.random.default_rng(724)
    est = (np.asfortranarray(0.3 * rng.standard_normal((nk, 5))), np.asfortranarray(0.3 * rng.standard_normal((nk, 5))))
    yield "warm_40_f32_D", W.group_args(T, fr, geom, est=est, max_iter=12, flags="D", lambda_=1e-3, filter_id=2,
                                        filter_strength=0.5, solver_id=0)


def save(name, a, ref):
    d = {}
    for k, v in a.items():
        d["arg_" + k] = np.asarray(v)
    for k in OUTPUTS:
        if ref.get(k) is not None:
            d["out_" + k] = np.asarray(ref[k])
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)


def load(path):
    """(args, expected outputs) of a fixture; scalar arguments come back as
    Python numbers and the flags as a string."""
    z = np.load(path, allow_pickle=False)
    a, out = {}, {}
    for k in z.files:
        v = z[k]
        if k.startswith("arg_"):
            key = k[4:]
            if v.ndim == 0:
                v = str(v) if v.dtype.kind == "U" else v.item()
            a[key] = v
        else:
            out[k[4:]] = v
    return a, out


if __name__ == "__main__":
    for name, a in cases():
        ref = O.align_group_to_template(a, threads=8)
        save(name, a, ref)
        print(f"{name}: iterations {ref['iterations']}, stop {ref['stop_criterion']}, "
              f"{os.path.getsize(os.path.join(HERE, name + '.npz')) / 1024:.0f} KiB")
