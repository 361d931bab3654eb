"""Edge shapes of a group against the FP64 oracle: a single frame (its last
knot is never used and its right neighbour is the 1e6 sentinel,
task_linesearch.cu:52-60), a single knot interval per frame, and the
largest group the reference accepts (group_size 1024, mexgateway.cpp:186-190:
513 knots per frame, one scan line per block).  Precision 'D' to the same
bar as tests/test_gpu_parity.py; 'S' must agree on the iteration count and
stay within the FP32 bar."""
import numpy as np
import pytest

import steadyscan_amd as ssa
from oracle import oracle as O
from steadyscan_amd import workloads as W

from test_gpu_parity import run_gpu

pytestmark = pytest.mark.gpu


def case(n_frames, size, nknots, seed, **over):
    T = W.make_template(size + 16, size + 16, max(8, size // 4), (1.5, 4.0), offset=10.0, seed=seed)
    dx, dy = W.make_motion(n_frames, size, 0.05, 0.0, (1.0, 1.0), seed=seed + 1)
    fr = W.make_frames(T, dx, dy, size, size, np.float32, True, seed=seed + 2)
    geom = ssa.init_settings(size, size, T.shape, nknots=nknots)
    return W.group_args(T, fr, geom, **over)


@pytest.mark.parametrize("flags", ["D", "S"])
@pytest.mark.parametrize("shape", ["one frame", "one interval", "group_size 1024"])
def test_edge_shapes_against_oracle(shape, flags):
    if shape == "one frame":
        a = case(1, 64, 9, seed=501, max_iter=15)
    elif shape == "one interval":
        a = case(6, 48, 2, seed=511, max_iter=15)
    else:
        # one scan line per knot interval leaves each knot weakly determined:
        # the normal equations are ill-conditioned, GPU and oracle agree to
        # ~1e-10 px after two iterations and both wander off with more
        a = case(3, 512, 513, seed=521, max_iter=2)
        assert a["group_size"] == 1024
    a["flags"] = flags
    ref = O.align_group_to_template(a, threads=16)
    out, info = run_gpu(a)
    label = f"{shape} [{flags}]"
    print(f"{label}: iterations gpu={int(out[3])} oracle={ref['iterations']} "
          f"stop {info['stop_criterion']}", flush=True)
    # every frame, the last one included (a one-frame group has no other)
    d = np.concatenate([np.ravel(out[0]) - ref["knots_x"], np.ravel(out[1]) - ref["knots_y"]])
    rms = float(np.sqrt(np.mean(d ** 2)))
    print(f"{label}: knot RMS vs FP64 oracle over all frames = {rms:.3e} px", flush=True)
    if flags == "D":
        assert int(out[3]) == ref["iterations"], label
        assert rms <= 1e-9, label
        # NaN where a frame has no masked pixel: with one interval the last
        # frame's right knot is the sentinel and its block leaves the template
        assert np.allclose(np.ravel(out[5]), ref["residrms"], rtol=1e-9, atol=1e-12, equal_nan=True), label
        assert np.allclose(np.ravel(out[2]), ref["frame_corr"], rtol=1e-9, atol=1e-12, equal_nan=True), label
        assert np.mean(np.asarray(out[4], bool) == ref["mask"].astype(bool)) >= 0.99999, label
    else:
        assert rms < 0.05, label
