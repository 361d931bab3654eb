"""Pins of the oracle's estimate_displacements restatement (CPU)."""
import numpy as np
import pytest

from oracle import driver as D
from oracle import oracle as O
from steadyscan_amd import workloads as W


def test_set_starts_follow_set_start_only():
    # steady_scan.m:97-99: starts every frames_per_set frames, sentinel nframes
    assert D.set_starts(1050, 500) == [0, 500, 1000, 1050]
    assert D.set_starts(500, 500) == [0, 500]
    assert D.set_starts(501, 500) == [0, 500, 501]


def test_extend_firstandlast_known_answer():
    x = np.arange(1.0, 10.0)                     # 9 knots, 8 blocks
    used = np.array([0, 0, 1, 1, 1, 0, 0, 0], bool)
    y = D.extend_firstandlast(x, used)
    # knots before block 3 (1-based) take x(3); knots after block 5 take x(5)
    assert np.array_equal(y, [3, 3, 3, 4, 5, 5, 5, 5, 5])
    assert np.array_equal(D.extend_firstandlast(x, np.zeros(8, bool)), x)


def test_single_set_equals_direct_oracle_call():
    # one set of 6 frames: frames 0..4 come from the group call, frame 5 is NaN
    T, fr, geom, _ = W.config0(n_frames=6, size=48, seed=5)
    args = dict(W.default_args(), flags="D", max_iter=6)
    r = D.estimate_displacements(fr, T, args, nknots=geom["nknots"], frames_per_set=500, threads=4)
    a = W.group_args(T, fr, geom, max_iter=6)
    ref = O.align_group_to_template(a, threads=4)
    nk = geom["nknots"]
    assert np.array_equal(r["dx_knots"][:, :5], ref["knots_x"].reshape(6, nk).T[:, :5])
    assert np.array_equal(r["corr"][:5], ref["frame_corr"][:5])
    assert np.all(np.isnan(r["dx_knots"][:, 5])) and np.isnan(r["corr"][5])
    assert np.all(r["iter_used"][:5] == ref["iterations"])


def test_sets_overlap_by_one_frame():
    # 10 frames in sets of 4: groups {0..4}, {4..8}, {8, 9}; frame 4 comes from
    # the second set (the first set's last frame is discarded)
    T, fr, geom, _ = W.config0(n_frames=10, size=48, seed=6)
    args = dict(W.default_args(), flags="D", max_iter=4)
    r = D.estimate_displacements(fr, T, args, nknots=geom["nknots"], frames_per_set=4, threads=4)
    a = W.group_args(T, fr, geom, first=4, count=5, max_iter=4)
    a["est_dx_knots"] = np.zeros((geom["nknots"], 10))  # every set reads the first columns
    a["est_dy_knots"] = np.zeros((geom["nknots"], 10))
    ref = O.align_group_to_template(a, threads=4)
    nk = geom["nknots"]
    assert np.array_equal(r["dx_knots"][:, 4:8], ref["knots_x"].reshape(5, nk).T[:, :4])
    assert np.all(np.isnan(r["dx_knots"][:, 9]))
