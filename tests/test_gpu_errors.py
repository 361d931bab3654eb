"""Forced failures through the C ABI.

Argument errors return SS_ERR_INVALID_ARGUMENT with the reference's messages
(group size, mexgateway.cpp:186-190; solver id, task.cpp:74-88; filter id,
task.cpp:54-70) before any device work.  Degenerate data (NaN frames, a flat
template whose normal equations are singular, a group without any masked
pixel) is not an error in the reference: its NaN comparisons are false, so
the line search accepts nothing and the call stops after one iteration with
the initial knots and NaN statistics.  The GPU path returns SS_OK with the
oracle's outcome there."""
import ctypes

import numpy as np
import pytest

import steadyscan_amd as ssa
from oracle import oracle as O
from steadyscan_amd import _native as N
from steadyscan_amd import workloads as W
from steadyscan_amd.workloads import ARG_ORDER

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def base():
    T, fr, geom, _ = W.config0(n_frames=6, size=48, seed=5)
    return T, fr, geom


def run_gpu(a):
    info = {}
    out = ssa.align_group_to_template(*[a[k] for k in ARG_ORDER], nargout=6, info=info)
    return out, info


@pytest.mark.parametrize("over,message", [
    (dict(group_size=1026), "groupsize is too big"),
    (dict(solver_id=3), "invalid solver id"),
    (dict(solver_id=-1), "invalid solver id"),
    (dict(filter_id=3), "invalid filter ID"),
    (dict(max_iter=0), "max_iterations"),
])
def test_invalid_arguments(base, over, message):
    T, fr, geom = base
    a = W.group_args(T, fr, geom, max_iter=4)
    a.update(over)
    with pytest.raises(ssa.SteadyScanError, match=message) as e:
        run_gpu(a)
    assert e.value.status == N.SS_ERR_INVALID_ARGUMENT
    assert e.value.identifier == "GPUMC:InvalidArgument"


def test_invalid_arguments_through_the_c_abi(base):
    # the raw C entry point: status code and ss_last_error, nothing launched
    T, fr, geom = base
    h = ssa.Handle(0)
    s = ssa.default_settings()
    s.solver_id = 7
    gi = N.SSGroupIn()
    go = N.SSGroupOut()
    a = W.group_args(T, fr, geom)
    keep = [np.asfortranarray(a[k], dtype=np.float64) for k in ("template_image", "xgrad", "ygrad", "tfrac",
                                                                  "basex", "basey", "est_dx_knots", "est_dy_knots")]
    bm = np.asfortranarray(a["bmData"], dtype=np.float32)
    mask = np.asfortranarray(a["base_mask"], dtype=np.uint8)
    mp = np.ascontiguousarray(a["min_points_per_block"].astype(np.int32))
    P = lambda x, t=ctypes.c_double: x.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
    gi.frame_h, gi.n_blocks = a["basex"].shape
    gi.n_frames, gi.group_size = a["nsetframes"], a["group_size"]
    gi.template_h, gi.template_w = T.shape
    gi.template_image, gi.xgrad, gi.ygrad = P(keep[0]), P(keep[1]), P(keep[2])
    gi.frac, gi.base_x, gi.base_y = P(keep[3]), P(keep[4]), P(keep[5])
    gi.base_mask = P(mask, ctypes.c_uint8)
    gi.data_type = N.SS_F32
    gi.data = ctypes.c_void_p(bm.ctypes.data)
    gi.init_x, gi.init_y = P(keep[6]), P(keep[7])
    gi.min_points = P(mp, ctypes.c_int32)
    st = N.lib().ss_align_group(h.ptr, ctypes.byref(gi), ctypes.byref(s), ctypes.byref(go))
    assert st == N.SS_ERR_INVALID_ARGUMENT
    assert b"invalid solver id" in N.lib().ss_last_error(h.ptr)
    s.solver_id = 2
    gi.group_size = 2048
    st = N.lib().ss_align_group(h.ptr, ctypes.byref(gi), ctypes.byref(s), ctypes.byref(go))
    assert st == N.SS_ERR_INVALID_ARGUMENT
    assert b"groupsize is too big" in N.lib().ss_last_error(h.ptr)
    h.close()


def degenerate(base, kind):
    T, fr, geom = base
    a = W.group_args(T, fr, geom, max_iter=8)
    if kind == "nan frames":
        bm = a["bmData"].astype(np.float64)
        bm[5:9, :] = np.nan
        a["bmData"] = np.asfortranarray(bm)
    elif kind == "flat template":
        Tf = np.full_like(T, 7.0)
        xg, yg = W.gradients(Tf)
        a.update(template_image=Tf, xgrad=xg, ygrad=yg)
    else:  # no pixel of any block is inside the scan's mask
        a["base_mask"] = np.zeros_like(a["base_mask"])
    return a


@pytest.mark.parametrize("flags", ["D", "S"])
@pytest.mark.parametrize("kind", ["nan frames", "flat template", "no masked pixels"])
def test_degenerate_data_follows_the_oracle(base, kind, flags):
    a = degenerate(base, kind)
    a["flags"] = flags
    ref = O.align_group_to_template(a, threads=4)
    out, info = run_gpu(a)
    print(f"{kind} [{flags}]: iterations gpu={int(out[3])} oracle={ref['iterations']}, "
          f"stop {info['stop_criterion']}", flush=True)
    assert int(out[3]) == ref["iterations"]
    assert info["stop_criterion"] == N.STOP_NAMES[ref["stop_criterion"]]
    assert np.array_equal(np.ravel(out[0]), ref["knots_x"]) and np.array_equal(np.ravel(out[1]), ref["knots_y"])
    assert np.array_equal(np.isnan(np.ravel(out[5])), np.isnan(ref["residrms"]))
    assert np.array_equal(np.isnan(np.ravel(out[2])), np.isnan(ref["frame_corr"]))
