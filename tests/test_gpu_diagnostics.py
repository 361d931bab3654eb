"""Diagnostic outputs of the mex call on the GPU against the FP64 oracle.

Output 11 (flag 'c', task.cpp:286-306): every frame's correlation after
every iteration, taken on the state the global correlation leaves, and
output 3 then holds the last iteration's values (task.cpp:392-394).
Output 12 (flag 'e', task.cpp:362-370): both shifts' errvals after every
iteration that continues.  Precision 'D'; the group's 120 block columns make
the global correlation's swap-back cover rows 0-1 and part of row 2 only."""
import numpy as np
import pytest

import steadyscan_amd as ssa
from oracle import oracle as O
from steadyscan_amd import workloads as W
from steadyscan_amd.workloads import ARG_ORDER

pytestmark = pytest.mark.gpu


def call(a, nargout=12):
    return ssa.align_group_to_template(*[a[k] for k in ARG_ORDER], nargout=nargout)


def case(filter_id, flags="Dce", seed=21):
    T, fr, geom, _ = W.config0(n_frames=10, size=96, seed=seed)
    a = W.group_args(T, fr, geom, max_iter=12, filter_id=filter_id)
    a["flags"] = flags
    return a


@pytest.mark.parametrize("filter_id", [0, 1])
def test_frame_correlations_and_errvals_against_oracle(filter_id):
    a = case(filter_id)
    nF, mi = a["nsetframes"], 12
    fw = a["basex"].shape[1] * nF
    assert fw % 256 != 0 and a["basex"].shape[0] * fw > 256  # a partial swap-back
    ref = O.align_group_to_template(a, threads=8)
    out = call(a)
    it = int(out[3])
    skipped = [k for k in range(it - 1) if not np.any(ref["iteration_errvals"][2 * k:2 * k + 2])]
    print(f"filter {filter_id}: iterations gpu={it} oracle={ref['iterations']}, "
          f"iterations without saved errvals {skipped}", flush=True)
    assert it == ref["iterations"]
    fci, ie = out[10], out[11]
    assert fci.shape == (nF, mi) and ie.shape == (2 * mi, nF)
    assert np.all(fci[:, it:] == 0) and np.all(ie[2 * (it - 1):, :] == 0)
    assert np.max(np.abs(fci - ref["frame_corr_iter"])) <= 1e-9
    assert np.allclose(ie, ref["iteration_errvals"], rtol=1e-9, atol=0)
    # output 3 is the last iteration's correlation, not the final state's
    assert np.array_equal(np.ravel(out[2]), fci[:, it - 1])
    assert np.max(np.abs(np.ravel(out[2]) - ref["frame_corr"])) <= 1e-9
    plain = call(case(filter_id, flags="D"), nargout=6)
    assert np.max(np.abs(np.ravel(plain[2]) - np.ravel(out[2]))) > 1e-3
    # the other outputs do not depend on the flags
    for i in (0, 1, 5):
        assert np.array_equal(plain[i], out[i])


def test_diagnostics_batched_equal_single_calls():
    calls = [case(1, seed=31), case(0, flags="De", seed=32), case(1, flags="Dc", seed=33)]
    res = ssa.align_groups(calls, with_mask=False)
    for a, r in zip(calls, res):
        single = call(a)
        assert np.array_equal(r["frame_corr"], np.ravel(single[2]))
        if "c" in a["flags"]:
            assert np.array_equal(r["frame_corr_iter"], single[10])
        else:
            assert "frame_corr_iter" not in r and single[10].size == 0
        if "e" in a["flags"]:
            assert np.array_equal(r["iteration_errvals"], single[11])
        else:
            assert "iteration_errvals" not in r and single[11].size == 0


def test_frame_correlations_in_precision_s():
    # FP32 pixels: the same outputs, consistent with each other
    a = case(1, flags="Sce")
    out = call(a)
    it = int(out[3])
    fci = out[10]
    assert np.all(np.isfinite(fci[:, :it])) and np.all(fci[:, it:] == 0)
    assert np.array_equal(np.ravel(out[2]), fci[:, it - 1])
    # saved after every iteration that continues; the iteration that turns
    # the filter off continues without saving (task.cpp:309-321)
    pairs = out[11][:2 * (it - 1)].reshape(it - 1, 2, -1)
    assert np.all(pairs[0] > 0)
    assert all(np.all(p > 0) or np.all(p == 0) for p in pairs)


def test_iteration_times_raw_and_half_iteration_knots():
    # outputs 9 and 10 ('t'): seconds per iteration and per stage by the
    # device clock; output 13
    # ('P'): the returned knots in the device layout; output 14: the knots
    # after each half-iteration, against the oracle's
    import ctypes
    import time
    a = case(1, flags="DtP")
    nF, mi = a["nsetframes"], 12
    nb = a["basex"].shape[1]
    P = 2 * (nb * nF + 1)
    ref = O.align_group_to_template(a, threads=8, with_iterations=True)
    t = time.perf_counter()
    out = call(a, nargout=15)
    wall = time.perf_counter() - t
    it = int(out[3])
    assert it == ref["iterations"]
    times = np.ravel(out[8])
    print(f"iteration times (ms): {np.round(times[:it] * 1e3, 3).tolist()}, call {wall * 1e3:.1f} ms", flush=True)
    assert times.shape == (mi,) and np.all(times[:it] > 0) and np.all(times[it:] == 0)
    assert times.sum() < wall
    raw = np.ravel(out[12])
    assert raw.shape == (P,)
    kx, ky = np.zeros((nb + 1) * nF), np.zeros((nb + 1) * nF)
    PD = ctypes.POINTER(ctypes.c_double)
    O.lib().orc_organize_out(raw.ctypes.data_as(PD), nF, nb, kx.ctypes.data_as(PD), ky.ctypes.data_as(PD))
    assert np.array_equal(kx, np.ravel(out[0])) and np.array_equal(ky, np.ravel(out[1]))
    rg = out[13]
    assert rg.shape == (P, 2 * mi)
    assert np.all(rg[:, 2 * it:] == 0)
    assert np.max(np.abs(rg - ref["rg_params"])) <= 1e-9
    ft = np.ravel(out[9])  # seconds per stage in the reference's slots
    print(f"function times (ms): {dict((i, round(v * 1e3, 3)) for i, v in enumerate(ft) if v)}", flush=True)
    assert ft.shape == (14,) and np.all(ft >= 0)
    assert all(ft[i] > 0 for i in (1, 7, 9, 10))
    assert np.all(np.delete(ft, [1, 7, 9, 10]) == 0)
    assert ft.sum() < wall


def test_diagnostics_through_the_server_equal_batched():
    # groups admitted while others run (the table is rebuilt around them)
    # return the same per-iteration outputs as a batched submission
    calls = [case(1, flags="Dce", seed=41), case(0, flags="Dc", seed=42), case(1, flags="De", seed=43)]
    ref = ssa.align_groups(calls, with_mask=False)
    srv = ssa.Server(max_groups=2)
    tickets, done = [], {}
    for a in calls:
        tickets.append(srv.submit(a))
        srv.step()
        done.update(srv.completed())
    while len(done) < len(calls):
        srv.step()
        done.update(srv.completed())
    for t, r in zip(tickets, ref):
        got = done[t]
        assert not isinstance(got, Exception), got
        for k in ("knots_x", "knots_y", "frame_corr", "residrms", "frame_corr_iter", "iteration_errvals"):
            if k in r:
                assert np.array_equal(got[k], r[k]), k
    srv.close()
