"""Continuous batching (ss_server_*): groups admitted while others run give
the same results as separate calls, bit for bit (every group's arithmetic
is independent of the mix and deterministic)."""
import numpy as np
import pytest

import steadyscan_amd as ssa
from steadyscan_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_calls():
    calls = []
    spec = [(9, 64, 8, "D"), (16, 64, 16, "S"), (5, 96, 8, "D"), (13, 64, 8, "S"), (7, 64, 8, "D"),
            (11, 96, 16, "D")]
    for i, (n, size, lpk, fl) in enumerate(spec):
        T, fr, geom, _ = W.config0(n_frames=n, size=size, seed=70 + i)
        geom = ssa.init_settings(size, size, T.shape, nknots=size // lpk + 1)
        calls.append(W.group_args(T, fr, geom, max_iter=12, flags=fl))
    return calls


def same(r, ref):
    for k in ("knots_x", "knots_y", "frame_corr", "residrms"):
        assert np.array_equal(r[k], ref[k]), k
    assert r["iterations"] == ref["iterations"]
    assert r["linesearch_tries"] == ref["linesearch_tries"]


def run_server(srv, calls, arrivals):
    """Submits calls[i] at step arrivals[i]; steps until everything is done."""
    results, tickets, step = {}, {}, 0
    order = sorted(range(len(calls)), key=lambda i: arrivals[i])
    k = 0
    while len(results) < len(calls):
        while k < len(order) and arrivals[order[k]] <= step:
            tickets[srv.submit(calls[order[k]])] = order[k]
            k += 1
        srv.step()
        for t, r in srv.completed().items():
            assert not isinstance(r, Exception), r
            results[tickets[t]] = r
        step += 1
        assert step < 10000
    return results


def test_server_results_equal_single_calls():
    calls = make_calls()
    singles = [ssa.align_groups([a], with_mask=False)[0] for a in calls]
    # arrivals spread over the first steps; at most 3 groups resident
    srv = ssa.Server(max_groups=3)
    res = run_server(srv, calls, arrivals=[0, 0, 1, 3, 3, 6])
    for i, ref in enumerate(singles):
        same(res[i], ref)
    reserved, in_use = srv.memory()
    assert in_use == 0 and reserved > 0
    srv.close()


def test_server_admits_by_memory_budget():
    # four groups of 32 frames at 128x128 (2 MiB of f32 frames each) under a
    # 6 MiB budget: at most two are resident, the others wait for memory
    calls = []
    for i in range(4):
        T, fr, geom, _ = W.config0(n_frames=32, size=128, seed=90 + i)
        calls.append(W.group_args(T, fr, geom, max_iter=8, flags="DS"[i % 2]))
    singles = [ssa.align_groups([a], with_mask=False)[0] for a in calls]
    srv = ssa.Server(hbm_budget_bytes=6 << 20, max_groups=64)
    tickets = {srv.submit(a): i for i, a in enumerate(calls)}
    results, waited = {}, False
    for _ in range(10000):
        nr, nq = srv.step()
        waited |= nq > 0 and nr < 64
        for t, r in srv.completed().items():
            assert not isinstance(r, Exception), r
            results[tickets[t]] = r
        if len(results) == len(calls):
            break
    assert waited
    for i, ref in enumerate(singles):
        same(results[i], ref)
    reserved, in_use = srv.memory()
    assert reserved <= 6 << 20 and in_use == 0
    srv.close()


def test_server_request_that_cannot_be_staged_completes_with_its_error():
    # A page-locked staging limit that host frames cannot fit in: that group
    # completes with SS_ERR_OUT_OF_MEMORY and its own message, never joins the
    # resident set, and leaves the pool as it was; a group with device-resident
    # frames (a few KB of staging) in the same step runs normally.
    from devbuf import DeviceBuffer
    calls = make_calls()[:2]
    singles = [ssa.align_groups([a], with_mask=False)[0] for a in calls]
    bm = calls[1]["bmData"]
    d = DeviceBuffer(bm.nbytes)
    d.upload(bm.ravel(order="F"))
    dev_call = dict(calls[1], bmData=(d.ptr.value, bm.dtype))
    srv = ssa.Server(staging_limit_bytes=64 << 10)
    assert calls[0]["bmData"].nbytes > 64 << 10
    t_host = srv.submit(calls[0])
    t_dev = srv.submit(dev_call)
    results = {}
    nr0 = None
    for _ in range(10000):
        nr, nq = srv.step()
        nr0 = nr if nr0 is None else nr0
        results.update(srv.completed())
        if len(results) == 2:
            break
    assert nr0 == 1  # only the device-frame group was admitted
    err = results[t_host]
    assert isinstance(err, ssa.SteadyScanError) and err.status == 2
    assert "staging" in str(err)
    same(results[t_dev], singles[1])
    reserved, in_use = srv.memory()
    assert in_use == 0
    # the server goes on: a later request that fits is admitted and runs
    t2 = srv.submit(dev_call)
    for _ in range(10000):
        srv.step()
        r = srv.completed()
        if r:
            same(r[t2], singles[1])
            break
    srv.close()
    d.free()
