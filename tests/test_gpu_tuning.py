"""The autotuner of the pixel kernels' launch variants: every variant gives
the same results bit for bit, and with tuning on a new group shape is tuned
once per handle on its first tick (the choice is the fastest variant per
kernel of the trial).  Tuning is off by default: the default picks the
variants by the size of each batch, with the same results."""
import numpy as np
import pytest

import steadyscan_amd as ssa
from steadyscan_amd import _native as N
from steadyscan_amd import workloads as W

pytestmark = pytest.mark.gpu


def same(r, ref):
    for k in ("knots_x", "knots_y", "frame_corr", "residrms", "mask"):
        assert np.array_equal(r[k], ref[k]), k
    assert r["iterations"] == ref["iterations"]
    assert r["linesearch_tries"] == ref["linesearch_tries"]


@pytest.fixture(scope="module")
def calls():
    out = []
    for flags in ("D", "S"):
        T, fr, geom, _ = W.config0(n_frames=10, size=64, seed=91)
        out.append(W.group_args(T, fr, geom, max_iter=10, flags=flags))
    T, fr, geom, _ = W.config1(n_frames=12)
    out.append(W.group_args(T, fr, geom, max_iter=4, flags="D"))
    out.append(dict(W.group_args(T, fr, geom, max_iter=4, flags="S"), template_bf16=True))
    return out


def test_every_variant_gives_the_same_results(calls):
    results = []
    for v in range(3):
        with ssa.Handle(0) as h:
            h.tuning_mode(v)
            results.append(ssa.align_groups(calls, handle=h))
            assert h.tuning() == []  # forced: nothing tuned
    for v in (1, 2):
        for r, ref in zip(results[v], results[0]):
            same(r, ref)
    # the default: a small batch (the small-batch variants and two
    # workgroups per column in the four-candidate line-search rounds)
    with ssa.Handle(0) as h:
        for r, ref in zip(ssa.align_groups(calls, handle=h), results[0]):
            same(r, ref)
        assert h.tuning() == []


def test_new_shapes_are_tuned_once(calls):
    with ssa.Handle(0) as h:
        h.tuning_mode(0)
        forced = ssa.align_groups(calls, handle=h)
    with ssa.Handle(0) as h:
        assert ssa.align_groups(calls[:1], handle=h) and h.tuning() == []  # off by default
        h.tuning_mode(-1)
        tuned = ssa.align_groups(calls, handle=h)
        for r, ref in zip(tuned, forced):
            same(r, ref)
        entries = h.tuning()
        assert len(entries) == len(calls)  # four shapes
        for e in entries:
            for name, k in e["kernels"].items():
                assert all(t > 0 for t in k["trial_us"]), (name, k)
                assert k["chosen"] == int(np.argmin(k["trial_us"]))
            print(e["precision"], e["frame_h"], e["n_blocks"], e["frames"],
                  {n: (k["chosen"], [round(t) for t in k["trial_us"]]) for n, k in e["kernels"].items()}, flush=True)
        # a second call of a known shape is not tuned again
        ssa.align_groups(calls[:1], handle=h)
        assert len(h.tuning()) == len(calls)


def test_server_admission_tunes_and_matches(calls):
    with ssa.Handle(0) as h:
        h.tuning_mode(0)
        ref = ssa.align_groups(calls[:2], handle=h, with_mask=False)
    with ssa.Handle(0) as h:
        h.tuning_mode(-1)
        srv = ssa.Server(handle=h, max_groups=2)
        t = [srv.submit(c) for c in calls[:2]]
        done = {}
        while len(done) < 2:
            srv.step()
            done.update(srv.completed())
        for ti, r in zip(t, ref):
            assert not isinstance(done[ti], Exception)
            for k in ("knots_x", "knots_y", "frame_corr", "residrms"):
                assert np.array_equal(done[ti][k], r[k])
        assert len(h.tuning()) == 2
        srv.close()


def test_tuning_mode_range():
    # -2 (the default: variants by batch size), -1 (tune) and the forced
    # variants 0..2 are accepted; anything else is an invalid argument and
    # leaves the mode unchanged
    with ssa.Handle(0) as h:
        for mode in (-2, -1, 0, 1, 2):
            h.tuning_mode(mode)
        for bad in (-3, 3):
            with pytest.raises(ssa.SteadyScanError) as e:
                h.tuning_mode(bad)
            assert e.value.status == N.SS_ERR_INVALID_ARGUMENT
