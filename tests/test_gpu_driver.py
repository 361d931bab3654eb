"""ss_estimate_displacements (the steady_scan set loop, steady_scan.m:157-288)
against the oracle's restatement of it (oracle/driver.py) on the same frame
stacks.  Precision 'D': knots within 1e-6 px (FP64 rounding can flip a
line-search try whose step is already below the movement threshold; seen:
3e-8 px), correlations and residuals within 1e-6 relative, identical NaN
patterns, iteration counts and failure flags."""
import numpy as np
import pytest

import steadyscan_amd as ssa
from oracle import driver as D
from steadyscan_amd import workloads as W

pytestmark = pytest.mark.gpu


def compare(r, ref):
    for k in ("dx_knots", "dy_knots"):
        a, b = np.asarray(r[k]), np.asarray(ref[k])
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        ok = ~np.isnan(a)
        assert np.allclose(a[ok], b[ok], rtol=0, atol=1e-6), (k, np.abs(a[ok] - b[ok]).max())
    for k in ("corr", "residrms"):
        a, b = np.asarray(r[k]), np.asarray(ref[k])
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        ok = ~np.isnan(a)
        assert np.allclose(a[ok], b[ok], rtol=1e-6, atol=1e-9), k
    assert np.array_equal(np.nan_to_num(r["iter_used"], nan=-1), np.nan_to_num(ref["iter_used"], nan=-1))
    assert np.array_equal(np.asarray(r["failed"], bool), np.asarray(ref["failed"], bool))


@pytest.mark.parametrize("case", ["sets", "subrec", "failures", "crop_u16"])
def test_driver_equals_oracle(case):
    args = dict(W.default_args(), max_iter=12)
    kw = dict(nknots=9, frames_per_set=16)
    if case == "sets":
        T, fr, geom, _ = W.config0(n_frames=40, size=64, seed=61)
    elif case == "subrec":
        # the template holds only the first 48 of the 64 scanned lines: the
        # last two blocks of every frame map outside it and their knots are
        # extended from the last used block
        T, fr, geom, _ = W.config0(n_frames=20, size=64, seed=62)
        T = np.asfortranarray(T[:48, :])
        kw.update(subrec=(1.0, 64.0, 1.0, 64.0))
    elif case == "failures":
        T, fr, geom, _ = W.config0(n_frames=20, size=64, seed=63)
        kw.update(corr_thresh=0.999)
    else:
        T, fr, geom, _ = W.config0(n_frames=24, size=64, seed=64)
        fr = np.asfortranarray(np.rint(fr).astype(np.uint16))
        kw.update(cropping=(2, 3, 0, 0))  # horizontal only: init_settings rejects vertical cropping
    ref = D.estimate_displacements(fr, T, dict(args, flags="D"), threads=8, **kw)
    r = ssa.estimate_displacements(fr, T, flags="D", **kw, **args)
    if case == "subrec":
        est = ~np.isnan(ref["corr"])
        # most frames keep blocks 6 and 7 outside the template: knots extended
        assert np.mean(ref["dx_knots"][6, est] == ref["dx_knots"][5, est]) > 0.5
    if case == "failures":
        assert ref["failed"][:-1].any()
    compare(r, ref)
