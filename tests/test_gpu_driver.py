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


@pytest.mark.parametrize("case", ["blur", "blur_u16", "mednorm"])
def test_driver_with_prefilter(case):
    """estimate_displacements prefilters each set's frames (steady_scan.m:228)."""
    args = dict(W.default_args(), max_iter=12)
    kw = dict(nknots=9, frames_per_set=16)
    T, fr, geom, _ = W.config0(n_frames=20, size=64, seed=65)
    if case == "blur_u16":
        fr = np.asfortranarray(np.rint(fr).astype(np.uint16))
    pre = dict(pregauss=0.0 if case == "mednorm" else 1.0, mednorm=case == "mednorm")
    from oracle import prefilter as P
    Tp = P.prefilter_foralignment(np.asarray(T, np.float64), pre["pregauss"], pre["mednorm"])
    ref = D.estimate_displacements(fr, Tp, dict(args, flags="D"), threads=8, **kw, **pre)
    r = ssa.estimate_displacements(fr, ssa.prefilter_foralignment(np.asarray(T, np.float64), **dict(
        gsd=pre["pregauss"], mednorm=pre["mednorm"])), flags="D", **kw, **pre, **args)
    compare(r, ref)


@pytest.mark.parametrize("case", ["plain", "prefilter_inpaint", "f32_inpaint_crop"])
def test_reconstruction_loop_equals_oracle(case):
    """ss_reconstruct (steady_scan.m:124-134) against the oracle's loop, on the
    same knots: set by set, crop + prefilter + invert_bilinear_sampling (+ fill).
    Without the prefilter the corrected frames are bit-exact; the prefilter
    (FFT vs circulant GEMM) and the fill (direct vs conjugate gradients) leave
    FP64 rounding-level differences that could move a Gauss-Seidel stop, so
    there the bar is 1e-4 relative (observed 1e-14)."""
    T, fr, geom, _ = W.config0(n_frames=7, size=48, seed=66)
    rng = np.random.default_rng(67)
    nk = 7
    dx = np.cumsum(rng.normal(0, 0.4, (nk, 7)), 0) + 1.5
    dy = np.cumsum(rng.normal(0, 0.4, (nk, 7)), 0) - 1.0
    dx[:, -1] = np.nan                          # the stack's last frame is never estimated
    kw = dict(nknots=nk, frames_per_set=3)
    if case == "plain":
        fr = np.asarray(fr, np.float64)
        opts = {}
    elif case == "prefilter_inpaint":
        fr = np.asarray(fr, np.float64)
        opts = dict(pregauss=0.8, ibs_inpaint=True)
    else:
        fr = np.asarray(fr, np.float32)
        kw.update(cropping=(2, 1, 0, 0))
        opts = dict(ibs_inpaint=True)
    th, tw = T.shape
    g = ssa.geometry_desc(fr.shape[0], fr.shape[1], T.shape, nknots=nk, cropping=kw.get("cropping"))
    dc, mk = ssa.reconstruct(fr, g, dx, dy, frames_per_set=3, **opts)
    rdc, rmk = D.reconstruct(fr, dx, dy, (th, tw), nknots=nk, cropping=kw.get("cropping", (0, 0, 0, 0)),
                             frames_per_set=3, **opts)
    np.testing.assert_array_equal(mk, rmk)
    assert np.isnan(dc[:, :, -1]).all() and not mk[:, :, -1].any()
    if case == "plain":
        np.testing.assert_array_equal(dc, rdc)
    else:
        both = ~np.isnan(rdc)
        assert np.array_equal(np.isnan(dc), np.isnan(rdc))
        err = np.abs(dc[both] - rdc[both]).max() / np.abs(rdc[both]).max()
        print(f"{case}: max rel diff {err:.2e}", flush=True)
        assert err <= 1e-4


def test_reconstruction_loop_over_several_launches():
    """More frames than one launch takes (4096 at most, in whole sets): the
    loop equals one invert_bilinear_sampling call per set, bit for bit."""
    n, per = 4300, 1000
    rng = np.random.default_rng(68)
    fr = 50 + 100 * rng.random((16, 16, n))
    nk = 3
    dx = np.cumsum(rng.normal(0, 0.2, (nk, n)), 0)
    dy = np.cumsum(rng.normal(0, 0.2, (nk, n)), 0)
    g = ssa.geometry_desc(16, 16, (16, 16), nknots=nk)
    geom = ssa.init_settings(16, 16, (16, 16), nknots=nk)
    dc, mk = ssa.reconstruct(fr, g, dx, dy, frames_per_set=per)
    for f0 in range(0, n, per):
        f1 = min(n, f0 + per)
        o, m = ssa.invert_bilinear_sampling(fr[:, :, f0:f1], geom, dx[:, f0:f1], dy[:, f0:f1])
        np.testing.assert_array_equal(mk[:, :, f0:f1], m)
        np.testing.assert_array_equal(dc[:, :, f0:f1], o.astype(np.float64))


def test_steady_scan_composes_the_driver_and_keeps_the_extra_channel_slots():
    """steady_scan (steady_scan.m:93-154): prefiltered template, estimation,
    reconstruction; extra channel k lands in slot k (the reference's loop
    writes data_corrected(:, :, frames, ch) for ch = 1..n_extra)."""
    T, fr, geom, _ = W.config0(n_frames=10, size=48, seed=68)
    fr = np.asarray(fr, np.float64)
    extra = np.stack([fr * 2.0, fr + 5.0], axis=3)
    args = dict(W.default_args(), max_iter=8)
    args.pop("flags", None)
    kw = dict(nknots=7, frames_per_set=4, pregauss=0.7, corr_thresh=0.5)
    r = ssa.steady_scan(fr, T, data_extra_channels=extra, **kw, **args)
    assert r["data_corrected"].shape == (48, 48, 10, 3)
    Tp = ssa.prefilter_foralignment(np.asarray(T, np.float64), 0.7, False)
    est = ssa.estimate_displacements(fr, Tp, nknots=7, frames_per_set=4, pregauss=0.7, corr_thresh=0.5, **args)
    for k in ("dx_knots", "dy_knots"):
        np.testing.assert_array_equal(r[k], est[k])
    g = ssa.geometry_desc(48, 48, T.shape, nknots=7)
    ch1, m1 = ssa.reconstruct(extra[:, :, :, 0], g, est["dx_knots"], est["dy_knots"], frames_per_set=4, pregauss=0.7)
    ch2, m2 = ssa.reconstruct(extra[:, :, :, 1], g, est["dx_knots"], est["dy_knots"], frames_per_set=4, pregauss=0.7)
    np.testing.assert_array_equal(r["data_corrected"][:, :, :, 0], ch1)   # the main channel is overwritten
    np.testing.assert_array_equal(r["data_corrected"][:, :, :, 1], ch2)
    assert not r["data_corrected"][:, :, :, 2].any() and not r["mask"][:, :, :, 2].any()
