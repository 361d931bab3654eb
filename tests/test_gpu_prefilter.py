"""GPU parity of prefilter_foralignment and imgapfill_leastsquares
(steady_scan.m:625-643, 525-560) against the oracle (oracle/prefilter.py).

Tolerances: the blur is the same circular convolution computed by FFT in
the oracle and by circulant GEMMs on the GPU, both in FP64: |difference| <=
1e-12 x max|frame| for FP64 output; single and integer outputs are the same
doubles rounded to the class, so they must agree except at a rounding tie
(at most 1 ulp / 1 count, observed none).  Median normalisation alone is
exact (exact selection, same arithmetic).  Gap filling: the GPU's conjugate
gradients (relative residual 1e-13) against the oracle's sparse direct
solve, |difference| <= 1e-9 x max|frame| in FP64 (1 ulp in FP32).
"""
import numpy as np
import pytest

import steadyscan_amd as ssa
from oracle import prefilter as P
from steadyscan_amd import workloads as W

pytestmark = pytest.mark.gpu


def frames(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    a = 200 + 800 * rng.random(shape)
    if np.issubdtype(np.dtype(dtype), np.integer):
        a = np.round(a)
    return a.astype(dtype)


@pytest.mark.parametrize("gsd", [0.6, 1.5, 3.0])
def test_blur_f64(gsd):
    a = frames((37, 50, 3), np.float64, 1)
    got = ssa.prefilter_foralignment(a, gsd, False)
    ref = P.prefilter_foralignment(a, gsd, False)
    err = np.abs(got - ref).max() / np.abs(a).max()
    print(f"gsd {gsd}: max rel err {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("dtype", [np.float32, np.uint16, np.int16])
def test_blur_keeps_the_class(dtype):
    a = frames((64, 48, 2), dtype, 2)
    got = ssa.prefilter_foralignment(a, 1.2, False)
    ref = P.prefilter_foralignment(a, 1.2, False)
    assert got.dtype == np.dtype(dtype)
    if dtype == np.float32:
        ulp = np.spacing(np.abs(ref).astype(np.float32))
        assert np.all(np.abs(got - ref) <= ulp)
    else:
        d = np.abs(got.astype(np.int64) - ref.astype(np.int64))
        print(f"{np.dtype(dtype).name}: {np.count_nonzero(d)} differing pixels", flush=True)
        assert d.max() <= 1 and np.count_nonzero(d) <= 2


def test_median_normalisation_is_exact():
    a = frames((30, 41, 3), np.float64, 3)     # odd count
    np.testing.assert_array_equal(ssa.prefilter_foralignment(a, 0, True), P.prefilter_foralignment(a, 0, True))
    b = frames((30, 40, 2), np.float32, 4)     # even count, single
    np.testing.assert_array_equal(ssa.prefilter_foralignment(b, 0, True), P.prefilter_foralignment(b, 0, True))
    c = a.copy()
    c[5, 5, 1] = np.nan                        # MATLAB median of data with NaN is NaN
    out = ssa.prefilter_foralignment(c, 0, True)
    assert np.isnan(out[:, :, 1]).all() and not np.isnan(out[:, :, 0]).any()


def test_blur_then_median_and_integer_median_rejected():
    a = frames((40, 40, 2), np.float64, 5)
    got = ssa.prefilter_foralignment(a, 1.0, True)
    ref = P.prefilter_foralignment(a, 1.0, True)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    with pytest.raises(ssa.SteadyScanError):
        ssa.prefilter_foralignment(frames((8, 8, 1), np.uint16, 6), 0, True)


def test_workload_frames_512():
    T, fr, geom, mot = W.config1(n_frames=3, size=512, seed=31)
    got = ssa.prefilter_foralignment(fr, 1.0, False)
    ref = P.prefilter_foralignment(fr, 1.0, False)
    d = np.abs(got.astype(np.int64) - ref.astype(np.int64))
    assert d.max() <= 1 and np.count_nonzero(d) <= 5
    t = ssa.prefilter_foralignment(T, 2.0, False)
    assert np.abs(t - P.prefilter_foralignment(T, 2.0, False)).max() <= 1e-12 * np.abs(T).max()


def gappy(shape, seed, band=0, frac=0.0):
    rng = np.random.default_rng(seed)
    a = 50 + 100 * rng.random(shape)
    if band:
        a[:band, :, :] = np.nan
        a[:, -band:, :] = np.nan
    if frac:
        a[rng.random(shape) < frac] = np.nan
    return a


@pytest.mark.parametrize("case", ["border", "holes", "both"])
def test_fill_gaps_f64(case):
    a = gappy((48, 56, 3), 7, band=4 if case != "holes" else 0, frac=0.15 if case != "border" else 0.0)
    got = ssa.imgapfill_leastsquares(a)
    ref = P.imgapfill_leastsquares(a)
    assert not np.isnan(got).any()
    np.testing.assert_array_equal(got[~np.isnan(a)], a[~np.isnan(a)])
    err = np.abs(got - ref).max() / np.nanmax(np.abs(a))
    print(f"{case}: max rel err {err:.2e}", flush=True)
    assert err <= 1e-9


def test_fill_gaps_f32_and_frames_without_data():
    a = gappy((33, 35, 3), 8, band=3, frac=0.1).astype(np.float32)
    a[:, :, 1] = np.nan                         # nothing known: left as NaN
    a[:, :, 2] = 5.0                            # nothing missing
    got = ssa.imgapfill_leastsquares(a)
    ref = P.imgapfill_leastsquares(a)
    assert got.dtype == np.float32
    assert np.isnan(got[:, :, 1]).all()
    np.testing.assert_array_equal(got[:, :, 2], a[:, :, 2])
    ulp = np.spacing(np.abs(ref[:, :, 0]))
    assert np.all(np.abs(got[:, :, 0] - ref[:, :, 0]) <= ulp)


def test_reconstruction_with_inpainting():
    T, fr, geom, mot = W.config0(n_frames=2, size=64, seed=33)
    rng = np.random.default_rng(34)
    nk = geom["nknots"]
    dx = np.cumsum(rng.normal(0, 0.5, (nk, 2)), 0) + 2.0
    dy = np.cumsum(rng.normal(0, 0.5, (nk, 2)), 0) - 2.0
    raw, mask = ssa.invert_bilinear_sampling(fr, geom, dx, dy)
    filled, mask2 = ssa.invert_bilinear_sampling(fr, geom, dx, dy, ibs_inpaint=True)
    assert np.isnan(raw).any() and not np.isnan(filled).any()
    np.testing.assert_array_equal(mask, mask2)
    assert filled.dtype == np.float32          # single frames: single data_corrected
    np.testing.assert_array_equal(filled[mask], raw[mask])
    ref = P.imgapfill_leastsquares(raw)
    assert np.all(np.abs(filled - ref) <= np.spacing(np.abs(ref)))


def test_fill_gaps_large_region():
    # half of a 256 x 256 frame missing: a large Dirichlet problem
    a = gappy((256, 256, 1), 9)
    a[:, 128:, 0] = np.nan
    a[100:140, 60:128, 0] = np.nan
    got = ssa.imgapfill_leastsquares(a)
    ref = P.imgapfill_leastsquares(a)
    err = np.abs(got - ref).max() / np.nanmax(np.abs(a))
    print(f"large region: max rel err {err:.2e}", flush=True)
    assert not np.isnan(got).any() and err <= 1e-9
