"""Pins of the oracle's prefilter_foralignment and imgapfill_leastsquares
(steady_scan.m:625-643, 525-560)."""
import numpy as np
import pytest

from oracle import prefilter as P


def direct_circular_blur(img, gsd):
    """The FFT product written out: a circular convolution on the zero-padded
    grid with the wrapped Gaussian, summed term by term."""
    h, w = img.shape
    g, (s1, s2) = P.gauss_kernel(h, w, gsd)
    pad = np.zeros((s1, s2))
    pad[:h, :w] = img
    out = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            yy = (y - np.arange(s1)) % s1
            xx = (x - np.arange(s2)) % s2
            out[y, x] = np.sum(pad * g[np.ix_(yy, xx)])
    return out


def test_pad_and_kernel():
    assert P.gauss_pad(1.0) == int(np.ceil(np.sqrt(2 * np.log(10000 / np.sqrt(2 * np.pi)))))
    g, (s1, s2) = P.gauss_kernel(10, 7, 1.5)
    assert (s1, s2) == (10 + P.gauss_pad(1.5), 7 + P.gauss_pad(1.5))
    assert abs(g.sum() - 1) < 1e-14
    assert g[0, 0] == g.max()                       # centred on the first pixel, wrapped
    np.testing.assert_allclose(g[1, 0], g[-1, 0])   # symmetric under wrapping
    # separable: g = gy gx^T
    np.testing.assert_allclose(g, np.outer(g[:, 0], g[0, :]) / g[0, 0], rtol=1e-12)


@pytest.mark.parametrize("gsd", [0.7, 1.5])
def test_fft_blur_equals_the_direct_circular_sum(gsd):
    rng = np.random.default_rng(1)
    img = rng.random((9, 13)) * 100
    out = P.prefilter_foralignment(img, gsd, False)
    np.testing.assert_allclose(out, direct_circular_blur(img, gsd), rtol=1e-12, atol=1e-11)


def test_blur_keeps_constants_far_from_the_padding_and_integer_class():
    img = np.full((40, 40), 7.0)
    out = P.prefilter_foralignment(img, 1.0, False)
    assert abs(out[20, 20] - 7.0) < 1e-9             # interior: the kernel sums to 1
    assert out[-1, -1] < 7.0                          # the zero padding darkens the far edge
    u = P.prefilter_foralignment(np.full((16, 16), 1001, np.uint16), 1.0, False)
    assert u.dtype == np.uint16 and u[8, 8] == 1001


def test_integer_class_rounds_half_away_and_saturates():
    v = np.array([-0.5, 0.5, 1.5, 2.5, -2.5, 70000.0, np.nan, 0.49999999999999994])
    np.testing.assert_array_equal(P.to_class(v, np.int16), [-1, 1, 2, 3, -3, 32767, 0, 0])
    np.testing.assert_array_equal(P.to_class(v, np.uint16), [0, 1, 2, 3, 0, 65535, 0, 0])


def test_median_normalisation():
    a = np.arange(1.0, 13.0).reshape(3, 4)[:, :, None] * np.array([1.0, 2.0])
    out = P.prefilter_foralignment(a, 0, True)
    np.testing.assert_allclose(out[:, :, 0], a[:, :, 0] / 6.5)
    np.testing.assert_allclose(out[:, :, 1], a[:, :, 1] / 13.0)
    odd = np.arange(1.0, 10.0).reshape(3, 3)
    np.testing.assert_allclose(P.prefilter_foralignment(odd, 0, True), odd / 5.0)


def test_gap_fill_1d_is_linear_interpolation():
    a = np.array([[1.0, np.nan, np.nan, np.nan, 9.0]])
    np.testing.assert_allclose(P.imgapfill_leastsquares(a), [[1, 3, 5, 7, 9]], atol=1e-12)


def test_gap_fill_reproduces_a_plane_away_from_the_border_and_is_harmonic():
    h, w = 12, 15
    y, x = np.mgrid[0:h, 0:w]
    plane = 3.0 + 0.5 * x - 0.25 * y
    a = plane.copy()
    a[4:8, 5:10] = np.nan
    np.testing.assert_allclose(P.imgapfill_leastsquares(a), plane, atol=1e-10)
    rng = np.random.default_rng(2)
    b = rng.random((h, w))
    miss = rng.random((h, w)) < 0.3
    b[miss] = np.nan
    f = P.imgapfill_leastsquares(b)
    assert not np.isnan(f).any()
    np.testing.assert_array_equal(f[~miss], b[~miss])
    for r, c in zip(*np.nonzero(miss)):
        nb = [(r + dr, c + dc) for dr, dc in ((-1, 0), (1, 0), (0, -1), (0, 1)) if 0 <= r + dr < h and 0 <= c + dc < w]
        assert abs(f[r, c] * len(nb) - sum(f[p] for p in nb)) < 1e-9   # discrete harmonic


def test_gap_fill_skips_frames_without_data_and_rounds_integers():
    a = np.full((4, 4, 2), np.nan)
    a[:, :, 1] = 2.0
    a[1, 1, 1] = np.nan
    f = P.imgapfill_leastsquares(a)
    assert np.isnan(f[:, :, 0]).all() and f[1, 1, 1] == 2.0
