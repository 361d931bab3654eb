"""GPU parity of invert_bilinear_sampling (steady_scan.m:290-404) against the
oracle's restatement on the same inputs.

The bar is bit-exact: the GPU sums the constraints in the oracle's order
(ascending data pixel per template pixel), evaluates the stopping energy in
the same fixed order and runs the same single-precision sweeps, so the
corrected frames, the masks and the per-frame iteration counts must be
identical (NaN in the same places).
"""
import numpy as np
import pytest

import steadyscan_amd as ssa
from oracle import oracle as O
from oracle import driver as OD
from steadyscan_amd import workloads as W
from devbuf import DeviceBuffer

pytestmark = pytest.mark.gpu


def oracle_geometry(nlines, linewidth, th, tw, nknots, cropping=(0, 0, 0, 0), subrec=None):
    (hc, wc), pk, tf, bx, by = OD._geometry(nlines, linewidth, th, tw, nknots, cropping, subrec, False, 0.3)
    s = (hc, wc)
    return dict(prevknotind=pk.reshape(s, order="F"), tfrac=tf.reshape(s, order="F"),
                basex=bx.reshape(s, order="F"), basey=by.reshape(s, order="F"))


def smooth_knots(rng, nk, n, step=0.3, offset=0.0):
    return offset + np.cumsum(rng.normal(0, step, (nk, n)), axis=0)


def compare(frames, dx, dy, th, tw, nknots, cropping=(0, 0, 0, 0), subrec=None, regfac=3e-3, aspect=1.0,
            chunk_frames=0, label=""):
    nl, lw, n = frames.shape
    geom = ssa.init_settings(nl, lw, (th, tw), nknots=nknots, cropping=cropping, subrec=subrec)
    info = {}
    out, mask = ssa.invert_bilinear_sampling(frames, geom, dx, dy, ibs_regfac=regfac,
                                             ibs_template_pixel_aspect_ratio=aspect, chunk_frames=chunk_frames,
                                             info=info)
    og = oracle_geometry(nl, lw, th, tw, nknots, cropping, subrec)
    c = cropping
    cropped = np.asarray(frames, dtype=np.float64)[c[2]:nl - c[3], c[0]:lw - c[1], :]
    ref, rmask, rit = O.invert_bilinear_sampling(cropped, dx, dy, og, (th, tw), regfac=regfac, aspect=aspect)
    diff = np.abs(out.astype(np.float64) - ref)
    ndiff = int(np.sum(~((out == ref) | (np.isnan(out) & np.isnan(ref)))))
    print(f"{label}: iterations gpu {info['iterations'].tolist()} oracle {rit.tolist()}; "
          f"{ndiff} differing pixels, max |diff| {np.nanmax(diff) if ndiff else 0.0}", flush=True)
    np.testing.assert_array_equal(info["iterations"], rit)
    np.testing.assert_array_equal(mask, rmask)
    np.testing.assert_array_equal(out, ref)
    return out, mask, info["iterations"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.uint16, np.int16])
def test_small_stack_all_frame_types(dtype):
    T, fr, geom, mot = W.config0(n_frames=4, size=64, seed=11)
    rng = np.random.default_rng(12)
    nk = geom["nknots"]
    fr = np.asarray(fr)
    if dtype in (np.uint16, np.int16):
        fr = np.clip(np.round(fr), 0, 30000)
    fr = fr.astype(dtype)
    compare(fr, smooth_knots(rng, nk, 4), smooth_knots(rng, nk, 4), 64, 64, nk, label=f"64x64 {np.dtype(dtype).name}")


def test_strong_regulariser_chains_the_check_interval():
    # more iterations per frame, so later frames check every 2..5 iterations
    T, fr, geom, mot = W.config0(n_frames=6, size=48, seed=13)
    rng = np.random.default_rng(14)
    nk = geom["nknots"]
    _, _, it = compare(np.asarray(fr, np.float64), smooth_knots(rng, nk, 6), smooth_knots(rng, nk, 6), 48, 48, nk,
                       regfac=1.0, label="regfac 1")
    assert it.max() >= 3


def test_chunks_carry_the_check_interval():
    T, fr, geom, mot = W.config0(n_frames=5, size=40, seed=15)
    rng = np.random.default_rng(16)
    nk = geom["nknots"]
    dx, dy = smooth_knots(rng, nk, 5), smooth_knots(rng, nk, 5)
    compare(np.asarray(fr, np.float64), dx, dy, 40, 40, nk, regfac=0.3, chunk_frames=2, label="chunks of 2")


def test_frame_off_the_template_is_skipped():
    T, fr, geom, mot = W.config0(n_frames=4, size=32, seed=17)
    rng = np.random.default_rng(18)
    nk = geom["nknots"]
    dx, dy = smooth_knots(rng, nk, 4), smooth_knots(rng, nk, 4)
    dx[:, 1] = 100.0
    dy[:, 3] = np.nan                      # a frame whose knots were never estimated
    out, mask, it = compare(np.asarray(fr, np.float64), dx, dy, 32, 32, nk, regfac=0.5, label="skipped frames")
    assert it[1] == 0 and it[3] == 0 and not mask[:, :, 1].any() and np.isnan(out[:, :, 3]).all()


def test_long_chain_with_skipped_frames_in_one_chunk():
    # 48 frames run at once: most learn their check interval late or never and
    # stop once every interval their predecessor still allows has stopped;
    # skipped frames in the chain pass the interval on
    n = 48
    T, fr, geom, mot = W.config0(n_frames=n, size=48, seed=43)
    rng = np.random.default_rng(44)
    nk = geom["nknots"]
    dx, dy = smooth_knots(rng, nk, n), smooth_knots(rng, nk, n)
    dx[:, 10] = 100.0
    dx[:, 11] = -100.0
    dy[:, 30] = np.nan
    _, _, it = compare(np.asarray(fr, np.float64), dx, dy, 48, 48, nk, regfac=0.3, label="48-frame chain")
    assert it[10] == 0 and it[11] == 0 and it[30] == 0 and len(set(it.tolist())) > 3


def test_rectangular_template_aspect_and_cropping():
    rng = np.random.default_rng(19)
    nl, lw = 40, 56
    fr = 20 + 50 * rng.random((nl, lw, 3))
    nk = 11
    dx, dy = smooth_knots(rng, nk, 3, 0.2), smooth_knots(rng, nk, 3, 0.2)
    compare(fr, dx, dy, 40, 50, nk, cropping=(2, 4, 0, 0), aspect=2.0, label="40x50 aspect 2 cropped")


@pytest.mark.parametrize("chain,chunk", [(5, 0), (5, 3), (4, 8), (1, 0)])
def test_chains_equal_separate_calls(chain, chunk):
    # chain_frames = k: one call over all frames gives what one call per k
    # frames gives (each restarts the check interval at 1), also when the
    # chunks of frames on the device do not line up with the chains
    n = 17
    T, fr, geom, mot = W.config0(n_frames=n, size=40, seed=45)
    rng = np.random.default_rng(46)
    nk = geom["nknots"]
    dx, dy = smooth_knots(rng, nk, n), smooth_knots(rng, nk, n)
    dx[:, 6] = 100.0                       # a skipped frame inside a chain
    fr = np.asarray(fr, np.float64)
    info = {}
    out, mask = ssa.invert_bilinear_sampling(fr, geom, dx, dy, ibs_regfac=0.3, chain_frames=chain,
                                             chunk_frames=chunk, info=info)
    its = []
    for f0 in range(0, n, chain):
        f1 = min(n, f0 + chain)
        inf = {}
        o, m = ssa.invert_bilinear_sampling(fr[:, :, f0:f1], geom, dx[:, f0:f1], dy[:, f0:f1], ibs_regfac=0.3, info=inf)
        np.testing.assert_array_equal(out[:, :, f0:f1], o)
        np.testing.assert_array_equal(mask[:, :, f0:f1], m)
        its += inf["iterations"].tolist()
    print(f"chain {chain} chunk {chunk}: iterations {info['iterations'].tolist()}", flush=True)
    np.testing.assert_array_equal(info["iterations"], its)


@pytest.mark.parametrize("nl,lw,th,tw", [(210, 40, 200, 37), (1110, 10, 1100, 8), (70, 70, 63, 65)])
def test_tall_and_odd_templates(nl, lw, th, tw):
    # row counts that are not multiples of the GPU walks' row blocks (62 rows
    # for the energy, 63 colour rows for the sweep), more blocks than waves
    # (1100 rows), and a template with an odd row count
    rng = np.random.default_rng(th + tw)
    fr = 20 + 50 * rng.random((nl, lw, 2))
    nk = nl // 16 + 1
    dx, dy = smooth_knots(rng, nk, 2, 0.2), smooth_knots(rng, nk, 2, 0.2)
    compare(fr, dx, dy, th, tw, nk, regfac=0.05, label=f"{th}x{tw} template")


def test_subrec_squeezes_several_pixels_per_cell():
    # a scan mapped onto a quarter-size template region: many data pixels per
    # cell, repeated pushgrid targets (the last value is kept)
    rng = np.random.default_rng(20)
    fr = 10 + 90 * rng.random((32, 32, 2))
    nk = 5
    dx, dy = smooth_knots(rng, nk, 2, 0.1), smooth_knots(rng, nk, 2, 0.1)
    compare(fr, dx, dy, 24, 24, nk, subrec=(6, 14, 6, 14), label="subrec")


def test_workload_size_512():
    T, fr, geom, mot = W.config1(n_frames=2, size=512, seed=21)
    rng = np.random.default_rng(22)
    nk = geom["nknots"]
    dx, dy = smooth_knots(rng, nk, 2, 0.3), smooth_knots(rng, nk, 2, 0.3)
    compare(np.asarray(fr), dx, dy, 512, 512, nk, label="512x512 uint16")


def test_device_resident_frames_and_outputs_match_the_host_path():
    T, fr, geom, mot = W.config0(n_frames=3, size=64, seed=23)
    rng = np.random.default_rng(24)
    nk = geom["nknots"]
    fr = np.clip(np.round(np.asarray(fr)), 0, 60000).astype(np.uint16)
    dx, dy = smooth_knots(rng, nk, 3), smooth_knots(rng, nk, 3)
    host_out, host_mask = ssa.invert_bilinear_sampling(fr, geom, dx, dy)
    n = 64 * 64 * 3
    d_in, d_out, d_mask = DeviceBuffer(2 * n), DeviceBuffer(4 * n), DeviceBuffer(n)
    try:
        d_in.upload(np.asfortranarray(fr).ravel(order="F"))
        ret = ssa.invert_bilinear_sampling((d_in.ptr.value, np.uint16, 3), geom, dx, dy,
                                           device_out=(d_out.ptr.value, d_mask.ptr.value))
        assert ret is None
        out = d_out.download(np.float32, n).reshape(64, 64, 3, order="F")
        mask = d_mask.download(np.uint8, n).reshape(64, 64, 3, order="F").astype(bool)
    finally:
        for b in (d_in, d_out, d_mask):
            b.free()
    np.testing.assert_array_equal(mask, host_mask)
    np.testing.assert_array_equal(out, host_out)


def test_bidirectional_scan_and_fractional_knot_spacing():
    # per-pixel geometry tables: odd lines scanned right to left, and 10 knots
    # over 64 lines (equal spacing in time, not on line boundaries)
    rng = np.random.default_rng(40)
    fr = 30 + 60 * rng.random((64, 64, 3))
    nk = 10
    dx, dy = smooth_knots(rng, nk, 3, 0.3), smooth_knots(rng, nk, 3, 0.3)
    nl, lw, n = fr.shape
    geom = ssa.init_settings(nl, lw, (64, 64), nknots=nk, isbidirectional=True)
    info = {}
    out, mask = ssa.invert_bilinear_sampling(fr, geom, dx, dy, info=info)
    (hc, wc), pk, tf, bx, by = OD._geometry(nl, lw, 64, 64, nk, (0, 0, 0, 0), None, True, 0.3)
    s = (hc, wc)
    og = dict(prevknotind=pk.reshape(s, order="F"), tfrac=tf.reshape(s, order="F"),
              basex=bx.reshape(s, order="F"), basey=by.reshape(s, order="F"))
    ref, rmask, rit = O.invert_bilinear_sampling(fr, dx, dy, og, (64, 64))
    np.testing.assert_array_equal(info["iterations"], rit)
    np.testing.assert_array_equal(mask, rmask)
    np.testing.assert_array_equal(out, ref)


def test_frame_1024():
    T, fr, geom, mot = W.config1(n_frames=1, size=1024, seed=41)
    rng = np.random.default_rng(42)
    nk = geom["nknots"]
    dx, dy = smooth_knots(rng, nk, 1, 0.3), smooth_knots(rng, nk, 1, 0.3)
    compare(np.asarray(fr), dx, dy, 1024, 1024, nk, label="1024x1024 uint16")
