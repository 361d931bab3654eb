"""GPU resampling (resample_frame + pushgrid, steady_scan.m:562-599) against
the oracle's orc_resample_frame, bit for bit: MATLAB's last-write-wins
indexed assignment with repeated targets, single-precision accumulation in
pushgrid's statement order, cropped frames, every element type, frames and
output in host or device memory, and calls that need several launches."""
import ctypes
from ctypes import POINTER, c_double, c_int32

import numpy as np
import pytest

import steadyscan_amd as ssa
from devbuf import DeviceBuffer
from oracle import oracle as O
from steadyscan_amd import workloads as W

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(POINTER(c_double))


def oracle_frames(frames, geom, kx, ky):
    """orc_resample_frame of every (cropped) frame."""
    g = geom["desc"]
    cl, cr, cb, ct = list(g.cropping)
    th, tw = g.template_h, g.template_w
    pk = np.ascontiguousarray(geom["prevknotind"].ravel(order="F"), np.int32)
    bx = np.ascontiguousarray(geom["basex"].ravel(order="F"))
    by = np.ascontiguousarray(geom["basey"].ravel(order="F"))
    tf = np.ascontiguousarray(geom["tfrac"].ravel(order="F"))
    n = frames.shape[2]
    out = np.zeros((th, tw, n), np.float32, order="F")
    for f in range(n):
        data = np.asarray(frames[cb:frames.shape[0] - ct, cl:frames.shape[1] - cr, f], np.float64)
        o = np.zeros(th * tw, np.float32)
        O.lib().orc_resample_frame(_p(np.ascontiguousarray(data.ravel(order="F"))), data.size,
                                   _p(np.ascontiguousarray(kx[:, f])), _p(np.ascontiguousarray(ky[:, f])),
                                   _p(bx), _p(by), pk.ctypes.data_as(POINTER(c_int32)), _p(tf), th, tw,
                                   o.ctypes.data_as(POINTER(ctypes.c_float)))
        out[:, :, f] = o.reshape(th, tw, order="F")
    return out


def knots(nk, n, amp, seed):
    """Knots with large jumps between neighbours: neighbouring lines land on
    the same template cells (repeated targets), and some pixels leave the
    template."""
    rng = np.random.default_rng(seed)
    return (np.asfortranarray(rng.uniform(-amp, amp, (nk, n))), np.asfortranarray(rng.uniform(-amp, amp, (nk, n))))


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.uint16, np.int16])
def test_resample_equals_oracle_cropped(dtype):
    H, Wd, nk, n = 64, 64, 9, 6
    geom = ssa.init_settings(H, Wd, (H, Wd - 3), nknots=nk, cropping=(2, 1, 0, 0))
    T = W.make_template(H, Wd, 20, (2.0, 4.0), offset=10.0, seed=5)
    rng = np.random.default_rng(6)
    fr = np.asfortranarray(np.repeat(T[:, :, None], n, axis=2) * rng.uniform(0.5, 1.5, (1, 1, n)))
    if dtype in (np.uint16, np.int16):
        fr = np.rint(fr)
    fr = np.asfortranarray(fr.astype(dtype))
    kx, ky = knots(nk, n, 4.0, 7)
    ref = oracle_frames(fr, geom, kx, ky)
    out = ssa.resample_frames(fr, geom, kx, ky)
    assert np.isnan(ref).any() and np.isfinite(ref).any()
    assert np.array_equal(out, ref, equal_nan=True)


def test_resample_repeated_targets_last_write_wins():
    # a compressing displacement (knots jump by several pixels within a few
    # lines) sends many data pixels into the same cells
    H, Wd, nk, n = 48, 40, 13, 3
    geom = ssa.init_settings(H, Wd, (H, Wd), nknots=nk)
    rng = np.random.default_rng(11)
    fr = np.asfortranarray(rng.uniform(0, 1000, (H, Wd, n)))
    kx = np.asfortranarray(np.tile(np.where(np.arange(nk) % 2, 3.5, -3.5)[:, None], (1, n)))
    ky = np.asfortranarray(np.tile(np.where(np.arange(nk) % 2, -2.25, 2.25)[:, None], (1, n)))
    ref = oracle_frames(fr, geom, kx, ky)
    out = ssa.resample_frames(fr, geom, kx, ky)
    assert np.array_equal(out, ref, equal_nan=True)


def test_resample_device_io_and_several_launches(handle):
    # 512x512 u16 frames: a launch holds about 45 frames of winner tables and
    # staging in 256 MiB, so 100 frames take three launches
    T, fr, geom, _ = W.config1(n_frames=100)
    nk = geom["nknots"]
    kx, ky = knots(nk, 100, 3.0, 21)
    kx[:, -1] = np.nan  # a frame without knots comes out NaN
    ref = oracle_frames(fr, geom, kx, ky)
    out = ssa.resample_frames(fr, geom, kx, ky, handle=handle)
    assert np.all(np.isnan(out[:, :, -1]))
    assert np.array_equal(out, ref, equal_nan=True)
    dfr = DeviceBuffer(fr.nbytes)
    dfr.upload(fr.ravel(order="F"))
    tp = geom["desc"].template_h * geom["desc"].template_w
    dout = DeviceBuffer(4 * tp * 100)
    ssa.resample_frames((dfr.ptr.value, np.uint16, 100), geom, kx, ky, device_out=dout.ptr.value, handle=handle)
    dev = dout.download(np.float32, tp * 100).reshape(out.shape, order="F")
    assert np.array_equal(dev, ref, equal_nan=True)
    dfr.free()
    dout.free()
