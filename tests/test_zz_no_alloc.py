"""ss_resample_noalloc (no allocation; static device work area) against
ss_resample through the C ABI, bit for bit, on two calls in a row (the
second smaller), plus its argument checks."""
import numpy as np
import pytest

import steadyscan_amd as ssa
from devbuf import DeviceBuffer
from steadyscan_amd import workloads as W

pytestmark = pytest.mark.gpu


def both(frames, geom, kx, ky, handle):
    """(ss_resample, ss_resample_noalloc) of device-resident frames, each
    into its own device output."""
    g = geom["desc"]
    n = frames.shape[2]
    tp = g.template_h * g.template_w
    dfr = DeviceBuffer(frames.nbytes)
    dfr.upload(frames.ravel(order="F"))
    outs = []
    for noalloc in (False, True):
        dout = DeviceBuffer(4 * tp * n)
        # stale values in the output must be overwritten everywhere
        dout.upload(np.full(tp * n, 12345.0, np.float32))
        ssa.resample_frames((dfr.ptr.value, frames.dtype, n), geom, kx, ky, device_out=dout.ptr.value,
                            handle=handle, noalloc=noalloc)
        outs.append(dout.download(np.float32, tp * n).reshape((g.template_h, g.template_w, n), order="F"))
        dout.free()
    dfr.free()
    return outs


def test_noalloc_equals_resample_on_two_calls(handle):
    # call 1: 30 frames of the 512x512 uint16 workload -- about 15 frames'
    # winner tables fit the 64 MiB work area, so this takes two launches
    T, fr, geom, _ = W.config1(n_frames=30)
    nk = geom["nknots"]
    rng = np.random.default_rng(3)
    kx = np.asfortranarray(rng.uniform(-3, 3, (nk, 30)))
    ky = np.asfortranarray(rng.uniform(-3, 3, (nk, 30)))
    ref, got = both(np.asfortranarray(fr), geom, kx, ky, handle)
    print("call 1: 30 x 512x512 u16, NaN pixels", int(np.isnan(ref).sum()), flush=True)
    assert np.isfinite(ref).any()
    assert np.array_equal(got, ref, equal_nan=True)
    # call 2, smaller: 5 cropped 64x64 float64 frames, bidirectional scan,
    # non-integer knot spacing; the work area still holds call 1's contents
    geom2 = ssa.init_settings(64, 64, (64, 61), nknots=8, cropping=(2, 1, 0, 0), isbidirectional=True)
    T2 = W.make_template(64, 64, 20, (2.0, 4.0), offset=10.0, seed=5)
    fr2 = np.asfortranarray(np.repeat(T2[:, :, None], 5, axis=2) * rng.uniform(0.5, 1.5, (1, 1, 5)))
    kx2 = np.asfortranarray(rng.uniform(-4, 4, (8, 5)))
    ky2 = np.asfortranarray(rng.uniform(-4, 4, (8, 5)))
    ref2, got2 = both(fr2, geom2, kx2, ky2, handle)
    print("call 2: 5 x 64x64 f64 cropped, NaN pixels", int(np.isnan(ref2).sum()), flush=True)
    assert np.isnan(ref2).any() and np.isfinite(ref2).any()
    assert np.array_equal(got2, ref2, equal_nan=True)


def test_noalloc_rejects_host_memory_and_oversized_templates(handle):
    geom = ssa.init_settings(64, 64, (64, 64), nknots=9)
    fr = np.zeros((64, 64, 2), np.float32, order="F")
    k = np.zeros((9, 2), order="F")
    with pytest.raises(ValueError):
        ssa.resample_frames(fr, geom, k, k, handle=handle, noalloc=True)
    # a 2100 x 2100 template needs 70.6 MB of winner tables per frame
    big = ssa.init_settings(64, 64, (2100, 2100), nknots=9)
    dfr = DeviceBuffer(fr.nbytes)
    with pytest.raises(ssa.SteadyScanError, match="static work area"):
        ssa.resample_frames((dfr.ptr.value, np.float32, 2), big, k, k, device_out=dfr.ptr.value, handle=handle,
                            noalloc=True)
    dfr.free()
