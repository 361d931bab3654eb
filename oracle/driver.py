"""CPU restatement of SteadyScan's driver: estimate_displacements
(steady_scan.m:157-288) over the FP64 oracle of align_group_to_template, and
the reconstruction loop (steady_scan.m:124-134) over the oracle of
invert_bilinear_sampling.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the library's
ss_estimate_displacements; never called by the product path.

Follows, line by line: the set starts (steady_scan.m:97-99; the short-last-set
merge at 101-106 rewrites set_size only, which estimate_displacements does not
read), the +1-frame set extension (:216), the block-matrix packing of every
frame after cropimage (:220-234, 830-834), the whole estimate arrays passed to
every set (:240), the unpacking of frames 1..nsetframes-1 (:246-260), the
blocks-used test (:263; blockmatrix2image is not defined in the tree, Q23, and
is taken as the inverse of image2blockmatrix), extend_firstandlast (:270-274,
615-623) and the failure test (:276-283, where numel(blocksused) makes the
minblockfrac term 1).
"""
import ctypes
from ctypes import POINTER, c_int32

import numpy as np

from . import oracle as O
from . import prefilter as P


def set_starts(nframes, frames_per_set):
    """0-based set starts plus the end sentinel (steady_scan.m:97-99)."""
    return list(range(0, nframes, frames_per_set)) + [nframes]


def extend_firstandlast(x, mask):
    """steady_scan.m:615-623 (mask has one entry per block, x one per knot)."""
    mask = np.asarray(mask, bool)
    if not mask.any():
        return x
    first = int(np.argmax(mask))
    last = len(mask) - 1 - int(np.argmax(mask[::-1]))
    y = np.array(x, dtype=float)
    y[:first] = x[first]
    y[last + 1:] = x[last]
    return y


def _geometry(nlines, linewidth, th, tw, nknots, cropping, subrec, bidirectional, ratio):
    g = O.OrcGeometryIn(nlines=nlines, linewidth=linewidth, template_h=th, template_w=tw, nknots=nknots,
                        bidirectional=1 if bidirectional else 0, min_blockpoints_ratio=ratio)
    for i, v in enumerate(cropping):
        g.cropping[i] = int(v)
    if subrec is not None:
        g.has_subrec = 1
        for i, v in enumerate(subrec):
            g.subrec[i] = float(v)
    hc, wc = c_int32(), c_int32()
    if O.lib().orc_geometry(ctypes.byref(g), None, None, None, None, ctypes.byref(hc), ctypes.byref(wc)):
        raise ValueError("invalid geometry")
    n = hc.value * wc.value
    pk = np.zeros(n, np.int32)
    tf, bx, by = np.zeros(n), np.zeros(n), np.zeros(n)
    O.lib().orc_geometry(ctypes.byref(g), pk.ctypes.data_as(POINTER(c_int32)), O._p(tf), O._p(bx), O._p(by),
                         ctypes.byref(hc), ctypes.byref(wc))
    return (hc.value, wc.value), pk, tf, bx, by


def _image2blockmatrix(x, pk, nb):
    """steady_scan.m:601-613 through the oracle."""
    fh = O.lib().orc_image2blockmatrix(None, pk.ctypes.data_as(POINTER(c_int32)), len(pk), nb, None, None)
    y = np.zeros(fh * nb)
    m = np.zeros(fh * nb, np.uint8)
    O.lib().orc_image2blockmatrix(O._p(np.ascontiguousarray(x, dtype=np.float64)), pk.ctypes.data_as(POINTER(c_int32)),
                                  len(pk), nb, O._p(y), m.ctypes.data_as(POINTER(ctypes.c_uint8)))
    return y.reshape(fh, nb, order="F"), m.reshape(fh, nb, order="F")


def estimate_displacements(frames, template, align_args, nknots=33, cropping=(0, 0, 0, 0), subrec=None,
                           bidirectional=False, min_blockpoints_ratio=0.3, frames_per_set=500,
                           corr_thresh=0.75, minblockfrac=0.25, frame_groupsize=2, est_dx=None, est_dy=None,
                           threads=8, pregauss=0.0, mednorm=False):
    """frames: nlines x linewidth x n; align_args: the scalar mex arguments
    (flags, max_iter, move_thresh, ...) as in workloads.default_args().
    Returns dict(dx_knots, dy_knots, corr, iter_used, residrms, failed)."""
    H, Wd, n = frames.shape
    th, tw = template.shape
    nb = nknots - 1
    (hc, wc), pk, tf, bx, by = _geometry(H, Wd, th, tw, nknots, cropping, subrec, bidirectional,
                                         min_blockpoints_ratio)
    tfrac_bm, base_mask = _image2blockmatrix(tf, pk, nb)
    basex_bm, _ = _image2blockmatrix(bx - 1, pk, nb)
    basey_bm, _ = _image2blockmatrix(by - 1, pk, nb)
    min_points = np.bincount(pk - 1, minlength=nknots)[:nknots] * min_blockpoints_ratio
    xgrad = np.full(template.shape, np.nan)
    ygrad = np.full(template.shape, np.nan)
    xgrad[:, :-1] = np.diff(template, axis=1)
    ygrad[:-1, :] = np.diff(template, axis=0)
    dx = np.full((nknots, n), np.nan)
    dy = np.full((nknots, n), np.nan)
    corr = np.full(n, np.nan)
    iters = np.full(n, np.nan)
    resid = np.full(n, np.nan)
    failed = np.zeros(n, bool)
    if est_dx is None or est_dy is None:
        est_dx = np.zeros((nknots, n))
        est_dy = np.zeros((nknots, n))
    starts = set_starts(n, frames_per_set)
    c0, c1, c2, c3 = cropping
    for iset in range(len(starts) - 1):
        ii = list(range(starts[iset], min(starts[iset + 1], n - 1) + 1))
        ns = len(ii)
        # setdata = prefilter_foralignment(frames(ii), pregauss, mednorm) (:228)
        setdata = P.prefilter_foralignment(frames[:, :, ii], pregauss, mednorm) if (pregauss > 0 or mednorm) \
            else frames[:, :, ii]
        bm = np.zeros((tfrac_bm.shape[0], nb * ns))
        for k, fi in enumerate(ii):
            img = np.asarray(setdata[c2:H - c3, c0:Wd - c1, k], dtype=np.float64)
            bm[:, k * nb:(k + 1) * nb], _ = _image2blockmatrix(img.ravel(order="F"), pk, nb)
        a = dict(align_args)
        a.update(template_image=np.asfortranarray(template, dtype=np.float64), xgrad=np.asfortranarray(xgrad),
                 ygrad=np.asfortranarray(ygrad), tfrac=np.asfortranarray(tfrac_bm),
                 basex=np.asfortranarray(basex_bm), basey=np.asfortranarray(basey_bm),
                 base_mask=np.asfortranarray(base_mask), bmData=np.asfortranarray(bm),
                 est_dx_knots=np.asfortranarray(est_dx), est_dy_knots=np.asfortranarray(est_dy),
                 group_size=nb * frame_groupsize, nsetframes=ns, min_points_per_block=min_points)
        r = O.align_group_to_template(a, threads=threads, with_mask=True)
        for sj in range(ns - 1):
            j = ii[sj]
            dx[:, j] = r["knots_x"][sj * nknots:(sj + 1) * nknots]
            dy[:, j] = r["knots_y"][sj * nknots:(sj + 1) * nknots]
            corr[j] = r["frame_corr"][sj]
            iters[j] = r["iterations"]
            resid[j] = r["residrms"][sj]
            fm = np.asarray(r["mask"], bool).reshape(tfrac_bm.shape[0], nb * ns, order="F")[:, sj * nb:(sj + 1) * nb]
            used = (fm & base_mask.astype(bool)).any(axis=0)
            if not used.any():
                failed[j] = True
                corr[j] = np.nan
                continue
            if not used[0] or not used[-1]:
                dx[:, j] = extend_firstandlast(dx[:, j], used)
                dy[:, j] = extend_firstandlast(dy[:, j], used)
            failed[j] = (not (r["frame_corr"][sj] >= corr_thresh)) or (len(used) / nb < minblockfrac)
    return dict(dx_knots=dx, dy_knots=dy, corr=corr, iter_used=iters, residrms=resid, failed=failed)


def reconstruct(frames, dx_knots, dy_knots, template_shape, nknots=33, cropping=(0, 0, 0, 0), subrec=None,
                bidirectional=False, frames_per_set=500, pregauss=0.0, mednorm=False, ibs_inpaint=False,
                ibs_regfac=3e-3, aspect=1.0):
    """The reconstruction loop of steady_scan (steady_scan.m:124-134): per set
    ii = set_start:set_start_next - 1, prefilter_foralignment(cropimage(frames))
    then invert_bilinear_sampling with the frames' knots (the energy-check
    interval restarts per call), inpainted if ibs_inpaint (in the frames'
    class: single for float32 frames, double otherwise).  Returns
    (data_corrected float64 th x tw x n, mask)."""
    H, Wd, n = frames.shape
    th, tw = template_shape
    (hc, wc), pk, tf, bx, by = _geometry(H, Wd, th, tw, nknots, cropping, subrec, bidirectional, 0.3)
    geom = dict(prevknotind=pk.reshape(hc, wc, order="F"), tfrac=tf.reshape(hc, wc, order="F"),
                basex=bx.reshape(hc, wc, order="F"), basey=by.reshape(hc, wc, order="F"))
    c0, c1, c2, c3 = cropping
    dc = np.zeros((th, tw, n))
    mk = np.zeros((th, tw, n), bool)
    starts = set_starts(n, frames_per_set)
    for iset in range(len(starts) - 1):
        ii = list(range(starts[iset], starts[iset + 1]))
        setdata = frames[c2:H - c3, c0:Wd - c1, ii]
        if pregauss > 0 or mednorm:
            setdata = P.prefilter_foralignment(setdata, pregauss, mednorm)
        out, m, _ = O.invert_bilinear_sampling(np.asarray(setdata, dtype=np.float64), dx_knots[:, ii],
                                               dy_knots[:, ii], geom, (th, tw), regfac=ibs_regfac, aspect=aspect)
        if ibs_inpaint:
            out = P.imgapfill_leastsquares(out if frames.dtype == np.float32 else out.astype(np.float64))
        dc[:, :, ii] = out
        mk[:, :, ii] = m
    return dc, mk
