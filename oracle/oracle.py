"""ctypes binding of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the parity checker.  The product path never
imports it.  `align_group_to_template` takes the mex's 26 arguments in the
reference's order (steady_scan.m:239-242) like the product's mirror does.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_char, c_double, c_int32, c_uint8, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
LIB_F32 = os.path.join(HERE, "liboracle_f32.so")  # FP32 pixels: a timed baseline, not a checker
REF_LIB = os.path.join(HERE, "_ref", "libref_paramorg.so")


class OrcIn(ctypes.Structure):
    _fields_ = [("frame_h", c_int32), ("n_blocks", c_int32), ("n_frames", c_int32),
                ("template_h", c_int32), ("template_w", c_int32), ("group_size", c_int32),
                ("template_image", POINTER(c_double)), ("xgrad", POINTER(c_double)),
                ("ygrad", POINTER(c_double)), ("frac", POINTER(c_double)),
                ("base_x", POINTER(c_double)), ("base_y", POINTER(c_double)),
                ("base_mask", POINTER(c_uint8)), ("data", POINTER(c_double)),
                ("init_x", POINTER(c_double)), ("init_y", POINTER(c_double)),
                ("min_points", POINTER(c_int32))]


class OrcSettings(ctypes.Structure):
    _fields_ = [("max_iterations", c_int32), ("move_threshold", c_double),
                ("halt_correlation", c_double), ("correlation_increase_threshold", c_double),
                ("filter_strength", c_double), ("n_filter_iterations", c_int32),
                ("filter_correlation_threshold", c_double), ("lambda_", c_double),
                ("linesearch_reduction_multiplier", c_double), ("solver_id", c_int32),
                ("filter_id", c_int32), ("precision", c_char)]


class OrcOut(ctypes.Structure):
    _fields_ = [("knots_x", POINTER(c_double)), ("knots_y", POINTER(c_double)),
                ("frame_corr", POINTER(c_double)), ("residrms", POINTER(c_double)),
                ("mask", POINTER(c_uint8)), ("iteration_corr", POINTER(c_double)),
                ("all_params", POINTER(c_double)), ("iterations", c_int32),
                ("stop_criterion", c_int32), ("half_iterations", c_int32),
                ("linesearch_tries", c_int32), ("frame_corr_iter", POINTER(c_double)),
                ("iteration_errvals", POINTER(c_double)), ("rg_params", POINTER(c_double))]


class OrcGeometryIn(ctypes.Structure):
    _fields_ = [("nlines", c_int32), ("linewidth", c_int32), ("template_h", c_int32),
                ("template_w", c_int32), ("nknots", c_int32), ("cropping", c_int32 * 4),
                ("subrec", c_double * 4), ("has_subrec", c_int32), ("bidirectional", c_int32),
                ("min_blockpoints_ratio", c_double)]


_lib = None
_ref = None
PD = POINTER(c_double)


def lib():
    global _lib
    if _lib is None:
        _lib = load(LIB)
    return _lib


def lib_f32():
    """The FP32-pixel copy of the oracle (the same source built with
    ORC_REAL=float): the benchmark's FP32 CPU baseline.  Not pinned and not
    used to check anything."""
    return load(LIB_F32)


def load(path):
    """The oracle library at `path` with its prototypes declared (the tests
    also load copies built with a planted mistake)."""
    L = ctypes.CDLL(path)
    L.orc_align_group.argtypes = [POINTER(OrcIn), POINTER(OrcSettings), POINTER(OrcOut), c_int32]
    L.orc_align_group.restype = c_int32
    L.orc_state_create.argtypes = [POINTER(OrcIn), POINTER(OrcSettings), POINTER(c_void_p)]
    L.orc_state_create.restype = c_int32
    L.orc_state_destroy.argtypes = [c_void_p]
    L.orc_state_get.argtypes = [c_void_p, PD, PD, PD]
    L.orc_state_set_p.argtypes = [c_void_p, PD]
    L.orc_state_set_delta.argtypes = [c_void_p, PD]
    L.orc_half_iteration_assemble.argtypes = [c_void_p, c_int32, POINTER(c_uint8), PD, PD, PD]
    L.orc_half_iteration_assemble.restype = c_int32
    L.orc_state_pixels.argtypes = [c_void_p, POINTER(c_uint8), PD, PD, PD]
    L.orc_half_iteration_linesearch.argtypes = [c_void_p, c_int32]
    L.orc_half_iteration_linesearch.restype = c_int32
    L.orc_global_correlation.argtypes = [c_void_p]
    L.orc_global_correlation.restype = c_double
    L.orc_organize_in.argtypes = [PD, PD, c_int32, c_int32, PD]
    L.orc_organize_out.argtypes = [PD, c_int32, c_int32, PD, PD]
    L.orc_solve_coo.argtypes = [c_int32, c_int32, POINTER(c_int32), POINTER(c_int32), PD, PD, PD, c_int32]
    L.orc_solve_coo.restype = c_int32
    L.orc_filter_apply.argtypes = [c_int32, c_double, PD, c_int32]
    L.orc_geometry.argtypes = [POINTER(OrcGeometryIn), POINTER(c_int32), PD, PD, PD,
                               POINTER(c_int32), POINTER(c_int32)]
    L.orc_geometry.restype = c_int32
    L.orc_image2blockmatrix.argtypes = [PD, POINTER(c_int32), c_int32, c_int32, PD, POINTER(c_uint8)]
    L.orc_image2blockmatrix.restype = c_int32
    L.orc_invert_bilinear_sampling.argtypes = [PD, c_int32, c_int32, c_int32, PD, PD, c_int32,
                                               POINTER(c_int32), PD, PD, PD, c_int32, c_int32, c_double,
                                               c_double, POINTER(ctypes.c_float), POINTER(c_uint8),
                                               POINTER(c_int32)]
    L.orc_invert_bilinear_sampling.restype = c_int32
    L.orc_ibs_neighbour.argtypes = [ctypes.c_long, c_int32, ctypes.c_long, ctypes.c_long,
                                    POINTER(ctypes.c_long)]
    L.orc_ibs_neighbour.restype = c_int32
    L.orc_resample_frame.argtypes = [PD, c_int32, PD, PD, PD, PD, POINTER(c_int32), PD, c_int32,
                                     c_int32, POINTER(ctypes.c_float)]
    PF = POINTER(ctypes.c_float)
    L.orc_ibs_system.argtypes = [PD, c_int32, PD, PD, POINTER(c_int32), PD, PD, PD, c_int32, c_int32,
                                 c_double, c_double, PD, PD, PF, POINTER(c_uint8)]
    L.orc_ibs_system.restype = c_int32
    L.orc_ibs_energy.argtypes = [PD, PD, PF, POINTER(c_uint8), c_int32, c_int32]
    L.orc_ibs_energy.restype = ctypes.c_float
    L.orc_ibs_interactions.argtypes = [PD, PD, POINTER(c_uint8), c_int32, c_int32, PF, PF]
    L.orc_ibs_sweep.argtypes = [PF, PF, c_int32, c_int32, PF]
    L.orc_ibs_check_interval.argtypes = [c_int32]
    L.orc_ibs_check_interval.restype = c_int32
    L.orc_frame_correlation.argtypes = [c_void_p, PD]
    return L


def ref_lib():
    """The reference's own ParameterOrganizer, built by oracle/build_ref.sh, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = ctypes.CDLL(REF_LIB)
        _ref.ref_organize_in.argtypes = [PD, PD, c_int32, c_int32, PD]
        _ref.ref_organize_out.argtypes = [PD, c_int32, c_int32, PD, PD]
    return _ref


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _p(a):
    return a.ctypes.data_as(PD)


class Inputs:
    """Holds converted arrays alive and exposes the OrcIn/OrcSettings pair."""

    def __init__(self, a):
        self.keep = []
        t, gx, gy = _f(a["template_image"]), _f(a["xgrad"]), _f(a["ygrad"])
        fr, bx, by = _f(a["tfrac"]), _f(a["basex"]), _f(a["basey"])
        bm = np.asfortranarray(np.asarray(a["base_mask"], dtype=np.uint8))
        d = _f(a["bmData"])
        ex, ey = _f(a["est_dx_knots"]), _f(a["est_dy_knots"])
        mp = np.ascontiguousarray(np.trunc(np.asarray(a["min_points_per_block"], dtype=np.float64)).astype(np.int32).ravel())
        self.keep = [t, gx, gy, fr, bx, by, bm, d, ex, ey, mp]
        self.nb = bx.shape[1]
        self.fh = bx.shape[0]
        self.nF = int(a["nsetframes"])
        self.i = OrcIn(frame_h=self.fh, n_blocks=self.nb, n_frames=self.nF,
                       template_h=t.shape[0], template_w=t.shape[1], group_size=int(a["group_size"]),
                       template_image=_p(t), xgrad=_p(gx), ygrad=_p(gy), frac=_p(fr), base_x=_p(bx),
                       base_y=_p(by), base_mask=bm.ctypes.data_as(POINTER(c_uint8)), data=_p(d),
                       init_x=_p(ex), init_y=_p(ey), min_points=mp.ctypes.data_as(POINTER(c_int32)))
        prec = "D"
        for ch in a.get("flags", "D"):
            if ch in "SD":
                prec = ch
        self.s = OrcSettings(max_iterations=int(a["max_iter"]), move_threshold=float(a["move_thresh"]),
                             halt_correlation=float(a["halt_correlation"]),
                             correlation_increase_threshold=float(a["correlation_increase_threshold"]),
                             filter_strength=float(a["filter_strength"]),
                             n_filter_iterations=int(a["max_filter_iterations"]),
                             filter_correlation_threshold=float(a["filter_deactivation_threshold"]),
                             lambda_=float(a["lambda_"]),
                             linesearch_reduction_multiplier=float(a["linesearch_reduction_multiplier"]),
                             solver_id=int(a["solver_id"]), filter_id=int(a["filter_id"]),
                             precision=prec.encode())


def align_group_to_template(a, threads=1, with_mask=True, with_iterations=False, L=None):
    """Runs the oracle on a dict of the 26 mex arguments; returns a dict.
    Flag 'c' adds frame_corr_iter (n_frames x max_iter, and frame_corr
    becomes the last iteration's values), flag 'e' iteration_errvals
    (2*max_iter x n_frames)."""
    inp = Inputs(a)
    nb, nF, fh = inp.nb, inp.nF, inp.fh
    kx = np.zeros((nb + 1) * nF)
    ky = np.zeros((nb + 1) * nF)
    fc = np.zeros(nF)
    rr = np.zeros(nF)
    o = OrcOut(knots_x=_p(kx), knots_y=_p(ky), frame_corr=_p(fc), residrms=_p(rr))
    mask = None
    if with_mask:
        mask = np.zeros((fh, nb * nF), dtype=np.uint8, order="F")
        o.mask = mask.ctypes.data_as(POINTER(c_uint8))
    mi = int(a["max_iter"])
    icorr = np.zeros(mi)
    o.iteration_corr = _p(icorr)
    allp = rg = None
    if with_iterations:
        allp = np.zeros((2 * (nb * nF + 1), mi), order="F")
        o.all_params = _p(allp)
        rg = np.zeros((2 * (nb * nF + 1), 2 * mi), order="F")
        o.rg_params = _p(rg)
    flags = a.get("flags", "D")
    fci = np.zeros((nF, mi), order="F") if "c" in flags else None
    if fci is not None:
        o.frame_corr_iter = _p(fci)
    ierr = np.zeros((2 * mi, nF), order="F") if "e" in flags else None
    if ierr is not None:
        o.iteration_errvals = _p(ierr)
    rc = (L or lib()).orc_align_group(byref(inp.i), byref(inp.s), byref(o), int(threads))
    if rc:
        raise RuntimeError(f"oracle returned status {rc}")
    return dict(knots_x=kx, knots_y=ky, frame_corr=fc, residrms=rr, mask=mask,
                iterations=o.iterations, stop_criterion=o.stop_criterion,
                half_iterations=o.half_iterations, linesearch_tries=o.linesearch_tries,
                iteration_corr=icorr[:min(o.iterations, mi)], all_params=allp, rg_params=rg,
                frame_corr_iter=fci, iteration_errvals=ierr)


class State:
    """Stage-level access to the oracle (one half-iteration at a time)."""

    def __init__(self, a, L=None):
        self.L = L if L is not None else lib()
        self.inp = Inputs(a)
        self.ptr = c_void_p()
        rc = self.L.orc_state_create(byref(self.inp.i), byref(self.inp.s), byref(self.ptr))
        if rc:
            raise RuntimeError(f"oracle state status {rc}")
        self.N = self.inp.nb * self.inp.nF + 1

    def __del__(self):
        if getattr(self, "ptr", None):
            self.L.orc_state_destroy(self.ptr)

    def get(self):
        p = np.zeros(2 * self.N)
        d = np.zeros(2 * self.N)
        e = np.zeros(2 * self.inp.nF)
        self.L.orc_state_get(self.ptr, _p(p), _p(d), _p(e))
        return p, d, e

    def set_p(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        self.L.orc_state_set_p(self.ptr, _p(p))

    def assemble(self, shift):
        used = np.zeros(self.N, np.uint8)
        jtr = np.zeros(2 * self.N)
        h = np.zeros(6 * self.N)
        dn = np.zeros(2 * self.N)
        n = self.L.orc_half_iteration_assemble(self.ptr, shift, used.ctypes.data_as(POINTER(c_uint8)),
                                              _p(jtr), _p(h), _p(dn))
        return dict(n_used=n, used=used.astype(bool), jtr=jtr, h=h.reshape(self.N, 6), delta=dn)

    def pixels(self):
        fh, nc = self.inp.fh, self.inp.nb * self.inp.nF
        m = np.zeros((fh, nc), np.uint8, order="F")
        x = np.zeros((fh, nc), order="F")
        y = np.zeros((fh, nc), order="F")
        d = np.zeros((fh, nc), order="F")
        self.L.orc_state_pixels(self.ptr, m.ctypes.data_as(POINTER(c_uint8)), _p(x), _p(y), _p(d))
        return m.astype(bool), x, y, d

    def linesearch(self, shift):
        return bool(self.L.orc_half_iteration_linesearch(self.ptr, shift))

    def global_correlation(self):
        return self.L.orc_global_correlation(self.ptr)

    def frame_correlation(self):
        c = np.zeros(self.inp.nF)
        self.L.orc_frame_correlation(self.ptr, _p(c))
        return c


def invert_bilinear_sampling(frames, dx_knots, dy_knots, geometry, template_shape, regfac=3e-3, aspect=1.0):
    """Oracle of invert_bilinear_sampling (steady_scan.m:290-404) for one call.
    frames: Hc x Wc x n cropped frames; knots: nknots x n; geometry: dict with
    image-shaped prevknotind, tfrac, basex, basey (as init_settings gives).
    Returns (corrected th x tw x n float32 with NaN outside the mask, mask, iterations)."""
    fr = np.asfortranarray(np.asarray(frames, dtype=np.float64))
    if fr.ndim == 2:
        fr = fr[:, :, None]
    Hc, Wc, n = fr.shape
    th, tw = template_shape
    dxk = np.asfortranarray(np.asarray(dx_knots, dtype=np.float64).reshape(-1, n, order="F"))
    dyk = np.asfortranarray(np.asarray(dy_knots, dtype=np.float64).reshape(-1, n, order="F"))
    pk = np.asfortranarray(np.asarray(geometry["prevknotind"], dtype=np.int32))
    tf, bx, by = (np.asfortranarray(np.asarray(geometry[k], dtype=np.float64)) for k in ("tfrac", "basex", "basey"))
    out = np.zeros((th, tw, n), dtype=np.float32, order="F")
    mask = np.zeros((th, tw, n), dtype=np.uint8, order="F")
    iters = np.zeros(n, dtype=np.int32)
    rc = lib().orc_invert_bilinear_sampling(_p(fr), Hc, Wc, n, _p(dxk), _p(dyk), dxk.shape[0],
                                            pk.ctypes.data_as(POINTER(c_int32)), _p(tf), _p(bx), _p(by), th, tw,
                                            float(regfac), float(aspect),
                                            out.ctypes.data_as(POINTER(ctypes.c_float)),
                                            mask.ctypes.data_as(POINTER(c_uint8)), iters.ctypes.data_as(POINTER(c_int32)))
    if rc:
        raise RuntimeError(f"oracle returned status {rc}")
    return out, mask.astype(bool), iters


def resample_frames(frames, geometry, dx_knots, dy_knots):
    """Oracle of resample_frame + pushgrid (steady_scan.m:562-599) for every
    frame: raw nlines x linewidth x n frames (cropped here by the geometry's
    cropping), knots nknots x n (or flat, frame after frame).  Returns
    template_h x template_w x n float32, NaN where no data landed."""
    g = geometry["desc"]
    cl, cr, cb, ct = list(g.cropping)
    th, tw = g.template_h, g.template_w
    fr = np.asarray(frames)
    if fr.ndim == 2:
        fr = fr[:, :, None]
    n = fr.shape[2]
    kx = np.asarray(dx_knots, dtype=np.float64).reshape(g.nknots, n, order="F")
    ky = np.asarray(dy_knots, dtype=np.float64).reshape(g.nknots, n, order="F")
    pk = np.ascontiguousarray(geometry["prevknotind"].ravel(order="F"), np.int32)
    bx = np.ascontiguousarray(geometry["basex"].ravel(order="F"))
    by = np.ascontiguousarray(geometry["basey"].ravel(order="F"))
    tf = np.ascontiguousarray(geometry["tfrac"].ravel(order="F"))
    out = np.zeros((th, tw, n), np.float32, order="F")
    for f in range(n):
        data = np.asarray(fr[cb:fr.shape[0] - ct, cl:fr.shape[1] - cr, f], np.float64)
        o = np.zeros(th * tw, np.float32)
        lib().orc_resample_frame(_p(np.ascontiguousarray(data.ravel(order="F"))), data.size,
                                 _p(np.ascontiguousarray(kx[:, f])), _p(np.ascontiguousarray(ky[:, f])),
                                 _p(bx), _p(by), pk.ctypes.data_as(POINTER(c_int32)), _p(tf), th, tw,
                                 o.ctypes.data_as(POINTER(ctypes.c_float)))
        out[:, :, f] = o.reshape(th, tw, order="F")
    return out
