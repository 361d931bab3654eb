"""Host-side mirror of SteadyScan's `align_group_to_template` interface.

`align_group_to_template` takes the 26 arguments of the mex call in the same
order and meaning (steady_scan.m:239-242, mex/mexgateway.cpp:31-174) and
returns its outputs (mexgateway.cpp:261-370), computed by the HIP library
through the C ABI of include/steadyscan.h.  Matrices are numpy arrays in
MATLAB orientation (rows x columns); they are passed column-major.

`init_settings`, `image2blockmatrix` and `resample_frame` mirror the
MATLAB helpers that build the call's geometry and the corrected frames
(steady_scan.m:748-828, 601-613, 562-599).
"""
import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_int32, c_uint8, c_void_p

import numpy as np

from . import _native as N


class SteadyScanError(RuntimeError):
    """An error of the library; `identifier` is the mex's GPUMC:* id."""

    def __init__(self, status, message):
        self.status = int(status)
        self.identifier = N.STATUS_IDS.get(self.status, "GPUMC:Unknown")
        super().__init__(f"{self.identifier}: {message}")


def _check(status, handle=None, what=""):
    if status != N.SS_OK:
        msg = ""
        if handle is not None and handle.ptr:
            m = N.lib().ss_last_error(handle.ptr)
            msg = m.decode() if m else ""
        raise SteadyScanError(status, msg or what or f"status {status}")


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _ptr(a, ctype):
    return a.ctypes.data_as(POINTER(ctype))


class Handle:
    """Owns a device context and its resident buffers (ss_create/ss_destroy)."""

    def __init__(self, device=0):
        self.ptr = c_void_p()
        self.device = int(device)
        _check(N.lib().ss_create(self.device, byref(self.ptr)), what="ss_create failed (no gfx950 device?)")

    def close(self):
        if self.ptr:
            N.lib().ss_destroy(self.ptr)
            self.ptr = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_template(self, template_image, xgrad, ygrad):
        t, gx, gy = _f64(template_image), _f64(xgrad), _f64(ygrad)
        tid = c_int32()
        _check(N.lib().ss_upload_template(self.ptr, t.shape[0], t.shape[1], _ptr(t, c_double),
                                          _ptr(gx, c_double), _ptr(gy, c_double), byref(tid)), self)
        return tid.value

    def upload_geometry(self, tfrac, basex, basey, base_mask):
        f, bx, by = _f64(tfrac), _f64(basex), _f64(basey)
        bm = np.asfortranarray(np.asarray(base_mask, dtype=np.uint8))
        gid = c_int32()
        _check(N.lib().ss_upload_geometry(self.ptr, f.shape[0], f.shape[1], _ptr(f, c_double),
                                          _ptr(bx, c_double), _ptr(by, c_double), _ptr(bm, c_uint8),
                                          byref(gid)), self)
        return gid.value

    def release(self, rid):
        _check(N.lib().ss_release(self.ptr, int(rid)), self)

    def tuning_mode(self, mode):
        """-2 (default): launch variants by the size of each batch; 0, 1,
        2: use that launch variant for every pixel kernel; -1: tune each new
        group shape on its first tick."""
        _check(N.lib().ss_tuning_mode(self.ptr, int(mode)), self)

    def tuning(self):
        """The autotuner's choices so far: one dict per group shape."""
        n = N.lib().ss_tuning_entries(self.ptr, None, 0)
        arr = (N.SSTuningEntry * max(n, 1))()
        n = N.lib().ss_tuning_entries(self.ptr, arr, n)
        kernels = ("assemble", "ls_round", "corr")
        out = []
        for e in arr[:n]:
            d = dict(precision=e.precision.decode(), mixed=bool(e.mixed), data_type=int(e.data_type),
                     frame_h=e.frame_h, n_blocks=e.n_blocks, frames=e.frames, kernels={})
            for k, name in enumerate(kernels):
                d["kernels"][name] = dict(
                    chosen=int(e.chosen[k]), trial_us=[float(e.trial_us[k][v]) for v in range(3)],
                    variants=[(int(e.samples[k][v]), int(e.waves[k][v])) for v in range(3)])
            out.append(d)
        return out


_handles = {}


def _default_handle(device):
    h = _handles.get(device)
    if h is None:
        h = Handle(device)
        _handles[device] = h
    return h


ROM_FILE = "steadyscan_rom.bin"  # where the 'F' flag saves a call


def save_rom(path, call):
    """Writes one call (a dict of the mex arguments with host arrays) to a
    ROM file (ss_rom_write)."""
    h_dummy = type("H", (), {"device": int(call.get("gpu_device_id", 0))})()
    gi, s, go = N.SSGroupIn(), N.SSSettings(), N.SSGroupOut()
    keep = []
    _prepare_call(call, h_dummy, s, gi, go, False, call.get("linesearch_candidates", 2), keep)
    rc = N.lib().ss_rom_write(str(path).encode(), byref(gi), byref(s))
    if rc:
        raise SteadyScanError(rc, "ss_rom_write failed (host arrays needed)")


def load_rom(path):
    """Reads a ROM file back into a call dict (arrays copied out)."""
    rom = c_void_p()
    gi, s = N.SSGroupIn(), N.SSSettings()
    rc = N.lib().ss_rom_read(str(path).encode(), byref(rom), byref(gi), byref(s))
    if rc:
        raise SteadyScanError(rc, f"ss_rom_read failed for {path}")
    try:
        fh, nb, nF, th, tw = gi.frame_h, gi.n_blocks, gi.n_frames, gi.template_h, gi.template_w

        def arr(p, shape, dt):
            n = int(np.prod(shape))
            return np.ctypeslib.as_array(ctypes.cast(p, POINTER(dt)), (n,)).copy().reshape(shape, order="F")

        dtype = {N.SS_F64: (np.float64, c_double), N.SS_F32: (np.float32, c_float),
                 N.SS_U16: (np.uint16, ctypes.c_uint16), N.SS_I16: (np.int16, ctypes.c_int16)}[gi.data_type]
        call = dict(
            flags=s.precision.decode() + ("c" if s.save_frame_correlations else ""), template_image=arr(gi.template_image, (th, tw), c_double),
            xgrad=arr(gi.xgrad, (th, tw), c_double), ygrad=arr(gi.ygrad, (th, tw), c_double),
            tfrac=arr(gi.frac, (fh, nb), c_double), basex=arr(gi.base_x, (fh, nb), c_double),
            basey=arr(gi.base_y, (fh, nb), c_double), base_mask=arr(gi.base_mask, (fh, nb), c_uint8).astype(bool),
            bmData=arr(gi.data, (fh, nb * nF), dtype[1]).astype(dtype[0]),
            est_dx_knots=arr(gi.init_x, (nb + 1, nF), c_double), est_dy_knots=arr(gi.init_y, (nb + 1, nF), c_double),
            group_size=gi.group_size, nsetframes=nF,
            min_points_per_block=arr(gi.min_points, (nb + 1,), c_int32).astype(np.float64),
            max_iter=s.max_iterations, move_thresh=s.move_threshold, halt_correlation=s.halt_correlation,
            correlation_increase_threshold=s.correlation_increase_threshold, filter_strength=s.filter_strength,
            max_filter_iterations=s.max_filter_iterations,
            filter_deactivation_threshold=s.filter_correlation_threshold, lambda_=s.lambda_,
            linesearch_reduction_multiplier=s.linesearch_reduction_multiplier, solver_id=s.solver_id,
            filter_id=s.filter_id, gpu_device_id=s.device_id, linesearch_candidates=s.linesearch_candidates,
            template_bf16=bool(s.template_bf16))
    finally:
        N.lib().ss_rom_free(rom)
    return call


def default_settings():
    s = N.SSSettings()
    N.lib().ss_settings_default(byref(s))
    return s


def parse_flags(flags):
    """Precision letter and diagnostics of the flags string (configuration.cpp:83-152)."""
    precision = None
    opts = set()
    for ch in flags:
        if ch in "SD":
            precision = ch
        elif ch in "MRrp":
            raise SteadyScanError(N.SS_ERR_INVALID_ARGUMENT, f"flag '{ch}' is not supported")
        else:
            opts.add(ch)
    if precision is None:
        raise SteadyScanError(N.SS_ERR_INVALID_ARGUMENT,
                              "invalid precision type! use 'S' for single-precision and 'D' for double-precision.")
    return precision, opts


_DTYPES = {np.dtype(np.float64): N.SS_F64, np.dtype(np.float32): N.SS_F32,
           np.dtype(np.uint16): N.SS_U16, np.dtype(np.int16): N.SS_I16}


def align_group_to_template(flags, template_image, xgrad, ygrad, tfrac, basex, basey, base_mask,
                            bmData, est_dx_knots, est_dy_knots, group_size, nsetframes,
                            min_points_per_block, max_iter, move_thresh, halt_correlation,
                            correlation_increase_threshold, filter_strength, max_filter_iterations,
                            filter_deactivation_threshold, lambda_, linesearch_reduction_multiplier,
                            solver_id, filter_id, gpu_device_id, *, nargout=6, handle=None,
                            linesearch_candidates=2, template_bf16=False, info=None):
    """The mex call (MexGateway<T>::execute, mexgateway.cpp:373-445).

    Returns the first `nargout` of
    (block_dx_knots, block_dy_knots, frame_corr, block_iter_used, block_mask,
     residrms, deviceInitTime, iterationCorrelations, iterationTimes,
     functionTimes, frameCorrelations, iterationErrvals, rawParameters,
     rgParams, allParams).  Flags 't'/'T' fill outputs 9 (seconds per
    iteration) and 10 (seconds per stage in the reference's 14 function-time
    slots: 1 assembly, 7 solve, 9 line search, 10 correlations), both from
    the device clock.  Flag 'c' fills output 11 (n_frames x max_iter) and
    makes frame_corr the last iteration's values, as the reference does
    (task.cpp:392-394); flag 'e' fills output 12 (2*max_iter x n_frames;
    the reference stores it in output 11's slot, mexgateway.cpp:287, which
    leaves output 12 unset).  Flag 'P' fills output 13 (the raw knots).
    Output 14 holds the knots after each half-iteration (the reference fills
    it in its timing mode only and returns uninitialised memory otherwise).
    `info`, if a dict, receives the stop criterion and work counters.
    `template_bf16` selects mixed precision (precision 'S' only): BF16
    template gradients for the Jacobian, FP32 residuals, one FP64
    iterative-refinement step of each knot update.
    """
    if nargout < 1 or nargout > 15:
        raise SteadyScanError(N.SS_ERR_INVALID_ARGUMENT, "invalid number of lhs arguments specified")
    precision, opts = parse_flags(flags)
    h = handle if handle is not None else _default_handle(int(gpu_device_id))
    s = default_settings()
    s.max_iterations = int(max_iter)
    s.move_threshold = float(move_thresh)
    s.halt_correlation = float(halt_correlation)
    s.correlation_increase_threshold = float(correlation_increase_threshold)
    s.filter_strength = float(filter_strength)
    s.max_filter_iterations = int(max_filter_iterations)
    s.filter_correlation_threshold = float(filter_deactivation_threshold)
    s.lambda_ = float(lambda_)
    s.linesearch_reduction_multiplier = float(linesearch_reduction_multiplier)
    s.solver_id = int(solver_id)
    s.filter_id = int(filter_id)
    s.device_id = int(gpu_device_id)
    s.precision = precision.encode()
    s.linesearch_candidates = int(linesearch_candidates)
    s.template_bf16 = int(bool(template_bf16))
    s.save_frame_correlations = int("c" in opts)

    gi = N.SSGroupIn()
    keep = []
    tmpl_id = template_image if isinstance(template_image, int) else 0
    geom_id = tfrac if isinstance(tfrac, int) else 0
    if tmpl_id:
        gi.template_id = tmpl_id
    else:
        t, gx, gy = _f64(template_image), _f64(xgrad), _f64(ygrad)
        keep += [t, gx, gy]
        gi.template_h, gi.template_w = t.shape
        gi.template_image, gi.xgrad, gi.ygrad = _ptr(t, c_double), _ptr(gx, c_double), _ptr(gy, c_double)
    if geom_id:
        gi.geometry_id = geom_id
        fh, nb = np.shape(basex) if not isinstance(basex, int) else (None, None)
    else:
        f, bx, by = _f64(tfrac), _f64(basex), _f64(basey)
        bm = np.asfortranarray(np.asarray(base_mask, dtype=np.uint8))
        keep += [f, bx, by, bm]
        gi.frac, gi.base_x, gi.base_y, gi.base_mask = (_ptr(f, c_double), _ptr(bx, c_double),
                                                       _ptr(by, c_double), _ptr(bm, c_uint8))
        fh, nb = bx.shape
    if isinstance(template_image, int):
        gi.template_h, gi.template_w = np.shape(xgrad)
    if fh is None:
        fh, nb = np.shape(base_mask)
    nF = int(nsetframes)
    gi.frame_h, gi.n_blocks, gi.n_frames, gi.group_size = int(fh), int(nb), nF, int(group_size)
    # frame data: host array (any of f64/f32/u16/i16) or (device_ptr, dtype) tuple
    if isinstance(bmData, tuple):
        ptr, dt = bmData
        gi.data = c_void_p(int(ptr))
        gi.data_type = _DTYPES[np.dtype(dt)]
        gi.data_on_device = 1
    else:
        d = np.asarray(bmData)
        if d.dtype not in _DTYPES:
            d = d.astype(np.float64)
        d = np.asfortranarray(d)
        if d.shape != (fh, nb * nF):
            raise SteadyScanError(N.SS_ERR_INVALID_ARGUMENT, f"bmData must be {fh} x {nb * nF}")
        keep.append(d)
        gi.data = c_void_p(d.ctypes.data)
        gi.data_type = _DTYPES[d.dtype]
    ex, ey = _f64(est_dx_knots), _f64(est_dy_knots)
    if ex.shape[0] != nb + 1 or ex.shape[1] < nF or ey.shape != ex.shape:
        raise SteadyScanError(N.SS_ERR_INVALID_ARGUMENT, "est knots must be (n_blocks+1) x >= nFrames")
    mp = np.ascontiguousarray(np.trunc(np.asarray(min_points_per_block, dtype=np.float64)).astype(np.int32).ravel())
    if mp.size < nb + 1:
        raise SteadyScanError(N.SS_ERR_INVALID_ARGUMENT, "min_points_per_block needs nknots entries")
    keep += [ex, ey, mp]
    gi.init_x, gi.init_y, gi.min_points = _ptr(ex, c_double), _ptr(ey, c_double), _ptr(mp, c_int32)

    go = N.SSGroupOut()
    kx = np.zeros((nb + 1) * nF)
    ky = np.zeros((nb + 1) * nF)
    fc = np.zeros(nF)
    rr = np.zeros(nF)
    go.knots_x, go.knots_y, go.frame_corr, go.residrms = (_ptr(kx, c_double), _ptr(ky, c_double),
                                                          _ptr(fc, c_double), _ptr(rr, c_double))
    mask = None
    if nargout >= 5:
        mask = np.zeros((fh, nb * nF), dtype=np.uint8, order="F")
        go.mask = _ptr(mask, c_uint8)
    icorr = np.zeros(int(max_iter)) if nargout >= 8 else None
    if icorr is not None:
        go.iteration_corr = _ptr(icorr, c_double)
    allp = np.zeros((2 * (nb * nF + 1), int(max_iter)), order="F") if nargout >= 15 else None
    if allp is not None:
        go.all_params = _ptr(allp, c_double)
    fci = np.zeros((nF, int(max_iter)), order="F") if (nargout >= 11 and "c" in opts) else None
    if fci is not None:
        go.frame_corr_iter = _ptr(fci, c_double)
    ierr = np.zeros((2 * int(max_iter), nF), order="F") if (nargout >= 12 and "e" in opts) else None
    if ierr is not None:
        go.iteration_errvals = _ptr(ierr, c_double)
    itimes = np.zeros(int(max_iter)) if (nargout >= 9 and ("t" in opts or "T" in opts)) else None
    if itimes is not None:
        go.iteration_times = _ptr(itimes, c_double)
    ftimes = np.zeros(N.FUNC_SLOTS) if (nargout >= 10 and ("t" in opts or "T" in opts)) else None
    if ftimes is not None:
        go.function_times = _ptr(ftimes, c_double)
    raw = np.zeros((2 * (nb * nF + 1), 1)) if (nargout >= 13 and "P" in opts) else None
    if raw is not None:
        go.raw_params = _ptr(raw, c_double)
    rgp = np.zeros((2 * (nb * nF + 1), 2 * int(max_iter)), order="F") if nargout >= 14 else None
    if rgp is not None:
        go.rg_params = _ptr(rgp, c_double)
    if "F" in opts:  # save the call's arguments (the 'F' flag, configuration.h:10-28)
        if N.lib().ss_rom_write(ROM_FILE.encode(), byref(gi), byref(s)) != 0:
            raise SteadyScanError(N.SS_ERR_INVALID_ARGUMENT, f"could not write {ROM_FILE}")
    _check(N.lib().ss_align_group(h.ptr, byref(gi), byref(s), byref(go)), h)
    del keep
    if info is not None:
        info.update(stop_criterion=N.STOP_NAMES[go.stop_criterion], iterations=go.iterations,
                    half_iterations=go.half_iterations, linesearch_rounds=go.linesearch_rounds,
                    linesearch_tries=go.linesearch_tries)
    if "E" in opts:
        print(f"Execution terminated:{N.STOP_NAMES[go.stop_criterion]}\niterations: {go.iterations}")
    outs = [kx.reshape(-1, 1), ky.reshape(-1, 1), fc.reshape(-1, 1), np.int32(go.iterations),
            None if mask is None else mask.astype(bool), rr.reshape(-1, 1), go.device_init_time,
            icorr if (icorr is not None and "C" in opts) else np.zeros((0, 0))]
    empty = np.zeros((0, 0))
    # 9, 10 ('t'/'T'), 11 ('c'), 12 ('e'), 13 ('P'), 14
    outs += [empty if itimes is None else itimes.reshape(-1, 1), empty if ftimes is None else ftimes.reshape(-1, 1),
             empty if fci is None else fci,
             empty if ierr is None else ierr, empty if raw is None else raw, empty if rgp is None else rgp]
    outs.append(allp if allp is not None else empty)
    return tuple(outs[:nargout])


def _prepare_call(a, h, s, gi, go, with_mask, linesearch_candidates, keep):
    """Fills settings s, group input gi and output go for one call (a dict
    keyed by the mex argument names); returns the result dict whose arrays
    go points into.  Arrays that must outlive the call are appended to keep."""
    precision, opts = parse_flags(a.get("flags", "D"))
    ds = default_settings()
    ctypes.memmove(ctypes.addressof(s), ctypes.addressof(ds), ctypes.sizeof(ds))
    s.max_iterations = int(a["max_iter"])
    s.move_threshold = float(a["move_thresh"])
    s.halt_correlation = float(a["halt_correlation"])
    s.correlation_increase_threshold = float(a["correlation_increase_threshold"])
    s.filter_strength = float(a["filter_strength"])
    s.max_filter_iterations = int(a["max_filter_iterations"])
    s.filter_correlation_threshold = float(a["filter_deactivation_threshold"])
    s.lambda_ = float(a["lambda_"])
    s.linesearch_reduction_multiplier = float(a["linesearch_reduction_multiplier"])
    s.solver_id = int(a["solver_id"])
    s.filter_id = int(a["filter_id"])
    s.device_id = h.device
    s.precision = precision.encode()
    s.linesearch_candidates = int(a.get("linesearch_candidates", linesearch_candidates))
    s.template_bf16 = int(bool(a.get("template_bf16", False)))
    s.save_frame_correlations = int("c" in opts)
    if isinstance(a["template_image"], int):  # resident template id
        gi.template_id = a["template_image"]
        gi.template_h, gi.template_w = np.shape(a["xgrad"])
    else:
        t, gx, gy = _f64(a["template_image"]), _f64(a["xgrad"]), _f64(a["ygrad"])
        keep += [t, gx, gy]
        gi.template_h, gi.template_w = t.shape
        gi.template_image, gi.xgrad, gi.ygrad = _ptr(t, c_double), _ptr(gx, c_double), _ptr(gy, c_double)
    if isinstance(a["tfrac"], int):  # resident geometry id
        gi.geometry_id = a["tfrac"]
        fh, nb = np.shape(a["basex"])
    else:
        f, bx, by = _f64(a["tfrac"]), _f64(a["basex"]), _f64(a["basey"])
        bm = np.asfortranarray(np.asarray(a["base_mask"], dtype=np.uint8))
        keep += [f, bx, by, bm]
        gi.frac, gi.base_x, gi.base_y, gi.base_mask = (_ptr(f, c_double), _ptr(bx, c_double),
                                                       _ptr(by, c_double), _ptr(bm, c_uint8))
        fh, nb = bx.shape
    if isinstance(a["bmData"], tuple):  # (device pointer, dtype)
        ptr, dt = a["bmData"]
        gi.data = c_void_p(int(ptr))
        gi.data_type = _DTYPES[np.dtype(dt)]
        gi.data_on_device = 1
    else:
        d = np.asarray(a["bmData"])
        if d.dtype not in _DTYPES:
            d = d.astype(np.float64)
        d = np.asfortranarray(d)
        keep.append(d)
        gi.data = c_void_p(d.ctypes.data)
        gi.data_type = _DTYPES[d.dtype]
    ex, ey = _f64(a["est_dx_knots"]), _f64(a["est_dy_knots"])
    mp = np.ascontiguousarray(np.trunc(np.asarray(a["min_points_per_block"], dtype=np.float64)).astype(np.int32).ravel())
    keep += [ex, ey, mp]
    nF = int(a["nsetframes"])
    gi.frame_h, gi.n_blocks, gi.n_frames, gi.group_size = fh, nb, nF, int(a["group_size"])
    gi.init_x, gi.init_y, gi.min_points = _ptr(ex, c_double), _ptr(ey, c_double), _ptr(mp, c_int32)
    r = dict(knots_x=np.zeros((nb + 1) * nF), knots_y=np.zeros((nb + 1) * nF),
             frame_corr=np.zeros(nF), residrms=np.zeros(nF),
             mask=None)
    mask_dev = isinstance(with_mask, int) and not isinstance(with_mask, bool)  # a device pointer
    if not mask_dev and bool(with_mask):
        r["mask"] = np.zeros((fh, nb * nF), dtype=np.uint8, order="F")
    go.knots_x, go.knots_y = _ptr(r["knots_x"], c_double), _ptr(r["knots_y"], c_double)
    go.frame_corr, go.residrms = _ptr(r["frame_corr"], c_double), _ptr(r["residrms"], c_double)
    if mask_dev:
        go.mask = ctypes.cast(c_void_p(with_mask), POINTER(c_uint8))
        go.mask_on_device = 1
    elif r["mask"] is not None:
        go.mask = _ptr(r["mask"], c_uint8)
    mi = int(a["max_iter"])
    if "c" in opts:  # per-iteration frame correlations (output 11)
        r["frame_corr_iter"] = np.zeros((nF, mi), order="F")
        go.frame_corr_iter = _ptr(r["frame_corr_iter"], c_double)
    if "e" in opts:  # per-iteration errvals (output 12)
        r["iteration_errvals"] = np.zeros((2 * mi, nF), order="F")
        go.iteration_errvals = _ptr(r["iteration_errvals"], c_double)
    if "t" in opts or "T" in opts:  # seconds per iteration (output 9)
        r["iteration_times"] = np.zeros(mi)
        go.iteration_times = _ptr(r["iteration_times"], c_double)
    if "P" in opts:  # raw knots (output 13)
        r["raw_params"] = np.zeros(2 * (nb * nF + 1))
        go.raw_params = _ptr(r["raw_params"], c_double)
    return r


def _collect(r, go):
    if r["mask"] is not None:
        r["mask"] = r["mask"].astype(bool)
    r.update(iterations=go.iterations, stop_criterion=N.STOP_NAMES[go.stop_criterion],
             half_iterations=go.half_iterations, linesearch_rounds=go.linesearch_rounds,
             linesearch_tries=go.linesearch_tries)
    return r


def align_groups(calls, handle=None, device=0, linesearch_candidates=2, with_mask=True):
    """Several independent align_group_to_template calls in one submission
    (ss_align_batch).  `calls` is a list of dicts keyed by the mex argument
    names (see workloads.ARG_ORDER); returns a list of dicts with the
    outputs of each call.  with_mask: True (host masks), False (none), or a
    list of device pointers, one per call (frame_h x n_blocks*n_frames
    bytes each), that receive the masks in place."""
    h = handle if handle is not None else _default_handle(device)
    n = len(calls)
    gins = (N.SSGroupIn * n)()
    sets = (N.SSSettings * n)()
    gouts = (N.SSGroupOut * n)()
    keep = []
    masks = with_mask if isinstance(with_mask, (list, tuple)) else [with_mask] * n
    results = [_prepare_call(a, h, sets[i], gins[i], gouts[i], masks[i], linesearch_candidates, keep)
               for i, a in enumerate(calls)]
    _check(N.lib().ss_align_batch(h.ptr, n, gins, sets, gouts), h)
    out = [_collect(r, go) for r, go in zip(results, gouts)]
    del keep
    return out


class Server:
    """Continuous batching of alignment groups on one device (ss_server_*):
    submit() queues a group (a dict of mex arguments, as for align_groups),
    step() admits queued groups as slots and the HBM budget allow, advances
    every resident group by one tick and retires finished ones; completed()
    returns {ticket: result dict} for the groups finished since the last
    call."""

    def __init__(self, handle=None, device=0, hbm_budget_bytes=0, max_groups=0, linesearch_candidates=2,
                 with_mask=False, staging_limit_bytes=0):
        self.h = handle if handle is not None else _default_handle(device)
        cfg = N.SSServerConfig()
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.max_groups = int(max_groups)
        cfg.staging_limit_bytes = int(staging_limit_bytes)
        self.ptr = c_void_p()
        _check(N.lib().ss_server_create(self.h.ptr, ctypes.byref(cfg), ctypes.byref(self.ptr)), self.h)
        self.K = linesearch_candidates
        self.with_mask = with_mask
        self._pending = {}  # ticket -> (structs, arrays, result)

    def submit(self, call, mask=None):
        """mask: None (the server's with_mask) or a device pointer that
        receives this group's mask in place."""
        gi, s, go = N.SSGroupIn(), N.SSSettings(), N.SSGroupOut()
        keep = []
        r = _prepare_call(call, self.h, s, gi, go, self.with_mask if mask is None else int(mask), self.K, keep)
        t = ctypes.c_int64()
        _check(N.lib().ss_server_submit(self.ptr, ctypes.byref(gi), ctypes.byref(s), ctypes.byref(go),
                                        ctypes.byref(t)), self.h)
        self._pending[t.value] = (gi, s, go, keep, r)
        return t.value

    def step(self):
        nr, nq = c_int32(), c_int32()
        _check(N.lib().ss_server_step(self.ptr, ctypes.byref(nr), ctypes.byref(nq)), self.h)
        return nr.value, nq.value

    def completed(self):
        out = {}
        cap = 64
        tickets = (ctypes.c_int64 * cap)()
        status = (c_int32 * cap)()
        messages = (ctypes.c_char_p * cap)()
        while True:
            n = N.lib().ss_server_completed(self.ptr, tickets, status, messages, cap)
            for i in range(n):
                gi, s, go, keep, r = self._pending.pop(tickets[i])
                if status[i] != 0:
                    out[tickets[i]] = SteadyScanError(status[i], (messages[i] or b"").decode())
                else:
                    out[tickets[i]] = _collect(r, go)
            if n < cap:
                return out

    def memory(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(N.lib().ss_server_memory(self.ptr, ctypes.byref(a), ctypes.byref(b)), self.h)
        return a.value, b.value

    def close(self):
        if self.ptr:
            N.lib().ss_server_destroy(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def estimate_displacements(frames, template, flags="D", nknots=33, cropping=None, subrec=None,
                           isbidirectional=False, min_blockpoints_ratio=0.3, frames_per_set=500,
                           min_frames_per_set=100, corr_thresh=0.75, minblockfrac=0.25, frame_groupsize=2,
                           est_dx_knots=None, est_dy_knots=None, handle=None, device=0,
                           linesearch_candidates=2, pregauss=0.0, mednorm=False, **align):
    """estimate_displacements of steady_scan.m:157-288 on a frame stack
    (nlines x linewidth x n, f64/f32/u16/i16, frame_list already applied):
    sets of frames_per_set frames (+1 overlap) are aligned in one batched
    submission (ss_estimate_displacements).  `align` overrides the scalar
    alignment settings (max_iter, move_thresh, ... as in
    workloads.default_args()).  Returns dict(dx_knots, dy_knots (nknots x n),
    corr, iter_used, residrms, failed)."""
    h = handle if handle is not None else _default_handle(device)
    fr = np.asarray(frames)
    if fr.dtype not in _DTYPES:
        fr = fr.astype(np.float64)
    fr = np.asfortranarray(fr)
    if fr.ndim == 2:
        fr = fr[:, :, None]
    H, Wd, n = fr.shape
    T = _f64(template)
    g = geometry_desc(H, Wd, T.shape, nknots=nknots, cropping=cropping, subrec=subrec,
                      isbidirectional=isbidirectional, min_blockpoints_ratio=min_blockpoints_ratio)
    d = N.SSDriverSettings()
    N.lib().ss_driver_settings_default(byref(d))
    d.frames_per_set, d.min_frames_per_set = int(frames_per_set), int(min_frames_per_set)
    d.corr_thresh, d.minblockfrac, d.frame_groupsize = float(corr_thresh), float(minblockfrac), int(frame_groupsize)
    d.pregauss, d.mednorm = float(pregauss), 1 if mednorm else 0
    precision, _ = parse_flags(flags)
    a = default_args_dict()
    a.update(align)
    s = default_settings()
    s.max_iterations = int(a["max_iter"])
    s.move_threshold = float(a["move_thresh"])
    s.halt_correlation = float(a["halt_correlation"])
    s.correlation_increase_threshold = float(a["correlation_increase_threshold"])
    s.filter_strength = float(a["filter_strength"])
    s.max_filter_iterations = int(a["max_filter_iterations"])
    s.filter_correlation_threshold = float(a["filter_deactivation_threshold"])
    s.lambda_ = float(a["lambda_"])
    s.linesearch_reduction_multiplier = float(a["linesearch_reduction_multiplier"])
    s.solver_id = int(a["solver_id"])
    s.filter_id = int(a["filter_id"])
    s.device_id = h.device
    s.precision = precision.encode()
    s.linesearch_candidates = int(linesearch_candidates)
    out = dict(dx_knots=np.zeros((nknots, n), order="F"), dy_knots=np.zeros((nknots, n), order="F"),
               corr=np.zeros(n), iter_used=np.zeros(n), residrms=np.zeros(n), failed=np.zeros(n, np.uint8))
    o = N.SSDisplacementsOut()
    o.dx_knots, o.dy_knots = _ptr(out["dx_knots"], c_double), _ptr(out["dy_knots"], c_double)
    o.corr, o.iter_used, o.residrms = (_ptr(out["corr"], c_double), _ptr(out["iter_used"], c_double),
                                       _ptr(out["residrms"], c_double))
    o.failed = _ptr(out["failed"], c_uint8)
    ex = ey = None
    if est_dx_knots is not None and est_dy_knots is not None:
        ex, ey = _f64(est_dx_knots), _f64(est_dy_knots)
        if ex.shape != (nknots, n) or ey.shape != (nknots, n):
            raise SteadyScanError(N.SS_ERR_INVALID_ARGUMENT, "est knots must be nknots x n")
    _check(N.lib().ss_estimate_displacements(h.ptr, byref(g), byref(d), byref(s), _ptr(T, c_double),
                                             _DTYPES[fr.dtype], c_void_p(fr.ctypes.data), n,
                                             _ptr(ex, c_double) if ex is not None else None,
                                             _ptr(ey, c_double) if ey is not None else None, byref(o)), h)
    out["failed"] = out["failed"].astype(bool)
    return out


def reconstruct(frames, geometry, dx_knots, dy_knots, frames_per_set=500, pregauss=0.0, mednorm=False,
                ibs_inpaint=False, ibs_regfac=3e-3, ibs_template_pixel_aspect_ratio=1.0, handle=None, device=0):
    """The reconstruction loop of steady_scan (steady_scan.m:124-134; ss_reconstruct):
    set by set, crop + prefilter + invert_bilinear_sampling (+ inpainting).
    frames: nlines x linewidth x n raw frames; knots nknots x n.  Returns
    (data_corrected float64 template_h x template_w x n, mask bool)."""
    h = handle if handle is not None else _default_handle(device)
    fr = np.asarray(frames)
    if fr.ndim == 2:
        fr = fr[:, :, None]
    if fr.dtype not in _DTYPES:
        fr = fr.astype(np.float64)
    fr = np.asfortranarray(fr)
    g = geometry["desc"] if isinstance(geometry, dict) else geometry
    n = fr.shape[2]
    d = N.SSDriverSettings()
    N.lib().ss_driver_settings_default(byref(d))
    d.frames_per_set = int(frames_per_set)
    d.pregauss, d.mednorm = float(pregauss), 1 if mednorm else 0
    st = N.SSIbsSettings()
    N.lib().ss_ibs_settings_default(byref(st))
    st.regfac = float(ibs_regfac)
    st.pixel_aspect = float(ibs_template_pixel_aspect_ratio)
    st.inpaint = 1 if ibs_inpaint else 0
    kx = _f64(np.reshape(dx_knots, (g.nknots, n), order="F"))
    ky = _f64(np.reshape(dy_knots, (g.nknots, n), order="F"))
    out = np.zeros((g.template_h, g.template_w, n), order="F")
    mask = np.zeros((g.template_h, g.template_w, n), np.uint8, order="F")
    _check(N.lib().ss_reconstruct(h.ptr, byref(g), byref(d), byref(st), _DTYPES[fr.dtype], c_void_p(fr.ctypes.data),
                                  n, _ptr(kx, c_double), _ptr(ky, c_double), _ptr(out, c_double),
                                  _ptr(mask, c_uint8)), h)
    return out, mask.astype(bool)


def steady_scan(data, template_image, est_dx_knots=None, est_dy_knots=None, data_extra_channels=None,
                flags="D", nknots=33, cropping=None, subrec=None, isbidirectional=False,
                min_blockpoints_ratio=0.3, frames_per_set=500, min_frames_per_set=100, corr_thresh=0.75,
                minblockfrac=0.25, frame_groupsize=2, pregauss=0.0, mednorm=False, ibs_inpaint=False,
                ibs_regfac=3e-3, ibs_template_pixel_aspect_ratio=1.0, handle=None, device=0, **align):
    """steady_scan (steady_scan.m:1-155) on a frame stack (nlines x linewidth x n,
    frame_list already applied): the template is prefiltered (:94), the
    displacements estimated set by set (:111, estimate_displacements) and the
    frames reconstructed set by set (:124-134); data_extra_channels
    (nlines x linewidth x n x n_extra) are reconstructed with the same knots
    (:136-154).  As in the reference, extra channel k (1-based) is written to
    channel slot k, so the first extra channel replaces the main channel's
    corrected frames and the last slot stays zero.  Returns dict(dx_knots,
    dy_knots, iter_used, corr, failed, residrms, data_corrected (th x tw x n x
    (1 + n_extra), float64), mask (same shape))."""
    h = handle if handle is not None else _default_handle(device)
    fr = np.asarray(data)
    if fr.ndim == 2:
        fr = fr[:, :, None]
    T = prefilter_foralignment(_f64(template_image), pregauss, mednorm, handle=h)
    est = estimate_displacements(fr, T, flags=flags, nknots=nknots, cropping=cropping, subrec=subrec,
                                 isbidirectional=isbidirectional, min_blockpoints_ratio=min_blockpoints_ratio,
                                 frames_per_set=frames_per_set, min_frames_per_set=min_frames_per_set,
                                 corr_thresh=corr_thresh, minblockfrac=minblockfrac, frame_groupsize=frame_groupsize,
                                 est_dx_knots=est_dx_knots, est_dy_knots=est_dy_knots, handle=h, pregauss=pregauss,
                                 mednorm=mednorm, **align)
    g = geometry_desc(fr.shape[0], fr.shape[1], T.shape, nknots=nknots, cropping=cropping, subrec=subrec,
                      isbidirectional=isbidirectional, min_blockpoints_ratio=min_blockpoints_ratio)
    extra = None if data_extra_channels is None else np.asarray(data_extra_channels)
    if extra is not None and extra.ndim == 3:
        extra = extra[:, :, :, None]
    nch = 1 + (0 if extra is None else extra.shape[3])
    n = fr.shape[2]
    dc = np.zeros(T.shape + (n, nch), order="F")
    mk = np.zeros(T.shape + (n, nch), bool, order="F")
    kw = dict(frames_per_set=frames_per_set, pregauss=pregauss, mednorm=mednorm, ibs_inpaint=ibs_inpaint,
              ibs_regfac=ibs_regfac, ibs_template_pixel_aspect_ratio=ibs_template_pixel_aspect_ratio, handle=h)
    dc[:, :, :, 0], mk[:, :, :, 0] = reconstruct(fr, g, est["dx_knots"], est["dy_knots"], **kw)
    for ch in range(nch - 1):
        dc[:, :, :, ch], mk[:, :, :, ch] = reconstruct(extra[:, :, :, ch], g, est["dx_knots"], est["dy_knots"], **kw)
    est.update(data_corrected=dc, mask=mk)
    return est


def default_args_dict():
    """Scalar alignment settings of default_alignment_settings
    (steady_scan.m:195-208, 836-867) keyed by the mex argument names."""
    return dict(max_iter=120, move_thresh=0.010, halt_correlation=0.995,
                correlation_increase_threshold=-1.0, filter_strength=0.8,
                max_filter_iterations=10, filter_deactivation_threshold=0.01, lambda_=0.0,
                linesearch_reduction_multiplier=0.5, solver_id=2, filter_id=1, gpu_device_id=0)


# ---------------------------------------------------------------------------
# geometry and corrected frames
# ---------------------------------------------------------------------------
def geometry_desc(nlines, linewidth, template_shape, nknots=33, cropping=None, subrec=None,
                  isbidirectional=False, min_blockpoints_ratio=0.3):
    g = N.SSGeometryDesc()
    g.nlines, g.linewidth = int(nlines), int(linewidth)
    g.template_h, g.template_w = int(template_shape[0]), int(template_shape[1])
    g.nknots = int(nknots)
    for i, v in enumerate(cropping if cropping is not None else (0, 0, 0, 0)):
        g.cropping[i] = int(v)
    if subrec is not None:
        g.has_subrec = 1
        for i, v in enumerate(subrec):
            g.subrec[i] = float(v)
    g.bidirectional = 1 if isbidirectional else 0
    g.min_blockpoints_ratio = float(min_blockpoints_ratio)
    return g


def init_settings(nlines, linewidth, template_shape, **kw):
    """init_settings + the block-matrix packing of estimate_displacements.

    Returns a dict with the image-shaped prevknotind/tfrac/basex/basey
    (steady_scan.m:813-824) and the block matrices tfrac_bm, basex_bm
    (= basex - 1), basey_bm, base_mask_bm and min_points_per_block
    (steady_scan.m:170-175)."""
    g = geometry_desc(nlines, linewidth, template_shape, **kw)
    hc, wc, fh = c_int32(), c_int32(), c_int32()
    _check(N.lib().ss_geometry_size(byref(g), byref(hc), byref(wc), byref(fh)), what="invalid geometry")
    hc, wc, fh = hc.value, wc.value, fh.value
    nb = g.nknots - 1
    out = dict(
        prevknotind=np.zeros((hc, wc), np.int32, order="F"), tfrac=np.zeros((hc, wc), order="F"),
        basex=np.zeros((hc, wc), order="F"), basey=np.zeros((hc, wc), order="F"),
        tfrac_bm=np.zeros((fh, nb), order="F"), basex_bm=np.zeros((fh, nb), order="F"),
        basey_bm=np.zeros((fh, nb), order="F"), base_mask_bm=np.zeros((fh, nb), np.uint8, order="F"),
        min_points_per_block=np.zeros(g.nknots, np.int32))
    _check(N.lib().ss_geometry_init(
        byref(g), _ptr(out["prevknotind"], c_int32), _ptr(out["tfrac"], c_double),
        _ptr(out["basex"], c_double), _ptr(out["basey"], c_double), _ptr(out["tfrac_bm"], c_double),
        _ptr(out["basex_bm"], c_double), _ptr(out["basey_bm"], c_double),
        _ptr(out["base_mask_bm"], c_uint8), _ptr(out["min_points_per_block"], c_int32)),
        what="invalid geometry")
    out["desc"] = g
    out["frame_h"] = fh
    out["nknots"] = g.nknots
    return out


def image2blockmatrix(frames, geometry):
    """Packs raw frames (H x W x n) into the block matrix frame_h x (nb*n)
    (steady_scan.m:229-234, 601-613)."""
    fr = np.asarray(frames)
    if fr.ndim == 2:
        fr = fr[:, :, None]
    if fr.dtype not in _DTYPES:
        fr = fr.astype(np.float64)
    fr = np.asfortranarray(fr)
    g = geometry["desc"]
    n = fr.shape[2]
    nb = g.nknots - 1
    out = np.zeros((geometry["frame_h"], nb * n), dtype=fr.dtype, order="F")
    _check(N.lib().ss_pack_frames(byref(g), _DTYPES[fr.dtype], c_void_p(fr.ctypes.data), n,
                                  c_void_p(out.ctypes.data)), what="pack failed")
    return out


def resample_frames(frames, geometry, dx_knots, dy_knots, device_out=None, handle=None, device=0, noalloc=False):
    """resample_frame for every frame (steady_scan.m:562-572): forward splat
    into template space, float32, NaN where no data landed.

    frames: nlines x linewidth x n raw frames, or a tuple (device_pointer,
    dtype, n) for frames resident in HBM (same layout).  With device_out (a
    device pointer to template_h x template_w x n float32) the result is
    written there and None is returned.  noalloc=True calls
    ss_resample_noalloc, which allocates nothing; it needs device frames and
    device_out."""
    g = geometry["desc"]
    ri = N.SSResampleIn()
    if isinstance(frames, tuple):
        ptr, dtype, n = frames
        ri.frames = c_void_p(int(ptr))
        ri.data_type = _DTYPES[np.dtype(dtype)]
        ri.data_on_device = 1
    else:
        fr = np.asarray(frames)
        if fr.ndim == 2:
            fr = fr[:, :, None]
        if fr.dtype not in _DTYPES:
            fr = fr.astype(np.float64)
        fr = np.asfortranarray(fr)
        if fr.shape[:2] != (g.nlines, g.linewidth):
            raise ValueError(f"frames are {fr.shape[:2]}, the geometry expects {(g.nlines, g.linewidth)}")
        n = fr.shape[2]
        ri.frames = c_void_p(fr.ctypes.data)
        ri.data_type = _DTYPES[fr.dtype]
    kx, ky = _f64(np.reshape(dx_knots, (g.nknots, n), order="F")), _f64(np.reshape(dy_knots, (g.nknots, n), order="F"))
    h = handle if handle is not None else _default_handle(device)
    ri.geometry = ctypes.pointer(g)
    ri.n_frames = n
    ri.knots_x, ri.knots_y = _ptr(kx, c_double), _ptr(ky, c_double)
    if noalloc and (device_out is None or not ri.data_on_device):
        raise ValueError("noalloc needs frames and output in device memory")
    if device_out is not None:
        ri.out_on_device = 1
        fn = N.lib().ss_resample_noalloc if noalloc else N.lib().ss_resample
        _check(fn(h.ptr, byref(ri), ctypes.cast(c_void_p(int(device_out)), POINTER(c_float))), h)
        return None
    out = np.zeros((g.template_h, g.template_w, n), dtype=np.float32, order="F")
    _check(N.lib().ss_resample(h.ptr, byref(ri), out.ctypes.data_as(POINTER(c_float))), h)
    return out


def invert_bilinear_sampling(frames, geometry, dx_knots, dy_knots, ibs_regfac=3e-3,
                             ibs_template_pixel_aspect_ratio=1.0, ibs_inpaint=False, max_iterations=0,
                             chunk_frames=0, work_bytes=0, chain_frames=0, device_out=None, handle=None, device=0,
                             info=None):
    """invert_bilinear_sampling (steady_scan.m:290-404) on the GPU.

    frames: nlines x linewidth x n raw frames (cropped here by the geometry's
    cropping, as steady_scan.m:132 crops them before the call), or a tuple
    (device_pointer, dtype, n) for frames resident in HBM (same layout);
    geometry: the dict of init_settings; dx_knots/dy_knots: nknots x n, one
    column per frame.  Returns (data_corrected float32 template_h x
    template_w x n with NaN outside the mask, mask bool); with
    device_out=(out_pointer, mask_pointer) the results are written to device
    memory instead and None is returned.  info, a dict if given, receives the
    per-frame Gauss-Seidel iteration counts under "iterations".  chain_frames
    = k > 0 restarts the energy-check interval every k frames, as separate
    calls of k frames would (the driver's sets in one launch)."""
    g = geometry["desc"]
    ri = N.SSResampleIn()
    if isinstance(frames, tuple):
        ptr, dtype, n = frames
        ri.frames = c_void_p(int(ptr))
        ri.data_type = _DTYPES[np.dtype(dtype)]
        ri.data_on_device = 1
    else:
        fr = np.asarray(frames)
        if fr.ndim == 2:
            fr = fr[:, :, None]
        if fr.dtype not in _DTYPES:
            fr = fr.astype(np.float64)
        fr = np.asfortranarray(fr)
        if fr.shape[:2] != (g.nlines, g.linewidth):
            raise ValueError(f"frames are {fr.shape[:2]}, the geometry expects {(g.nlines, g.linewidth)}")
        n = fr.shape[2]
        ri.frames = c_void_p(fr.ctypes.data)
        ri.data_type = _DTYPES[fr.dtype]
    kx, ky = _f64(np.reshape(dx_knots, (g.nknots, n), order="F")), _f64(np.reshape(dy_knots, (g.nknots, n), order="F"))
    h = handle if handle is not None else _default_handle(device)
    ri.geometry = ctypes.pointer(g)
    ri.n_frames = n
    ri.knots_x, ri.knots_y = _ptr(kx, c_double), _ptr(ky, c_double)
    st = N.SSIbsSettings()
    N.lib().ss_ibs_settings_default(byref(st))
    st.regfac = float(ibs_regfac)
    st.pixel_aspect = float(ibs_template_pixel_aspect_ratio)
    st.max_iterations = int(max_iterations)
    st.chunk_frames = int(chunk_frames)
    st.work_bytes = int(work_bytes)
    st.chain_frames = int(chain_frames)
    iters = np.zeros(n, dtype=np.int32)
    if device_out is not None:
        st.outputs_on_device = 1
        _check(N.lib().ss_invert_bilinear_sampling(h.ptr, byref(ri), byref(st),
                                                   ctypes.cast(c_void_p(int(device_out[0])), POINTER(c_float)),
                                                   ctypes.cast(c_void_p(int(device_out[1])), POINTER(c_uint8)),
                                                   _ptr(iters, c_int32)), h)
        if ibs_inpaint:  # in place, single precision
            _check(N.lib().ss_fill_gaps(h.ptr, N.SS_F32, c_void_p(int(device_out[0])), c_void_p(int(device_out[0])),
                                        g.template_h, g.template_w, n, 1), h)
        if info is not None:
            info["iterations"] = iters
        return None
    out = np.zeros((g.template_h, g.template_w, n), dtype=np.float32, order="F")
    mask = np.zeros((g.template_h, g.template_w, n), dtype=np.uint8, order="F")
    _check(N.lib().ss_invert_bilinear_sampling(h.ptr, byref(ri), byref(st), out.ctypes.data_as(POINTER(c_float)),
                                               _ptr(mask, c_uint8), _ptr(iters, c_int32)), h)
    if info is not None:
        info["iterations"] = iters
    if ibs_inpaint:
        # data_corrected has the class of the frames (steady_scan.m:312): single
        # frames stay single, everything else is filled in double
        single = ri.data_type == N.SS_F32
        out = imgapfill_leastsquares(out if single else out.astype(np.float64), handle=h)
    return out, mask.astype(bool)


def prefilter_foralignment(data, gsd, mednorm, handle=None, device=0):
    """prefilter_foralignment (steady_scan.m:625-643) on the GPU: Gaussian blur
    (gsd > 0) and median normalisation of every frame of an H x W (x n) array;
    the result keeps the element type (MATLAB's class)."""
    a = np.asarray(data)
    squeeze = a.ndim == 2
    if squeeze:
        a = a[:, :, None]
    if a.dtype not in _DTYPES:
        a = a.astype(np.float64)
    a = np.asfortranarray(a)
    out = np.empty_like(a, order="F")
    h = handle if handle is not None else _default_handle(device)
    _check(N.lib().ss_prefilter(h.ptr, _DTYPES[a.dtype], c_void_p(a.ctypes.data), c_void_p(out.ctypes.data),
                                a.shape[0], a.shape[1], a.shape[2], float(gsd), 1 if mednorm else 0, 0), h)
    return out[:, :, 0] if squeeze else out


def imgapfill_leastsquares(a, handle=None, device=0):
    """imgapfill_leastsquares (steady_scan.m:525-560) on the GPU: the NaN pixels
    of each frame filled with the minimiser of the summed squared adjacent
    differences (float32 or float64 frames; other types are taken as float64)."""
    x = np.asarray(a)
    squeeze = x.ndim == 2
    if squeeze:
        x = x[:, :, None]
    if x.dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
        x = x.astype(np.float64)
    x = np.asfortranarray(x)
    out = np.empty_like(x, order="F")
    h = handle if handle is not None else _default_handle(device)
    _check(N.lib().ss_fill_gaps(h.ptr, _DTYPES[x.dtype], c_void_p(x.ctypes.data), c_void_p(out.ctypes.data),
                                x.shape[0], x.shape[1], x.shape[2], 0), h)
    return out[:, :, 0] if squeeze else out
