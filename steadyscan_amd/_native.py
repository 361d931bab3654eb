"""ctypes declarations of include/steadyscan.h and include/steadyscan_synth.h."""
import ctypes
import os
from ctypes import POINTER, c_char, c_char_p, c_double, c_int32, c_uint8, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsteadyscan_amd.so")
SYNTH_PATH = os.path.join(HERE, "libss_synth.so")

SS_OK, SS_ERR_INVALID_ARGUMENT, SS_ERR_OUT_OF_MEMORY, SS_ERR_DEVICE, SS_ERR_NUMERICAL, SS_ERR_NOT_AVAILABLE = range(6)
SS_F64, SS_F32, SS_U16, SS_I16 = range(4)

# report ids of the mex (mex/reporter/reporter.cpp:3-37)
STATUS_IDS = {
    SS_ERR_INVALID_ARGUMENT: "GPUMC:InvalidArgument",
    SS_ERR_OUT_OF_MEMORY: "GPUMC:CudaError",
    SS_ERR_DEVICE: "GPUMC:CudaError",
    SS_ERR_NUMERICAL: "GPUMC:Unknown",
    SS_ERR_NOT_AVAILABLE: "GPUMC:CudaError",
}

FUNC_SLOTS = 14  # output 10: the reference's function-time slots (timedmainiteration.cpp:28, task.cpp:244)
STOP_NAMES = ["CorrelationIncreaseThreshold", "HaltCorrelationReached", "MovementThreshold",
              "IterationLimit", "Error"]


class SSSettings(ctypes.Structure):
    _fields_ = [
        ("max_iterations", c_int32),
        ("move_threshold", c_double),
        ("halt_correlation", c_double),
        ("correlation_increase_threshold", c_double),
        ("filter_strength", c_double),
        ("max_filter_iterations", c_int32),
        ("filter_correlation_threshold", c_double),
        ("lambda_", c_double),
        ("linesearch_reduction_multiplier", c_double),
        ("solver_id", c_int32),
        ("filter_id", c_int32),
        ("device_id", c_int32),
        ("precision", c_char),
        ("linesearch_candidates", c_int32),
        ("template_bf16", c_int32),
        ("save_frame_correlations", c_int32),
    ]


class SSGroupIn(ctypes.Structure):
    _fields_ = [
        ("frame_h", c_int32), ("n_blocks", c_int32), ("n_frames", c_int32), ("group_size", c_int32),
        ("template_h", c_int32), ("template_w", c_int32),
        ("template_image", POINTER(c_double)), ("xgrad", POINTER(c_double)), ("ygrad", POINTER(c_double)),
        ("frac", POINTER(c_double)), ("base_x", POINTER(c_double)), ("base_y", POINTER(c_double)),
        ("base_mask", POINTER(c_uint8)),
        ("data_type", c_int32), ("data", c_void_p), ("data_on_device", c_int32),
        ("init_x", POINTER(c_double)), ("init_y", POINTER(c_double)),
        ("min_points", POINTER(c_int32)),
        ("template_id", c_int32), ("geometry_id", c_int32),
    ]


class SSGroupOut(ctypes.Structure):
    _fields_ = [
        ("knots_x", POINTER(c_double)), ("knots_y", POINTER(c_double)),
        ("frame_corr", POINTER(c_double)), ("iterations", c_int32),
        ("mask", POINTER(c_uint8)), ("residrms", POINTER(c_double)),
        ("device_init_time", c_double),
        ("iteration_corr", POINTER(c_double)), ("all_params", POINTER(c_double)),
        ("stop_criterion", c_int32), ("half_iterations", c_int32),
        ("linesearch_rounds", c_int32), ("linesearch_tries", c_int32),
        ("mask_on_device", c_int32),
        ("frame_corr_iter", POINTER(c_double)), ("iteration_errvals", POINTER(c_double)),
        ("iteration_times", POINTER(c_double)), ("raw_params", POINTER(c_double)),
        ("rg_params", POINTER(c_double)), ("function_times", POINTER(c_double)),
    ]


class SSDriverSettings(ctypes.Structure):
    _fields_ = [("frames_per_set", c_int32), ("min_frames_per_set", c_int32), ("corr_thresh", c_double),
                ("minblockfrac", c_double), ("frame_groupsize", c_int32), ("pregauss", c_double),
                ("mednorm", c_int32), ("pad", c_int32)]


class SSDisplacementsOut(ctypes.Structure):
    _fields_ = [("dx_knots", POINTER(c_double)), ("dy_knots", POINTER(c_double)), ("corr", POINTER(c_double)),
                ("iter_used", POINTER(c_double)), ("residrms", POINTER(c_double)), ("failed", POINTER(c_uint8))]


class SSTuningEntry(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_char), ("mixed", c_int32), ("data_type", c_int32),
                ("frame_h", c_int32), ("n_blocks", c_int32), ("frames", c_int32),
                ("chosen", c_int32 * 3), ("trial_us", (c_double * 3) * 3),
                ("samples", (c_int32 * 3) * 3), ("waves", (c_int32 * 3) * 3)]


class SSServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_groups", c_int32), ("reserved", c_int32),
                ("staging_limit_bytes", ctypes.c_int64)]


class SSGeometryDesc(ctypes.Structure):
    _fields_ = [
        ("nlines", c_int32), ("linewidth", c_int32),
        ("template_h", c_int32), ("template_w", c_int32),
        ("nknots", c_int32),
        ("cropping", c_int32 * 4),
        ("subrec", c_double * 4),
        ("has_subrec", c_int32),
        ("bidirectional", c_int32),
        ("min_blockpoints_ratio", c_double),
    ]


class SSResampleIn(ctypes.Structure):
    _fields_ = [
        ("geometry", POINTER(SSGeometryDesc)),
        ("n_frames", c_int32),
        ("data_type", c_int32),
        ("frames", c_void_p),
        ("data_on_device", c_int32),
        ("knots_x", POINTER(c_double)),
        ("knots_y", POINTER(c_double)),
        ("out_on_device", c_int32),
    ]


class SSIbsSettings(ctypes.Structure):
    _fields_ = [
        ("regfac", c_double), ("pixel_aspect", c_double), ("inpaint", c_int32),
        ("max_iterations", c_int32), ("chunk_frames", c_int32), ("outputs_on_device", c_int32),
        ("work_bytes", ctypes.c_int64), ("chain_frames", c_int32), ("reserved", c_int32),
    ]


class SynthTemplateDesc(ctypes.Structure):
    _fields_ = [
        ("h", c_int32), ("w", c_int32), ("n_blobs", c_int32),
        ("sigma_lo", c_double), ("sigma_hi", c_double),
        ("amp_lo", c_double), ("amp_hi", c_double), ("offset", c_double),
        ("n_background", c_int32), ("background_sigma", c_double),
        ("background_amp_lo", c_double), ("background_amp_hi", c_double),
        ("target_mean", c_double), ("seed", c_uint64),
    ]


class SynthMotionDesc(ctypes.Structure):
    _fields_ = [
        ("n_frames", c_int32), ("nlines", c_int32), ("frame_rate", c_double),
        ("walk_sigma", c_double), ("walk_revert", c_double),
        ("sin_amp_x", c_double), ("sin_amp_y", c_double), ("sin_freq", c_double),
        ("sin_phase_x", c_double), ("sin_phase_y", c_double),
        ("jump_prob", c_double), ("jump_size", c_double), ("seed", c_uint64),
    ]


# exported by libsteadyscan_amd.so: (name, restype, argtypes)
API = [
    ("ss_settings_default", None, [POINTER(SSSettings)]),
    ("ss_create", c_int32, [c_int32, POINTER(c_void_p)]),
    ("ss_destroy", None, [c_void_p]),
    ("ss_last_error", c_char_p, [c_void_p]),
    ("ss_upload_template", c_int32, [c_void_p, c_int32, c_int32, POINTER(c_double), POINTER(c_double),
                                     POINTER(c_double), POINTER(c_int32)]),
    ("ss_upload_geometry", c_int32, [c_void_p, c_int32, c_int32, POINTER(c_double), POINTER(c_double),
                                     POINTER(c_double), POINTER(c_uint8), POINTER(c_int32)]),
    ("ss_release", c_int32, [c_void_p, c_int32]),
    ("ss_align_group", c_int32, [c_void_p, POINTER(SSGroupIn), POINTER(SSSettings), POINTER(SSGroupOut)]),
    ("ss_align_batch", c_int32, [c_void_p, c_int32, POINTER(SSGroupIn), POINTER(SSSettings), POINTER(SSGroupOut)]),
    ("ss_geometry_size", c_int32, [POINTER(SSGeometryDesc), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
    ("ss_geometry_init", c_int32, [POINTER(SSGeometryDesc), POINTER(c_int32), POINTER(c_double), POINTER(c_double),
                                   POINTER(c_double), POINTER(c_double), POINTER(c_double), POINTER(c_double),
                                   POINTER(c_uint8), POINTER(c_int32)]),
    ("ss_pack_frames", c_int32, [POINTER(SSGeometryDesc), c_int32, c_void_p, c_int32, c_void_p]),
    ("ss_resample", c_int32, [c_void_p, POINTER(SSResampleIn), POINTER(ctypes.c_float)]),
    ("ss_resample_noalloc", c_int32, [c_void_p, POINTER(SSResampleIn), POINTER(ctypes.c_float)]),
    ("ss_ibs_settings_default", None, [POINTER(SSIbsSettings)]),
    ("ss_prefilter", c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_double, c_int32,
                               c_int32]),
    ("ss_fill_gaps", c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32]),
    ("ss_reconstruct", c_int32, [c_void_p, POINTER(SSGeometryDesc), POINTER(SSDriverSettings), POINTER(SSIbsSettings),
                                 c_int32, c_void_p, c_int32, POINTER(c_double), POINTER(c_double), POINTER(c_double),
                                 POINTER(c_uint8)]),
    ("ss_invert_bilinear_sampling", c_int32, [c_void_p, POINTER(SSResampleIn), POINTER(SSIbsSettings),
                                              POINTER(ctypes.c_float), POINTER(c_uint8), POINTER(c_int32)]),
    ("ss_profile_enable", c_int32, [c_void_p, c_int32]),
    ("ss_profile_read", c_int32, [c_void_p, POINTER(ctypes.c_int64), POINTER(c_double), POINTER(c_double)]),
    ("ss_driver_settings_default", None, [POINTER(SSDriverSettings)]),
    ("ss_estimate_displacements", c_int32, [c_void_p, POINTER(SSGeometryDesc), POINTER(SSDriverSettings),
                                            POINTER(SSSettings), POINTER(c_double), c_int32, c_void_p, c_int32,
                                            POINTER(c_double), POINTER(c_double), POINTER(SSDisplacementsOut)]),
    ("ss_rom_write", c_int32, [c_char_p, POINTER(SSGroupIn), POINTER(SSSettings)]),
    ("ss_rom_read", c_int32, [c_char_p, POINTER(c_void_p), POINTER(SSGroupIn), POINTER(SSSettings)]),
    ("ss_rom_free", None, [c_void_p]),
    ("ss_tuning_mode", c_int32, [c_void_p, c_int32]),
    ("ss_tuning_entries", c_int32, [c_void_p, POINTER(SSTuningEntry), c_int32]),
    ("ss_server_create", c_int32, [c_void_p, POINTER(SSServerConfig), POINTER(c_void_p)]),
    ("ss_server_destroy", None, [c_void_p]),
    ("ss_server_submit", c_int32, [c_void_p, POINTER(SSGroupIn), POINTER(SSSettings), POINTER(SSGroupOut),
                                   POINTER(ctypes.c_int64)]),
    ("ss_server_step", c_int32, [c_void_p, POINTER(c_int32), POINTER(c_int32)]),
    ("ss_server_completed", c_int32, [c_void_p, POINTER(ctypes.c_int64), POINTER(c_int32), POINTER(c_char_p),
                                      c_int32]),
    ("ss_server_memory", c_int32, [c_void_p, POINTER(ctypes.c_int64), POINTER(ctypes.c_int64)]),
]

KERNEL_CLASSES = ["assemble", "solve", "ls_round", "ls_decide", "corr", "final"]

SYNTH_API = [
    ("ss_synth_template", c_int32, [POINTER(SynthTemplateDesc), POINTER(c_double)]),
    ("ss_synth_motion", c_int32, [POINTER(SynthMotionDesc), POINTER(c_double), POINTER(c_double)]),
    ("ss_synth_frames", c_int32, [POINTER(c_double), c_int32, c_int32, c_int32, c_int32, c_int32,
                                  POINTER(c_double), POINTER(c_double), c_int32, c_uint64, c_int32,
                                  c_void_p, c_int32]),
]

_lib = None
_synth = None


def _bind(lib, table):
    for name, res, args in table:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """The HIP library.  Raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "(python -m steadyscan_amd.build) first")
        _lib = _bind(ctypes.CDLL(LIB_PATH), API)
    return _lib


def synth():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"{SYNTH_PATH} is missing: run python -m steadyscan_amd.build first")
        _synth = _bind(ctypes.CDLL(SYNTH_PATH), SYNTH_API)
    return _synth
