"""MI355X-native implementation of SteadyScan's align_group_to_template path.

The compute runs in the HIP library ``libsteadyscan_amd.so`` (C ABI in
include/steadyscan.h); this package is its host-side mirror of the
reference's mex interface.  Importing the package does not touch the GPU.
"""
from .api import (Handle, Server, SteadyScanError, align_group_to_template, align_groups,  # noqa: F401
                  default_settings, estimate_displacements, geometry_desc, image2blockmatrix,
                  imgapfill_leastsquares, init_settings, invert_bilinear_sampling, load_rom, parse_flags,
                  prefilter_foralignment, reconstruct, resample_frames, save_rom, steady_scan)

__all__ = ["Handle", "Server", "SteadyScanError", "align_group_to_template", "align_groups",
           "default_settings", "estimate_displacements", "geometry_desc", "image2blockmatrix",
           "imgapfill_leastsquares", "init_settings", "invert_bilinear_sampling", "load_rom", "parse_flags",
           "prefilter_foralignment", "reconstruct", "resample_frames", "save_rom", "steady_scan"]
