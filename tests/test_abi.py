"""The C ABI library loads and exports every function include/*.h declares;
host-only entry points behave like the reference's argument checks."""
import ctypes
import os
import re

import numpy as np
import pytest

import steadyscan_amd as ssa
from steadyscan_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b((?:ss|ss_synth)_[a-z0-9_]+)\s*\(", text))


def exported(path):
    lib = ctypes.CDLL(path)
    return lib


@pytest.mark.parametrize("header,path", [("steadyscan.h", N.LIB_PATH),
                                         ("steadyscan_synth.h", N.SYNTH_PATH)])
def test_every_declared_symbol_is_exported(header, path):
    names = declared(header)
    assert names
    lib = exported(path)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_covers_header():
    assert declared("steadyscan.h") == {n for n, _, _ in N.API}
    assert declared("steadyscan_synth.h") == {n for n, _, _ in N.SYNTH_API}


@pytest.mark.parametrize("cname,pyt", [("ss_settings", N.SSSettings), ("ss_group_in", N.SSGroupIn),
                                       ("ss_group_out", N.SSGroupOut), ("ss_geometry_desc", N.SSGeometryDesc),
                                       ("ss_resample_in", N.SSResampleIn),
                                       ("ss_server_config", N.SSServerConfig),
                                       ("ss_tuning_entry", N.SSTuningEntry),
                                       ("ss_ibs_settings", N.SSIbsSettings),
                                       ("ss_driver_settings", N.SSDriverSettings),
                                       ("ss_displacements_out", N.SSDisplacementsOut)])
def test_struct_layout_matches_header(cname, pyt, tmp_path):
    """Offsets of every field of the ctypes mirror equal the C compiler's."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = [f[0] for f in pyt._fields_]
    cfields = [("lambda" if f == "lambda_" else f) for f in fields]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "steadyscan.h"', "int main(void){",
           f'printf("%zu\\n", sizeof({cname}));']
    src += [f'printf("%zu\\n", offsetof({cname}, {f}));' for f in cfields]
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(pyt)
    assert vals[1:] == [getattr(pyt, f).offset for f in fields]


def test_settings_default_matches_default_alignment_settings():
    s = ssa.default_settings()
    assert (s.max_iterations, s.move_threshold, s.halt_correlation) == (120, 0.010, 0.995)
    assert (s.correlation_increase_threshold, s.filter_strength, s.max_filter_iterations) == (-1, 0.8, 10)
    assert (s.filter_correlation_threshold, s.linesearch_reduction_multiplier) == (0.01, 0.5)
    assert (s.solver_id, s.filter_id, s.precision) == (2, 1, b"D")


def test_flags_parsing():
    assert ssa.parse_flags("D")[0] == "D"
    assert ssa.parse_flags("SCE")[0] == "S"
    assert ssa.parse_flags("SD")[0] == "D"  # the last precision letter wins
    with pytest.raises(ssa.SteadyScanError):
        ssa.parse_flags("CE")
    with pytest.raises(ssa.SteadyScanError):
        ssa.parse_flags("DM")


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    assert N.lib().ss_create(0, ctypes.byref(h)) == N.SS_ERR_NOT_AVAILABLE
    assert not h
