"""The GPU path against the golden fixtures (tests/golden/*.npz): the FP64
oracle's outputs on small seeded groups, committed, so this test runs no
oracle.  Precision 'D' to the parity bar of tests/test_gpu_parity.py (knots
within 1e-9 px, equal iteration counts and stop criteria), 'S' to its FP32
bar."""
import os

import numpy as np
import pytest

import steadyscan_amd as ssa
from steadyscan_amd import _native as N
from steadyscan_amd.workloads import ARG_ORDER

from test_golden import FIXTURES, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_gpu_matches_golden(path):
    a, want = load(path)
    info = {}
    out = ssa.align_group_to_template(*[a[k] for k in ARG_ORDER], nargout=12, info=info)
    precision = ssa.parse_flags(a["flags"])[0]
    kx, ky = np.ravel(out[0]), np.ravel(out[1])
    d = np.concatenate([kx - want["knots_x"], ky - want["knots_y"]])
    rms = float(np.sqrt(np.mean(d ** 2)))
    name = os.path.basename(path)
    print(f"{name}: iterations gpu={int(out[3])} golden={int(want['iterations'])}, knot RMS {rms:.3e} px", flush=True)
    if precision == "D":
        assert int(out[3]) == int(want["iterations"])
        assert info["stop_criterion"] == N.STOP_NAMES[int(want["stop_criterion"])]
        assert np.max(np.abs(d)) <= 1e-9
        for i, k in ((5, "residrms"), (2, "frame_corr")):
            assert np.allclose(np.ravel(out[i]), want[k], rtol=1e-9, atol=1e-12, equal_nan=True), k
        assert np.mean(np.asarray(out[4], bool) == want["mask"].astype(bool)) >= 0.99999
        if "frame_corr_iter" in want:
            assert np.max(np.abs(out[10] - want["frame_corr_iter"])) <= 1e-9
        if "iteration_errvals" in want:
            assert np.allclose(out[11], want["iteration_errvals"], rtol=1e-9, atol=0)
    else:
        assert rms <= 0.01
        assert np.allclose(np.ravel(out[5]), want["residrms"], rtol=1e-3, atol=1e-4)
