"""The golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) hold small seeded groups with all 26 mex
arguments and the FP64 oracle's outputs.  The oracle must still reproduce
them bit for bit: a change to the oracle that moves any output shows here."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.load(path)


def test_fixtures_present():
    assert len(FIXTURES) >= 4


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    a, want = load(path)
    assert a["bmData"].size and a["template_image"].ndim == 2
    ref = O.align_group_to_template(a, threads=4)
    for k, v in want.items():
        got = np.asarray(ref[k])
        assert got.shape == v.shape and np.array_equal(got, v, equal_nan=True), k
