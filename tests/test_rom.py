"""ROM capture/replay format (host only; no GPU needed)."""
import struct

import numpy as np
import pytest

import steadyscan_amd as ssa
from steadyscan_amd import workloads as W


@pytest.fixture
def call():
    T, fr, geom, _ = W.config0(n_frames=5, size=32, seed=3)
    a = W.group_args(T, fr, geom, max_iter=7, flags="S", lambda_=0.5)
    a["bmData"] = np.asfortranarray(np.rint(a["bmData"]).astype(np.uint16))
    a["est_dx_knots"] = np.asfortranarray(np.arange(a["est_dx_knots"].size, dtype=float).reshape(a["est_dx_knots"].shape))
    return a


def test_round_trip_is_exact(call, tmp_path):
    p = tmp_path / "call.rom"
    ssa.save_rom(p, call)
    back = ssa.load_rom(p)
    for k in ("template_image", "xgrad", "ygrad", "tfrac", "basex", "basey", "bmData", "est_dx_knots", "est_dy_knots"):
        a, b = np.asarray(call[k]), back[k]
        assert a.shape == b.shape and a.dtype == b.dtype, k
        assert np.array_equal(np.nan_to_num(a, nan=7.5), np.nan_to_num(b, nan=7.5)), k
        assert np.array_equal(np.isnan(a.astype(float)), np.isnan(b.astype(float))), k
    assert np.array_equal(np.asarray(call["base_mask"], bool), back["base_mask"])
    assert np.array_equal(np.trunc(call["min_points_per_block"]), back["min_points_per_block"])
    for k in ("max_iter", "move_thresh", "halt_correlation", "correlation_increase_threshold", "filter_strength",
              "max_filter_iterations", "filter_deactivation_threshold", "lambda_",
              "linesearch_reduction_multiplier", "solver_id", "filter_id", "group_size", "nsetframes"):
        assert back[k] == call[k], k
    assert back["flags"] == "S"


def test_format_header_and_records(call, tmp_path):
    p = tmp_path / "call.rom"
    ssa.save_rom(p, call)
    raw = p.read_bytes()
    assert raw[:8] == b"SSROM001"
    bom, count = struct.unpack("<II", raw[8:16])
    assert bom == 0x01020304 and count == 29
    tags, off = [], 16
    for _ in range(count):
        tag = raw[off:off + 4].decode()
        code, rows, cols = struct.unpack("<III", raw[off + 4:off + 16])
        size = {1: 8, 2: 4, 3: 2, 4: 2, 5: 1, 6: 4}[code] * rows * cols
        if tag == "TMPL":
            t = np.frombuffer(raw[off + 16:off + 16 + size], "<f8").reshape(rows, cols, order="F")
            assert np.array_equal(t, call["template_image"])
        if tag == "DATA":
            assert code == 3 and (rows, cols) == call["bmData"].shape
        tags.append(tag)
        off += 16 + size
    assert off == len(raw)
    assert tags[0] == "PREC" and "DATA" in tags and "MINP" in tags


def test_corrupt_file_is_rejected(call, tmp_path):
    p = tmp_path / "call.rom"
    ssa.save_rom(p, call)
    raw = bytearray(p.read_bytes())
    raw[3] = ord("X")
    p.write_bytes(bytes(raw))
    with pytest.raises(ssa.SteadyScanError):
        ssa.load_rom(p)
    p.write_bytes(bytes(p.read_bytes()[:100]))
    with pytest.raises(ssa.SteadyScanError):
        ssa.load_rom(p)


def _records(raw):
    """(tag, header offset, code, rows, cols, payload size) of each record."""
    out, off = [], 16
    for _ in range(struct.unpack("<I", raw[12:16])[0]):
        code, rows, cols = struct.unpack("<III", raw[off + 4:off + 16])
        size = {1: 8, 2: 4, 3: 2, 4: 2, 5: 1, 6: 4}[code] * rows * cols
        out.append((raw[off:off + 4].decode(), off, code, rows, cols, size))
        off += 16 + size
    return out


def test_record_larger_than_the_file_is_rejected_without_allocating(call, tmp_path):
    p = tmp_path / "call.rom"
    ssa.save_rom(p, call)
    raw = bytearray(p.read_bytes())
    _, off, _, _, _, _ = next(r for r in _records(raw) if r[0] == "DATA")
    # 2^32-1 x 2^32-1 cells of 2 bytes: the size wraps 64 bits unless checked
    raw[off + 8:off + 16] = struct.pack("<II", 0xFFFFFFFF, 0xFFFFFFFF)
    p.write_bytes(bytes(raw))
    with pytest.raises(ssa.SteadyScanError):
        ssa.load_rom(p)
    raw[off + 8:off + 16] = struct.pack("<II", 1 << 20, 1 << 10)  # 2 GiB claimed in a small file
    p.write_bytes(bytes(raw))
    with pytest.raises(ssa.SteadyScanError):
        ssa.load_rom(p)


@pytest.mark.parametrize("tag", ["BASX", "ESTY", "MINP", "XGRD"])
def test_records_whose_shapes_disagree_are_rejected(call, tmp_path, tag):
    p = tmp_path / "call.rom"
    ssa.save_rom(p, call)
    raw = bytearray(p.read_bytes())
    _, off, code, rows, cols, _ = next(r for r in _records(raw) if r[0] == tag)
    # the same cell count in another shape keeps the file well formed
    assert rows > 1 or cols > 1
    raw[off + 8:off + 16] = struct.pack("<II", rows * cols, 1) if cols > 1 else struct.pack("<II", 1, rows)
    p.write_bytes(bytes(raw))
    with pytest.raises(ssa.SteadyScanError):
        ssa.load_rom(p)


def test_frame_correlation_flag_round_trips(call, tmp_path):
    # flag 'c' changes output 3 (the last iteration's correlations), so a
    # replay must keep it
    call["flags"] = "Sc"
    p = tmp_path / "call_c.rom"
    ssa.save_rom(p, call)
    assert ssa.load_rom(p)["flags"] == "Sc"
