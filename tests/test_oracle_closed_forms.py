"""Closed-form pins of three oracle functions (CPU only).

- Scan geometry (init_settings, steady_scan.m:787-827) beyond integer
  unidirectional spacing: bidirectional scans (odd lines reversed, :792-795),
  non-integer knot spacing (knot times by linspace over the scanned time,
  :806-811), horizontal cropping (knot times from the uncropped lines,
  :801-804) and a given subrec (:819-824).  Every expected value below is
  written from the definitions (pixel scan time, knot time, bin, fraction),
  not from the oracle's code.
- ILU(0) of solver id 0 (cusparse csrilu02 + csrsv2 on the x-then-y
  ordered normal equations, cusparsesolver.cu:119-158): a dense textbook
  ILU(0) (IKJ order, zero fill outside the matrix's pattern) in numpy.
- Frame correlation (hd_frameCorrelation, task_correlation.cu:101-183):
  np.corrcoef of the warped template (sampled here by numpy from the state's
  coordinates) and the image over each frame's masked pixels; after the
  global correlation (the per-iteration correlations of flag 'c',
  task.cpp:286-306), the same statistic with the pixels past the global
  correlation's partial swap-back (task_correlation.cu:97) holding the
  residual.
- Per-iteration errvals (flag 'e', task.cpp:362-370): the masked mean
  squared residual in numpy at the knots each half-iteration started from,
  and the rows and save points against a half-iteration replay.

Each check is also run against copies of the oracle built with one planted
mistake (ORC_PLANT in oracle/ss_oracle.c), and must reject every one of them.
"""
import ctypes
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from ctypes import POINTER, c_double, c_int32

import numpy as np
import pytest

import steadyscan_amd as ssa
from oracle import oracle as O
from steadyscan_amd import workloads as W

ORACLE_DIR = os.path.dirname(os.path.abspath(O.__file__))
PLANTS = {1: "bidirectional flips even lines", 2: "knot times from cropped lines",
          3: "linspace divides by n", 4: "ILU(0) on unsorted rows", 5: "ILU(0) wrong pivot",
          6: "frame correlation means over all pixels", 7: "frame correlation of the residual",
          8: "global correlation swaps every pixel back", 9: "iteration errvals rows swapped"}


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    """Copies of the oracle, each with one planted mistake, built with the
    oracle's own flags."""
    cc = shutil.which("gcc") or "gcc"
    d = tmp_path_factory.mktemp("planted")
    src = os.path.join(ORACLE_DIR, "ss_oracle.c")

    def build(k):
        out = str(d / f"liboracle_plant{k}.so")
        subprocess.run([cc, "-O1", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", f"-DORC_PLANT={k}",
                        "-shared", "-o", out, src, "-lm"], check=True, capture_output=True)
        return k, O.load(out)

    with ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build, PLANTS))


def _p(a):
    return a.ctypes.data_as(POINTER(c_double))


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------
def oracle_geometry(L, H, Wd, th, tw, nk, crop=(0, 0, 0, 0), bidi=False, subrec=None):
    g = O.OrcGeometryIn(nlines=H, linewidth=Wd, template_h=th, template_w=tw, nknots=nk,
                        bidirectional=1 if bidi else 0, min_blockpoints_ratio=0.3)
    for i, v in enumerate(crop):
        g.cropping[i] = v
    if subrec is not None:
        g.has_subrec = 1
        for i, v in enumerate(subrec):
            g.subrec[i] = v
    hc, wc = c_int32(), c_int32()
    assert L.orc_geometry(ctypes.byref(g), None, None, None, None, ctypes.byref(hc), ctypes.byref(wc)) == 0
    n = hc.value * wc.value
    pk = np.zeros(n, np.int32)
    tf, bx, by = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = L.orc_geometry(ctypes.byref(g), pk.ctypes.data_as(POINTER(c_int32)), _p(tf), _p(bx), _p(by),
                        ctypes.byref(hc), ctypes.byref(wc))
    shp = (hc.value, wc.value)
    return None if rc else dict(prevknotind=pk.reshape(shp, order="F"), tfrac=tf.reshape(shp, order="F"),
                                basex=bx.reshape(shp, order="F"), basey=by.reshape(shp, order="F"))


def product_geometry(H, Wd, th, tw, nk, crop=(0, 0, 0, 0), bidi=False, subrec=None):
    return ssa.init_settings(H, Wd, (th, tw), nknots=nk, cropping=crop, isbidirectional=bidi, subrec=subrec)


def scan_time(H, Wd, bidi):
    """Scan time of every raw pixel in units of the frame time: pixel (i, j)
    is the (i*W + j)-th pixel scanned (j reversed on odd 0-based lines of a
    bidirectional scan), sampled at the middle of its dwell."""
    i, j = np.meshgrid(np.arange(H), np.arange(Wd), indexing="ij")
    jj = np.where((i % 2 == 1) & bidi, Wd - 1 - j, j)
    return (i * Wd + jj + 0.5) / (H * Wd)


def knot_times(H, Wd, nk, crop, bidi):
    """Knot times from their definition: at integer lines per knot, the
    start of the knot's first (uncropped) line; otherwise evenly spaced over
    the scan time of the cropped data widened by half a dwell at each end.
    The first knot is moved to 0 and the last to 1."""
    cl, cr = crop[0], crop[1]
    L = H / (nk - 1)
    if L == int(L):
        t = np.array([k * int(L) / H for k in range(nk)], float)
    else:
        tc = scan_time(H, Wd, bidi)[:, cl:Wd - cr]
        lo, hi = tc.min() - 0.5 / (H * Wd), tc.max() + 0.5 / (H * Wd)
        t = lo + np.arange(nk) * (hi - lo) / (nk - 1)
    t[0], t[-1] = 0.0, 1.0
    return t


def geometry_ok(g, H, Wd, th, tw, nk, crop=(0, 0, 0, 0), bidi=False, subrec=None):
    """True when prevknotind/tfrac/basex/basey match the closed forms."""
    if g is None:
        return False
    cl, cr = crop[0], crop[1]
    t = scan_time(H, Wd, bidi)[:, cl:Wd - cr]
    tk = knot_times(H, Wd, nk, crop, bidi)
    pk = g["prevknotind"]
    if pk.shape != t.shape or pk.min() < 1 or pk.max() > nk - 1:
        return False
    lo, hi = tk[pk - 1], tk[pk]
    # the pixel lies in its knot interval and tfrac is its position there
    if not (np.all(lo <= t + 1e-13) and np.all(t < hi + 1e-13)):
        return False
    if not np.allclose(g["tfrac"], (t - lo) / (hi - lo), rtol=0, atol=1e-12):
        return False
    # template coordinates: subrec spread linearly over the raw line width
    # and the (vertically uncropped) lines; the default subrec maps a frame
    # cropped to the template's size onto the template pixel for pixel
    Hc, Wc = t.shape
    if subrec is None:
        subrec = ((1 - cl, tw + cr, 1, th) if (Hc, Wc) == (th, tw) else (1, tw, 1, th))
    i, j = np.meshgrid(np.arange(Hc), np.arange(Wc), indexing="ij")
    bx = subrec[0] + (j + cl) / (Wd - 1) * (subrec[1] - subrec[0])
    by = subrec[2] + i / (Hc - 1) * (subrec[3] - subrec[2])
    return bool(np.allclose(g["basex"], bx, rtol=0, atol=1e-12) and np.allclose(g["basey"], by, rtol=0, atol=1e-12))


GEOMETRY_CASES = [
    # H, W, template, nknots, crop, bidirectional, subrec
    (16, 8, (16, 8), 5, (0, 0, 0, 0), True, None),       # bidirectional, 4 lines per knot
    (17, 10, (17, 10), 4, (0, 0, 0, 0), True, None),     # bidirectional, 16/3 lines per knot
    (15, 9, (15, 9), 5, (0, 0, 0, 0), False, None),      # 3.75 lines per knot
    (60, 48, (60, 48), 8, (0, 0, 0, 0), False, None),    # 60/7 lines per knot
    (16, 12, (16, 9), 5, (2, 1, 0, 0), False, None),     # cropping, integer spacing
    (15, 12, (15, 9), 5, (2, 1, 0, 0), True, None),      # cropping, non-integer, bidirectional
    (32, 32, (32, 32), 9, (0, 0, 0, 0), False, (-3.0, 40.0, 2.0, 30.0)),  # subrec
    (20, 16, (40, 40), 6, (0, 0, 0, 0), False, None),    # frame smaller than the template
]


@pytest.mark.parametrize("case", GEOMETRY_CASES)
def test_geometry_closed_forms(case):
    H, Wd, (th, tw), nk, crop, bidi, subrec = case
    assert geometry_ok(oracle_geometry(O.lib(), H, Wd, th, tw, nk, crop, bidi, subrec),
                       H, Wd, th, tw, nk, crop, bidi, subrec)
    assert geometry_ok(product_geometry(H, Wd, th, tw, nk, crop, bidi, subrec),
                       H, Wd, th, tw, nk, crop, bidi, subrec)


def test_geometry_closed_forms_exact_cases():
    # two cases written out by hand: bidirectional integer spacing, and
    # 3.75 lines per knot (knots at t = k/4)
    H, Wd, nk = 16, 8, 5
    g = oracle_geometry(O.lib(), H, Wd, H, Wd, nk, bidi=True)
    i, j = np.meshgrid(np.arange(H), np.arange(Wd), indexing="ij")
    jj = np.where(i % 2 == 1, Wd - 1 - j, j)
    assert np.array_equal(g["prevknotind"], i // 4 + 1)
    assert np.allclose(g["tfrac"], ((i % 4) * Wd + jj + 0.5) / 32, rtol=0, atol=1e-13)
    g = oracle_geometry(O.lib(), 15, 9, 15, 9, 5)
    t = (np.arange(135).reshape(15, 9) + 0.5) / 135
    assert np.allclose((g["prevknotind"] - 1 + g["tfrac"]) / 4, t, rtol=0, atol=1e-13)


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_geometry_checks_reject_planted_mistakes(planted, plant):
    rejected = [not geometry_ok(oracle_geometry(planted[plant], H, Wd, th, tw, nk, crop, bidi, subrec),
                                H, Wd, th, tw, nk, crop, bidi, subrec)
                for H, Wd, (th, tw), nk, crop, bidi, subrec in GEOMETRY_CASES]
    assert any(rejected), PLANTS[plant]


# ---------------------------------------------------------------------------
# ILU(0), solver id 0
# ---------------------------------------------------------------------------
def coupled_system(u, seed):
    """Random SPD system with the pattern of the normal equations in the
    compacted x-then-y order (x_k couples to x_{k-1..k+1} and y_{k-1..k+1}),
    as COO with unused INT_MAX slots, in random order."""
    rng = np.random.default_rng(seed)
    n = 2 * u
    J = []
    for k in range(u):
        for _ in range(6):
            row = np.zeros(n)
            row[[k, u + k]] = rng.standard_normal(2)
            if k + 1 < u:
                row[[k + 1, u + k + 1]] = rng.standard_normal(2)
            J.append(row)
    H = np.array(J).T @ np.array(J)
    rr, cc = np.nonzero(H)
    rows = np.concatenate([rr, np.full(len(rr) // 3, 2**31 - 1)]).astype(np.int32)
    cols = np.concatenate([cc, np.zeros(len(rr) // 3)]).astype(np.int32)
    vals = np.concatenate([H[rr, cc], np.full(len(rr) // 3, 7.0)])
    perm = rng.permutation(len(rows))
    return H, rows[perm], cols[perm], vals[perm], rng.standard_normal(n)


def dense_ilu0_solve(H, b):
    """Textbook ILU(0): Gaussian elimination in IKJ order that updates only
    entries inside the matrix's nonzero pattern, then L z = b (unit lower)
    and U x = z."""
    A = H.astype(np.float64).copy()
    P = H != 0
    n = len(A)
    for i in range(1, n):
        for k in range(i):
            if not P[i, k]:
                continue
            A[i, k] = A[i, k] / A[k, k]
            for j in range(k + 1, n):
                if P[i, j] and P[k, j]:
                    A[i, j] = A[i, j] - A[i, k] * A[k, j]
    z = np.zeros(n)
    for i in range(n):
        s = b[i]
        for k in range(i):
            if P[i, k]:
                s = s - A[i, k] * z[k]
        z[i] = s
    x = np.zeros(n)
    for i in range(n - 1, -1, -1):
        s = z[i]
        for j in range(i + 1, n):
            if P[i, j]:
                s = s - A[i, j] * x[j]
        x[i] = s / A[i, i]
    return x


def oracle_ilu0(L, r, c, v, b):
    x = np.zeros(len(b))
    L.orc_solve_coo(len(b), len(r), r.ctypes.data_as(POINTER(c_int32)), c.ctypes.data_as(POINTER(c_int32)),
                    _p(np.ascontiguousarray(v)), _p(np.ascontiguousarray(b)), _p(x), 0)
    return x


ILU_CASES = [(2, 20), (5, 21), (12, 22), (40, 23)]


@pytest.mark.parametrize("u,seed", ILU_CASES)
def test_ilu0_equals_dense_textbook(u, seed):
    H, r, c, v, b = coupled_system(u, seed)
    x = oracle_ilu0(O.lib(), r, c, v, b)
    xd = dense_ilu0_solve(H, b)
    assert np.allclose(x, xd, rtol=1e-12, atol=1e-12 * np.abs(xd).max())
    # it is an approximate solve of this coupled system (fill was dropped),
    # and the ordering matters: the same system in interleaved (x_k, y_k)
    # order has another incomplete factorisation
    assert not np.allclose(x, np.linalg.solve(H, b), rtol=1e-6) or u < 3
    perm = np.ravel(np.column_stack([np.arange(u), u + np.arange(u)]))
    xi = np.empty_like(xd)
    xi[perm] = dense_ilu0_solve(H[np.ix_(perm, perm)], b[perm])
    assert not np.allclose(xi, xd, rtol=1e-6) or u < 3


@pytest.mark.parametrize("plant", [4, 5])
def test_ilu0_check_rejects_planted_mistakes(planted, plant):
    rejected = []
    for u, seed in ILU_CASES:
        H, r, c, v, b = coupled_system(u, seed)
        x = oracle_ilu0(planted[plant], r, c, v, b)
        xd = dense_ilu0_solve(H, b)
        rejected.append(not np.allclose(x, xd, rtol=1e-12, atol=1e-12 * np.abs(xd).max()))
    assert any(rejected), PLANTS[plant]


# ---------------------------------------------------------------------------
# frame correlation
# ---------------------------------------------------------------------------
def bilinear_np(T, x, y):
    ix, iy = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - ix, y - iy
    return ((1 - fx) * ((1 - fy) * T[iy, ix] + fy * T[iy + 1, ix]) +
            fx * ((1 - fy) * T[iy, ix + 1] + fy * T[iy + 1, ix + 1]))


def correlation_case(seed):
    """A small group in a state whose knots push part of every frame off the
    template (so the mask is not full)."""
    T = W.make_template(40, 40, 15, (2.0, 4.0), offset=10.0, seed=seed)
    dx, dy = W.make_motion(4, 32, 0.05, 0.0, (1.0, 1.0), seed=seed + 1)
    fr = W.make_frames(T, dx, dy, 32, 32, np.float64, True, seed=seed + 2)
    geom = ssa.init_settings(32, 32, T.shape, nknots=5, subrec=(3.0, 38.0, 4.0, 37.0))
    a = W.group_args(T, fr, geom, max_iter=3)
    rng = np.random.default_rng(seed)
    N = 4 * 4 + 1
    p = np.concatenate([rng.uniform(-5, 5, N), rng.uniform(-5, 5, N)])
    return T, a, p


def frame_correlation_ok(L, seed):
    T, a, p = correlation_case(seed)
    st = O.State(a, L=L)
    st.set_p(p)
    st.assemble(0)
    mask, x, y, _ = st.pixels()
    corr = st.frame_correlation()
    I = np.asarray(a["bmData"], np.float64)
    nb = a["basex"].shape[1]
    assert 0.02 < 1 - mask.mean() < 0.6  # a partial mask
    for f in range(a["nsetframes"]):
        cols = slice(f * nb, (f + 1) * nb)
        m = mask[:, cols]
        warped = bilinear_np(T, x[:, cols][m], y[:, cols][m])
        expected = np.corrcoef(warped, I[:, cols][m])[0, 1]
        if not abs(corr[f] - expected) <= 1e-12:
            return False
    return True


@pytest.mark.parametrize("seed", [61, 62, 63])
def test_frame_correlation_is_pearson_of_masked_pixels(seed):
    assert frame_correlation_ok(O.lib(), seed)


@pytest.mark.parametrize("plant", [6, 7])
def test_frame_correlation_check_rejects_planted_mistakes(planted, plant):
    assert not all(frame_correlation_ok(planted[plant], s) for s in (61, 62, 63)), PLANTS[plant]


# ---------------------------------------------------------------------------
# per-frame correlations (hd_frameCorrelation) and per-iteration errvals
# ---------------------------------------------------------------------------
def pearson(a, b):
    va, vb = a - a.mean(), b - b.mean()
    return (va * vb).sum() / np.sqrt((va * va).sum() * (vb * vb).sum())


def swap_back_ok(L):
    # 3 frames of 3 blocks x 192 rows: 9 columns, 1728 pixels, so the global
    # correlation's swap-back (task_correlation.cu:97, one thread per column
    # rounded up to 256) restores row-major pixels 0..255 only: rows 0-27
    # and the first 4 columns of row 28
    T = W.make_template(24, 24, 6, (1.5, 3.0), offset=5.0, seed=11)
    dx, dy = W.make_motion(3, 24, 0.05, 0.0, (0.5, 0.5), seed=12)
    fr = W.make_frames(T, dx, dy, 24, 24, np.float64, True, seed=13)
    geom = ssa.init_settings(24, 24, T.shape, nknots=4)
    a = W.group_args(T, fr, geom, max_iter=5)
    p = np.random.default_rng(11).uniform(-0.6, 0.6, 2 * (3 * 3 + 1))
    fh, nb, nF = a["basex"].shape[0], a["basex"].shape[1], a["nsetframes"]
    fw = nb * nF
    assert fh * fw > 256 and 256 % fw != 0
    I = np.asarray(a["bmData"], dtype=np.float64)
    st = O.State(a, L=L)
    st.set_p(p)
    st.assemble(0)
    m, _, _, d = st.pixels()
    fc = st.frame_correlation()
    for f in range(nF):
        cols = slice(f * nb, (f + 1) * nb)
        sel = m[:, cols]
        # warped template = image - difference image
        if not abs(fc[f] - pearson((I - d)[:, cols][sel], I[:, cols][sel])) < 1e-12:
            return False
    st.global_correlation()
    fcq = st.frame_correlation()
    rm = np.arange(fh)[:, None] * fw + np.arange(fw)[None, :]  # row-major index
    prefix = rm < min(-(-fw // 256) * 256, fh * fw)
    v = np.where(prefix, I - d, d)
    for f in range(nF):
        cols = slice(f * nb, (f + 1) * nb)
        sel = m[:, cols]
        if not abs(fcq[f] - pearson(v[:, cols][sel], I[:, cols][sel])) < 1e-12:
            return False
        assert abs(fcq[f] - fc[f]) > 0.05  # the residual, not the template, past the prefix
    return True


def test_frame_correlation_after_global_correlation():
    assert swap_back_ok(O.lib())


def test_swap_back_check_rejects_planted_mistake(planted):
    assert not swap_back_ok(planted[8]), PLANTS[8]


def errval_np(T, a, p, f):
    fh, nb, nF = a["basex"].shape[0], a["basex"].shape[1], a["nsetframes"]
    N = nb * nF + 1
    th, tw = T.shape
    e, n = 0.0, 0
    for c in range(f * nb, (f + 1) * nb):
        b = c % nb
        f_ = a["tfrac"][:, b]
        x = a["basex"][:, b] + p[c] * (1 - f_) + p[c + 1] * f_
        y = a["basey"][:, b] + p[c + N] * (1 - f_) + p[c + N + 1] * f_
        m = a["base_mask"][:, b].astype(bool) & (x >= 0) & (x < tw - 1) & (y >= 0) & (y < th - 1)
        r = np.asarray(a["bmData"], dtype=np.float64)[m, c] - bilinear_np(T, x[m], y[m])
        e += (r * r).sum()
        n += m.sum()
    return e / n


def iteration_errvals_ok(L):
    T, fr, geom, _ = W.config0(n_frames=4, size=32, seed=17)
    a = W.group_args(T, fr, geom, max_iter=8, filter_id=0)
    a["flags"] = "De"
    r = O.align_group_to_template(a, with_iterations=True, L=L)
    it, E, allp = r["iterations"], r["iteration_errvals"], r["all_params"]
    nF, mi = a["nsetframes"], 8
    assert E.shape == (2 * mi, nF) and it >= 3
    # saved after every iteration that continues: all but the last
    assert np.all(E[2 * (it - 1):, :] == 0) and np.all(E[:2 * (it - 1), :] > 0)
    assert np.all(r["rg_params"][:, 2 * it:] == 0)
    assert np.array_equal(r["rg_params"][:, 2 * it - 1], allp[:, it - 1])  # the last half-iteration's p
    # replay: half-iteration by half-iteration on a state
    st = O.State(a, L=L)
    for k in range(1, it):
        for shift in range(2):
            start = st.get()[0]
            st.assemble(shift)
            st.linesearch(shift)
            ev = st.get()[2]
            # output 14: the parameters after this half-iteration
            assert np.array_equal(st.get()[0], r["rg_params"][:, shift + 2 * (k - 1)])
            # the front errval of frames > 0: the masked mean squared residual
            # at the knots the half-iteration started from (frame 0's slot
            # takes the line search's accepted value, task_linesearch.cu:91)
            for f in range(1, nF):
                e = errval_np(T, a, start, f)
                if not abs(E[2 * (k - 1) + shift, f] - e) < 1e-9 * e:
                    return False
        if not (np.array_equal(E[2 * (k - 1), :], ev[:nF]) and np.array_equal(E[2 * (k - 1) + 1, :], ev[nF:])):
            return False
        assert np.array_equal(st.get()[0], allp[:, k - 1])
    return True


def test_iteration_errvals():
    assert iteration_errvals_ok(O.lib())


def test_iteration_errvals_check_rejects_planted_mistake(planted):
    assert not iteration_errvals_ok(planted[9]), PLANTS[9]
