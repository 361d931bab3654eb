"""Pins of the oracle's invert_bilinear_sampling (steady_scan.m:290-404).

The checks build the reference's system independently, as dense matrices
from its own definitions (bilinear_map_linear_constraints, steady_scan.m:438-499;
imageadjdiffquadmat, :406-436; pushgrid, :575-599), and compare the oracle's
stencil, energy, sweep and stopping loop against them.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from oracle import driver as OD

PF = ctypes.POINTER(ctypes.c_float)
PU8 = ctypes.POINTER(ctypes.c_uint8)
PI32 = ctypes.POINTER(ctypes.c_int32)
N, NE, E, SE, S, SW, W, NW = range(8)
SLOT = {(-1, 0): N, (-1, 1): NE, (0, 1): E, (1, 1): SE, (1, 0): S, (1, -1): SW, (0, -1): W, (-1, -1): NW}


def geometry(nlines, linewidth, th, tw, nknots, subrec=None):
    (hc, wc), pk, tf, bx, by = OD._geometry(nlines, linewidth, th, tw, nknots, (0, 0, 0, 0), subrec, False, 0.3)
    shape = (hc, wc)
    return dict(prevknotind=pk.reshape(shape, order="F"), tfrac=tf.reshape(shape, order="F"),
                basex=bx.reshape(shape, order="F"), basey=by.reshape(shape, order="F"), nknots=nknots)


def positions(geom, dx, dy):
    pk = geom["prevknotind"].ravel(order="F") - 1
    tf = geom["tfrac"].ravel(order="F")
    x = geom["basex"].ravel(order="F") + dx[pk] * (1 - tf) + dx[pk + 1] * tf
    y = geom["basey"].ravel(order="F") + dy[pk] * (1 - tf) + dy[pk + 1] * tf
    return x, y


def reference_system(frame, dx, dy, geom, th, tw, regfac, aspect=1.0):
    """Dense ww, wy (bilinear constraints), G'G and the pushgrid resample."""
    x, y = positions(geom, dx, dy)
    d = np.asarray(frame, dtype=float).ravel(order="F")
    eps = np.finfo(float).eps
    n = th * tw
    ww = np.zeros((n, n))
    wy = np.zeros(n)
    m = (x > 1 + eps) & (x < tw - eps) & (y > 1 + eps) & (y < th - eps)
    for xx, yy, v in zip(x[m], y[m], d[m]):
        xi, yi = np.floor(xx), np.floor(yy)
        fx, fy = xx - xi, yy - yi
        c0 = int(yi - 1 + th * (xi - 1))
        cn = np.array([c0, c0 + 1, c0 + th, c0 + th + 1])
        w = np.array([(1 - fx) * (1 - fy), (1 - fx) * fy, fx * (1 - fy), fx * fy])
        wy[cn] += w * v
        ww[np.ix_(cn, cn)] += np.outer(w, w)
    # G: one row per adjacent pair, [1 -1] horizontally, [vfac -vfac] vertically
    vfac = 1.0 / aspect
    rows = []
    for c in range(tw):
        for r in range(th):
            i = r + c * th
            if c + 1 < tw:
                g = np.zeros(n); g[i], g[i + th] = 1, -1; rows.append(g)
    for c in range(tw):
        for r in range(th):
            i = r + c * th
            if r + 1 < th:
                g = np.zeros(n); g[i], g[i + 1] = vfac, -vfac; rows.append(g)
    G = np.array(rows)
    # pushgrid: z(ind) = z(ind) + v keeps the last value of a repeated index
    z = np.zeros(n, np.float32)
    nn = np.zeros(n, np.float32)
    pm = (x >= 1) & (x < tw) & (y >= 1) & (y < th)
    xs, ys, ds = x[pm], y[pm], d[pm]
    xi, yi = np.floor(xs), np.floor(ys)
    fx, fy = xs - xi, ys - yi
    mi = (yi - 1 + th * (xi - 1)).astype(int)
    for arr, terms in ((z, [(0, (1 - fx) * (1 - fy) * ds), (1, (1 - fx) * fy * ds), (th, fx * (1 - fy) * ds),
                            (th + 1, fx * fy * ds)]),
                       (nn, [(1, (1 - fx) * fy), (th, fx * (1 - fy)), (0, (1 - fx) * (1 - fy)), (th + 1, fx * fy)])):
        for off, val in terms:
            old = arr.copy()
            for q in range(len(mi)):
                arr[mi[q] + off] = old[mi[q] + off] + np.float32(val[q])
    with np.errstate(invalid="ignore", divide="ignore"):
        iv = z / nn
    return ww, wy, regfac * (G.T @ G), iv


def oracle_system(frame, dx, dy, geom, th, tw, regfac, aspect=1.0):
    n = th * tw
    qm = np.zeros(n * 9)
    wy = np.zeros(n)
    iv = np.zeros(n, np.float32)
    mask = np.zeros(n, np.uint8)
    fr = np.asfortranarray(frame, dtype=float)
    any_ = O.lib().orc_ibs_system(O._p(fr), fr.size, O._p(np.ascontiguousarray(dx, dtype=float)),
                                  O._p(np.ascontiguousarray(dy, dtype=float)),
                                  np.asfortranarray(geom["prevknotind"], dtype=np.int32).ctypes.data_as(PI32),
                                  O._p(np.asfortranarray(geom["tfrac"])), O._p(np.asfortranarray(geom["basex"])),
                                  O._p(np.asfortranarray(geom["basey"])), th, tw, regfac, aspect, O._p(qm), O._p(wy),
                                  iv.ctypes.data_as(PF), mask.ctypes.data_as(PU8))
    return any_, qm.reshape(n, 9), wy, iv, mask.astype(bool)


def stencil_to_dense(qm, th, tw):
    n = th * tw
    Q = np.zeros((n, n))
    for i in range(n):
        r, c = i % th, i // th
        Q[i, i] = qm[i, 0]
        for (dr, dc), s in SLOT.items():
            rr, cc = r + dr, c + dc
            if 0 <= rr < th and 0 <= cc < tw:
                Q[i, rr + cc * th] = qm[i, 1 + s]
    return Q


def masked_reference(ww, wy, reg, iv):
    sww = ww.sum(axis=0)
    mask = (sww > 0) & ~np.isnan(iv)
    Q = ww + reg
    Q[~mask, :] = 0
    Q[:, ~mask] = 0
    return Q, np.where(mask, wy, 0.0), mask, (sww > 0).any()


def case(seed, th=10, tw=10, nknots=6, amp=0.6, nlines=None, linewidth=None):
    rng = np.random.default_rng(seed)
    nlines = nlines or th
    linewidth = linewidth or tw
    geom = geometry(nlines, linewidth, th, tw, nknots)
    frame = 50 + 20 * rng.random((nlines, linewidth))
    dx = rng.uniform(-amp, amp, nknots)
    dy = rng.uniform(-amp, amp, nknots)
    return geom, frame, dx, dy


def neighbour(i, d, nrows, npix):
    j = ctypes.c_long()
    m = O.lib().orc_ibs_neighbour(i, d, nrows, npix, ctypes.byref(j))
    return m, j.value


def test_find_functions_and_their_quirks():
    nrows, ncols = 4, 3
    npix = nrows * ncols
    # interior pixel (row 1, column 1) -> index 5: all eight neighbours
    assert [neighbour(5, d, nrows, npix) for d in range(8)] == [
        (1, 4), (1, 8), (1, 9), (1, 10), (1, 6), (1, 2), (1, 1), (1, 0)]
    # top row: no N, no NE; find_NW has no top-row test and wraps to the
    # bottom row two columns back (0-based: 8 - 4 - 1 = 3)
    assert neighbour(4, N, nrows, npix)[0] == 0 and neighbour(4, NE, nrows, npix)[0] == 0
    assert neighbour(8, NW, nrows, npix) == (1, 3)
    assert neighbour(4, NW, nrows, npix)[0] == 0            # 1-based index 0 is out
    # find_SE uses I < npixels: the SE neighbour that is the last pixel is dropped
    i = (nrows - 2) + (ncols - 2) * nrows
    assert neighbour(i, SE, nrows, npix)[0] == 0
    assert neighbour(i - 1, SE, nrows, npix) == (1, i - 1 + nrows + 1)
    # bottom row: no S, SE, SW; first column: no W, NW, SW; last column: no E, NE, SE
    assert all(neighbour(nrows - 1 + nrows, d, nrows, npix)[0] == 0 for d in (SE, S, SW))
    assert all(neighbour(1, d, nrows, npix)[0] == 0 for d in (W, NW, SW))
    assert all(neighbour(1 + 2 * nrows, d, nrows, npix)[0] == 0 for d in (E, NE, SE))


@pytest.mark.parametrize("seed,th,tw,aspect", [(1, 10, 10, 1.0), (2, 8, 12, 2.0), (3, 9, 7, 0.5)])
def test_system_matches_the_dense_definition(seed, th, tw, aspect):
    geom, frame, dx, dy = case(seed, th, tw, nknots=5)
    any_, qm, wy, iv, mask = oracle_system(frame, dx, dy, geom, th, tw, 3e-3, aspect)
    ww, wyr, reg, ivr = reference_system(frame, dx, dy, geom, th, tw, 3e-3, aspect)
    Qr, Lr, mr, anyr = masked_reference(ww, wyr, reg, ivr)
    assert bool(any_) == anyr
    np.testing.assert_array_equal(mask, mr)
    np.testing.assert_array_equal(np.isnan(iv), np.isnan(ivr))
    np.testing.assert_allclose(iv[~np.isnan(iv)], ivr[~np.isnan(ivr)], rtol=2e-6)
    np.testing.assert_allclose(wy, wyr, rtol=1e-13, atol=1e-12)
    Q = stencil_to_dense(qm, th, tw)
    np.testing.assert_allclose(Q, Qr, rtol=1e-13, atol=1e-15)
    assert np.array_equal(Q, Q.T)                       # exactly symmetric


def test_resample_keeps_the_last_value_of_a_repeated_index():
    # two data pixels with the same integer cell: pushgrid's z(ind) = z(ind) + v
    # keeps only the later one; a 4x4 frame squeezed by a large negative dx
    th = tw = 6
    geom = geometry(4, 4, th, tw, 3, subrec=(2, 3, 2, 5))   # four columns within one template unit
    frame = np.arange(1.0, 17.0).reshape(4, 4, order="F")
    dx = np.zeros(3)
    dy = np.zeros(3)
    _, _, _, iv, _ = oracle_system(frame, dx, dy, geom, th, tw, 3e-3)
    _, _, _, ivr = reference_system(frame, dx, dy, geom, th, tw, 3e-3)
    np.testing.assert_array_equal(np.isnan(iv), np.isnan(ivr))
    np.testing.assert_allclose(iv[~np.isnan(iv)], ivr[~np.isnan(ivr)], rtol=1e-6)
    # an accumulating splat would give a different value somewhere
    x, y = positions(geom, dx, dy)
    cells = (np.floor(y) - 1 + th * (np.floor(x) - 1)).astype(int)
    assert len(np.unique(cells)) < len(cells)


def oracle_energy(qm, wy, x, mask, th, tw):
    return O.lib().orc_ibs_energy(O._p(np.ascontiguousarray(qm.ravel())), O._p(wy),
                                  np.ascontiguousarray(x, dtype=np.float32).ctypes.data_as(PF),
                                  np.ascontiguousarray(mask, dtype=np.uint8).ctypes.data_as(PU8), th, tw)


@pytest.mark.parametrize("seed", [4, 5])
def test_energy_is_the_quadratic_form(seed):
    th, tw = 10, 10
    geom, frame, dx, dy = case(seed, th, tw)
    _, qm, wy, iv, mask = oracle_system(frame, dx, dy, geom, th, tw, 3e-3)
    ww, wyr, reg, ivr = reference_system(frame, dx, dy, geom, th, tw, 3e-3)
    Qr, Lr, mr, _ = masked_reference(ww, wyr, reg, ivr)
    rng = np.random.default_rng(seed)
    x = (60 + 10 * rng.standard_normal(th * tw)).astype(np.float32)
    xd = x.astype(float)
    e = oracle_energy(qm, wy, x, mask, th, tw)
    lin = np.float32(Lr[mr].astype(np.float32) @ x[mr])
    er = xd[mr] @ Qr[np.ix_(mr, mr)] @ xd[mr] - 2 * float(lin)
    assert abs(e - er) <= 4 * np.spacing(np.float32(abs(er)))


def interactions(qm, wy, mask, th, tw):
    n = th * tw
    Lk = np.zeros(n, np.float32)
    Rm = np.zeros(n * 8, np.float32)
    O.lib().orc_ibs_interactions(O._p(np.ascontiguousarray(qm.ravel())), O._p(wy),
                                 np.ascontiguousarray(mask, dtype=np.uint8).ctypes.data_as(PU8), th, tw,
                                 Lk.ctypes.data_as(PF), Rm.ctypes.data_as(PF))
    return Lk, Rm


def sweep(Lk, Rm, th, tw, x):
    x = np.ascontiguousarray(x, dtype=np.float32).copy()
    O.lib().orc_ibs_sweep(Lk.ctypes.data_as(PF), Rm.ctypes.data_as(PF), th, tw, x.ctypes.data_as(PF))
    return x


def colour_order(th, tw):
    """A B C D = 1-based (odd,odd) (odd,even) (even,odd) (even,even) row/col."""
    order = []
    for rp, cp in ((0, 0), (0, 1), (1, 0), (1, 1)):
        order += [r + c * th for c in range(cp, tw, 2) for r in range(rp, th, 2)]
    return order


def quirked(Q, th, tw):
    """find_SE's strict bound drops the coupling of pixel (th-2, tw-2) to the
    last pixel from that pixel's update (not the reverse one)."""
    Qq = Q.copy()
    Qq[(th - 2) + (tw - 2) * th, th * tw - 1] = 0
    return Qq


@pytest.mark.parametrize("seed", [6, 7])
def test_one_sweep_is_gauss_seidel_in_colour_order(seed):
    th, tw = 10, 10
    geom, frame, dx, dy = case(seed, th, tw)
    _, qm, wy, iv, mask = oracle_system(frame, dx, dy, geom, th, tw, 3e-3)
    Lk, Rm = interactions(qm, wy, mask, th, tw)
    ww, wyr, reg, ivr = reference_system(frame, dx, dy, geom, th, tw, 3e-3)
    Qr, Lr, mr, _ = masked_reference(ww, wyr, reg, ivr)
    Qq = quirked(Qr, th, tw)
    x0 = np.where(np.isnan(iv), 0, iv).astype(np.float32)
    x1 = sweep(Lk, Rm, th, tw, x0)
    xr = x0.astype(float)
    for i in colour_order(th, tw):
        xr[i] = 0.0 if Qq[i, i] == 0 else (Lr[i] - (Qq[i] @ xr - Qq[i, i] * xr[i])) / Qq[i, i]
    np.testing.assert_allclose(x1, xr, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("reg", [3e-3, 0.5])
def test_sweeps_converge_to_the_quirked_solution(reg):
    th, tw = 10, 10
    geom, frame, dx, dy = case(8, th, tw)
    _, qm, wy, iv, mask = oracle_system(frame, dx, dy, geom, th, tw, reg)
    Lk, Rm = interactions(qm, wy, mask, th, tw)
    ww, wyr, regm, ivr = reference_system(frame, dx, dy, geom, th, tw, reg)
    Qr, Lr, mr, _ = masked_reference(ww, wyr, regm, ivr)
    sol = np.linalg.solve(quirked(Qr, th, tw)[np.ix_(mr, mr)], Lr[mr])
    x = np.where(np.isnan(iv), 0, iv).astype(np.float32)
    for _ in range(20000 if reg < 0.1 else 2000):
        x = sweep(Lk, Rm, th, tw, x)
    np.testing.assert_allclose(x[mr], sol, rtol=2e-4, atol=2e-3 * np.abs(sol).max())
    assert np.all(x[~mr] == 0)


def reference_loop(frames, dxk, dyk, geom, th, tw, reg):
    """The frame loop of steady_scan.m:326-398 over the oracle's pinned steps."""
    n = th * tw
    out = np.full((n, frames.shape[2]), np.nan, np.float32)
    masks = np.zeros((n, frames.shape[2]), bool)
    iters = []
    n_iter_error = 1
    for f in range(frames.shape[2]):
        any_, qm, wy, iv, mask = oracle_system(frames[:, :, f], dxk[:, f], dyk[:, f], geom, th, tw, reg)
        if not any_:
            iters.append(0)
            continue
        masks[:, f] = mask
        Lk, Rm = interactions(qm, wy, mask, th, tw)
        x = np.where(np.isnan(iv), 0, iv).astype(np.float32)
        xo = np.zeros(n, np.float32)
        eo = oracle_energy(qm, wy, x, mask, th, tw)
        it = 0
        while True:
            x = sweep(Lk, Rm, th, tw, x)
            it += 1
            if it % n_iter_error == 0:
                e = oracle_energy(qm, wy, x, mask, th, tw)
                if eo <= e:
                    break
                eo, xo = e, x.copy()
        n_iter_error = int(np.floor(max(1.0, min(10.0, float(np.sqrt(np.float32(it))))) + 0.5))
        xo[~mask] = np.nan
        out[:, f] = xo
        iters.append(it)
    return out.reshape(th, tw, -1, order="F"), masks.reshape(th, tw, -1, order="F"), np.array(iters)


def test_frame_loop_carries_the_check_interval_and_skips_empty_frames():
    th, tw = 10, 10
    rng = np.random.default_rng(9)
    geom = geometry(th, tw, th, tw, 6)
    nf = 5
    frames = 50 + 20 * rng.random((th, tw, nf))
    dxk = rng.uniform(-0.6, 0.6, (6, nf))
    dyk = rng.uniform(-0.6, 0.6, (6, nf))
    dxk[:, 2] = 40.0                                   # frame 2 lies entirely off the template
    out, mask, it = O.invert_bilinear_sampling(frames, dxk, dyk, geom, (th, tw), regfac=2.0)
    ro, rm, rit = reference_loop(frames, dxk, dyk, geom, th, tw, 2.0)
    print("iterations", it)
    assert it[2] == 0 and np.all(np.isnan(out[:, :, 2])) and not mask[:, :, 2].any()
    assert (it[[0, 1, 3, 4]] > 3).all()                # later frames check every 2+ iterations
    np.testing.assert_array_equal(it, rit)
    np.testing.assert_array_equal(mask, rm)
    np.testing.assert_array_equal(out, ro)


def test_integer_positions_start_at_the_minimiser_and_return_zeros():
    # zero displacement on an identity geometry: every template pixel is one
    # data pixel with weight 1, so the resampled frame that initialises x is
    # the frame itself and already the minimiser (tiny regulariser).  The
    # first check then finds no decrease and the reference returns xo, which
    # it initialised to zeros (steady_scan.m:365-385).
    th = tw = 12
    geom = geometry(12, 12, th, tw, 4)
    rng = np.random.default_rng(4)
    frame = 100 + 30 * rng.random((12, 12))
    _, qm, wy, iv, mask = oracle_system(frame, np.zeros(4), np.zeros(4), geom, th, tw, 1e-9)
    sel = mask.reshape(th, tw, order="F")
    assert sel[1:-1, 1:-1].all()
    np.testing.assert_allclose(iv.reshape(th, tw, order="F")[sel], frame[sel], rtol=1e-6)
    Lk, Rm = interactions(qm, wy, mask, th, tw)
    x1 = sweep(Lk, Rm, th, tw, np.where(np.isnan(iv), 0, iv))
    np.testing.assert_allclose(x1[mask], frame.ravel(order="F")[mask], rtol=1e-5)
    out, omask, it = O.invert_bilinear_sampling(frame[:, :, None], np.zeros((4, 1)), np.zeros((4, 1)), geom,
                                                (th, tw), regfac=1e-9)
    assert it[0] == 1
    assert np.all(out[:, :, 0][omask[:, :, 0]] == 0)


def test_check_interval():
    # int32(max(1, min(10, sqrt(single(iter))))) rounds half away from zero
    got = [O.lib().orc_ibs_check_interval(i) for i in (0, 1, 2, 3, 6, 7, 12, 13, 90, 91, 500)]
    assert got == [1, 1, 1, 2, 2, 3, 3, 4, 9, 10, 10]
