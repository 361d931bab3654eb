"""Oracle of prefilter_foralignment and imgapfill_leastsquares.

TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg as
the parity checker; the product path never imports it.

prefilter_foralignment (steady_scan.m:625-643): a circular Gaussian blur
computed as MATLAB does it, with 2-D FFTs of the frame zero-padded by L
pixels and of a wrapped Gaussian (numpy's FFT stands in for MATLAB's FFTW:
the two agree to rounding, which is parity-unpinned; for single-precision
frames MATLAB transforms in single precision and this restatement in double,
also parity-unpinned at that rounding level), then division by each frame's
median.  The result keeps the input's class: integer frames are
rounded half away from zero and saturated, as MATLAB's assignment into an
integer array does.

imgapfill_leastsquares (steady_scan.m:525-560): the missing (NaN) pixels of
each frame are the minimiser of the sum of squared adjacent differences with
the known pixels fixed; the reference solves Q c = -L with MATLAB's sparse
backslash, here scipy's sparse direct solver (SuperLU) on the same matrices.
"""
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

INT_RANGE = {np.dtype(np.uint16): (0, 65535), np.dtype(np.int16): (-32768, 32767)}


def gauss_pad(gsd):
    """L = ceil(sqrt(2 gsd^2 log(10000 / (sqrt(2 pi) gsd)))) (steady_scan.m:628)."""
    return int(math.ceil(math.sqrt((2 * gsd ** 2) * math.log(10000 / (math.sqrt(2 * math.pi) * gsd)))))


def gauss_kernel(h, w, gsd):
    """The wrapped, normalised Gaussian on the padded grid (steady_scan.m:629-632)."""
    L = gauss_pad(gsd)
    s1, s2 = h + L, w + L
    x, y = np.meshgrid(np.arange(1, s2 + 1), np.arange(1, s1 + 1))
    g2d = np.exp(-(np.minimum(x - 1, s2 + 1 - x) ** 2 + np.minimum(y - 1, s1 + 1 - y) ** 2) / (2 * gsd ** 2))
    return g2d / np.sum(g2d), (s1, s2)


def to_class(u, dtype):
    """Assignment of double values into an array of class dtype."""
    dtype = np.dtype(dtype)
    if dtype in INT_RANGE:
        lo, hi = INT_RANGE[dtype]
        v = np.where(np.isnan(u), 0.0, u)
        t = np.trunc(v)
        v = t + np.sign(v) * (np.abs(v - t) >= 0.5)   # round half away from zero, exactly
        return np.clip(v, lo, hi).astype(dtype)
    return u.astype(dtype)


def median_of(frame):
    """MATLAB median over all elements: the middle value, or the mean (a + b) / 2
    of the two middle values (MATLAB's exact averaging formula is
    parity-unpinned at rounding level)."""
    v = np.sort(np.asarray(frame, dtype=np.float64).ravel())
    n = v.size
    if n % 2:
        return v[n // 2]
    return (v[n // 2 - 1] + v[n // 2]) / 2


def prefilter_foralignment(data, gsd, mednorm):
    """data: H x W x n (or H x W); returns an array of the same class and shape."""
    a = np.asarray(data)
    squeeze = a.ndim == 2
    if squeeze:
        a = a[:, :, None]
    h, w, n = a.shape
    out = a.copy()
    if gsd > 0:
        g2d, (s1, s2) = gauss_kernel(h, w, gsd)
        G = np.fft.fft2(g2d)
        for k in range(n):
            u = np.real(np.fft.ifft2(np.fft.fft2(a[:, :, k].astype(np.float64), s=(s1, s2)) * G))
            out[:, :, k] = to_class(u[:h, :w], a.dtype)
    if mednorm:
        if out.dtype in INT_RANGE:
            raise NotImplementedError("median normalisation of integer frames is not restated")
        med = np.array([median_of(out[:, :, k]) for k in range(n)])
        out = (out / med.reshape(1, 1, n).astype(out.dtype)).astype(out.dtype)
    return out[:, :, 0] if squeeze else out


def adjdiff_quadmat(w, h, vfac=1.0):
    """imageadjdiffquadmat (steady_scan.m:406-436): G'G as a sparse matrix."""
    x, y = np.meshgrid(np.arange(1, w + 1), np.arange(1, h + 1))
    x = x.ravel(order="F")
    y = y.ravel(order="F")
    xp = np.concatenate([np.stack([x, x + 1], 1), np.stack([x, x], 1)])
    yp = np.concatenate([np.stack([y, y], 1), np.stack([y, y + 1], 1)])
    keep = np.all((xp <= w) & (yp <= h), axis=1)
    xp, yp = xp[keep], yp[keep]
    m = xp.shape[0]
    ind = yp + h * (xp - 1) - 1
    vert = xp[:, 0] == xp[:, 1]
    coef = np.where(vert, vfac, 1.0)
    G = sp.csr_matrix((np.concatenate([coef, -coef]), (np.concatenate([np.arange(m)] * 2), np.concatenate([ind[:, 0], ind[:, 1]]))),
                      shape=(m, w * h))
    return (G.T @ G).tocsr()


def imgapfill_leastsquares(a):
    """a: h x w x n float array with NaN at the missing pixels."""
    a = np.asarray(a)
    squeeze = a.ndim == 2
    if squeeze:
        a = a[:, :, None]
    h, w, n = a.shape
    flat = a.reshape(h * w, n, order="F")
    mask = ~np.isnan(flat)
    if mask.all():
        return a[:, :, 0] if squeeze else a
    GtG = adjdiff_quadmat(w, h)
    b = flat.astype(np.float64)
    for k in range(n):
        if not mask[:, k].any():
            continue
        miss = np.flatnonzero(~mask[:, k])
        if miss.size == 0:
            continue
        a0 = b[:, k].copy()
        a0[miss] = 0
        Q = GtG[miss][:, miss].tocsc()
        Lt = GtG[miss] @ a0
        b[miss, k] = -spla.spsolve(Q, Lt)
    out = b.reshape(h, w, n, order="F")
    if np.issubdtype(a.dtype, np.integer):
        out = np.round(out)
    out = out.astype(a.dtype)
    return out[:, :, 0] if squeeze else out
