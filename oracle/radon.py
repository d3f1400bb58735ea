"""ADRT core, restated (oracle / test infrastructure only).

Layout (adrtcore.py:9-20): single-quadrant data after level m is a dense array of shape (N + 2^m - 1, N); row
index h + 2^m - 1 holds offset h in [-(2^m - 1), N), column s is the slope.  Entries with h < -(s mod 2^m) are
structural zeros.  The full transform stacks the four quadrants, each of shape (2N - 1, N).

Sums follow the reference order (a zero-initialised output, then the half-0 term, then the half-1 term;
adrtcore.py:74-82 and 96-100) so float results are bit-identical, and integer inputs stay exact.
"""

import numpy as np


def quadrant(q, f):
    """T_q f for q = 0..3 (adrtcore.py:29-39): (0) f(j, N-1-i), (1) f(N-1-i, j), (2) f, (3) f(N-1-j, N-1-i)."""
    f = np.asarray(f)
    if q == 0:
        out = f.T[::-1, :]
    elif q == 1:
        out = f[::-1, :]
    elif q == 2:
        out = f
    elif q == 3:
        out = f.T[::-1, ::-1]
    else:
        raise ValueError(q)
    return np.ascontiguousarray(out)


def quadrant_T(q, g):
    """T_q^T = T_q^{-1} (adrtcore.py:42-52)."""
    g = np.asarray(g)
    if q == 0:
        out = g[::-1, :].T
    elif q == 1:
        out = g[::-1, :]
    elif q == 2:
        out = g
    elif q == 3:
        out = g[::-1, ::-1].T
    else:
        raise ValueError(q)
    return np.ascontiguousarray(out)


def _level_maps(n, m):
    """Index maps of level m: for every output column j = l*2w + 2t + b its two source columns and t, b."""
    N = 1 << n
    w = 1 << (m - 1)
    col = np.arange(N)
    base = (col // (2 * w)) * (2 * w)
    t = (col % (2 * w)) >> 1
    b = col & 1
    return N, w, base + t, base + w + t, t, b


def level(n, m, f):
    """S_m: (N + w - 1, N) -> (N + 2w - 1, N), w = 2^(m-1) (adrtcore.py:63-83).

    g(h, l2w + 2t + b) = f(h, l2w + t) [h >= -t]  +  f(h + t + b, l2w + w + t) [-t <= h + t + b < N].
    """
    N, w, c0, c1, t, b = _level_maps(n, m)
    f = np.asarray(f)
    assert f.shape == (N + w - 1, N), f.shape
    Hout = N + 2 * w - 1
    h = np.arange(Hout)[:, None] - (2 * w - 1)  # offset of every output row
    g = np.zeros((Hout, N), dtype=f.dtype)
    # half-0 term: source row h (same offset), valid for h >= -t
    ok0 = h >= -t[None, :]
    r0 = np.clip(h + (w - 1), 0, N + w - 2)
    g = g + np.where(ok0, f[r0, c0[None, :]], 0)
    # half-1 term: source offset h + t + b, valid for -t <= h + t + b < N
    hs = h + t[None, :] + b[None, :]
    ok1 = (hs >= -t[None, :]) & (hs < N)
    r1 = np.clip(hs + (w - 1), 0, N + w - 2)
    g = g + np.where(ok1, f[r1, c1[None, :]], 0)
    return g.astype(f.dtype, copy=False)


def level_T(n, m, g):
    """S_m^T: (N + 2w - 1, N) -> (N + w - 1, N) (adrtcore.py:86-101).

    f(h, l2w + t)     = g(h, l2w + 2t) + g(h, l2w + 2t + 1)                 for h in [-t, N)
    f(h, l2w + w + t) = g(h - t, l2w + 2t) + g(h - t - 1, l2w + 2t + 1)     for h in [-t, N)
    """
    N = 1 << n
    w = 1 << (m - 1)
    g = np.asarray(g)
    assert g.shape == (N + 2 * w - 1, N), g.shape
    Hin = N + w - 1
    f = np.zeros((Hin, N), dtype=g.dtype)
    h = np.arange(Hin)[:, None] - (w - 1)
    col = np.arange(N)
    base = (col // (2 * w)) * (2 * w)
    half = (col % (2 * w)) // w
    t = col % w
    e0 = base + 2 * t  # slope column 2t of the strip
    e1 = e0 + 1
    valid = h >= -t[None, :]
    off = np.where(half == 1, t, 0)[None, :]  # half 1 reads offsets shifted by t (b = 0) and t + 1 (b = 1)
    ra = h - off + (2 * w - 1)
    rb = h - off - half[None, :] + (2 * w - 1)
    ra = np.clip(ra, 0, N + 2 * w - 2)
    rb = np.clip(rb, 0, N + 2 * w - 2)
    f = f + np.where(valid, g[ra, e0[None, :]], 0)
    f = f + np.where(valid, g[rb, e1[None, :]], 0)
    return f.astype(g.dtype, copy=False)


def sq_adrt(n, x):
    """R^n_n of one quadrant (adrtcore.py:104-109)."""
    g = np.asarray(x)
    for m in range(1, n + 1):
        g = level(n, m, g)
    return g


def sq_adrt_T(n, g):
    """(R^n_n)^T (adrtcore.py:112-116)."""
    f = np.asarray(g)
    for m in range(n, 0, -1):
        f = level_T(n, m, f)
    return f


def adrt(n, x, dtype=float):
    """Full ADRT (4, 2N-1, N) (adrtcore.py:122-128).  dtype=float mirrors the reference; an integer dtype is exact."""
    N = 1 << n
    x = np.asarray(x, dtype=dtype).reshape(N, N)
    return np.stack([sq_adrt(n, quadrant(q, x)) for q in range(4)])


def adrt_T(n, b, dtype=float):
    """Transpose R^T b = sum_q T_q^T (sq_adrt_T b_q), summed in quadrant order (adrtcore.py:131-137)."""
    N = 1 << n
    b = np.asarray(b, dtype=dtype).reshape(4, 2 * N - 1, N)
    x = np.zeros((N, N), dtype=dtype)
    for q in range(4):
        x = x + quadrant_T(q, sq_adrt_T(n, b[q]))
    return x


def mask_full(n):
    """Structurally nonzero entries h >= -s of the (4, 2N-1, N) output (adrtcore.py:140-146)."""
    N = 1 << n
    h = np.arange(2 * N - 1)[:, None] - (N - 1)
    s = np.arange(N)[None, :]
    return np.broadcast_to(h >= -s, (4, 2 * N - 1, N)).copy()


def delta_s(n, s, t):
    """Digital-line offset delta_s(t) = sum_k bit_{n-1-k}(s) floor((t + 2^k) / 2^(k+1)) (adrtcore.py:179-185)."""
    d = 0
    for k in range(n):
        if (s >> (n - 1 - k)) & 1:
            d += (t + (1 << k)) >> (k + 1)
    return d


def delta_table(n):
    N = 1 << n
    return np.array([[delta_s(n, s, t) for t in range(N)] for s in range(N)], dtype=np.int64)


def dense(op, in_shape):
    """Dense matrix of a linear map by columns of the identity (adrtcore.py:152-161)."""
    nin = int(np.prod(in_shape))
    cols = []
    for k in range(nin):
        e = np.zeros(nin)
        e[k] = 1.0
        cols.append(np.asarray(op(e.reshape(in_shape)), dtype=float).ravel())
    return np.stack(cols, axis=1)
