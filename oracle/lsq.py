"""Least-squares cross-check by CGLS, restated (oracle / test infrastructure only).  Follows pinv.py:23-48."""

import numpy as np

from . import radon


def cgls(n, b, tol=1e-12, maxit=10000, x0=None, fwd=None, bwd=None):
    """min ||R x - b||_2 by CGLS; stops when ||R^T r|| <= tol ||R^T b||_0 (pinv.py:23-48).

    ``fwd``/``bwd`` replace the ADRT pair (e.g. with the GPU kernels); they default to the oracle."""
    N = 1 << n
    fwd = fwd or (lambda v: radon.adrt(n, v))
    bwd = bwd or (lambda v: radon.adrt_T(n, v))
    b = np.asarray(b, dtype=float).reshape(4, 2 * N - 1, N)
    x = np.zeros((N, N)) if x0 is None else np.array(x0, dtype=float)
    r = b - fwd(x)
    s = bwd(r)
    p = s.copy()
    gamma = float(np.vdot(s, s).real)
    g0 = np.sqrt(gamma)
    hist = [g0]
    for _ in range(maxit):
        q = fwd(p)
        alpha = gamma / float(np.vdot(q, q).real)
        x += alpha * p
        r -= alpha * q
        s = bwd(r)
        gnew = float(np.vdot(s, s).real)
        hist.append(np.sqrt(gnew))
        if np.sqrt(gnew) <= tol * g0:
            break
        p = s + (gnew / gamma) * p
        gamma = gnew
    return x, np.array(hist)


def pinv_dense(n):
    """Dense Moore-Penrose pseudoinverse of the ADRT (pinv.py:85-88), small n only."""
    N = 1 << n
    A = radon.dense(lambda x: radon.adrt(n, x), (N, N))
    return np.linalg.pinv(A)
