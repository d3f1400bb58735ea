"""Least-squares comparators, restated (oracle / test infrastructure only): CGLS (pinv.py:23-48) and the normal-equation
Chebyshev lstsq_cheb with its Gram (patches/adrt-gram-freq.patch:411-545)."""

import numpy as np

from . import radon


def cgls(n, b, tol=1e-12, maxit=10000, x0=None, fwd=None, bwd=None, callback=None):
    """min ||R x - b||_2 by CGLS; stops when ||R^T r|| <= tol ||R^T b||_0 (pinv.py:23-48).

    ``fwd``/``bwd`` replace the ADRT pair (e.g. with the GPU kernels); they default to the oracle.
    ``callback(it, x, ||R^T r||)`` after every iteration, before the stop test (pinv.py:43-44)."""
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
    for it in range(maxit):
        q = fwd(p)
        alpha = gamma / float(np.vdot(q, q).real)
        x += alpha * p
        r -= alpha * q
        s = bwd(r)
        gnew = float(np.vdot(s, s).real)
        hist.append(np.sqrt(gnew))
        if callback is not None:
            callback(it, x, np.sqrt(gnew))
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


def gram(n, a):
    """R^T R a through the spatial ADRT pair: the reference against which the patch tests its frequency-row Gram
    (patches/adrt-gram-freq.patch:592-594: truncate(bdrt(adrt(a))).sum(axis=-3)), in this repository's layout."""
    N = 1 << n
    a = np.asarray(a, dtype=float).reshape(N, N)
    return radon.adrt_T(n, radon.adrt(n, a))


def cheb_degree(lo, hi, tol):
    """Degree of lstsq_cheb for the interval [lo, hi] and tolerance tol (patch:525-532)."""
    kappa = hi / lo
    rho = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
    return max(1, int(np.ceil(np.log(tol / 2) / np.log(rho))))


def power_lam_max(n, op=None, iters=50):
    """Upper bound of lambda_max(R^T R) as lstsq_cheb estimates it (patch:512-520): 50 normalised power steps from
    default_rng(0).standard_normal((N, N)), then theta = <v, Gv> enlarged by ||Gv - theta v||."""
    N = 1 << n
    op = op or (lambda v: gram(n, v))
    v = np.random.default_rng(0).standard_normal((N, N))
    v /= np.linalg.norm(v)
    for _ in range(iters):
        w = op(v)
        v = w / np.linalg.norm(w)
    w = op(v)
    theta = float(np.vdot(v, w))
    return theta + float(np.linalg.norm(w - theta * v))


def lstsq_cheb(n, b, lam_min, lam_max=None, tol=1e-10, degree=None, op=None):
    """x_D = c_D(R^T R) R^T b, the fixed-degree Chebyshev semi-iteration on the normal equations over
    [lam_min, lam_max] (patches/adrt-gram-freq.patch:474-545), with y = R^T b = adrt_T(b) in this layout."""
    N = 1 << n
    op = op or (lambda v: gram(n, v))
    y = radon.adrt_T(n, np.asarray(b, dtype=float).reshape(4, 2 * N - 1, N))
    if lam_max is None:
        lam_max = power_lam_max(n, op)
    lo, hi = float(lam_min), float(lam_max)
    if not 0 < lo < hi:
        raise ValueError("need 0 < lam_min < lam_max")
    deg = cheb_degree(lo, hi, tol) if degree is None else int(degree)
    c, a = (hi - lo) / (hi + lo), 2.0 / (hi + lo)
    x_prev, x = np.zeros_like(y), a * y
    om = 1.0
    for k in range(1, deg):
        om = 1.0 / (1.0 - c * c / 2.0) if k == 1 else 1.0 / (1.0 - c * c * om / 4.0)
        r = y - op(x)
        x_prev, x = x, x + om * (a * r) + (om - 1.0) * (x - x_prev)
    return x
