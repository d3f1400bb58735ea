"""Host-side setup arithmetic of the inverses: Chebyshev constants, symmetrised bounds and degrees, the explicit bound,
the level-1 Gram stencil used by the outer-interval eigensolve, and the outer interval itself.

Follows fastspec.py:56-59 (cheb_coeffs), fastops.py:170-192 (_sym, _degrees), fastcpp.py:128-134 (explicit bound),
fasthybrid.py:81-98 (G_1) and fastops.py:222-245 (_outer_interval).  Everything here is setup: it runs once per plan.
"""

import numpy as np


def cheb_coeffs(lam_max, tau):
    """(c, a) for spec(M + tau) in [tau, lam_max + tau] (fastspec.py:56-59)."""
    lo = tau
    hi = np.asarray(lam_max, dtype=float) + tau
    return (hi - lo) / (hi + lo), 2.0 / (hi + lo)


def sym(lam):
    """Lambda_k <- max(Lambda_k, Lambda_{-k}) so the packed pairs stay exact (fastops.py:170-172)."""
    lam = np.asarray(lam, dtype=float)
    return np.maximum(lam, lam[(-np.arange(len(lam))) % len(lam)])


def degrees(lam_max, tau, eps_in):
    """Symmetrised, clipped bounds and inner degrees max(2, ceil(log(eps_in/2) / log rho_k)) (fastops.py:188-192)."""
    lam = np.maximum(sym(lam_max), 1.0)
    kap = 1.0 + lam / tau
    rate = (np.sqrt(kap) - 1) / (np.sqrt(kap) + 1)
    deg = np.maximum(2, np.ceil(np.log(eps_in / 2.0) / np.log(np.maximum(rate, 1e-12)))).astype(int)
    return lam, deg


def explicit_lam_max(n, oms):
    """min(N^2, 2N(1+n)/|w|), N^2 at w = 0 (fastcpp.py:128-134)."""
    N = 1 << n
    w = np.abs(np.asarray(oms, dtype=float))
    lam = np.full(w.shape, float(N) ** 2)
    nz = w > 0
    lam[nz] = np.minimum(float(N) ** 2, 2.0 * N * (1 + n) / w[nz])
    return lam


def g1_apply(x):
    """G_1 x = 8x + J(Kx) + (J(K x^T))^T, K = tridiag(1, 2, 1), J = pair swap (fasthybrid.py:81-98)."""
    x = np.asarray(x, dtype=float)

    def k_rows(v):
        y = 2.0 * v
        y[1:] += v[:-1]
        y[:-1] += v[1:]
        return y

    def j_cols(v):
        y = np.empty_like(v)
        y[:, 0::2] = v[:, 1::2]
        y[:, 1::2] = v[:, 0::2]
        return y

    return 8.0 * x + j_cols(k_rows(x)) + j_cols(k_rows(x.T)).T


def pi_apply(x):
    """2 x 2 cell mean projector (fasthybrid.py:101-105)."""
    N = x.shape[0]
    c = x.reshape(N // 2, 2, N // 2, 2).mean(axis=(1, 3))
    return np.repeat(np.repeat(c, 2, axis=0), 2, axis=1)


def outer_interval(N, G, Pinv=None, P=None, eps=1e-10, tol=1e-3, maxiter=300, margins=(0.99, 1.01)):
    """Enclosure [lo, hi] of spec(P G) by Lanczos (eigsh SA / LA) plus the Weinstein residual bound, widened by
    ``margins``, and the outer degree D = ceil(log(eps/2) / log rho) + 2 (fastops.py:222-245)."""
    import scipy.sparse.linalg as spla

    A = spla.LinearOperator((N * N, N * N), matvec=lambda v: G(v.reshape(N, N)).ravel(), dtype=float)
    if Pinv is None:
        lo, ulo = spla.eigsh(A, k=1, which="SA", tol=tol, maxiter=maxiter)
        hi, uhi = spla.eigsh(A, k=1, which="LA", tol=tol, maxiter=maxiter)

        def res(th, u):
            return np.linalg.norm(A.matvec(u[:, 0]) - th * u[:, 0])

    else:
        Mo = spla.LinearOperator((N * N, N * N), matvec=lambda v: Pinv(v.reshape(N, N)).ravel(), dtype=float)
        Mi = spla.LinearOperator((N * N, N * N), matvec=lambda v: P(v.reshape(N, N)).ravel(), dtype=float)
        lo, ulo = spla.eigsh(A, k=1, M=Mo, Minv=Mi, which="SA", tol=tol, maxiter=maxiter)
        hi, uhi = spla.eigsh(A, k=1, M=Mo, Minv=Mi, which="LA", tol=tol, maxiter=maxiter)

        def res(th, u):
            u = u[:, 0]
            r = A.matvec(u) - th * Mo.matvec(u)
            return np.sqrt(abs(r @ Mi.matvec(r))) / np.sqrt(abs(u @ Mo.matvec(u)))

    lo, hi = float(lo[0]), float(hi[0])
    lo_, hi_ = (lo - res(lo, ulo)) * margins[0], (hi + res(hi, uhi)) * margins[1]
    kap = hi_ / lo_
    D = int(np.ceil(np.log(eps / 2.0) / np.log((np.sqrt(kap) - 1) / (np.sqrt(kap) + 1)))) + 2
    return lo_, hi_, D
