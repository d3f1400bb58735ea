"""Per-frequency row kernels and the row engine, restated (oracle / test infrastructure only).

A row is one offset frequency w_k (z = exp(i w_k)) of one packed quadrant pair: a length-N complex vector over
columns t (image side) or slopes s (data side).  Level m (w = 2^(m-1)) of the butterfly, per strip l:
    forward  (cpp/oblique.cpp:22-33, fastspec.py:21-35):  u = z^t v[l2w+w+t];  o[l2w+2t] = v[l2w+t] + u;
                                                          o[l2w+2t+1] = v[l2w+t] + z u
    adjoint  (cpp/oblique.cpp:36-50, fastspec.py:38-52):  o[l2w+t] = g0 + g1;  o[l2w+w+t] = conj(z^t) (g0 + conj(z) g1)
with g0, g1 = g[l2w+2t], g[l2w+2t+1].  ``B lo..hi`` applies level lo first, ``BT lo..hi`` level hi first.
Twiddles z^t come from the same left-to-right product recurrence as cpp/oblique.cpp:16-19.
"""

import numpy as np


def powers(z, count):
    """pw[:, t] = z^t for t < count by the recurrence pw[t] = pw[t-1] * z (cpp/oblique.cpp:16-19)."""
    z = np.asarray(z, dtype=complex)
    pw = np.empty((z.size, max(count, 1)), dtype=complex)
    pw[:, 0] = 1.0
    for t in range(1, count):
        pw[:, t] = pw[:, t - 1] * z
    return pw


def bfly(n, z, X, lo, hi):
    """Forward levels lo..hi on every row of X (K, N) (cpp/oblique.cpp:22-33; fasthybrid.py:23-37)."""
    X = np.asarray(X, dtype=complex)
    K, N = X.shape
    z = np.asarray(z, dtype=complex).reshape(K)
    pw = powers(z, N // 2 + 1)
    g = X
    for m in range(lo, hi + 1):
        w = 1 << (m - 1)
        v = g.reshape(K, N // (2 * w), 2, w)
        u = pw[:, None, :w] * v[:, :, 1, :]
        out = np.empty((K, N // (2 * w), w, 2), dtype=complex)
        out[..., 0] = v[:, :, 0, :] + u
        out[..., 1] = v[:, :, 0, :] + u * z[:, None, None]
        g = out.reshape(K, N)
    return g


def bflyT(n, z, G, lo, hi):
    """Adjoint levels hi..lo on every row of G (K, N) (cpp/oblique.cpp:36-50; fasthybrid.py:40-54)."""
    G = np.asarray(G, dtype=complex)
    K, N = G.shape
    z = np.asarray(z, dtype=complex).reshape(K)
    pwc = np.conj(powers(z, N // 2 + 1))
    zc = np.conj(z)
    f = G
    for m in range(hi, lo - 1, -1):
        w = 1 << (m - 1)
        v = f.reshape(K, N // (2 * w), w, 2)
        g0, g1 = v[..., 0], v[..., 1]
        out = np.empty((K, N // (2 * w), 2, w), dtype=complex)
        out[:, :, 0, :] = g0 + g1
        out[:, :, 1, :] = pwc[:, None, :w] * (g0 + zc[:, None, None] * g1)
        f = out.reshape(K, N)
    return f


def cheb_coeffs(lam_max, tau):
    """(c, a) of the Chebyshev semi-iteration for M + tau, spec(M) in [0, lam_max] (fastspec.py:56-59)."""
    lo, hi = tau, np.asarray(lam_max, dtype=float) + tau
    return (hi - lo) / (hi + lo), 2.0 / (hi + lo)


def cheb_filter(n, z, Y, deg, a, c, tau, lo, hi):
    """Z[k] = p_k(M_k) Y[k] with deg[k] - 1 Chebyshev steps (cpp/oblique.cpp:74-104; fastops.py:118-138).

    Returns (Z, steps) where steps counts butterfly level-rows exactly as cheb_rows does.
    """
    Y = np.asarray(Y, dtype=complex)
    K, N = Y.shape
    deg = np.asarray(deg).astype(np.int64).reshape(K)
    a = np.asarray(a, dtype=float).reshape(K)
    c = np.asarray(c, dtype=float).reshape(K)
    z = np.asarray(z, dtype=complex).reshape(K)
    Z = a[:, None] * Y
    Zp = np.zeros_like(Y)
    om = np.ones(K)
    steps = 0
    dmax = int(deg.max()) if K else 0
    for j in range(1, dmax):
        act = np.nonzero(deg > j)[0]
        if act.size == 0:
            break
        ca = c[act]
        omj = 1.0 / (1.0 - ca * ca / 2.0) if j == 1 else 1.0 / (1.0 - ca * ca * om[act] / 4.0)
        Za = Z[act]
        Mz = bflyT(n, z[act], bfly(n, z[act], Za, lo, hi), lo, hi)
        r = Y[act] - Mz - tau * Za
        Zn = Za + omj[:, None] * (a[act][:, None] * r) + (omj - 1.0)[:, None] * (Za - Zp[act])
        Zp[act] = Za
        Z[act] = Zn
        om[act] = omj
        steps += act.size * 2 * (hi - lo + 1)
    return Z, steps


def level_matrix(n, z, lo, hi):
    """Dense matrix A of the forward levels lo..hi for one frequency: A[s, t] (columns are images of e_t)."""
    N = 1 << n
    return bfly(n, np.full(N, z), np.eye(N, dtype=complex), lo, hi).T


def dense_filter(n, z, tau, lo, hi):
    """Exact Tikhonov filter F = (M + tau)^{-1}, M = A^H A, by eigh with lam clipped at 0 (fastops.py:28-39)."""
    A = level_matrix(n, z, lo, hi)
    lam, V = np.linalg.eigh(A.conj().T @ A)
    lam = np.maximum(lam, 0)
    return (V / (lam + tau)) @ V.conj().T


def cert_rowsum(n, omega, lo):
    """Certified upper bound of lambda_max(M) for levels lo..n by the row-sum recursion
    r_m(a) <= |1+z^a| r(2a) + max(|1+z^(a+1)| r(2a+1), |1+z^(a-1)| r(2a-1))  (fasthybrid.py:134-149, cpp/oblique.cpp:108-130).
    |1+z^a| is evaluated as 2|cos(w a / 2)| as the C++ kernel does."""
    A = 1 << n
    r = np.ones(2 * A + 1)
    for m in range(lo, n + 1):
        amax = (1 << (n - m)) - 1
        av = np.arange(-amax, amax + 1)
        f = lambda e: 2.0 * np.abs(np.cos(0.5 * omega * e.astype(float)))
        t0 = f(av) * r[2 * av + A]
        tp = f(av + 1) * r[2 * av + 1 + A]
        tm = f(av - 1) * r[2 * av - 1 + A]
        rn = np.zeros(2 * A + 1)
        rn[av + A] = t0 + np.where(tp > tm, tp, tm)
        r = rn
    return float(r[A])


class Rows:
    """Row engine with the RowEngine semantics (fastops.py:20-167): rows r = 0..copies*L-1 use frequency r mod L;
    ``run(Y, ops)`` applies ('B'|'BT'|'P', lo, hi) left to right; rows whose degree exceeds ``dense_cap`` use the
    exact filter in P.  ``count[0]`` accumulates level-row steps like the reference."""

    def __init__(self, n, z, deg, lam_max, tau, threads=1, copies=2, dense_cap=None, p_levels=(1, None)):
        self.n = n
        self.N = N = 1 << n
        self.L = len(z)
        self.R = copies * self.L
        self.tau = float(tau)
        self.z = np.asarray(z, dtype=complex)
        self.deg = np.asarray(deg).astype(np.int64)
        self.lam_max = np.asarray(lam_max, dtype=float)
        self.c, self.a = cheb_coeffs(self.lam_max, self.tau)
        lo = p_levels[0]
        hi = p_levels[1] or n
        self.dense = {}
        if dense_cap is not None:
            for k in np.nonzero(self.deg > dense_cap)[0]:
                self.dense[int(k)] = dense_filter(n, self.z[k], self.tau, lo, hi)
        self.n_dense = len(self.dense)
        self.count = [0]

    def run(self, Y, ops):
        Y = np.asarray(Y, dtype=complex)
        R = Y.shape[0]
        k = np.arange(R) % self.L
        zr = self.z[k]
        X = Y.copy()
        steps = 0
        for op, lo, hi in ops:
            if op == "B":
                X = bfly(self.n, zr, X, lo, hi)
                steps += R * (hi - lo + 1)
            elif op == "BT":
                X = bflyT(self.n, zr, X, lo, hi)
                steps += R * (hi - lo + 1)
            elif op == "P":
                isd = np.array([int(kk) in self.dense for kk in k], dtype=bool)
                out = np.empty_like(X)
                rest = np.nonzero(~isd)[0]
                if rest.size:
                    out[rest], s = cheb_filter(
                        self.n, zr[rest], X[rest], self.deg[k[rest]], self.a[k[rest]], self.c[k[rest]], self.tau, lo, hi
                    )
                    steps += s
                for r in np.nonzero(isd)[0]:
                    out[r] = self.dense[int(k[r])] @ X[r]
                steps += int(isd.sum()) * self.N // max(1, self.n)
                X = out
            else:
                raise ValueError(op)
        self.count[0] += steps
        return X
