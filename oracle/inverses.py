"""Host orchestration of the oblique inverses L_p and the hybrid, restated (oracle / test infrastructure only).

Follows fastops.py:170-338 (packing, degrees, outer Chebyshev, op lists), fastspec.py:142-174 (torus maps
Tq / TqT / iota), fasthybrid.py:81-130 (G_1 stencil, lorr, cheb_solve) and fastcpp.py:102-134 (bounds).
The row engine is pluggable exactly like fastops._Base.engine_cls.
"""

import numpy as np

from . import rows as rw


# ------------------------------------------------------------------ torus maps (fastspec.py:142-174)
def _refl(N, L):
    return (N - 1 - np.arange(L)) % L


def Tq(N, q, X):
    """Torus image X (L, L) -> strip (L, N) of quadrant q."""
    L = 2 * N
    r = _refl(N, L)
    if q == 2:
        return X[:, :N]
    if q == 1:
        return X[r, :N]
    if q == 0:
        return X.T[r, :N]
    if q == 3:
        return X[np.ix_(r, r)].T[:, :N]
    raise ValueError(q)


def TqT(N, q, S):
    """Adjoint of Tq: strip (L, N) -> torus image (L, L)."""
    L = 2 * N
    r = _refl(N, L)
    X = np.zeros((L, L), dtype=S.dtype)
    X[:, :N] = S
    if q == 2:
        return X
    if q == 1:
        return X[r]
    if q == 0:
        return X[r].T
    if q == 3:
        return X.T[:, r][r]
    raise ValueError(q)


def iota(N, bq):
    """Offset h in [-(N-1), N) of a (2N-1, N) quadrant goes to torus row h mod L (fastspec.py:170-174)."""
    L = 2 * N
    S = np.zeros((L, N))
    S[np.arange(-(N - 1), N) % L] = bq
    return S


# ------------------------------------------------------------------ level-1 stencils (fasthybrid.py:81-130)
def K_rows(x):
    y = 2.0 * x
    y[1:] += x[:-1]
    y[:-1] += x[1:]
    return y


def J_cols(x):
    y = np.empty_like(x)
    y[:, 0::2] = x[:, 1::2]
    y[:, 1::2] = x[:, 0::2]
    return y


def G1_apply(x):
    """G_1 x = 8x + J(Kx) + (J(K x^T))^T, spec in [4, 16] (fasthybrid.py:95-98)."""
    return 8.0 * x + J_cols(K_rows(x)) + J_cols(K_rows(x.T)).T


def Pi_apply(x):
    N = x.shape[0]
    c = x.reshape(N // 2, 2, N // 2, 2).mean(axis=(1, 3))
    return np.repeat(np.repeat(c, 2, axis=0), 2, axis=1)


def lorr_inv(y):
    """(4I + 8 Pi)^{-1} y = (y - Pi y)/4 + Pi y / 12 (fasthybrid.py:108-111)."""
    p = Pi_apply(y)
    return (y - p) / 4.0 + p / 12.0


def cheb_solve(op, y, lo, hi, deg):
    """Fixed-degree Chebyshev semi-iteration for op^{-1} y, spec(op) in [lo, hi] (fasthybrid.py:114-126)."""
    c = (hi - lo) / (hi + lo)
    a = 2.0 / (hi + lo)
    xp = np.zeros_like(y)
    x = a * y
    om = 1.0
    for k in range(1, deg):
        om = 1.0 / (1.0 - c * c / 2.0) if k == 1 else 1.0 / (1.0 - c * c * om / 4.0)
        r = y - op(x)
        xn = x + om * (a * r) + (om - 1.0) * (x - xp)
        xp, x = x, xn
    return x


def g1_inv(y, deg=14):
    return cheb_solve(G1_apply, y, 4.0, 16.0, deg)


# ------------------------------------------------------------------ spectral bounds (fastcpp.py:102-134, fastops.py:170-192)
def sym(lam):
    """max(lam_k, lam_{-k}) (fastops.py:170-172)."""
    L = len(lam)
    return np.maximum(lam, lam[(-np.arange(L)) % L])


def degrees(lam_max, tau, eps_in):
    """lam = max(sym(lam), 1); d_k = max(2, ceil(log(eps_in/2) / log rho_k)) (fastops.py:188-192)."""
    lam = np.maximum(sym(np.asarray(lam_max, dtype=float)), 1.0)
    kap = 1.0 + lam / tau
    rate = (np.sqrt(kap) - 1) / (np.sqrt(kap) + 1)
    deg = np.maximum(2, np.ceil(np.log(eps_in / 2.0) / np.log(np.maximum(rate, 1e-12)))).astype(int)
    return lam, deg


def power_bound(n, z, lo, iters=40, safety=1.5, streams="cpp", apply=None):
    """safety x Rayleigh quotient after ``iters`` normalised power steps of M_k (levels lo..n).

    streams='cpp': real part from default_rng(0), imaginary part from default_rng(1) (fastcpp.py:102-118);
    streams='numpy': both parts drawn in turn from one default_rng(0) (fastops.py:266-272, 309-316).
    ``apply(V)`` overrides the Gram application (the reference C++ kernels for a bit-level match)."""
    N = 1 << n
    K = len(z)
    if streams == "cpp":
        V = np.random.default_rng(0).standard_normal((K, N)) + 1j * np.random.default_rng(1).standard_normal((K, N))
    else:
        rng = np.random.default_rng(0)
        V = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    M = apply or (lambda V: rw.bflyT(n, z, rw.bfly(n, z, V, lo, n), lo, n))
    for _ in range(iters):
        V = M(V)
        V /= np.linalg.norm(V, axis=1, keepdims=True)
    return safety * np.real(np.sum(np.conj(V) * M(V), axis=1))


def explicit_bound(n, oms):
    """min(N^2, 2N(1+n)/|w|), N^2 at w = 0 (fastcpp.py:128-134)."""
    N = 1 << n
    w = np.abs(oms)
    lam = np.full(len(oms), float(N) ** 2)
    nz = w > 0
    lam[nz] = np.minimum(float(N) ** 2, 2.0 * N * (1 + n) / w[nz])
    return lam


def cert_bound(n, oms, lo):
    return np.array([rw.cert_rowsum(n, float(om), lo) for om in oms])


# ------------------------------------------------------------------ the inverses (fastops.py:175-338)
class _Packed:
    PAIRS = ((0, 1), (2, 3))
    engine_cls = rw.Rows

    def _common(self, n, tau_fac, eps_in, threads):
        self.n = n
        N = self.N = 1 << n
        L = self.L = 2 * N
        self.tau = tau_fac * N
        self.eps_in = eps_in
        self.threads = threads
        self.oms = 2 * np.pi * np.fft.fftfreq(L)
        self.z = np.exp(1j * self.oms)

    def _engine(self, lam_max, dense_cap, p_levels):
        self.lam_max, self.deg = degrees(lam_max, self.tau, self.eps_in)
        self.eng = self.engine_cls(
            self.n, self.z, self.deg, self.lam_max, self.tau, self.threads, copies=2, dense_cap=dense_cap, p_levels=p_levels
        )

    def pack(self, strip):
        """Y (2L, N): pair i rows = fft over offsets of strip(qa) + i strip(qb) (fastops.py:212-214)."""
        L, N = self.L, self.N
        Y = np.empty((2 * L, N), dtype=complex)
        for i, (qa, qb) in enumerate(self.PAIRS):
            Y[i * L : (i + 1) * L] = np.fft.fft(strip(qa) + 1j * strip(qb), axis=0)
        return Y

    def unpack(self, Z):
        """ifft, Re -> quadrant qa, Im -> quadrant qb, TqT, window (fastops.py:216-220)."""
        L, N = self.L, self.N
        out = np.zeros((L, L))
        for i, (qa, qb) in enumerate(self.PAIRS):
            W = np.fft.ifft(Z[i * L : (i + 1) * L], axis=0)
            out += TqT(N, qa, W.real) + TqT(N, qb, W.imag)
        return out[:N, :N]

    def _packed(self, strip, ops):
        return self.unpack(self.eng.run(self.pack(strip), ops))

    def _image_strip(self, x):
        L, N = self.L, self.N
        X = np.zeros((L, L))
        X[:N, :N] = x
        return lambda q: Tq(N, q, X)

    def _data_strip(self, b):
        return lambda q: iota(self.N, b[q])

    def set_outer(self, lo, hi, deg):
        self.lo, self.hi, self.outer_deg = float(lo), float(hi), int(deg)

    def outer_interval(self, G, Pinv=None, P=None, eps=1e-10, margins=(0.99, 1.01), tol=1e-3, maxiter=300):
        """eigsh SA/LA + Weinstein residual enclosure, margins, D = ceil(log(eps/2)/log rho) + 2 (fastops.py:222-245)."""
        import scipy.sparse.linalg as spla

        N = self.N
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
        self.set_outer(lo_, hi_, D)

    def outer(self, y, G, P=lambda v: v):
        """Fixed-degree outer Chebyshev (fastops.py:247-258)."""
        c = (self.hi - self.lo) / (self.hi + self.lo)
        a = 2.0 / (self.hi + self.lo)
        xp = np.zeros_like(y)
        x = a * P(y)
        om = 1.0
        for k in range(1, self.outer_deg):
            om = 1.0 / (1.0 - c * c / 2.0) if k == 1 else 1.0 / (1.0 - c * c * om / 4.0)
            r = P(y - G(x))
            xn = x + om * (a * r) + (om - 1.0) * (x - xp)
            xp, x = x, xn
        return x


class Lp(_Packed):
    """L_p (fastops.py:261-288).  ``outer=(lo, hi, D)`` skips the eigsh setup."""

    def __init__(self, n, tau_fac=0.2, eps_in=0.1, eps=1e-10, threads=1, lam_max=None, power_iters=40, safety=1.5,
                 dense_cap=None, outer=None, streams="numpy"):
        self._common(n, tau_fac, eps_in, threads)
        if lam_max is None:
            lam_max = power_bound(n, self.z, 1, power_iters, safety, streams)
        self._engine(lam_max, dense_cap, (1, None))
        if outer is None:
            self.outer_interval(self.frame_window, eps=eps)
        else:
            self.set_outer(*outer)
        self.eng.count = [0]

    def ops_frame(self):
        n = self.n
        return [("P", 1, n), ("B", 1, n), ("BT", 1, n)]

    def ops_back(self):
        n = self.n
        return [("BT", 1, n), ("P", 1, n)]

    def frame_window(self, x):
        return self._packed(self._image_strip(x), self.ops_frame())

    def backproject(self, b):
        return self._packed(self._data_strip(b), self.ops_back())

    def P(self, y):
        return y

    def apply(self, b):
        return self.outer(self.backproject(b), self.frame_window)


class Hybrid(_Packed):
    """Hybrid inverse with explicit level 1 (fastops.py:291-338).  ``outer=(lo, hi, D)`` skips the eigsh setup."""

    def __init__(self, n, tau_fac=0.2, eps_in=0.1, eps=1e-10, threads=1, lam_max=None, g1_deg=14, power_iters=40,
                 safety=1.5, dense_cap=None, precond="g1", outer=None, streams="numpy"):
        self._common(n, tau_fac, eps_in, threads)
        self.g1_deg = g1_deg
        self.precond = precond
        if lam_max is None:
            lam_max = power_bound(n, self.z, 2, power_iters, safety, streams)
        self._engine(lam_max, dense_cap, (2, None))
        if outer is None:
            Pinv = G1_apply if precond == "g1" else (lambda y: 4.0 * y + 8.0 * Pi_apply(y))
            self.outer_interval(self.Gh, Pinv=Pinv, P=self.P, eps=eps)
        else:
            self.set_outer(*outer)
        self.eng.count = [0]

    def ops_gh(self):
        n = self.n
        return [("B", 1, 1), ("P", 2, n), ("B", 2, n), ("BT", 2, n), ("BT", 1, 1)]

    def ops_back(self):
        n = self.n
        return [("BT", 2, n), ("P", 2, n), ("BT", 1, 1)]

    def P(self, y):
        return g1_inv(y, self.g1_deg) if self.precond == "g1" else lorr_inv(y)

    def Gh(self, x):
        return self._packed(self._image_strip(x), self.ops_gh())

    def backproject(self, b):
        return self._packed(self._data_strip(b), self.ops_back())

    def apply(self, b):
        return self.outer(self.backproject(b), self.Gh, self.P)
