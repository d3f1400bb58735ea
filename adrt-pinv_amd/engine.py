"""HipEngine: the GPU row engine behind the fastops.RowEngine plug-in surface, and the spectral bounds.

``HipEngine(n, z, deg, lam_max, tau, threads, copies, dense_cap, p_levels)`` and ``run(Y, ops)`` keep the constructor
and call of fastops.RowEngine (fastops.py:20-167) and fastcpp.CppEngine (fastcpp.py:49-99), so the class can be
assigned as ``engine_cls`` of the reference's FastLp / FastHybrid.  Rows live on the GPU only for the duration of a
``run``; ``run_device`` takes device buffers for callers that keep rows resident.

``power_lam_max``, ``cert_lam_max`` and ``explicit_lam_max`` mirror fastcpp.py:102-134 on the GPU kernels.
"""

import ctypes

import numpy as np

from . import lib as L
from . import spectral


def level_steps(ops, rows, period, deg, dense_rows, n, N):
    """Level-row count of one run, accumulated like RowEngine / CppEngine (fastops.py:137-154; fastcpp.py:99)."""
    steps = 0
    reps, rem = divmod(rows, period)
    deg = np.asarray(deg, dtype=np.int64)
    dense = np.zeros(period, dtype=bool)
    if dense_rows is not None and len(dense_rows):
        dense[np.asarray(dense_rows, dtype=np.int64)] = True
    mult = np.full(period, reps, dtype=np.int64)
    mult[:rem] += 1
    for op, lo, hi in ops:
        if op in ("B", "BT"):
            steps += rows * (hi - lo + 1)
        else:
            d = np.where(dense, 0, np.maximum(deg - 1, 0))
            steps += int(np.sum(mult * d)) * 2 * (hi - lo + 1)
            steps += int(np.sum(mult[dense])) * N // max(1, n)
    return steps


def level_matrices(n, z, ks, lo, hi):
    """A_k[s, t] = (B_{lo..hi} e_t)[s] for the frequencies ks, from the GPU butterflies applied to the identity."""
    N = 1 << n
    ks = np.asarray(ks, dtype=np.int64)
    X = np.tile(np.eye(N, dtype=complex), (ks.size, 1))
    zr = np.ascontiguousarray(np.repeat(np.asarray(z)[ks], N))
    L.check(L.load().adrt_bfly_rows(n, X.shape[0], L.ptr(zr), L.ptr(X), lo, hi, 0), "bfly_rows")
    return X.reshape(ks.size, N, N).transpose(0, 2, 1)


def pass_matrices(n, z, ks, F, kind):
    """The whole op list of a dense row as one matrix per frequency (the ops are linear and row-local):
    hybrid G_h [B1, P, B', B'^H, B1^H] -> B1^H A'^H A' F B1,  backprojection [B'^H, P, B1^H] -> B1^H F A'^H;
    L_p frame [P, B, B^H] -> A^H A F,  backprojection [B^H, P] -> F A^H  (op lists of fastops.py:281-335)."""
    H = lambda M: np.conj(M.transpose(0, 2, 1))
    if kind == "hybrid":
        B1 = level_matrices(n, z, ks, 1, 1)
        Ap = level_matrices(n, z, ks, 2, n)
        return H(B1) @ (H(Ap) @ Ap) @ F @ B1, H(B1) @ F @ H(Ap)
    A = level_matrices(n, z, ks, 1, n)
    return H(A) @ A @ F, F @ H(A)


def dense_tables(n, z, tau, ks, lo, hi):
    """Exact Tikhonov filters F_k = (M_k + tau)^{-1}, M_k = A_k^H A_k over levels lo..hi, by eigh with the
    eigenvalues clipped at 0 (fastops.py:28-39).  A_k comes from the GPU butterflies applied to the identity."""
    N = 1 << n
    ks = np.asarray(ks, dtype=np.int64)
    if ks.size == 0:
        return np.zeros((0, N, N), dtype=complex)
    out = np.empty((ks.size, N, N), dtype=complex)
    chunk = max(1, (1 << 22) // (N * N))
    for i0 in range(0, ks.size, chunk):
        kk = ks[i0 : i0 + chunk]
        X = np.tile(np.eye(N, dtype=complex), (kk.size, 1))  # row t of block i = e_t
        zr = np.repeat(np.asarray(z)[kk], N)
        L.check(L.load().adrt_bfly_rows(n, X.shape[0], L.ptr(np.ascontiguousarray(zr)), L.ptr(X), lo, hi, 0),
                "bfly_rows")
        A = X.reshape(kk.size, N, N).transpose(0, 2, 1)  # A[s, t] = (B e_t)[s]
        M = np.conj(A.transpose(0, 2, 1)) @ A
        lam, V = np.linalg.eigh(M)
        lam = np.maximum(lam, 0)
        out[i0 : i0 + kk.size] = (V / (lam[:, None, :] + tau)) @ np.conj(V.transpose(0, 2, 1))
    return out


class HipEngine:
    """GPU row engine with the RowEngine constructor and ``run`` (fastops.py:20-167).

    ``threads`` is accepted and ignored (the GPU runs all rows at once).  ``precision`` ('fp64' | 'fp32' | 'fp32_plain')
    selects the row arithmetic; 'fp64' is the drop-in default."""

    def __init__(self, n, z, deg, lam_max, tau, threads=1, copies=2, dense_cap=None, p_levels=(1, None),
                 precision="fp64"):
        self.n = n
        self.N = N = 1 << n
        self.L = len(z)
        self.R = copies * self.L
        self.copies = copies
        self.tau = float(tau)
        self.z = np.ascontiguousarray(z, dtype=complex)
        self.deg = np.ascontiguousarray(deg, dtype=np.int32)
        self.lam_max = np.asarray(lam_max, dtype=float)
        self.c, self.a = spectral.cheb_coeffs(self.lam_max, self.tau)
        self.c = np.ascontiguousarray(self.c, dtype=float)
        self.a = np.ascontiguousarray(self.a, dtype=float)
        self.precision = precision
        self.p_levels = (p_levels[0], p_levels[1] or n)
        self.dense_k = np.nonzero(self.deg > dense_cap)[0].astype(np.int32) if dense_cap is not None else np.zeros(0, np.int32)
        self.n_dense = int(self.dense_k.size)
        self._dense_F = None
        self._plan = None
        self.count = [0]

    @property
    def dense_F(self):
        """Host row-major exact filters of the dense frequencies (built on first use)."""
        if self._dense_F is None and self.n_dense:
            self._dense_F = dense_tables(self.n, self.z, self.tau, self.dense_k, *self.p_levels)
        return self._dense_F

    def _rows_plan(self):
        if self._plan is None:
            h = L.P()
            F = np.ascontiguousarray(self.dense_F) if self.n_dense else None
            L.check(
                L.load().adrt_rows_plan_create(
                    ctypes.byref(h), self.n, self.L, L.ptr(self.z), L.ptr(self.deg), L.ptr(self.a), L.ptr(self.c),
                    self.tau, self.n_dense, L.ptr(self.dense_k) if self.n_dense else None, L.ptr(F),
                    L.PRECISION[self.precision],
                ),
                "rows_plan_create",
            )
            self._plan = h.value
        return self._plan

    def steps(self, ops, rows=None):
        return level_steps(ops, self.R if rows is None else rows, self.L, self.deg, self.dense_k, self.n, self.N)

    def run_device(self, dY, dZ, rows, ops, stream=None):
        """Rows on the device: dY -> dZ (may alias), complex128 (fp64) or complex64 (fp32)."""
        o = L.ops_array(ops)
        L.check(L.load().adrt_rows_run(self._rows_plan(), L.ptr(dY), L.ptr(dZ), rows, L.ptr(o), len(ops),
                                       stream.handle if stream else None), "rows_run")
        self.count[0] += self.steps(ops, rows)

    def run(self, Y, ops):
        """Y: (rows, N) complex, rows = copies * L for the drop-in use; returns a new array (Y is not modified)."""
        ct = np.complex128 if self.precision == "fp64" else np.complex64
        Y = np.ascontiguousarray(Y, dtype=ct)
        rows = Y.shape[0]
        d = L.DeviceArray(Y.shape, ct)
        d.upload(Y)
        self.run_device(d, d, rows, ops)
        out = d.download()
        d.free()
        return out.astype(np.complex128, copy=False)

    def close(self):
        if getattr(self, "_plan", None):
            L.load().adrt_rows_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- spectral bounds (fastcpp.py:102-134)
def power_lam_max(n, z, m_lo, iters=40, safety=1.5, threads=1):
    """safety x Rayleigh quotient after ``iters`` normalised power steps of M_k over levels m_lo..n, every frequency,
    from V = N(0,1)_{rng(0)} + i N(0,1)_{rng(1)} (fastcpp.py:102-118); the Gram runs on the GPU butterflies."""
    N = 1 << n
    K = len(z)
    zc = np.ascontiguousarray(z, dtype=complex)
    V = np.random.default_rng(0).standard_normal((K, N)) + 1j * np.random.default_rng(1).standard_normal((K, N))
    lb = L.load()

    def M(V):
        X = np.array(V, dtype=complex, order="C")
        L.check(lb.adrt_bfly_rows(n, K, L.ptr(zc), L.ptr(X), m_lo, n, 0), "bfly_rows")
        L.check(lb.adrt_bfly_rows(n, K, L.ptr(zc), L.ptr(X), m_lo, n, 1), "bfly_rows")
        return X

    for _ in range(iters):
        V = M(V)
        V /= np.linalg.norm(V, axis=1, keepdims=True)
    return safety * np.real(np.sum(np.conj(V) * M(V), axis=1))


def cert_lam_max(n, oms, m_lo=1, threads=1):
    """Certified bounds by the row-sum recursion on the GPU (fastcpp.py:121-125)."""
    om = np.ascontiguousarray(oms, dtype=float)
    out = np.empty(om.size)
    L.check(L.load().adrt_cert_rowsum(n, om.size, L.ptr(om), m_lo, L.ptr(out)), "cert_rowsum")
    return out


explicit_lam_max = spectral.explicit_lam_max
