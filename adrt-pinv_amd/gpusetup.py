"""Dense low-frequency tables built on the GPU (setup only): rocSOLVER batched Cholesky solves and rocBLAS batched
complex GEMMs, reached through ctypes.

For every dense frequency k (degree above dense_cap) the reference builds F_k = (M_k + tau)^{-1} with
M_k = A_k^H A_k over the P levels, by eigh with the eigenvalues clipped at 0 (fastops.py:28-39).  Here F_k comes from
the Cholesky factor of M_k + tau I and two triangular solves against the identity.  Both give (M_k + tau)^{-1}: the
clip only lifts round-off eigenvalues of the PSD Gram (|lambda| ~ 1e-13 lambda_max) to 0, which moves F by far less
than its condition number (1 + lambda_max / tau <= 1 + 5N) times the unit round-off.  The device plan also wants the
whole op list of such a row as one matrix (dense.hip): for the hybrid G_h = B1^H M' F B1 and the backprojection
B1^H F A'^H; for L_p M F and F A^H.  All matrices live on the device in column-major order, which is the row-major
transpose the plan stores.  The level matrices come from the package's own butterflies applied to the identity: the
rows it returns are the columns of A, i.e. A in column-major order.

Tables are cached per (kind, n, tau, dense frequencies) for the life of the process (``clear_tables`` drops the
cache), and plans borrow them, so every plan of one size shares one copy in HBM.  Only the pass matrices stay, and
only one per conjugate pair of frequencies (``table_slots``): a third of the F, W_op, W_back of every frequency.
"""

import ctypes

import numpy as np

from . import lib as L

_OP_N, _OP_T, _OP_C = 111, 112, 113
_FILL_U = 121


class _Dz(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class RocLinalg:
    """A rocBLAS handle plus the batched routines the setup needs."""

    _inst = None

    @classmethod
    def get(cls):
        if cls._inst is None:
            cls._inst = cls()
        return cls._inst

    def __init__(self):
        self.rb = ctypes.CDLL("librocblas.so")
        self.rs = ctypes.CDLL("librocsolver.so")
        h = ctypes.c_void_p()
        st = self.rb.rocblas_create_handle(ctypes.byref(h))
        if st != 0:
            raise L.HipError(f"rocblas_create_handle failed ({st})")
        self.h = h
        I, P, S = ctypes.c_int, ctypes.c_void_p, ctypes.c_int64
        self.rb.rocblas_zgemm_strided_batched.argtypes = [P, I, I, I, I, I, P, P, I, S, P, I, S, P, P, I, S, I]
        self.rb.rocblas_zaxpy_strided_batched.argtypes = [P, I, P, P, I, S, P, I, S, I]
        self.rs.rocsolver_zpotrf_strided_batched.argtypes = [P, I, I, P, I, S, P, I]
        self.rs.rocsolver_zpotrs_strided_batched.argtypes = [P, I, I, I, P, I, S, P, I, S, I]
        self.one, self.zero = _Dz(1.0, 0.0), _Dz(0.0, 0.0)
        self._ones = {}

    def gemm(self, opA, opB, N, A, B, C, batch):
        """C_b = op(A_b) op(B_b) for N x N column-major matrices, batch-strided by N^2."""
        s = N * N
        st = self.rb.rocblas_zgemm_strided_batched(self.h, opA, opB, N, N, N, ctypes.byref(self.one), A, N, s, B, N, s,
                                                   ctypes.byref(self.zero), C, N, s, batch)
        if st != 0:
            raise L.HipError(f"rocblas_zgemm_strided_batched failed ({st})")

    def add_diag(self, N, A, alpha, batch):
        """A_b += alpha I."""
        if N not in self._ones:
            self._ones[N] = L.DeviceArray.from_host(np.ones(N, dtype=np.complex128))
        al = _Dz(float(alpha), 0.0)
        st = self.rb.rocblas_zaxpy_strided_batched(self.h, N, ctypes.byref(al), self._ones[N].ptr, 1, 0, A, N + 1,
                                                   N * N, batch)
        if st != 0:
            raise L.HipError(f"rocblas_zaxpy_strided_batched failed ({st})")

    def potrf(self, N, A, info, batch):
        """In place: the upper Cholesky factor of the Hermitian positive definite A_b."""
        st = self.rs.rocsolver_zpotrf_strided_batched(self.h, _FILL_U, N, A, N, N * N, info, batch)
        if st != 0:
            raise L.HipError(f"rocsolver_zpotrf_strided_batched failed ({st})")

    def potrs(self, N, A, B, batch):
        """B_b <- (U_b^H U_b)^{-1} B_b with the factor from potrf."""
        st = self.rs.rocsolver_zpotrs_strided_batched(self.h, _FILL_U, N, N, A, N, N * N, B, N, N * N, batch)
        if st != 0:
            raise L.HipError(f"rocsolver_zpotrs_strided_batched failed ({st})")


def _level_device(n, z, ks, lo, hi):
    """Column-major A_k over levels lo..hi for the frequencies ks, on the device (K, N, N)."""
    N = 1 << n
    X = np.tile(np.eye(N, dtype=complex), (len(ks), 1))
    zr = np.ascontiguousarray(np.repeat(np.asarray(z)[ks], N))
    L.check(L.load().adrt_bfly_rows(n, X.shape[0], L.ptr(zr), L.ptr(X), lo, hi, 0), "bfly_rows")
    return L.DeviceArray.from_host(X.reshape(len(ks), N, N))


_TABLES = {}


def clear_tables():
    """Drops the cached tables (plans that borrow them keep their own references)."""
    _TABLES.clear()


def dense_tables(n, z, ks, tau, kind):
    """(F, W_op, W_back, w) for the dense frequencies ks, from the cache or built by ``dense_setup``."""
    key = (kind, int(n), float(tau), tuple(int(k) for k in ks))
    if key not in _TABLES:
        _TABLES[key] = dense_setup(n, z, ks, tau, kind)
    return _TABLES[key]


def table_bytes(tables):
    return sum(t.nbytes for t in tables if t is not None)


def table_slots(z, ks):
    """Table index of each dense frequency (adrt_plan_desc.dense_w).  Frequencies k and L - k whose twiddles are exact
    conjugates have conjugate pass matrices (every op has real coefficients), and the plan applies the first one's
    conjugated to the other: the first of such a pair in the order of ks, and every frequency without a partner, get
    the next index; the partner gets -1.  This is the plan's own pairing rule (adrt_plan_create), so every table built
    is read."""
    z = np.asarray(z)
    L = z.size
    slot = {int(k): i for i, k in enumerate(ks)}
    w = np.full(len(ks), -1, dtype=np.int32)
    used = np.zeros(len(ks), dtype=bool)
    t = 0
    for i, k in enumerate(int(k) for k in ks):
        if used[i]:
            continue
        used[i] = True
        w[i] = t
        t += 1
        kp = (L - k) % L
        j = slot.get(kp, -1)
        if j >= 0 and j != i and not used[j] and z[kp].real == z[k].real and z[kp].imag == -z[k].imag:
            used[j] = True
    return w


def dense_setup(n, z, ks, tau, kind, chunk_bytes=1 << 30, keep_F=False):
    """Device arrays of W_op,k and W_back,k in column-major order, and w, the table index of each dense frequency
    (``table_slots``; None below n = 4, where every frequency has its own table).  Only frequencies with w >= 0 get
    tables, in the order of w.  F_k, an intermediate here (device plans never read it once they have the pass
    matrices), is returned for the same frequencies with keep_F, else None."""
    N = 1 << n
    ks = np.asarray(ks, dtype=np.int64)
    w = table_slots(z, ks) if n >= 4 else None
    own = ks if w is None else ks[w >= 0]
    K = own.size
    R = RocLinalg.get()
    lo = 2 if kind == "hybrid" else 1
    F = L.DeviceArray((K, N, N), np.complex128) if keep_F else None
    Wop = L.DeviceArray((K, N, N), np.complex128)
    Wback = L.DeviceArray((K, N, N), np.complex128)
    per = max(1, int(chunk_bytes // (7 * N * N * 16)))
    nbytes = N * N * 16
    lib = L.load()
    for i0 in range(0, K, per):
        kk = own[i0 : i0 + per]
        b = kk.size
        off = i0 * nbytes
        Ap = _level_device(n, z, kk, lo, n)
        M = L.DeviceArray((b, N, N), np.complex128)
        T1 = L.DeviceArray((b, N, N), np.complex128)
        T2 = L.DeviceArray((b, N, N), np.complex128)
        Ftmp = None if keep_F else L.DeviceArray((b, N, N), np.complex128)
        info = L.DeviceArray((b,), np.int32)
        R.gemm(_OP_C, _OP_N, N, Ap.ptr, Ap.ptr, M.ptr, b)  # M = A^H A
        L.check(lib.adrt_memcpy(T2.ptr, M.ptr, M.nbytes, 3, None), "d2d")
        R.add_diag(N, T2.ptr, tau, b)  # M + tau I
        R.potrf(N, T2.ptr, info.ptr, b)
        Fc = F.ptr + off if keep_F else Ftmp.ptr
        L.check(lib.adrt_memset(Fc, 0, b * nbytes, None), "memset")
        R.add_diag(N, Fc, 1.0, b)  # I
        R.potrs(N, T2.ptr, Fc, b)  # F = (M + tau I)^{-1}
        L.check(lib.adrt_device_sync(), "sync")
        if np.any(info.download() != 0):
            raise L.HipError("rocsolver_zpotrf: shifted Gram not positive definite")
        chk = np.empty((b, N, N), np.complex128)
        L.check(lib.adrt_memcpy(chk.ctypes.data, Fc, chk.nbytes, 2, None), "d2h")
        if not np.isfinite(chk).all():
            raise L.HipError("non-finite dense filter from the GPU solver")
        if kind == "hybrid":
            B1 = _level_device(n, z, kk, 1, 1)
            R.gemm(_OP_N, _OP_N, N, M.ptr, Fc, T1.ptr, b)  # M' F
            R.gemm(_OP_N, _OP_N, N, T1.ptr, B1.ptr, T2.ptr, b)  # M' F B1
            R.gemm(_OP_C, _OP_N, N, B1.ptr, T2.ptr, Wop.ptr + off, b)  # B1^H M' F B1
            R.gemm(_OP_N, _OP_C, N, Fc, Ap.ptr, T1.ptr, b)  # F A'^H
            R.gemm(_OP_C, _OP_N, N, B1.ptr, T1.ptr, Wback.ptr + off, b)  # B1^H F A'^H
        else:
            R.gemm(_OP_N, _OP_N, N, M.ptr, Fc, Wop.ptr + off, b)  # M F
            R.gemm(_OP_N, _OP_C, N, Fc, Ap.ptr, Wback.ptr + off, b)  # F A^H
        L.check(lib.adrt_device_sync(), "sync")
    return F, Wop, Wback, w


def to_row_major(dev):
    """Host row-major copy of device column-major matrices (K, N, N)."""
    return np.ascontiguousarray(dev.download().transpose(0, 2, 1))
