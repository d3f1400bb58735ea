"""ctypes binding of the reference's own C++ row engine, built from cpp/oblique.cpp into oracle/_ref by
``make -C oracle ref`` (oracle / test infrastructure and CPU baseline only).

Mirrors the binding of fastcpp.py:19-32 and the engine of fastcpp.py:49-99: B / BT copy the rows and call
``bfly_rows`` in place; P applies dense rows with a matrix product and the rest with ``cheb_rows``.
"""

import ctypes
import os

import numpy as np

from . import rows as rw

_REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref")
V4_FLAGS = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")


def cpu_info():
    """(model name, flags) of this host from /proc/cpuinfo (empty when unreadable)."""
    model, flags = "", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not model and line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                elif not flags and line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                if model and flags:
                    break
    except OSError:
        pass
    return model, flags


def _pick():
    """The AVX-512 (x86-64-v4) build where this CPU runs it, else the x86-64-v3 one: the nearest to the reference's
    -march=native that a library built on another host can be."""
    v4 = os.path.join(_REF, "liboblique_v4.so")
    if os.path.exists(v4) and all(f in cpu_info()[1] for f in V4_FLAGS):
        return v4, "x86-64-v4"
    return os.path.join(_REF, "liboblique.so"), "x86-64-v3"


LIB_PATH, MARCH = _pick()
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise OSError(f"reference row engine not built: {LIB_PATH} (run make -C oracle ref)")
        L = ctypes.CDLL(LIB_PATH)
        P, I = ctypes.c_void_p, ctypes.c_int
        L.bfly_rows.argtypes = [I, I, P, P, I, I, I, I]
        L.bfly_rows.restype = None
        L.cheb_rows.argtypes = [I, I, P, P, P, P, ctypes.c_double, P, P, I, I, I, P]
        L.cheb_rows.restype = None
        L.cert_rowsum.argtypes = [I, I, P, I, P, I]
        L.cert_rowsum.restype = None
        L.omp_max_threads.restype = I
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def bfly_rows(n, z, X, lo, hi, transpose, threads=1):
    X = np.array(X, dtype=complex, order="C")
    zc = np.ascontiguousarray(z, dtype=complex)
    lib().bfly_rows(n, X.shape[0], _p(zc), _p(X), lo, hi, int(transpose), threads)
    return X


def cheb_rows(n, z, deg, a, c, tau, Y, lo, hi, threads=1):
    Y = np.ascontiguousarray(Y, dtype=complex)
    Z = np.empty_like(Y)
    steps = ctypes.c_longlong(0)
    lib().cheb_rows(
        n, Y.shape[0], _p(np.ascontiguousarray(z, dtype=complex)), _p(np.ascontiguousarray(deg, dtype=np.int32)),
        _p(np.ascontiguousarray(a, dtype=float)), _p(np.ascontiguousarray(c, dtype=float)), float(tau), _p(Y), _p(Z),
        lo, hi, threads, ctypes.byref(steps),
    )
    return Z, steps.value


def cert_rowsum(n, omega, lo, threads=1):
    om = np.ascontiguousarray(omega, dtype=float)
    out = np.empty(om.size)
    lib().cert_rowsum(n, om.size, _p(om), lo, _p(out), threads)
    return out


def power_bound(n, z, lo, iters=40, safety=1.5, threads=1):
    """fastcpp.power_lam_max (fastcpp.py:102-118) on the reference kernels."""
    from .inverses import power_bound as pb

    def M(V):
        X = bfly_rows(n, z, V, lo, n, 0, threads)
        return bfly_rows(n, z, X, lo, n, 1, threads)

    return pb(n, z, lo, iters, safety, streams="cpp", apply=M)


class RefRows(rw.Rows):
    """The reference C++ kernels behind the row-engine surface (fastcpp.CppEngine semantics)."""

    def __init__(self, n, z, deg, lam_max, tau, threads=1, copies=2, dense_cap=None, p_levels=(1, None)):
        super().__init__(n, z, deg, lam_max, tau, threads, copies, dense_cap, p_levels)
        self.omp = max(1, int(threads))
        lib()

    def run(self, Y, ops):
        Y = np.asarray(Y, dtype=complex)
        R = Y.shape[0]
        k = np.arange(R) % self.L
        zr = np.ascontiguousarray(self.z[k])
        X = np.array(Y, order="C")
        steps = 0
        isd = np.array([int(kk) in self.dense for kk in k], dtype=bool)
        rest = np.nonzero(~isd)[0]
        for op, lo, hi in ops:
            if op in ("B", "BT"):
                lib().bfly_rows(self.n, R, _p(zr), _p(X), lo, hi, int(op == "BT"), self.omp)
                steps += R * (hi - lo + 1)
            else:
                out = np.empty_like(X)
                if rest.size:
                    kr = k[rest]
                    out[rest], s = cheb_rows(self.n, zr[rest], self.deg[kr], self.a[kr], self.c[kr], self.tau, X[rest],
                                             lo, hi, self.omp)
                    steps += s
                dr = np.nonzero(isd)[0]
                if dr.size:
                    F = np.array([self.dense[int(k[r])] for r in dr])
                    out[dr] = np.matmul(F, X[dr][:, :, None])[:, :, 0]
                steps += int(dr.size) * self.N // max(1, self.n)
                X = out
        self.count[0] += steps
        return X
