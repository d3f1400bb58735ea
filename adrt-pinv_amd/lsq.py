"""Least-squares comparators on the GPU.

``cgls`` has the signature and return value of pinv.cgls (pinv.py:23-48).  The iteration runs in libadrt_hip
(adrt_cgls): forward and transpose ADRT level kernels in float64 plus device-side scalars, batched over sinograms.
``cgls_batch`` takes a stack of sinograms, each of which stops on its own test.

``gram`` and ``lstsq_cheb`` are adrt.core.gram / adrt.core.lstsq_cheb of the reference's adrt patch
(patches/adrt-gram-freq.patch:411-545) with the same arguments, validation and results, on data in this repository's
sinogram layout (4, 2N-1, N) (the adrt package's own layout is a permutation of it).  Both run through a device plan of
the normal operator (ADRT_KIND_NORMAL): offset FFT of the packed quadrant pairs, the row Gram (levels 1..n forward
then adjoint) in the register row kernel, inverse FFT and quadrant sum; lstsq_cheb is that plan's fixed-degree
Chebyshev apply on [lam_min, lam_max].  float32 inputs run the rows in complex64 and return float32, as the patch's
float32 path does.
"""

import ctypes
import operator

import numpy as np

from . import lib as L


def cgls_device(n, d_b, d_x, S, tol=1e-12, maxit=10000, x0=False, hist_len=0, stream=None, callback=None):
    """On device arrays: d_b (S, 4, 2N-1, N), d_x (S, N, N).  Returns (iters (S,), hist (S, hist_len) or None).
    ``callback(it, d_x_ptr, gnorm (S,), ran (S,))`` is called after every iteration (adrt_cgls_cb)."""
    lib = L.load()
    work = L.DeviceArray((max(1, int(lib.adrt_cgls_work_bytes(n, S, hist_len))),), np.uint8)
    iters = np.zeros(S, dtype=np.int32)
    hist = np.zeros((S, hist_len)) if hist_len else None
    errors = []

    def relay(it, dx, gn, ran, S_, user):
        if errors:
            return
        try:  # an exception cannot cross the C frame: keep it and raise it after the call
            callback(it, dx, np.ctypeslib.as_array(gn, (S_,)).copy(), np.ctypeslib.as_array(ran, (S_,)).astype(bool))
        except BaseException as e:
            errors.append(e)

    cb = L.CGLS_CALLBACK(relay) if callback is not None else None
    L.check(lib.adrt_cgls_cb(n, L.ptr(d_b), L.ptr(d_x), S, float(tol), int(maxit), 1 if x0 else 0, work.ptr,
                             iters.ctypes.data, hist.ctypes.data if hist_len else None, int(hist_len),
                             stream.handle if stream else None, ctypes.cast(cb, L.P) if cb else None, None), "cgls")
    if errors:
        raise errors[0]
    return iters, hist


def cgls_batch(n, B, tol=1e-12, maxit=10000, X0=None):
    """CGLS for each sinogram of B (S, 4, 2N-1, N); returns (X (S, N, N), iters (S,), hists: list of arrays)."""
    N = 1 << n
    B = np.ascontiguousarray(np.asarray(B, dtype=float).reshape(-1, 4, 2 * N - 1, N))
    S = B.shape[0]
    st = L.Stream()
    d_b = L.DeviceArray.from_host(B, st)
    d_x = L.DeviceArray.from_host(np.ascontiguousarray(X0, dtype=float).reshape(S, N, N), st) if X0 is not None \
        else L.DeviceArray((S, N, N))
    hl = int(maxit) + 1
    iters, hist = cgls_device(n, d_b, d_x, S, tol, maxit, X0 is not None, hl, st)
    X = d_x.download(st)
    return X, iters, [hist[s, : iters[s] + 1] for s in range(S)]


def cgls(n, b, tol=1e-12, maxit=10000, x0=None, callback=None):
    """min ||R x - b||_2 (pinv.cgls): returns (x, hist) with hist[j] = ||R^T r_j||.  ``callback(it, x, gnorm)`` is
    called after every iteration, the stopping one included, as pinv.py:43-44 does; ``x`` is a host copy of the
    iterate (the reference hands over its live array), so a callback that keeps it needs no copy of its own."""
    if callback is None:
        X, iters, hists = cgls_batch(n, b, tol, maxit, None if x0 is None else np.asarray(x0)[None])
        return X[0], hists[0]
    N = 1 << n
    B = np.ascontiguousarray(np.asarray(b, dtype=float).reshape(1, 4, 2 * N - 1, N))
    st = L.Stream()
    d_b = L.DeviceArray.from_host(B, st)
    d_x = L.DeviceArray.from_host(np.ascontiguousarray(x0, dtype=float).reshape(1, N, N), st) if x0 is not None \
        else L.DeviceArray((1, N, N))
    host = np.empty((N, N))

    def each(it, dx, gnorm, ran):
        if ran[0]:
            L.check(L.load().adrt_memcpy(host.ctypes.data, dx, host.nbytes, 2, None), "d2h")
            callback(it, host.copy(), float(gnorm[0]))

    hl = int(maxit) + 1
    iters, hist = cgls_device(n, d_b, d_x, 1, tol, maxit, x0 is not None, hl, st, callback=each)
    return d_x.download(st)[0], hist[0, : iters[0] + 1]


# ---------------------------------------------------------------- normal operator R^T R and lstsq_cheb

MAX_BATCH = 64
_plans = {}


class NormalPlan:
    """Device plan of R^T R (ADRT_KIND_NORMAL) for one N and precision, up to ``max_batch`` images per call."""

    def __init__(self, n, precision="fp64", max_batch=MAX_BATCH):
        self.n, self.N, self.precision, self.max_batch = n, 1 << n, precision, int(max_batch)
        L2 = 2 * self.N
        self._z = np.ascontiguousarray(np.exp(1j * 2 * np.pi * np.fft.fftfreq(L2)))
        d = L.PlanDesc()
        d.kind = L.KIND_NORMAL
        d.n = n
        d.precision = L.PRECISION[precision]
        d.precond = L.PRECOND["none"]
        d.g1_deg = 1
        d.outer_deg = 1
        d.lo, d.hi, d.tau = 1.0, 2.0, 0.0
        d.z = L.ptr(self._z)
        d.deg = d.a = d.c = None
        d.n_dense = 0
        d.max_batch = self.max_batch
        d.refine = 0
        d.dense_on_device = 0
        h = L.P()
        L.check(L.load().adrt_plan_create(ctypes.byref(h), ctypes.byref(d)), "plan_create")
        self._plan = h.value
        self.d_img = L.DeviceArray((self.max_batch, self.N, self.N))
        self.d_out = L.DeviceArray((self.max_batch, self.N, self.N))
        self.d_sino = L.DeviceArray((self.max_batch, 4, 2 * self.N - 1, self.N))

    def _run(self, fn, src, d_src, S):
        lib = L.load()
        L.check(lib.adrt_memcpy(d_src.ptr, src.ctypes.data, src.nbytes, 1, None), "h2d")
        L.check(fn(self._plan, d_src.ptr, self.d_out.ptr, S, None), "plan call")
        out = np.empty((S, self.N, self.N))
        L.check(lib.adrt_memcpy(out.ctypes.data, self.d_out.ptr, out.nbytes, 2, None), "d2h")
        return out

    def gram(self, A):
        """R^T R of each image of A (S, N, N), S <= max_batch."""
        return self._run(L.load().adrt_plan_operator, np.ascontiguousarray(A, dtype=float), self.d_img, A.shape[0])

    def apply(self, B, lo, hi, D):
        """Chebyshev degree D on [lo, hi] for each sinogram of B (S, 4, 2N-1, N)."""
        L.check(L.load().adrt_plan_set_outer(self._plan, float(lo), float(hi), int(D)), "plan_set_outer")
        return self._run(L.load().adrt_plan_apply, np.ascontiguousarray(B, dtype=float), self.d_sino, B.shape[0])

    def close(self):
        if getattr(self, "_plan", None):
            L.load().adrt_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _plan(n, precision):
    key = (n, precision)
    if key not in _plans:
        _plans[key] = NormalPlan(n, precision)
    return _plans[key]


def _precision(dtype):
    if dtype == np.float32:
        return "fp32", np.float32
    if dtype == np.float64:
        return "fp64", np.float64
    raise TypeError(f"unsupported array dtype {dtype}")


def gram(a):
    """R^T R a for square images a (..., N, N), N a power of two, float32 or float64 (patch:411-471): the same shape and
    dtype.  ValueError for a non-square or non-power-of-two shape, TypeError for other dtypes."""
    a = np.asarray(a)
    if a.ndim < 2 or a.shape[-1] != a.shape[-2]:
        raise ValueError(f"array must be square, but had shape {a.shape}")
    N = a.shape[-1]
    if N < 1 or N & (N - 1):
        raise ValueError(f"array size must be a power of two, but was {N}")
    prec, dt = _precision(a.dtype)
    if N < 2:
        raise ValueError("the device engine needs N >= 2 (the 1 x 1 normal operator is 4 I)")
    n = N.bit_length() - 1
    flat = a.reshape(-1, N, N)
    P = _plan(n, prec)
    out = np.empty(flat.shape, dtype=dt)
    for s0 in range(0, flat.shape[0], P.max_batch):
        chunk = flat[s0 : s0 + P.max_batch]
        out[s0 : s0 + chunk.shape[0]] = P.gram(chunk)
    return out.reshape(a.shape)


def lstsq_cheb(b, *, lam_min, lam_max=None, tol=1e-10, degree=None):
    """Least-squares inverse of the ADRT by a fixed-degree Chebyshev polynomial (patch:474-545):
    x_D = c_D(R^T R) R^T b on [lam_min, lam_max], the degree fixed from the interval and tol unless ``degree`` is
    given.  lam_max defaults to 50 normalised power steps of R^T R from default_rng(0).standard_normal((N, N))
    enlarged by the residual norm.  b: (..., 4, 2N-1, N) in this repository's layout; returns (..., N, N)."""
    b = np.asarray(b)
    if b.ndim < 3 or b.shape[-3] != 4 or b.shape[-2] != 2 * b.shape[-1] - 1:
        raise ValueError(f"array must have the shape of an ADRT output, but had shape {b.shape}")
    N = b.shape[-1]
    if N < 1 or N & (N - 1):
        raise ValueError(f"array size must be a power of two, but was {N}")
    prec, dt = _precision(b.dtype)
    if N < 2:
        raise ValueError("the device engine needs N >= 2")
    n = N.bit_length() - 1
    if lam_max is None:
        v = np.random.default_rng(0).standard_normal((N, N)).astype(dt)
        v /= np.linalg.norm(v)
        for _ in range(50):
            w = gram(v)
            v = w / np.linalg.norm(w)
        w = gram(v)
        theta = float(np.vdot(v, w))
        lam_max = theta + float(np.linalg.norm(w - theta * v))
    lo, hi = float(lam_min), float(lam_max)
    if not 0 < lo < hi:
        raise ValueError("need 0 < lam_min < lam_max")
    if degree is None:
        kappa = hi / lo
        rho = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
        deg = max(1, int(np.ceil(np.log(tol / 2) / np.log(rho))))
    else:
        deg = operator.index(degree)
        if deg < 1:
            raise ValueError("degree must be >= 1")
    flat = b.reshape(-1, 4, 2 * N - 1, N)
    P = _plan(n, prec)
    out = np.empty((flat.shape[0], N, N), dtype=dt)
    for s0 in range(0, flat.shape[0], P.max_batch):
        chunk = flat[s0 : s0 + P.max_batch]
        out[s0 : s0 + chunk.shape[0]] = P.apply(chunk, lo, hi, deg)
    return out.reshape(b.shape[:-3] + (N, N))
