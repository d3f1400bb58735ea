"""Least-squares cross-check on the GPU: ``cgls`` with the signature and return value of pinv.cgls (pinv.py:23-48).

The iteration runs in libadrt_hip (adrt_cgls): forward and transpose ADRT level kernels in float64 plus device-side
scalars, batched over sinograms.  ``cgls_batch`` takes a stack of sinograms, each of which stops on its own test.
"""

import ctypes

import numpy as np

from . import lib as L


def cgls_device(n, d_b, d_x, S, tol=1e-12, maxit=10000, x0=False, hist_len=0, stream=None):
    """On device arrays: d_b (S, 4, 2N-1, N), d_x (S, N, N).  Returns (iters (S,), hist (S, hist_len) or None)."""
    lib = L.load()
    work = L.DeviceArray((max(1, int(lib.adrt_cgls_work_bytes(n, S, hist_len))),), np.uint8)
    iters = np.zeros(S, dtype=np.int32)
    hist = np.zeros((S, hist_len)) if hist_len else None
    L.check(lib.adrt_cgls(n, L.ptr(d_b), L.ptr(d_x), S, float(tol), int(maxit), 1 if x0 else 0, work.ptr,
                          iters.ctypes.data, hist.ctypes.data if hist_len else None, int(hist_len),
                          stream.handle if stream else None), "cgls")
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
    """min ||R x - b||_2 (pinv.cgls): returns (x, hist) with hist[j] = ||R^T r_j||.  ``callback`` is not supported
    (the iteration runs on the device without host round trips of x)."""
    if callback is not None:
        raise NotImplementedError("cgls on the GPU does not call back per iteration")
    X, iters, hists = cgls_batch(n, b, tol, maxit, None if x0 is None else np.asarray(x0)[None])
    return X[0], hists[0]
