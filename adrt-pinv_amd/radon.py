"""Forward ADRT and its transpose on the GPU (adrtcore.adrt / adrt_T, adrtcore.py:122-137).

Integer images stay exact: int32 in, int32 out for the forward transform; int32 in, int64 out for the transpose (the
transpose of a forward ADRT of a [0, 255] image exceeds 2^31 at N = 2048).  float32 / float64 keep the reference's
summation order, so float64 results equal adrtcore's bit for bit.
"""

import numpy as np

from . import lib as L

_OUT_T = {0: np.int64, 1: np.float32, 2: np.float64}


def _n_of(N):
    n = int(N).bit_length() - 1
    if N != 1 << n:
        raise ValueError(f"N must be a power of two, got {N}")
    return n


def forward_device(n, dtype, d_x, d_b, S=1, stream=None, work=None):
    """d_x (S, N, N) -> d_b (S, 4, 2N-1, N) of ``dtype`` (int32 | float32 | float64)."""
    code = L.DTYPES[np.dtype(dtype)]
    w = work
    if w is None:
        w = L.DeviceArray((max(1, int(L.load().adrt_work_bytes_fwd(n, code, S))),), np.uint8)
    L.check(L.load().adrt_forward(n, code, L.ptr(d_x), L.ptr(d_b), L.ptr(w), S, stream.handle if stream else None),
            "adrt_forward")
    return w


def transpose_device(n, dtype, d_b, d_x, S=1, stream=None, work=None):
    """d_b (S, 4, 2N-1, N) of ``dtype`` -> d_x (S, N, N) (int64 for int32 input)."""
    code = L.DTYPES[np.dtype(dtype)]
    w = work
    if w is None:
        w = L.DeviceArray((max(1, int(L.load().adrt_work_bytes_T(n, code, S))),), np.uint8)
    L.check(L.load().adrt_transpose(n, code, L.ptr(d_b), L.ptr(d_x), L.ptr(w), S, stream.handle if stream else None),
            "adrt_transpose")
    return w


def adrt(x, dtype=None):
    """R x for one image (N, N) or a batch (S, N, N); dtype defaults to x's (int32 / float32 / float64)."""
    x = np.asarray(x)
    dtype = np.dtype(dtype or (x.dtype if x.dtype in (np.int32, np.float32) else np.float64))
    single = x.ndim == 2
    xb = np.ascontiguousarray(x.reshape((1,) + x.shape) if single else x, dtype=dtype)
    S, N = xb.shape[0], xb.shape[-1]
    n = _n_of(N)
    dx = L.DeviceArray.from_host(xb)
    db = L.DeviceArray((S, 4, 2 * N - 1, N), dtype)
    forward_device(n, dtype, dx, db, S)
    out = db.download()
    return out[0] if single else out


def adrt_T(b, dtype=None):
    """R^T b for one sinogram (4, 2N-1, N) or a batch (S, 4, 2N-1, N)."""
    b = np.asarray(b)
    dtype = np.dtype(dtype or (b.dtype if b.dtype in (np.int32, np.float32) else np.float64))
    single = b.ndim == 3
    bb = np.ascontiguousarray(b.reshape((1,) + b.shape) if single else b, dtype=dtype)
    S, N = bb.shape[0], bb.shape[-1]
    n = _n_of(N)
    db = L.DeviceArray.from_host(bb)
    dx = L.DeviceArray((S, N, N), _OUT_T[L.DTYPES[dtype]])
    transpose_device(n, dtype, db, dx, S)
    out = dx.download()
    return out[0] if single else out
