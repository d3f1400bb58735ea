"""adrt_forward_noalloc (the forward ADRT that allocates nothing) against adrt_forward with a workspace, through the
C ABI: the same bits on two calls in a row, the second smaller, including calls that run in several chunks of the
static scratch."""

import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu

DT = {0: np.int32, 1: np.float32, 2: np.float64}


def _forward(L, n, dtype, x, noalloc):
    """(S, N, N) -> (S, 4, 2N-1, N) through one of the two entry points; the output starts as all-ones bytes, so an
    entry a level does not write shows."""
    lib = L.load()
    S, N = x.shape[0], 1 << n
    dt = DT[dtype]
    d_x = L.DeviceArray.from_host(np.ascontiguousarray(x, dtype=dt))
    d_b = L.DeviceArray((S, 4, 2 * N - 1, N), dt)
    L.check(lib.adrt_memset(d_b.ptr, 0xFF, d_b.nbytes, None), "memset")
    if noalloc:
        L.check(lib.adrt_forward_noalloc(n, dtype, d_x.ptr, d_b.ptr, S, None), "forward_noalloc")
    else:
        work = L.DeviceArray((max(1, int(lib.adrt_work_bytes_fwd(n, dtype, S))),), np.uint8)
        L.check(lib.adrt_forward(n, dtype, d_x.ptr, d_b.ptr, work.ptr, S, None), "forward")
    L.check(lib.adrt_device_sync(), "sync")
    out = d_b.download()
    d_x.free()
    d_b.free()
    return out


def _images(n, S, dtype, seed):
    rng = np.random.default_rng(seed)
    N = 1 << n
    if dtype == 0:
        return rng.integers(0, 256, (S, N, N), dtype=np.int32)
    return rng.standard_normal((S, N, N))


@pytest.mark.parametrize("dtype", [0, 1, 2])
@pytest.mark.parametrize("n,S1,S2", [(1, 3, 1), (6, 9, 4), (8, 5, 2)])
def test_noalloc_forward_same_bits_two_calls(gpu, dtype, n, S1, S2):
    L = gpu
    for S, seed in ((S1, 10 + n), (S2, 20 + n)):
        x = _images(n, S, dtype, seed)
        want = _forward(L, n, dtype, x, noalloc=False)
        got = _forward(L, n, dtype, x, noalloc=True)
        assert np.array_equal(got, want), (n, S, dtype)


@pytest.mark.parametrize("dtype,S1,S2", [(0, 3, 1), (2, 2, 1)])
def test_noalloc_forward_in_chunks_n2048(gpu, dtype, S1, S2):
    """N = 2048: the static scratch holds two int32 sinograms' levels (one float64 sinogram's), so the first call
    runs in chunks; the second, smaller call reuses the scratch the first left behind."""
    L = gpu
    lib = L.load()
    n = 11
    per = 4 * (2048 + 1024 - 1) * 2048 * (8 if dtype == 2 else 4)
    assert lib.adrt_noalloc_scratch_bytes() // per < S1
    for S, seed in ((S1, 31), (S2, 32)):
        x = _images(n, S, dtype, seed)
        want = _forward(L, n, dtype, x, noalloc=False)
        got = _forward(L, n, dtype, x, noalloc=True)
        assert np.array_equal(got, want), (S, dtype)


def test_noalloc_forward_equals_oracle(gpu):
    """Integer images through the no-allocation entry point equal the oracle's forward ADRT exactly."""
    from oracle import radon as orad

    n = 5
    x = _images(n, 2, 0, 7)
    got = _forward(gpu, n, 0, x, noalloc=True)
    for s in range(2):
        assert np.array_equal(got[s], orad.adrt(n, x[s]).astype(np.int32))
