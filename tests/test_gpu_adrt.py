"""Forward / transpose ADRT on the GPU: bit-exact against the oracle (adrtcore restatement) on integer images."""

import numpy as np
import pytest

from conftest import pkg
from oracle import radon as orad

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8])
def test_forward_int32_exact(gpu, n):
    N = 1 << n
    x = np.random.default_rng(n).integers(0, 256, (N, N), dtype=np.int32)
    R = pkg("radon")
    out = R.adrt(x)
    assert out.dtype == np.int32
    assert np.array_equal(out, orad.adrt(n, x, dtype=np.int64))


@pytest.mark.parametrize("n", [1, 3, 5, 7])
def test_forward_float64_bitexact(gpu, n):
    N = 1 << n
    x = np.random.default_rng(n).standard_normal((N, N))
    out = pkg("radon").adrt(x)
    assert np.array_equal(out, orad.adrt(n, x))


@pytest.mark.parametrize("n", [1, 2, 4, 6, 8])
def test_transpose_int_exact(gpu, n):
    N = 1 << n
    b = np.random.default_rng(n).integers(0, 256, (4, 2 * N - 1, N), dtype=np.int32)
    out = pkg("radon").adrt_T(b)
    assert out.dtype == np.int64
    assert np.array_equal(out, orad.adrt_T(n, b, dtype=np.int64))


@pytest.mark.parametrize("n", [2, 5, 7])
def test_transpose_float64_bitexact(gpu, n):
    N = 1 << n
    b = np.random.default_rng(n).standard_normal((4, 2 * N - 1, N))
    assert np.array_equal(pkg("radon").adrt_T(b), orad.adrt_T(n, b))


def test_batch_and_large_identity(gpu):
    """N = 2048: <R x, b> = <x, R^T b> exactly in int64, and R^T R x of a [0,255] image needs int64 (max > 2^31)."""
    n = 11
    N = 1 << n
    R = pkg("radon")
    rng = np.random.default_rng(11)
    x = rng.integers(0, 256, (N, N), dtype=np.int32)
    bx = R.adrt(x)
    assert bx.max() <= 255 * N and bx.min() >= 0
    b = rng.integers(0, 4, bx.shape, dtype=np.int32)
    lhs = int(np.sum(bx.astype(np.int64) * b))
    rhs = int(np.sum(x.astype(np.int64) * R.adrt_T(b)))
    assert lhs == rhs
    # <R^T R x, 1> = <R x, R 1>
    big = R.adrt_T(bx)
    ones = R.adrt(np.ones((N, N), np.int32))
    assert int(big.sum()) == int(np.sum(bx.astype(np.int64) * ones.astype(np.int64)))
    # a constant 255 image: R^T R x reaches ~4 * 255 * N^2 > 2^31, which only the int64 accumulation holds
    full = R.adrt_T(R.adrt(np.full((N, N), 255, np.int32)))
    assert full.max() > 2**31 - 1
    assert np.array_equal(full, 255 * R.adrt_T(ones))


def test_batched_forward_matches_single(gpu):
    n = 5
    N = 1 << n
    X = np.random.default_rng(0).integers(0, 256, (3, N, N), dtype=np.int32)
    R = pkg("radon")
    out = R.adrt(X)
    for s in range(3):
        assert np.array_equal(out[s], orad.adrt(n, X[s], dtype=np.int64))
