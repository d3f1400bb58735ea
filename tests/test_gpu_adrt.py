"""Forward / transpose ADRT on the GPU: bit-exact against the oracle (adrtcore restatement) on integer images."""

import numpy as np
import pytest

from conftest import other_package, pkg, relmax
from oracle import layout as olay
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


@pytest.mark.parametrize("n", [1, 4, 8])
def test_layout_permute_matches_reference_maps(gpu, n):
    """The device permutation, both directions, on batches of 5 and 1 sinograms, equals the reference's placement
    with the reference's maps (skimage_adrt_pkg.py:17-51) bit for bit; the fingerprints use the GPU transform."""
    LY = pkg("layout")
    L = pkg("lib")
    N = 1 << n
    core = lambda x: orad.adrt(n, x)
    pk = other_package(n, core)
    print(f"layout n={n}", flush=True)
    lm = LY.LayoutMap(n, pk)
    maps = olay.layout_maps(n, core, pk)
    rng = np.random.default_rng(20 + n)
    bp = rng.standard_normal((5, 4, 2 * N - 1, N))  # noisy data: every entry set, structural ones included
    want = np.stack([olay.pkg_to_core(b, n, maps) for b in bp])
    assert np.array_equal(lm.pkg_to_core(bp), want)
    assert np.array_equal(lm.pkg_to_core(bp[2]), want[2])
    bc = rng.standard_normal((5, 4, 2 * N - 1, N))
    assert np.array_equal(lm.core_to_pkg(bc), np.stack([olay.core_to_pkg(b, n, maps) for b in bc]))
    d = L.DeviceArray((1,), np.float64)
    assert L.load().adrt_layout_permute(n, lm._to_core.ptr, d.ptr, lm._to_pkg.ptr, 0, None) == 0  # S = 0: nothing
    assert L.load().adrt_layout_permute(n, lm._to_core.ptr, d.ptr, d.ptr, 1, None) == -1  # aliasing refused
    # the module-level functions keep one map per size until clear_maps()
    assert np.array_equal(LY.pkg_to_core(bp[1], n, pk), want[1])
    assert len(LY._maps) >= 1
    LY.clear_maps()
    assert not LY._maps


def test_inverse_of_package_layout_data(gpu):
    """Data in the other layout, moved to the core's on the GPU, inverts like the core's own data."""
    LY = pkg("layout")
    I = pkg("inverses")
    C = pkg("plancache")
    R = pkg("radon")
    n, N = 6, 64
    pk = other_package(n, lambda x: R.adrt(x))
    lm = LY.LayoutMap(n, pk)
    x = pkg("phantom").ellipses(31, N)
    lam, outer = C.hybrid_setup(n, dense_cap=8)
    H = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=8)
    xp = H.apply(lm.pkg_to_core(pk(x)))
    assert relmax(xp, H.apply(R.adrt(x))) < 1e-12
    assert np.linalg.norm(xp - x) / np.linalg.norm(x) < 1e-6


def test_layout_permute_round_trip_at_n2048(gpu):
    """At the largest serving size (N=2048, 33.5M entries per sinogram) the device permutation matches the maps'
    placement and core_to_pkg undoes pkg_to_core on every entry a line reaches."""
    LY = pkg("layout")
    R = pkg("radon")
    n, N = 11, 2048
    print(f"layout round trip n={n}", flush=True)
    lm = LY.LayoutMap(n, other_package(n, lambda x: R.adrt(x)))
    rng = np.random.default_rng(5)
    bp = rng.standard_normal((2, 4, 2 * N - 1, N))
    bc = lm.pkg_to_core(bp)
    want = np.zeros((2, lm.M))
    want[:, lm.core_idx] = bp.reshape(2, -1)[:, lm.pkg_idx]
    assert np.array_equal(bc.reshape(2, -1), want)
    back = lm.core_to_pkg(bc).reshape(2, -1)
    reached = np.zeros(lm.M, bool)
    reached[lm.pkg_idx] = True
    assert np.array_equal(back[:, reached], bp.reshape(2, -1)[:, reached])
    assert not back[:, ~reached].any()
