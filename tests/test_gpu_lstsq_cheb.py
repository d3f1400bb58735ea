"""The normal operator R^T R and the normal-equation Chebyshev lstsq_cheb on the GPU (adrt.core.gram / lstsq_cheb of
patches/adrt-gram-freq.patch:411-545), against the oracle and the patch's own checks (tests/test_gram.py and
tests/test_lstsq_cheb.py of the patch, patch:592-714), on data in this repository's layout."""

import numpy as np
import pytest

from conftest import pkg, relmax

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_gram_float64(gpu, n):
    from oracle import lsq as olsq

    N = 1 << n
    a = np.random.default_rng(N).standard_normal((N, N))
    ref = olsq.gram(n, a)
    out = pkg("lsq").gram(a)
    assert out.shape == a.shape and out.dtype == np.float64
    assert np.allclose(out, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("n", [2, 4, 5])
def test_gram_float32(gpu, n):
    from oracle import lsq as olsq

    N = 1 << n
    a = np.random.default_rng(N).standard_normal((N, N)).astype(np.float32)
    ref = olsq.gram(n, a.astype(np.float64))
    out = pkg("lsq").gram(a)
    assert out.dtype == np.float32
    # the patch's float32 tolerance (patch:607-612)
    assert np.allclose(out, ref, rtol=0, atol=2e-5 * np.abs(ref).max())


def test_gram_batch_and_symmetry(gpu):
    from oracle import lsq as olsq

    G = pkg("lsq").gram
    a = np.random.default_rng(0).standard_normal((3, 2, 16, 16))
    out = G(a)
    for i in range(3):
        for j in range(2):
            assert np.allclose(out[i, j], olsq.gram(4, a[i, j]), rtol=0, atol=1e-10 * np.abs(out).max())
    u, v = np.random.default_rng(1).standard_normal((2, 16, 16))
    assert np.isclose(np.vdot(G(u), v), np.vdot(u, G(v)))
    assert np.vdot(G(u), u) > 0


@pytest.mark.parametrize("n", [2, 3, 4])
def test_lstsq_cheb_matches_dense_lstsq(gpu, n):
    from oracle import radon

    N = 1 << n
    rng = np.random.default_rng(N)
    A = radon.dense(lambda x: radon.adrt(n, x), (N, N))
    eig = np.linalg.eigvalsh(A.T @ A)
    b = radon.adrt(n, rng.standard_normal((N, N))) + 0.1 * rng.standard_normal((4, 2 * N - 1, N))
    ref = np.linalg.lstsq(A, b.ravel(), rcond=None)[0].reshape(N, N)
    out = pkg("lsq").lstsq_cheb(b, lam_min=0.99 * eig[0], lam_max=1.01 * eig[-1], tol=1e-10)
    assert np.linalg.norm(out - ref) <= 1e-8 * np.linalg.norm(ref)


def test_lstsq_cheb_estimated_lam_max(gpu):
    from oracle import radon

    x = np.random.default_rng(3).standard_normal((8, 8))
    out = pkg("lsq").lstsq_cheb(radon.adrt(3, x), lam_min=15.0)
    assert np.linalg.norm(out - x) <= 1e-8 * np.linalg.norm(x)


def test_lstsq_cheb_vs_oracle_n64(gpu):
    """Same interval and degree as the oracle's restatement: the device polynomial equals the host one."""
    from oracle import lsq as olsq
    from oracle import radon

    n, N = 6, 64
    ph = pkg("phantom")
    b = ph.add_noise(radon.adrt(n, ph.ellipses(5, N)), 0.01, 6)
    kw = dict(lam_min=0.99 * 50.3, lam_max=3.5 * N * N, tol=1e-8)
    ref = olsq.lstsq_cheb(n, b, **kw)
    out = pkg("lsq").lstsq_cheb(b, **kw)
    assert relmax(out, ref) < 1e-10


def test_lstsq_cheb_recovers_n256(gpu):
    """At N = 256 R has full column rank, so the least-squares solution of R x = b for b = R x is x: the degree-D
    polynomial reproduces x to about its tolerance (lam_min below the measured 88.69, exp_hybrid_c2.py:16)."""
    from oracle import radon

    n, N = 8, 256
    x = pkg("phantom").ellipses(7, N)
    out = pkg("lsq").lstsq_cheb(radon.adrt(n, x), lam_min=0.95 * 88.69, tol=1e-9)
    assert np.linalg.norm(out - x) <= 1e-7 * np.linalg.norm(x)


@pytest.mark.parametrize("n", [9, 10, 11])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gram_and_backprojection_large_n(gpu, n, dtype):
    """At N = 512 .. 2048 (the offset FFTs of every size: two-step at n = 9, three-step at n = 10, 11) the frequency
    path equals the spatial one: gram(a) = R^T R a and the normal plan's backprojection = R^T b, against the GPU level
    kernels (bit-exact against the oracle's ADRT, test_gpu_adrt.py)."""
    import ctypes

    R = pkg("radon")
    lsq = pkg("lsq")
    N = 1 << n
    rng = np.random.default_rng(n)
    a = rng.standard_normal((N, N))
    ref = R.adrt_T(R.adrt(a))
    tol = 1e-11 if dtype == np.float64 else 2e-5
    out = lsq.gram(a.astype(dtype))
    assert out.dtype == dtype
    assert np.abs(out - ref).max() <= tol * np.abs(ref).max()
    b = rng.standard_normal((1, 4, 2 * N - 1, N))
    P = lsq._plan(n, "fp64" if dtype == np.float64 else "fp32")
    L = gpu
    lib = L.load()
    L.check(lib.adrt_memcpy(P.d_sino.ptr, b.ctypes.data, b.nbytes, 1, None), "h2d")
    L.check(lib.adrt_plan_backproject(P._plan, P.d_sino.ptr, P.d_out.ptr, 1, None), "backproject")
    y = np.empty((N, N))
    L.check(lib.adrt_memcpy(y.ctypes.data, P.d_out.ptr, y.nbytes, 2, None), "d2h")
    yref = R.adrt_T(b[0])
    assert np.abs(y - yref).max() <= tol * np.abs(yref).max()
