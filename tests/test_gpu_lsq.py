"""CGLS on the GPU (pinv.cgls, pinv.py:23-48) against the oracle's CGLS and the dense pseudoinverse, and as the
cross-check of the hybrid inverse."""

import numpy as np
import pytest

from conftest import pkg, relmax

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [2, 3, 4])
def test_cgls_vs_dense_pinv(gpu, n):
    from oracle import lsq

    N = 1 << n
    b = np.random.default_rng(n).standard_normal((4, 2 * N - 1, N))
    x_ref = (lsq.pinv_dense(n) @ b.ravel()).reshape(N, N)
    x, hist = pkg("lsq").cgls(n, b, tol=1e-14, maxit=2000)
    assert relmax(x, x_ref) < 1e-10
    assert hist[-1] <= 1e-14 * hist[0] or len(hist) == 2001


@pytest.mark.parametrize("n", [5, 6])
def test_cgls_vs_oracle(gpu, n):
    from oracle import lsq

    ph = pkg("phantom")
    R = pkg("radon")
    N = 1 << n
    b = ph.add_noise(R.adrt(ph.ellipses(7 + n, N)), 0.01, 17 + n)
    x_o, h_o = lsq.cgls(n, b, tol=1e-10)
    x_g, h_g = pkg("lsq").cgls(n, b, tol=1e-10)
    assert abs(h_g[0] - h_o[0]) <= 1e-12 * h_o[0]
    assert abs(len(h_g) - len(h_o)) <= 2, (len(h_g), len(h_o))
    assert relmax(x_g, x_o) < 1e-8
    # the history is the reference's ||R^T r_j||: the leading entries agree closely (later ones drift apart by
    # rounding, as CG's do between any two summation orders), and both end below the tolerance
    assert np.allclose(h_g[:8], h_o[:8], rtol=1e-8)
    assert h_g[-1] <= 1e-10 * h_g[0] and h_o[-1] <= 1e-10 * h_o[0]


def test_cgls_batch_equals_single_and_zero(gpu):
    lsq = pkg("lsq")
    ph = pkg("phantom")
    R = pkg("radon")
    n, N = 5, 32
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(40 + i, N)), 0.01, 50 + i) for i in range(3)] +
                 [np.zeros((4, 2 * N - 1, N))])
    X, iters, hists = lsq.cgls_batch(n, B, tol=1e-9, maxit=500)
    for i in range(3):
        x1, h1 = lsq.cgls(n, B[i], tol=1e-9, maxit=500)
        assert np.array_equal(X[i], x1) and len(h1) == iters[i] + 1
    assert iters[3] == 0 and not X[3].any()
    # a warm start from the solution starts with a tiny normal residual and stays at the solution
    X2, it2, h2 = lsq.cgls_batch(n, B[:1], tol=1e-3, maxit=500, X0=X[:1])
    assert h2[0][0] < 1e-8 * hists[0][0] and relmax(X2[0], X[0]) < 1e-7


def test_cgls_cross_checks_hybrid(gpu):
    """Noiseless data: the least-squares solution and the hybrid inverse both return the image."""
    I = pkg("inverses")
    ph = pkg("phantom")
    R = pkg("radon")
    n, N = 6, 64
    x = ph.ellipses(3, N)
    b = R.adrt(x)
    x_ls, _ = pkg("lsq").cgls(n, b, tol=1e-12)
    H = I.FastHybridHip(n, dense_cap=10)
    x_h = H.apply(b)
    assert relmax(x_ls, x) < 1e-8
    assert relmax(x_h, x_ls) < 1e-7
