"""CPU pins of the oracle (tests/golden fixtures, closed forms, invariants, brute force, the reference C++ kernels).

Each check is chosen so that a plausible mistake in the restatement (a dropped term, a wrong sign or index, a
transposed operand) fails it; none of them retypes the formula under test.
"""

import numpy as np
import pytest

from conftest import golden, other_layout, other_package, relmax
from oracle import inverses as oiv
from oracle import layout as olay
from oracle import lsq, radon, refcpp
from oracle import rows as orw


# ------------------------------------------------------------------ forward ADRT (test_core.py:17-45)
@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_transpose_identity(n):
    N = 1 << n
    rng = np.random.default_rng(0)
    x = rng.standard_normal((N, N))
    b = rng.standard_normal((4, 2 * N - 1, N))
    lhs = np.sum(radon.adrt(n, x) * b)
    rhs = np.sum(x * radon.adrt_T(n, b))
    assert abs(lhs - rhs) < 1e-9 * max(1, abs(lhs))


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_digital_lines(n):
    """Row (h, s) of the single-quadrant ADRT is the indicator of {(h + delta_s(t), t)}; rows outside E^n_n vanish."""
    N = 1 << n
    A = radon.dense(lambda x: radon.sq_adrt(n, x), (N, N))
    for h in range(-(N - 1), N):
        for s in range(N):
            row = A[(h + N - 1) * N + s].reshape(N, N)
            expect = np.zeros((N, N))
            if h >= -s:
                for t in range(N):
                    i = h + radon.delta_s(n, s, t)
                    if 0 <= i < N:
                        expect[i, t] = 1
            assert np.array_equal(row, expect), (h, s)


@pytest.mark.parametrize("n", [2, 3, 5])
def test_integer_exact_and_quadrants(n):
    """int64 forward equals float64 forward; every quadrant is the square ADRT of the permuted image."""
    N = 1 << n
    x = np.random.default_rng(n).integers(0, 256, (N, N))
    bi = radon.adrt(n, x, dtype=np.int64)
    assert np.array_equal(bi, radon.adrt(n, x).astype(np.int64))
    # quadrant 2 of an impulse at (i, j): the lines through it are exactly the (h, s) with h + delta_s(j) = i
    for i, j in ((0, 0), (N - 1, 1), (N // 2, N - 1)):
        e = np.zeros((N, N), dtype=np.int64)
        e[i, j] = 1
        q2 = radon.adrt(n, e, dtype=np.int64)[2]
        for s in range(N):
            h = i - radon.delta_s(n, s, j)
            col = np.zeros(2 * N - 1, dtype=np.int64)
            if h >= -s:
                col[h + N - 1] = 1
            assert np.array_equal(q2[:, s], col)
    assert bi.sum() == 4 * N * x.sum()  # every pixel lies on exactly one line per slope and quadrant


def test_mask_full():
    n = 3
    N = 1 << n
    M = radon.mask_full(n)
    x = np.random.default_rng(0).standard_normal((N, N))
    b = radon.adrt(n, x)
    assert np.all(b[~M] == 0)
    assert M.sum() == 4 * (N * N + N * (N - 1) // 2)  # column s holds offsets -s..N-1


# ------------------------------------------------------------------ butterflies (lab_hybrid.py:54-56, exp_kernel_resonance.py)
@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_butterfly_closed_form(n):
    """bfly over levels 1..n is A_k(s, t) = exp(i w_k delta_s(t)) (lab_frame.py:37)."""
    N = 1 << n
    D = radon.delta_table(n)
    for w in (0.0, 0.3, -1.7, np.pi):
        A = orw.level_matrix(n, np.exp(1j * w), 1, n)
        assert np.abs(A - np.exp(1j * w * D)).max() < 1e-12


@pytest.mark.parametrize("lo,hi", [(1, 4), (2, 4), (1, 1), (3, 4), (2, 3)])
def test_adjointness_and_levels(lo, hi):
    n = 4
    N = 1 << n
    rng = np.random.default_rng(lo * 10 + hi)
    z = np.exp(1j * rng.uniform(-np.pi, np.pi, 5))
    u = rng.standard_normal((5, N)) + 1j * rng.standard_normal((5, N))
    v = rng.standard_normal((5, N)) + 1j * rng.standard_normal((5, N))
    lhs = np.sum(np.conj(orw.bfly(n, z, u, lo, hi)) * v, axis=1)
    rhs = np.sum(np.conj(u) * orw.bflyT(n, z, v, lo, hi), axis=1)
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(lhs).max()
    # levels compose: B lo..hi = B (m+1)..hi o B lo..m
    if hi > lo:
        m = lo
        two = orw.bfly(n, z, orw.bfly(n, z, u, lo, m), m + 1, hi)
        assert np.abs(two - orw.bfly(n, z, u, lo, hi)).max() < 1e-12


def test_gram_at_zero_frequency():
    """A_0 is all ones, so lambda_max(M_0) = N^2 (levels 1..n); for levels 2..n the Gram is block-diagonal."""
    n = 4
    N = 1 << n
    A = orw.level_matrix(n, 1.0 + 0j, 1, n)
    assert np.allclose(A, 1.0)
    assert abs(np.linalg.eigvalsh(A.conj().T @ A)[-1] - N * N) < 1e-9


@pytest.mark.parametrize("n,lo", [(4, 1), (4, 2), (5, 2)])
def test_cheb_filter_eigenvector(n, lo):
    """On an eigenvector of M_k the filter acts as the scalar Chebyshev polynomial p_d(lambda), and
    |1 - (lambda + tau) p_d(lambda)| <= 2 rho^d on [0, Lambda]."""
    N = 1 << n
    tau = 0.2 * N
    z = np.exp(1j * 0.4)
    A = orw.level_matrix(n, z, lo, n)
    lam, V = np.linalg.eigh(A.conj().T @ A)
    Lmax = lam[-1] * 1.2
    c, a = orw.cheb_coeffs(np.array([Lmax]), tau)
    d = 9
    Y = V.T.copy()  # rows = eigenvectors
    Z, steps = orw.cheb_filter(n, np.full(N, z), Y, np.full(N, d), np.full(N, a[0]), np.full(N, c[0]), tau, lo, n)
    p = np.sum(np.conj(Y) * Z, axis=1)
    assert np.abs(Z - p[:, None] * Y).max() < 1e-10 * np.abs(Z).max()
    kap = 1 + Lmax / tau
    rho = (np.sqrt(kap) - 1) / (np.sqrt(kap) + 1)
    assert np.all(np.abs(1 - (lam + tau) * p.real) <= 2 * rho ** (d - 1) + 1e-9)
    assert steps == N * (d - 1) * 2 * (n - lo + 1)


def test_dense_filter_residual():
    n, tau = 4, 3.2
    z = np.exp(1j * 0.1)
    A = orw.level_matrix(n, z, 2, n)
    F = orw.dense_filter(n, z, tau, 2, n)
    assert np.abs((A.conj().T @ A + tau * np.eye(1 << n)) @ F - np.eye(1 << n)).max() < 1e-10


@pytest.mark.parametrize("n", [3, 4, 5])
def test_cert_is_upper_bound(n):
    """cert >= the true maximum row sum of |M'| >= lambda_max(M') (exp_cert_lambda_upper.py:27-31)."""
    N = 1 << n
    for lo in (1, 2):
        for w in (0.05, 0.9, 2.5):
            A = orw.level_matrix(n, np.exp(1j * w), lo, n)
            M = A.conj().T @ A
            rowsum = np.abs(M).sum(axis=1).max()
            assert orw.cert_rowsum(n, w, lo) >= rowsum - 1e-9
    assert abs(orw.cert_rowsum(n, 0.0, 1) - N * N) < 1e-9  # at w = 0 the recursion is exact: (2^m)^2


# ------------------------------------------------------------------ level-1 stencils (fasthybrid.py:9-11)
def test_g1_dense_and_spectrum():
    n = 3
    N = 1 << n
    G = radon.dense(oiv.G1_apply, (N, N))
    K = 2 * np.eye(N) + np.eye(N, k=1) + np.eye(N, k=-1)
    J = np.kron(np.eye(N // 2), np.array([[0, 1], [1, 0]]))
    assert np.abs(G - (8 * np.eye(N * N) + np.kron(K, J) + np.kron(J, K))).max() == 0
    ev = np.linalg.eigvalsh(G)
    assert ev[0] >= 4 - 1e-9 and ev[-1] <= 16 + 1e-9
    y = np.random.default_rng(0).standard_normal((N, N))
    assert np.linalg.norm(oiv.G1_apply(oiv.g1_inv(y)) - y) <= 2 * (1 / 3) ** 13 * np.linalg.norm(y) * 1.01
    assert np.abs(4 * oiv.lorr_inv(y) + 8 * oiv.Pi_apply(oiv.lorr_inv(y)) - y).max() < 1e-12


# ------------------------------------------------------------------ golden fixtures from the reference
@pytest.mark.parametrize("fixture", ["rows_n4_cap-1", "rows_n4_cap8", "rows_n5_cap8"])
def test_rows_fixture(fixture):
    d = golden(fixture)
    n = int(d["n"])
    cap = int(d["dense_cap"])
    OPS = {
        "B_1n": [("B", 1, n)], "BT_1n": [("BT", 1, n)], "B_2n": [("B", 2, n)], "BT_2n": [("BT", 2, n)],
        "P_1n": [("P", 1, n)], "P_2n": [("P", 2, n)],
        "gh": [("B", 1, 1), ("P", 2, n), ("B", 2, n), ("BT", 2, n), ("BT", 1, 1)],
        "hyb_back": [("BT", 2, n), ("P", 2, n), ("BT", 1, 1)],
        "lp_frame": [("P", 1, n), ("B", 1, n), ("BT", 1, n)], "lp_back": [("BT", 1, n), ("P", 1, n)],
    }
    for p in (1, 2):
        for name, ops in OPS.items():
            e = orw.Rows(n, d["z"], d["deg"], d["lam_max"], float(d["tau"]), dense_cap=None if cap < 0 else cap,
                         p_levels=(p, None))
            Z = e.run(d["Y"], ops)
            assert relmax(Z, d[f"{name}_p{p}"]) < 1e-12, name
            assert e.count[0] == int(d[f"{name}_p{p}_count"]), name


@pytest.mark.parametrize("fixture", ["inverse_n3_cap-1", "inverse_n4_cap-1", "inverse_n4_cap8", "inverse_n6_phantom"])
def test_inverse_fixture(fixture):
    d = golden(fixture)
    n = int(d["n"])
    cap = int(d["dense_cap"])
    cap = None if cap < 0 else cap
    H = oiv.Hybrid(n, lam_max=d["h_lam_max"], dense_cap=cap, outer=(d["h_lo"], d["h_hi"], d["h_D"]))
    assert np.array_equal(H.deg, d["h_deg"])
    assert relmax(H.Gh(d["x"]), d["h_gh"]) < 1e-12
    assert relmax(H.backproject(d["b"]), d["h_back"]) < 1e-12
    assert relmax(H.apply(d["b"]), d["h_apply"]) < 1e-12
    F = oiv.Lp(n, lam_max=d["l_lam_max"], dense_cap=cap, outer=(d["l_lo"], d["l_hi"], d["l_D"]))
    assert relmax(F.frame_window(d["x"]), d["l_frame"]) < 1e-12
    assert relmax(F.apply(d["b"]), d["l_apply"]) < 1e-12
    # the numpy power iteration of fastops (one default_rng(0) for both parts) reproduces the stored bounds
    if n <= 4:
        z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(2 << n))
        lam, _ = oiv.degrees(oiv.power_bound(n, z, 2, streams="numpy"), 0.2 * (1 << n), 0.1)
        assert relmax(lam, d["h_lam_max"]) < 1e-12


@pytest.mark.parametrize("n", [3, 4])
def test_left_inverse_invariant(n):
    """L R x = x for any consistent filter (fastops.py:8-9)."""
    N = 1 << n
    x = np.random.default_rng(7).standard_normal((N, N))
    b = radon.adrt(n, x)
    H = oiv.Hybrid(n)
    assert np.linalg.norm(H.apply(b) - x) / np.linalg.norm(x) < 1e-9
    F = oiv.Lp(n)
    assert np.linalg.norm(F.apply(b) - x) / np.linalg.norm(x) < 1e-9


def test_packing_equals_per_quadrant():
    """The packed pass (two quadrants per complex row) equals the sum of the four per-quadrant real passes."""
    n = 3
    N, L = 1 << n, 2 << n
    d = golden("inverse_n3_cap-1")
    H = oiv.Hybrid(n, lam_max=d["h_lam_max"], outer=(d["h_lo"], d["h_hi"], d["h_D"]))
    x = d["x"]
    X = np.zeros((L, L))
    X[:N, :N] = x
    out = np.zeros((L, L))
    for q in range(4):
        S = np.fft.fft(oiv.Tq(N, q, X), axis=0)
        Z = H.eng.run(np.vstack([S, np.zeros_like(S)]), H.ops_gh())[:L]
        out += oiv.TqT(N, q, np.real(np.fft.ifft(Z, axis=0)))
    assert relmax(out[:N, :N], H.Gh(x)) < 1e-12


def test_cgls_vs_dense_pinv():
    """CGLS converges to the dense Moore-Penrose solution (pinv.py:91-104)."""
    for n in (1, 2, 3):
        N = 1 << n
        b = np.random.default_rng(n).standard_normal((4, 2 * N - 1, N))
        xd = (lsq.pinv_dense(n) @ b.ravel()).reshape(N, N)
        x, _ = lsq.cgls(n, b, tol=1e-14)
        assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)


# ------------------------------------------------------------------ the reference C++ kernels (oracle/_ref), when built
needs_ref = pytest.mark.skipif(not refcpp.available(), reason="oracle/_ref not built (make -C oracle ref)")


@needs_ref
@pytest.mark.parametrize("lo,hi,tr", [(1, 6, 0), (1, 6, 1), (2, 6, 0), (2, 6, 1), (3, 5, 0), (1, 1, 1)])
def test_reference_bfly(lo, hi, tr):
    n, N, L = 6, 64, 128
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(L))
    V = np.random.default_rng(1).standard_normal((L, N)) * (1 + 1j)
    a = refcpp.bfly_rows(n, z, V, lo, hi, tr)
    b = (orw.bflyT if tr else orw.bfly)(n, z, V, lo, hi)
    assert relmax(a, b) < 1e-13


@needs_ref
def test_reference_cheb_and_cert():
    n, N, L = 6, 64, 128
    rng = np.random.default_rng(2)
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(L))
    deg = rng.integers(1, 30, L)
    lam = np.abs(rng.standard_normal(L)) * N * N
    c, a = orw.cheb_coeffs(lam, 0.2 * N)
    V = rng.standard_normal((L, N)) + 1j * rng.standard_normal((L, N))
    Z1, s1 = refcpp.cheb_rows(n, z, deg, a, c, 0.2 * N, V, 2, n)
    Z2, s2 = orw.cheb_filter(n, z, V, deg, a, c, 0.2 * N, 2, n)
    assert relmax(Z1, Z2) < 1e-13 and s1 == s2
    om = 2 * np.pi * np.fft.fftfreq(L)
    assert np.array_equal(refcpp.cert_rowsum(n, om, 2), oiv.cert_bound(n, om, 2))


# ---- normal-equation Chebyshev (patches/adrt-gram-freq.patch:474-545) and its Gram


@pytest.mark.parametrize("n", [2, 3, 4])
def test_lstsq_cheb_matches_dense_lstsq(n):
    """The patch's own check (tests/test_lstsq_cheb.py, patch:694-701): against numpy's least squares on the dense ADRT
    with an interval 1% outside the extreme eigenvalues of R^T R."""
    N = 1 << n
    rng = np.random.default_rng(N)
    A = radon.dense(lambda x: radon.adrt(n, x), (N, N))
    eig = np.linalg.eigvalsh(A.T @ A)
    b = radon.adrt(n, rng.standard_normal((N, N))) + 0.1 * rng.standard_normal((4, 2 * N - 1, N))
    ref = np.linalg.lstsq(A, b.ravel(), rcond=None)[0].reshape(N, N)
    out = lsq.lstsq_cheb(n, b, lam_min=0.99 * eig[0], lam_max=1.01 * eig[-1], tol=1e-10)
    assert np.linalg.norm(out - ref) <= 1e-8 * np.linalg.norm(ref)


def test_lstsq_cheb_estimated_lam_max():
    """patch:704-709: lam_min = 15 (below lambda_min = 15.47 at N = 8) and the power estimate recover R^-1 R x = x."""
    n = 3
    x = np.random.default_rng(3).standard_normal((8, 8))
    out = lsq.lstsq_cheb(n, radon.adrt(n, x), lam_min=15.0)
    assert np.linalg.norm(out - x) <= 1e-8 * np.linalg.norm(x)
    with pytest.raises(ValueError):
        lsq.lstsq_cheb(n, radon.adrt(n, np.zeros((8, 8))), lam_min=10.0, lam_max=5.0)


@pytest.mark.parametrize("n", [0, 1, 2, 3])
def test_gram_dense_and_power_bound(n):
    """R^T R a equals the dense A^T A a; it is symmetric positive; the power estimate bounds lambda_max from above to
    within its residual; the degree is the smallest D with 2 rho^D <= tol."""
    N = 1 << n
    A = radon.dense(lambda x: radon.adrt(n, x), (N, N))
    rng = np.random.default_rng(n)
    u, v = rng.standard_normal((2, N, N))
    assert np.allclose(lsq.gram(n, u).ravel(), A.T @ (A @ u.ravel()), rtol=0, atol=1e-12 * np.abs(A).sum() ** 2)
    assert np.isclose(np.vdot(lsq.gram(n, u), v), np.vdot(u, lsq.gram(n, v)))
    assert np.vdot(lsq.gram(n, u), u) > 0
    lam = np.linalg.eigvalsh(A.T @ A)[-1]
    est = lsq.power_lam_max(n)
    assert lam <= est * (1 + 1e-12) and est <= lam * (1 + 1e-3)
    lo, hi, tol = 1.0, 50.0, 1e-10
    D = lsq.cheb_degree(lo, hi, tol)
    rho = (np.sqrt(hi / lo) - 1) / (np.sqrt(hi / lo) + 1)
    assert 2 * rho**D <= tol < 2 * rho ** (D - 1)


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_layout_maps_recover_a_known_permutation(n):
    """The reference's fingerprint maps (skimage_adrt_pkg.py:17-51), given a package whose layout is a known
    permutation of the core's, move fresh transforms of new images between the layouts exactly and pair every entry
    that a line reaches once."""
    N = 1 << n
    core = lambda x: radon.adrt(n, x)
    pk = other_package(n, core)
    maps = olay.layout_maps(n, core, pk)
    ci, pi = maps
    # integer images: every line sum is exact, so entries that sum one line along different quadrant paths are equal
    # and the maps (which pair such duplicates by index, not by position) must move the data exactly
    x = np.random.default_rng(11).integers(-1000, 1000, (N, N)).astype(float)
    bc, bp = core(x), pk(x)
    assert np.array_equal(olay.pkg_to_core(bp, n, maps), bc)
    # the package's entries are a permutation of the core's, so core_to_pkg restores every entry, zeros included
    assert np.array_equal(olay.core_to_pkg(bc, n, maps), bp)
    # float images: duplicates differ only by the order of their additions
    x = np.random.default_rng(12).standard_normal((N, N))
    assert relmax(olay.pkg_to_core(pk(x), n, maps), core(x)) < 1e-14
    reached = np.flatnonzero(core(np.ones((N, N))).ravel())
    assert np.array_equal(np.sort(ci), reached) and len(set(pi.tolist())) == len(pi)
    assert np.array_equal(np.sort(pi), np.sort(np.argsort(other_layout(n))[reached]))


# ---- setup functions: spectral bounds (fastcpp.py:102-134) and the outer interval (fastops.py:222-245)
def _lam_max_dense(n, z, lo):
    return np.array([np.linalg.eigvalsh((lambda A: A.conj().T @ A)(orw.level_matrix(n, zz, lo, n)))[-1] for zz in z])


@pytest.mark.parametrize("n", [3, 4, 5])
def test_power_bound_brackets_lambda_max(n):
    """The power bound with fastcpp's streams (real part rng(0), imaginary part rng(1)) is 1.5 x a Rayleigh quotient
    after 40 normalised steps: Rayleigh <= lambda_max(M_k) always (a Hermitian Gram), within 3% of it after 40 steps
    at N <= 32 (median under 1%; the slowest frequencies have close top eigenvalues), so 1.5 x Rayleigh >=
    lambda_max: an upper bound
    in practice.  A dropped normalisation, a wrong step count or safety factor fails one side."""
    N = 1 << n
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(2 * N))
    for lo in (1, 2):
        lam = _lam_max_dense(n, z, lo)
        pb = oiv.power_bound(n, z, lo, streams="cpp")
        ray = pb / 1.5
        assert np.all(ray <= lam * (1 + 1e-12)), lo
        assert np.all(ray >= lam * 0.97), lo
        assert np.median((lam - ray) / lam) < 1e-2, lo
        assert np.all(pb >= lam), lo


@pytest.mark.parametrize("n", [3, 4, 5])
def test_explicit_and_cert_bounds_are_upper_bounds(n):
    """explicit min(N^2, 2N(1+n)/|w|) >= lambda_max(M_k) over levels 1..n (equality at w = 0, where M_0 = all-ones
    Gram with lambda_max = N^2); the row-sum certificate >= lambda_max over levels m_lo..n."""
    N = 1 << n
    oms = 2 * np.pi * np.fft.fftfreq(2 * N)
    z = np.exp(1j * oms)
    lam1 = _lam_max_dense(n, z, 1)
    ex = oiv.explicit_bound(n, oms)
    assert np.all(ex >= lam1 * (1 - 1e-12))
    assert abs(ex[0] - N * N) == 0 and abs(lam1[0] - N * N) < 1e-9 * N * N
    for lo in (1, 2):
        lam = lam1 if lo == 1 else _lam_max_dense(n, z, 2)
        assert np.all(oiv.cert_bound(n, oms, lo) >= lam * (1 - 1e-12)), lo


@needs_ref
@pytest.mark.parametrize("n", [4, 6])
def test_power_bound_on_reference_kernels(n):
    """The power iteration over the oracle's butterflies equals the same iteration over the reference's own C++
    kernels (bfly_rows of cpp/oblique.cpp, the Gram fastcpp.power_lam_max applies) to rounding."""
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(2 << n))
    for lo in (1, 2):
        assert relmax(refcpp.power_bound(n, z, lo), oiv.power_bound(n, z, lo, streams="cpp")) < 1e-12


@pytest.mark.parametrize("fixture", ["inverse_n3_cap-1", "inverse_n4_cap-1", "inverse_n4_cap8", "inverse_n6_phantom"])
def test_outer_interval_matches_reference_fixture(fixture):
    """eigsh (SA / LA, tol 1e-3), the Weinstein residual enclosure, the 0.99 / 1.01 margins and the degree formula
    reproduce the reference's own lo, hi and D (FastHybrid with the G_1-generalised problem, FastLp plain).  ARPACK
    draws its start vector from an internal generator that advances with every call in a process, so only the first
    eigsh of a fresh process repeats the reference's digits; later calls land within eigsh's tol 1e-3 (lo, after the
    residual subtraction, within 4e-3; hi within 5e-5 in 12 orders tried).  The bounds are 8e-3 / 1e-3: a lost margin
    (0.99 or 1.01) is 1% and fails."""
    d = golden(fixture)
    n = int(d["n"])
    cap = int(d["dense_cap"])
    cap = None if cap < 0 else cap
    H = oiv.Hybrid(n, lam_max=d["h_lam_max"], dense_cap=cap)
    assert abs(H.lo / d["h_lo"] - 1) < 8e-3 and abs(H.hi / d["h_hi"] - 1) < 1e-3, (H.lo, H.hi)
    assert abs(H.outer_deg - int(d["h_D"])) <= 1
    F = oiv.Lp(n, lam_max=d["l_lam_max"], dense_cap=cap)
    assert abs(F.lo / d["l_lo"] - 1) < 8e-3 and abs(F.hi / d["l_hi"] - 1) < 1e-3, (F.lo, F.hi)
    assert abs(F.outer_deg - int(d["l_D"])) <= 1
