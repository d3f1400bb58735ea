"""Setup functions and drop-in paths on the GPU.

* The spectral bounds (fastcpp.py:102-134) and their dispatch by ``lam_bound`` (fastcpp.py:137-169) against the oracle
  and the reference's own C++ kernels; the outer interval (fastops.py:222-245) against the reference's stored lo / hi /
  D.
* The layer-1 symbols bound by name the way fastcpp._load does (fastcpp.py:19-32) under the oracle's host, and
  ``HipEngine`` as the ``engine_cls`` of the oracle's inverses (fastops.py:177, 193-203), against the reference's
  fixtures.
* The row-kernel timing ring, the own-queue stream cap and the CGLS callback (pinv.py:23, 43-44).
"""

import ctypes

import numpy as np
import pytest

from conftest import golden, pkg, relmax

pytestmark = pytest.mark.gpu


def _zo(n):
    oms = 2 * np.pi * np.fft.fftfreq(2 << n)
    return np.exp(1j * oms), oms


# ---------------------------------------------------------------- spectral bounds
@pytest.mark.parametrize("n", [4, 6, 7])
def test_power_lam_max_vs_oracle(gpu, n):
    """engine.power_lam_max (GPU Gram, rng(0) / rng(1) start vectors, 40 normalised steps, x 1.5) equals the oracle's
    power bound with fastcpp's streams; the oracle is pinned against dense eigenvalues and the reference kernels."""
    from oracle import inverses as oiv

    E = pkg("engine")
    z, _ = _zo(n)
    for lo in (1, 2):
        assert relmax(E.power_lam_max(n, z, lo), oiv.power_bound(n, z, lo, streams="cpp")) < 1e-12, lo


@pytest.mark.parametrize("n", [3, 5, 8])
def test_cert_lam_max_vs_reference(gpu, n):
    """The GPU row-sum certificate equals the oracle's recursion and, where built, the reference's cert_rowsum."""
    from oracle import inverses as oiv
    from oracle import refcpp

    E = pkg("engine")
    _, oms = _zo(n)
    for lo in (1, 2, n):
        g = E.cert_lam_max(n, oms, lo)
        assert relmax(g, oiv.cert_bound(n, oms, lo)) < 1e-13, lo
        if refcpp.available():
            assert relmax(g, refcpp.cert_rowsum(n, oms, lo)) < 1e-13, lo
    # above the top level no level runs: the reference returns the initial 1
    assert np.array_equal(E.cert_lam_max(n, oms, n + 1), np.ones(oms.size))


@pytest.mark.parametrize("bound", ["explicit", "cert", "power"])
def test_lp_lam_bound_dispatch(gpu, bound):
    """FastLpHip(lam_bound=...) takes the bound FastLpCpp takes (fastcpp.py:147-152) and gives the oracle's inverse on
    the same bound and outer interval."""
    from oracle import inverses as oiv

    I = pkg("inverses")
    n = 5
    z, oms = _zo(n)
    raw = {"explicit": lambda: oiv.explicit_bound(n, oms), "cert": lambda: oiv.cert_bound(n, oms, 1),
           "power": lambda: oiv.power_bound(n, z, 1, streams="cpp")}[bound]()
    F = I.FastLpHip(n, lam_bound=bound, dense_cap=8)
    lam, deg = oiv.degrees(raw, F.tau, F.eps_in)
    assert relmax(F.lam_max, lam) < 1e-12 and np.array_equal(F.deg, deg)
    R = oiv.Lp(n, lam_max=raw, dense_cap=8, outer=(F.lo, F.hi, F.outer_deg))
    b = np.random.default_rng(4).standard_normal((4, 2 * (1 << n) - 1, 1 << n))
    assert relmax(F.apply(b), R.apply(b)) < 1e-12
    F.close()


@pytest.mark.parametrize("bound", ["cert", "power"])
def test_hybrid_lam_bound_dispatch(gpu, bound):
    """FastHybridHip takes 'power' and 'cert' over levels 2..n (fastcpp.py:164-168); like FastHybridCpp it has no
    'explicit' bound (a KeyError there too)."""
    from oracle import inverses as oiv

    I = pkg("inverses")
    n = 5
    z, oms = _zo(n)
    raw = oiv.cert_bound(n, oms, 2) if bound == "cert" else oiv.power_bound(n, z, 2, streams="cpp")
    H = I.FastHybridHip(n, lam_bound=bound, dense_cap=8)
    lam, deg = oiv.degrees(raw, H.tau, H.eps_in)
    assert relmax(H.lam_max, lam) < 1e-12 and np.array_equal(H.deg, deg)
    R = oiv.Hybrid(n, lam_max=raw, dense_cap=8, outer=(H.lo, H.hi, H.outer_deg))
    b = np.random.default_rng(5).standard_normal((4, 2 * (1 << n) - 1, 1 << n))
    assert relmax(H.apply(b), R.apply(b)) < 1e-12
    H.close()
    with pytest.raises(KeyError):
        I.FastHybridHip(n, lam_bound="explicit")


@pytest.mark.parametrize("fixture", ["inverse_n4_cap-1", "inverse_n4_cap8", "inverse_n6_phantom"])
def test_outer_interval_vs_reference_fixture(gpu, fixture):
    """The product's own outer interval (eigsh on the GPU operator, Weinstein enclosure, margins, degree) lands on the
    reference's stored lo / hi / D, within what eigsh's tolerance and ARPACK's per-process start vectors leave (the
    oracle's pin of the same computation gives the bounds: lo 8e-3, hi 1e-3, D +-1)."""
    I = pkg("inverses")
    d = golden(fixture)
    n = int(d["n"])
    cap = int(d["dense_cap"])
    cap = None if cap < 0 else cap
    H = I.FastHybridHip(n, lam_max=d["h_lam_max"], dense_cap=cap)
    assert abs(H.lo / d["h_lo"] - 1) < 8e-3 and abs(H.hi / d["h_hi"] - 1) < 1e-3, (H.lo, H.hi)
    assert abs(H.outer_deg - int(d["h_D"])) <= 1
    F = I.FastLpHip(n, lam_max=d["l_lam_max"], dense_cap=cap)
    assert abs(F.lo / d["l_lo"] - 1) < 8e-3 and abs(F.hi / d["l_hi"] - 1) < 1e-3, (F.lo, F.hi)
    assert abs(F.outer_deg - int(d["l_D"])) <= 1
    H.close()
    F.close()


# ---------------------------------------------------------------- drop-in paths
def _bind_like_fastcpp(path):
    """ctypes.CDLL + the argtypes fastcpp._load declares (fastcpp.py:22-31), on another library file."""
    lib = ctypes.CDLL(path)
    P, I = ctypes.c_void_p, ctypes.c_int
    lib.bfly_rows.argtypes = [I, I, P, P, I, I, I, I]
    lib.bfly_rows.restype = None
    lib.cheb_rows.argtypes = [I, I, P, P, P, P, ctypes.c_double, P, P, I, I, I, P]
    lib.cheb_rows.restype = None
    lib.cert_rowsum.argtypes = [I, I, P, I, P, I]
    lib.cert_rowsum.restype = None
    lib.omp_max_threads.restype = I
    return lib


@pytest.fixture
def hip_as_reference_lib(gpu, monkeypatch):
    """oracle.refcpp (the fastcpp.CppEngine binding) pointed at libadrt_hip.so instead of the reference build."""
    from oracle import refcpp

    lib = _bind_like_fastcpp(gpu.LIB_PATH)
    monkeypatch.setattr(refcpp, "_lib", lib)
    monkeypatch.setattr(refcpp, "available", lambda: True)
    return refcpp


@pytest.mark.parametrize("fixture", ["inverse_n4_cap8", "inverse_n6_phantom"])
def test_layer1_symbols_under_reference_host(hip_as_reference_lib, fixture):
    """The four symbols, bound by name with fastcpp's argtypes, run the reference engine's calls (B / BT in place,
    dense rows by matmul, the rest by cheb_rows with a steps counter) under the oracle's host and reproduce the
    reference's inverses."""
    from oracle import inverses as oiv

    refcpp = hip_as_reference_lib
    d = golden(fixture)
    n = int(d["n"])
    cap = int(d["dense_cap"])
    cap = None if cap < 0 else cap

    class H(oiv.Hybrid):
        engine_cls = refcpp.RefRows

    class F(oiv.Lp):
        engine_cls = refcpp.RefRows

    h = H(n, lam_max=d["h_lam_max"], dense_cap=cap, threads=4, outer=(d["h_lo"], d["h_hi"], d["h_D"]))
    assert relmax(h.apply(d["b"]), d["h_apply"]) < 1e-12
    assert relmax(h.Gh(d["x"]), d["h_gh"]) < 1e-12
    f = F(n, lam_max=d["l_lam_max"], dense_cap=cap, threads=4, outer=(d["l_lo"], d["l_hi"], d["l_D"]))
    assert relmax(f.apply(d["b"]), d["l_apply"]) < 1e-12
    assert refcpp.lib().omp_max_threads() > 0


def test_layer1_row_engine_fixture(hip_as_reference_lib):
    """RowEngine.run fixtures (every op list, both level ranges) through the fastcpp-style binding, with the
    reference's level-row counts."""
    refcpp = hip_as_reference_lib
    d = golden("rows_n5_cap8")
    n = int(d["n"])
    ops = {"gh": [("B", 1, 1), ("P", 2, n), ("B", 2, n), ("BT", 2, n), ("BT", 1, 1)],
           "hyb_back": [("BT", 2, n), ("P", 2, n), ("BT", 1, 1)],
           "lp_frame": [("P", 1, n), ("B", 1, n), ("BT", 1, n)], "lp_back": [("BT", 1, n), ("P", 1, n)]}
    for p in (1, 2):
        for name, o in ops.items():
            e = refcpp.RefRows(n, d["z"], d["deg"], d["lam_max"], float(d["tau"]), threads=2,
                               dense_cap=int(d["dense_cap"]), p_levels=(p, None))
            assert relmax(e.run(d["Y"], o), d[f"{name}_p{p}"]) < 1e-12, (name, p)
            assert e.count[0] == int(d[f"{name}_p{p}_count"]), (name, p)


@pytest.mark.parametrize("fixture", ["inverse_n4_cap-1", "inverse_n4_cap8", "inverse_n6_phantom"])
def test_hip_engine_as_engine_cls(gpu, fixture):
    """HipEngine assigned as engine_cls of the oracle's FastHybrid / FastLp restatements (the reference's plug-in
    point, fastops.py:177) reproduces the reference's outputs and counts."""
    from oracle import inverses as oiv

    E = pkg("engine")
    d = golden(fixture)
    n = int(d["n"])
    cap = int(d["dense_cap"])
    cap = None if cap < 0 else cap

    class H(oiv.Hybrid):
        engine_cls = E.HipEngine

    class F(oiv.Lp):
        engine_cls = E.HipEngine

    h = H(n, lam_max=d["h_lam_max"], dense_cap=cap, outer=(d["h_lo"], d["h_hi"], d["h_D"]))
    assert h.eng.n_dense == int(d["h_n_dense"])
    assert relmax(h.apply(d["b"]), d["h_apply"]) < 1e-12
    assert relmax(h.backproject(d["b"]), d["h_back"]) < 1e-12
    f = F(n, lam_max=d["l_lam_max"], dense_cap=cap, outer=(d["l_lo"], d["l_hi"], d["l_D"]))
    assert relmax(f.apply(d["b"]), d["l_apply"]) < 1e-12
    assert relmax(f.frame_window(d["x"]), d["l_frame"]) < 1e-12
    # the level-row counter advances like the reference engine's (one backprojection + D - 1 operator passes)
    R = oiv.Hybrid(n, lam_max=d["h_lam_max"], dense_cap=cap, outer=(d["h_lo"], d["h_hi"], d["h_D"]))
    h.eng.count = [0]
    R.eng.count = [0]
    h.apply(d["b"])
    R.apply(d["b"])
    assert h.eng.count[0] == R.eng.count[0]


# ---------------------------------------------------------------- timing ring, own queues, CGLS callback
@pytest.mark.parametrize("prec,refine", [("fp32", 1), ("fp64", 0), ("fp32", 50)])
def test_timing_ring_same_bits_and_one_launch_per_pass(gpu, prec, refine):
    """With every row-kernel launch stamping an event pair, an apply gives the bits it gives without; it stamps
    exactly one launch per pass (D); the FP64 refinement count is min(refine, D - 1)."""
    L = gpu
    lib = L.load()
    I = pkg("inverses")
    d = golden("inverse_n6_phantom")
    H = I.FastHybridHip(6, lam_max=d["h_lam_max"], outer=(float(d["h_lo"]), float(d["h_hi"]), int(d["h_D"])),
                        dense_cap=10, precision=prec, refine=refine, max_batch=2)
    D = H.outer_deg
    B = np.stack([d["b"], -0.5 * d["b"]])
    want = H.apply_batch(B)
    evs = [(L.Event(), L.Event()) for _ in range(D)]
    b0 = (L.P * D)(*[a.handle for a, _ in evs])
    b1 = (L.P * D)(*[z.handle for _, z in evs])
    L.check(lib.adrt_plan_time_rows_ring(H._plan, b0, b1, D), "ring")
    try:
        got = H.apply_batch(B)
        stamped = L.LL()
        L.check(lib.adrt_plan_time_rows_count(H._plan, ctypes.byref(stamped)), "count")
    finally:
        L.check(lib.adrt_plan_time_rows_ring(H._plan, None, None, 0), "ring off")
    assert np.array_equal(got, want)
    assert stamped.value == D
    assert all(a.elapsed_ms(z) > 0 for a, z in evs)
    assert lib.adrt_plan_refine_steps(H._plan) == (min(refine, D - 1) if prec == "fp32" else 0)
    H.close()


def test_own_queue_stream_cap(gpu):
    """At most ADRT_MAX_OWN_QUEUES own-queue streams live at once: one more fails with -4 and creates nothing;
    destroying one frees its place."""
    import gc

    L = gpu
    lib = L.load()
    gc.collect()  # streams of earlier tests' servers that are only waiting for collection
    c0 = lib.adrt_own_queue_count()
    assert 0 <= c0 <= L.MAX_OWN_QUEUES
    made = [L.Stream(own_queue=True) for _ in range(L.MAX_OWN_QUEUES - c0)]
    assert lib.adrt_own_queue_count() == L.MAX_OWN_QUEUES
    h = L.P()
    assert lib.adrt_stream_create_own_queue(ctypes.byref(h)) == -4 and h.value is None
    assert "own-queue" in lib.adrt_last_error().decode()
    if made:
        made[0].close()
        assert lib.adrt_own_queue_count() == L.MAX_OWN_QUEUES - 1
        made[0] = L.Stream(own_queue=True)
    for s in made:
        s.sync()
        s.close()
    assert lib.adrt_own_queue_count() == c0


def test_cgls_callback_matches_reference_semantics(gpu):
    """cgls(callback=...) calls back after every iteration, the stopping one included, with (it, x, ||R^T r||) as
    pinv.py:43-44 does: the same iterations and values as the oracle's CGLS, and the same result as without one."""
    from oracle import lsq as olsq

    lsq = pkg("lsq")
    ph = pkg("phantom")
    R = pkg("radon")
    n, N = 5, 32
    b = ph.add_noise(R.adrt(ph.ellipses(61, N)), 0.01, 62)
    seen, ref = [], []
    x, hist = lsq.cgls(n, b, tol=1e-8, callback=lambda it, xx, g: seen.append((it, xx, g)))
    olsq.cgls(n, b, tol=1e-8, callback=lambda it, xx, g: ref.append((it, xx.copy(), g)))
    x0, hist0 = lsq.cgls(n, b, tol=1e-8)
    assert np.array_equal(x, x0) and np.array_equal(hist, hist0)
    assert [s[0] for s in seen] == list(range(len(hist) - 1))
    assert np.array_equal(np.array([s[2] for s in seen]), hist[1:])
    assert np.array_equal(seen[-1][1], x)
    assert abs(len(seen) - len(ref)) <= 2
    for (i, xg, gg), (j, xo, go) in zip(seen[:6], ref[:6]):
        assert i == j and abs(gg / go - 1) < 1e-8 and relmax(xg, xo) < 1e-8

    def stop(it, xx, g):
        raise StopIteration

    with pytest.raises(StopIteration):  # an exception in the callback reaches the caller
        lsq.cgls(n, b, tol=1e-8, callback=stop)
