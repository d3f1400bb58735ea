"""Whole inverses on the GPU against the reference's FastLp / FastHybrid outputs at identical setup scalars."""

import numpy as np
import pytest

from conftest import golden, pkg, relmax

pytestmark = pytest.mark.gpu


def _hyb(d, **kw):
    I = pkg("inverses")
    cap = int(d["dense_cap"])
    return I.FastHybridHip(int(d["n"]), lam_max=d["h_lam_max"], dense_cap=None if cap < 0 else cap,
                           outer=(float(d["h_lo"]), float(d["h_hi"]), int(d["h_D"])), **kw)


def _lp(d, **kw):
    I = pkg("inverses")
    cap = int(d["dense_cap"])
    return I.FastLpHip(int(d["n"]), lam_max=d["l_lam_max"], dense_cap=None if cap < 0 else cap,
                       outer=(float(d["l_lo"]), float(d["l_hi"]), int(d["l_D"])), **kw)


@pytest.mark.parametrize("fixture", ["inverse_n3_cap-1", "inverse_n4_cap-1", "inverse_n4_cap8", "inverse_n6_phantom"])
def test_hybrid_vs_reference(gpu, fixture):
    d = golden(fixture)
    H = _hyb(d)
    assert np.array_equal(H.deg, d["h_deg"])
    assert H.eng.n_dense == int(d["h_n_dense"])
    assert relmax(H.Gh(d["x"]), d["h_gh"]) < 1e-12
    assert relmax(H.backproject(d["b"]), d["h_back"]) < 1e-12
    assert relmax(H.P(d["x"]), d["h_P"]) < 1e-12
    assert relmax(H.apply(d["b"]), d["h_apply"]) < 1e-12


@pytest.mark.parametrize("fixture", ["inverse_n3_cap-1", "inverse_n4_cap-1", "inverse_n4_cap8", "inverse_n6_phantom"])
def test_lp_vs_reference(gpu, fixture):
    d = golden(fixture)
    F = _lp(d)
    assert np.array_equal(F.deg, d["l_deg"])
    assert relmax(F.frame_window(d["x"]), d["l_frame"]) < 1e-12
    assert relmax(F.backproject(d["b"]), d["l_back"]) < 1e-12
    assert relmax(F.apply(d["b"]), d["l_apply"]) < 1e-12


def test_batch_equals_single(gpu):
    """Every sinogram of a batch gives the same result as alone (batching must not couple sinograms)."""
    d = golden("inverse_n6_phantom")
    H = _hyb(d, max_batch=4)
    rng = np.random.default_rng(3)
    B = np.stack([d["b"] + 0.01 * rng.standard_normal(d["b"].shape) for _ in range(4)])
    xb = H.apply_batch(B)
    for s in range(4):
        assert relmax(xb[s], H.apply(B[s])) < 1e-14


def test_left_inverse(gpu):
    """L R x = x for a consistent filter (fastops.py:8-9), through the GPU forward ADRT."""
    R = pkg("radon")
    I = pkg("inverses")
    ph = pkg("phantom")
    n = 6
    x = ph.ellipses(5, 1 << n)
    b = R.adrt(x)
    d = golden("inverse_n6_phantom")
    H = I.FastHybridHip(n, lam_max=d["h_lam_max"], outer=(float(d["h_lo"]), float(d["h_hi"]), int(d["h_D"])))
    assert np.linalg.norm(H.apply(b) - x) / np.linalg.norm(x) < 1e-9


@pytest.mark.parametrize("kind", ["hybrid", "lp"])
@pytest.mark.parametrize("n,cap", [(6, 8), (7, 10)])
def test_dense_rows_vs_oracle(gpu, kind, n, cap):
    """Dense low-frequency rows (exact Tikhonov filter, run as a batched GEMM of the pass matrices) against the
    oracle's sequential ops with its own eigh tables."""
    from oracle import inverses as oiv
    from oracle import radon as orad

    I = pkg("inverses")
    ph = pkg("phantom")
    N = 1 << n
    x = ph.ellipses(11, N)
    b = ph.add_noise(orad.adrt(n, x), 0.01, 12)
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(2 * N))
    if kind == "hybrid":
        lam = oiv.power_bound(n, z, 2, streams="numpy")
        R = oiv.Hybrid(n, lam_max=lam, dense_cap=cap, outer=(0.16, 0.96, 30))
        G = I.FastHybridHip(n, lam_max=lam, dense_cap=cap, outer=(0.16, 0.96, 30), max_batch=2)
        assert relmax(G.Gh(x), R.Gh(x)) < 1e-12
    else:
        lam = oiv.power_bound(n, z, 1, streams="numpy")
        R = oiv.Lp(n, lam_max=lam, dense_cap=cap, outer=(1.0, 4.0, 25))
        G = I.FastLpHip(n, lam_max=lam, dense_cap=cap, outer=(1.0, 4.0, 25), max_batch=2)
        assert relmax(G.frame_window(x), R.frame_window(x)) < 1e-12
    assert G.eng.n_dense == R.eng.n_dense > 0
    assert relmax(G.backproject(b), R.backproject(b)) < 1e-12
    assert relmax(G.apply(b), R.apply(b)) < 1e-12
    xb = G.apply_batch(np.stack([b, 2 * b]))
    assert relmax(xb[1], 2 * xb[0]) < 1e-12  # linear in b


def test_lorr_and_none_preconditioners(gpu):
    from oracle import inverses as oiv
    from oracle import radon as orad

    d = golden("inverse_n6_phantom")
    n = 6
    I = pkg("inverses")
    b = orad.adrt(n, d["x"])
    for pc, outer in (("lorr", (0.15, 0.95, 30)), ("none", (0.5, 15.0, 40))):
        R = oiv.Hybrid(n, lam_max=d["h_lam_max"], precond=pc, outer=outer)
        G = I.FastHybridHip(n, lam_max=d["h_lam_max"], precond=pc, outer=outer)
        assert relmax(G.P(d["x"]), R.P(d["x"])) < 1e-13  # fastops applies lorr_inv for any precond but 'g1'
        assert relmax(G.apply(b), R.apply(b)) < 1e-12, pc


def test_fp32_and_refinement(gpu):
    """FP32 row engine against FP64; the FP64 refinement steps pull the FP32 result onto the FP64 one."""
    I = pkg("inverses")
    ph = pkg("phantom")
    R = pkg("radon")
    n = 7
    N = 1 << n
    H = I.FastHybridHip(n, dense_cap=12, max_batch=2)
    outer = (H.lo, H.hi, H.outer_deg)
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(20 + i, N)), 0.01, 30 + i) for i in range(2)])
    x64 = H.apply_batch(B)
    err = {}
    for refine in (0, 3):
        H32 = I.FastHybridHip(n, lam_max=H.lam_max, dense_cap=12, precision="fp32", max_batch=2, outer=outer,
                              refine=refine)
        x32 = H32.apply_batch(B)
        err[refine] = max(np.linalg.norm(x32[i] - x64[i]) / np.linalg.norm(x64[i]) for i in range(2))
    # FP32 alone is ~5e-7 from FP64 at N = 128; three FP64 outer steps take about a decade off that (each contracts
    # the FP32 part of the error by the outer iteration's rate)
    assert err[0] < 1e-5
    assert err[3] < 1e-6 and err[3] < err[0] / 5, err


def test_graph_replay_matches_eager(gpu):
    """A captured apply replays to the eager result, with eager work and syncs between replays."""
    L = gpu
    I = pkg("inverses")
    d = golden("inverse_n6_phantom")
    H = I.FastHybridHip(6, lam_max=d["h_lam_max"], outer=(float(d["h_lo"]), float(d["h_hi"]), int(d["h_D"])),
                        max_batch=3, dense_cap=10)
    B = np.stack([d["b"], 0.5 * d["b"], -d["b"]])
    eager = H.apply_batch(B)
    st = L.Stream()
    db = L.DeviceArray.from_host(B, st)
    dx = L.DeviceArray((3, 64, 64))
    with L.Graph(st) as g:
        H.apply_device(db, dx, 3, st, count=False)
    for rep in range(3):
        dx.zero(st)
        g.launch(st)
        st.sync()
        got = dx.download(st)
        assert relmax(got, eager) < 1e-13, rep
        H.apply(d["b"])  # eager work between replays


def test_fp32_batch_headline_config(gpu):
    """The benchmarked configuration at a smaller batch: N = 256, dense_cap 16, ellipse phantoms + 1% noise, the FP32
    plan (TwoSum rows, one FP64 refinement step) within 1e-5 of the FP64 plan for every sinogram, and a batch equal to
    its sinograms run one at a time (same bits)."""
    I = pkg("inverses")
    ph = pkg("phantom")
    R = pkg("radon")
    n, S, cap = 8, 6, 16
    N = 1 << n
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(1000 + i, N)), 0.01, 2000 + i) for i in range(S)])
    H64 = I.FastHybridHip(n, dense_cap=cap, max_batch=S)
    x64 = H64.apply_batch(B)
    H32 = I.FastHybridHip(n, lam_max=H64.lam_max, dense_cap=cap, precision="fp32", max_batch=S,
                          outer=(H64.lo, H64.hi, H64.outer_deg), refine=1)
    x32 = H32.apply_batch(B)
    for i in range(S):
        assert np.linalg.norm(x32[i] - x64[i]) / np.linalg.norm(x64[i]) < 1e-5, i
    assert np.array_equal(x32[2], H32.apply(B[2]))


@pytest.mark.parametrize("n,big", [(7, 256), (8, 64), (9, 32)])
def test_fp32_paired_inverse_fft(gpu, n, big):
    """FP32 plans whose max_batch gives the paired inverse FFT two blocks per CU take it: it writes U = S0 + S3 (the
    transposed strips, mirrored columns) and V = S1 + S2 (mirrored rows) instead of the four quadrant strips, and the
    combine adds U^T + V.  Against a plan of the same setup at max_batch 4 (four strips): the operator G_h and the
    inverse differ in bits but agree to FP32 rounding (a strip summed at the wrong row or column would not), both are
    within 1e-5 of FP64, and the big plan gives a sinogram the same bits alone and in a batch."""
    def relerr(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    I = pkg("inverses")
    ph = pkg("phantom")
    R = pkg("radon")
    N = 1 << n
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(1100 + i, N)), 0.01, 2100 + i) for i in range(3)])
    x = ph.ellipses(1200, N)
    H64 = I.FastHybridHip(n, dense_cap=16, max_batch=3)
    kw = dict(lam_max=H64.lam_max, dense_cap=16, precision="fp32", outer=(H64.lo, H64.hi, H64.outer_deg), refine=1)
    Hb = I.FastHybridHip(n, max_batch=big, **kw)
    Hs = I.FastHybridHip(n, max_batch=4, **kw)
    g64, gb, gs = H64.Gh(x), Hb.Gh(x), Hs.Gh(x)
    assert not np.array_equal(gb, gs)
    assert relerr(gb, gs) < 1e-6 and relerr(gb, g64) < 1e-5 and relerr(gs, g64) < 1e-5
    x64, xb, xs = H64.apply_batch(B), Hb.apply_batch(B), Hs.apply_batch(B)
    for i in range(3):
        assert relerr(xb[i], x64[i]) < 1e-5 and relerr(xs[i], x64[i]) < 1e-5, i
        assert relerr(xb[i], xs[i]) < 2e-6, i
    assert np.array_equal(Hb.apply(B[1]), xb[1])
    for P in (H64, Hb, Hs):
        P.close()


@pytest.mark.parametrize("prec,n,S", [("fp32", 8, 3), ("fp64", 8, 3), ("fp32", 9, 3), ("fp32", 8, 24),
                                       ("fp64", 8, 12), ("fp32", 7, 9)])
def test_dense_placement_same_bits(gpu, prec, n, S):
    """The dense low-frequency GEMM gives the same bits whichever launch computes it (a launch of its own, or 1, 2, 4
    MFMA waves per SIMD inside the row kernel's launch) and whatever its blocking: every placement runs the same tile
    arithmetic in the same order.  Batch 3 puts FP64 passes on the MFMA path too (6 rows > 4) with one row tile per
    frequency; batches 9, 12 and 24 (18, 24, 48 rows) take the blocked tiles (two row tiles per dense wave inside the
    row launch, 2 x 2 tiles per wave in a launch of its own), with a partly filled last block."""
    I = pkg("inverses")
    ph = pkg("phantom")
    R = pkg("radon")
    lib = gpu.load()
    cap = 12 if n == 7 else 16
    N = 1 << n
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(300 + i, N)), 0.01, 400 + i) for i in range(S)])
    H = I.FastHybridHip(n, dense_cap=cap, precision=prec, max_batch=S, refine=1)
    assert H.eng.n_dense > 0
    p = gpu.CONFIG_PRECISION[prec]
    old = lib.adrt_dense_config_get(n, p)
    try:
        out = {}
        for m in (-1, 1, 2, 4):
            gpu.check(lib.adrt_dense_config_set(n, p, m), "dense_config_set")
            out[m] = H.apply_batch(B)
        for m in (1, 2, 4):
            assert np.array_equal(out[m], out[-1]), m
    finally:
        lib.adrt_dense_config_set(n, p, old)
        H.close()


@pytest.mark.parametrize("prec,n", [("fp32", 10), ("fp64", 8)])
def test_one_sinogram_dense_placement_same_bits(gpu, prec, n):
    """Passes over one sinogram take the dense placement of their own class (precision + 2, tuned apart because the W
    stream dominates there); every placement gives the same bits, equal to the sinogram's result in a batch."""
    I = pkg("inverses")
    C = pkg("plancache")
    ph = pkg("phantom")
    R = pkg("radon")
    lib = gpu.load()
    N = 1 << n
    lam, outer = C.hybrid_setup(n, dense_cap=16)
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(960 + i, N)), 0.01, 970 + i) for i in range(3)])
    H = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=16, precision=prec, max_batch=3, refine=1)
    cls = gpu.CONFIG_PRECISION[prec] + 2
    try:
        xb = H.apply_batch(B)
        for m in (-1, 1, 2, 4):
            gpu.check(lib.adrt_dense_config_set(n, cls, m), "dense_config_set")
            assert np.array_equal(H.apply(B[1]), xb[1]), m
    finally:
        lib.adrt_dense_config_set(n, cls, 0)
        H.close()


@pytest.mark.parametrize("prec,n,S", [("fp64", 8, 20), ("fp32", 8, 20), ("fp64", 10, 17)])
def test_conjugate_dense_pairs_same_bits(gpu, prec, n, S):
    """Dense frequencies k and L - k have conjugate pass matrices; the partner's rows always run as
    conj(conj(y) W_k).  With few rows per frequency (one sinogram) the two share one unit and one stream of W_k,
    with many (20 or 17 sinograms: 40 or 34 rows) each has a unit of its own.  A sinogram gives the same bits either
    way, and the FP64 result stays within 1e-12 of the plan that reads every table (the oracle tests of the FP64
    inverses pin that)."""
    I = pkg("inverses")
    C = pkg("plancache")
    ph = pkg("phantom")
    R = pkg("radon")
    N = 1 << n
    lam, outer = C.hybrid_setup(n, dense_cap=16)
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(900 + i, N)), 0.01, 950 + i) for i in range(S)])
    H = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=16, precision=prec, max_batch=S, refine=1)
    try:
        assert H.eng.n_dense > 1
        x = H.apply_batch(B)
        for i in (0, S - 1):
            assert np.array_equal(x[i], H.apply(B[i])), i
    finally:
        H.close()


@pytest.mark.parametrize("n,loge,S", [(10, 3, 8), (10, 2, 8), (8, 3, 64)])
def test_row_groups_per_block_same_bits(gpu, n, loge, S):
    """Blocks of float rows run several row groups of their frequency on one set-up once the batch gives the launch
    enough wavefronts (up to four groups at N=1024 with 8 sinograms, two at N=256 with 64); one sinogram alone runs one
    group per block.  Every sinogram must give the same bits either way.  refine 0 runs the backprojection in float:
    its op list starts with the adjoint levels over the whole row, which write the whole exchange buffer, so with rows
    of two or four wavefronts (N=1024 with 8 or 4 values per lane) a block's rows need the barrier between them."""
    I = pkg("inverses")
    C = pkg("plancache")
    ph = pkg("phantom")
    R = pkg("radon")
    lib = gpu.load()
    N = 1 << n
    lam, outer = C.hybrid_setup(n, dense_cap=16)
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(700 + i, N)), 0.01, 800 + i) for i in range(S)])
    old = lib.adrt_rows_config_get(n, 1)
    H = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=16, precision="fp32", max_batch=S, refine=0)
    try:
        assert lib.adrt_rows_config_set(n, 1, loge) == 0
        x = H.apply_batch(B)
        for i in (0, S // 2, S - 1):
            assert np.array_equal(x[i], H.apply(B[i])), i
    finally:
        lib.adrt_rows_config_set(n, 1, old)
        H.close()


def test_plan_batch_limits(gpu):
    """A batch above the plan's max_batch is an error, not an overrun; an empty batch is a no-op."""
    L = gpu
    I = pkg("inverses")
    d = golden("inverse_n4_cap-1")
    H = I.FastHybridHip(4, lam_max=d["h_lam_max"], outer=(float(d["h_lo"]), float(d["h_hi"]), int(d["h_D"])),
                        max_batch=2)
    db = L.DeviceArray((3, 4, 31, 16))
    dx = L.DeviceArray((3, 16, 16))
    with pytest.raises(L.HipError, match="max_batch"):
        H.apply_device(db, dx, 3)
    H.apply_device(db, dx, 0)


def test_fp32_with_double_g1(gpu):
    """The FP32 plan with the G_1 inner iteration in double (g1_fp64): within 1e-5 of the FP64 plan, different bits
    from the float-G_1 plan (the option takes effect), and a sinogram alone equal to itself in the batch."""
    I = pkg("inverses")
    ph = pkg("phantom")
    R = pkg("radon")
    n, S, cap = 7, 3, 12
    N = 1 << n
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(70 + i, N)), 0.01, 80 + i) for i in range(S)])
    H64 = I.FastHybridHip(n, dense_cap=cap, max_batch=S)
    x64 = H64.apply_batch(B)
    kw = dict(lam_max=H64.lam_max, dense_cap=cap, precision="fp32", max_batch=S, refine=1,
              outer=(H64.lo, H64.hi, H64.outer_deg))
    Hd = I.FastHybridHip(n, g1_fp64=True, **kw)
    Hf = I.FastHybridHip(n, **kw)
    xd, xf = Hd.apply_batch(B), Hf.apply_batch(B)
    for i in range(S):
        assert np.linalg.norm(xd[i] - x64[i]) / np.linalg.norm(x64[i]) < 1e-5, i
    assert not np.array_equal(xd, xf)
    assert np.array_equal(Hd.apply(B[1]), xd[1])


@pytest.mark.parametrize("n,cap,loge,dense", [(7, 12, 0, 0), (8, 16, 2, -1), (8, 16, 4, 2), (6, 8, 3, 1)])
def test_fp32_with_double_explicit_level(gpu, n, cap, loge, dense):
    """The FP32 plan with the explicit level 1 (and the FFTs and combine around it) in double (l1_fp64): the row
    kernel reads and writes complex128 rows, with levels 2..n and the Chebyshev filter in float.  At refine 0 (where the
    backprojection takes that kernel too) and refine 1: within 1e-5 of the FP64 plan, different bits from the
    float-level plan, a sinogram alone equal to itself in the batch; under the row mappings E = 4, 8, 16 and with the
    dense tiles inside the row launch (1, 2 waves per SIMD) or in a launch of their own (complex128 rows there too)."""
    I = pkg("inverses")
    L = pkg("lib")
    ph = pkg("phantom")
    R = pkg("radon")
    lib = L.load()
    S = 3
    N = 1 << n
    print(f"l1_fp64 n={n} cap={cap} loge={loge} dense={dense}", flush=True)
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(90 + i, N)), 0.01, 95 + i) for i in range(S)])
    H64 = I.FastHybridHip(n, dense_cap=cap, max_batch=S)
    assert H64.eng.n_dense > 0
    x64 = H64.apply_batch(B)
    old = lib.adrt_rows_config_get(n, 1), lib.adrt_dense_config_get(n, 1)
    try:
        assert lib.adrt_rows_config_set(n, 1, loge) == 0 and lib.adrt_dense_config_set(n, 1, dense) == 0
        for refine in (0, 1):
            kw = dict(lam_max=H64.lam_max, dense_cap=cap, precision="fp32", max_batch=S, refine=refine,
                      outer=(H64.lo, H64.hi, H64.outer_deg))
            Hx = I.FastHybridHip(n, l1_fp64=True, **kw)
            Hf = I.FastHybridHip(n, **kw)
            xx, xf = Hx.apply_batch(B), Hf.apply_batch(B)
            for i in range(S):
                e = np.linalg.norm(xx[i] - x64[i]) / np.linalg.norm(x64[i])
                assert e < 1e-5, (refine, i, e)
            assert not np.array_equal(xx, xf)
            assert np.array_equal(Hx.apply(B[1]), xx[1])
            Hx.close()
            Hf.close()
    finally:
        lib.adrt_rows_config_set(n, 1, old[0])
        lib.adrt_dense_config_set(n, 1, old[1])
    H64.close()


def test_double_explicit_level_needs_fp32_hybrid(gpu):
    """l1_fp64 is refused for an FP64 plan and an uncompensated FP32 plan; L_p (no explicit level) has no such
    keyword."""
    I = pkg("inverses")
    L = pkg("lib")
    n = 6
    H = I.FastHybridHip(n, dense_cap=8)
    kw = dict(lam_max=H.lam_max, outer=(H.lo, H.hi, H.outer_deg), l1_fp64=True)
    for extra in (dict(), dict(precision="fp32_plain")):
        with pytest.raises(L.HipError, match="l1_fp64"):
            I.FastHybridHip(n, dense_cap=8, **kw, **extra)
    with pytest.raises(TypeError):
        I.FastLpHip(n, precision="fp32", l1_fp64=True)
    H.close()
