"""Row kernels on the GPU against the oracle and the reference's RowEngine fixtures (fastops.py:156-167)."""

import ctypes

import numpy as np
import pytest

from conftest import golden, pkg, relmax
from oracle import rows as orw

pytestmark = pytest.mark.gpu

OPS = lambda n: {
    "B_1n": [("B", 1, n)], "BT_1n": [("BT", 1, n)], "B_2n": [("B", 2, n)], "BT_2n": [("BT", 2, n)],
    "B_11": [("B", 1, 1)], "BT_11": [("BT", 1, 1)], "P_1n": [("P", 1, n)], "P_2n": [("P", 2, n)],
    "gh": [("B", 1, 1), ("P", 2, n), ("B", 2, n), ("BT", 2, n), ("BT", 1, 1)],
    "hyb_back": [("BT", 2, n), ("P", 2, n), ("BT", 1, 1)],
    "lp_frame": [("P", 1, n), ("B", 1, n), ("BT", 1, n)],
    "lp_back": [("BT", 1, n), ("P", 1, n)],
}


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 8, 9, 11])
@pytest.mark.parametrize("levels", ["full", "upper", "one", "mid"])
@pytest.mark.parametrize("transpose", [0, 1])
def test_bfly_rows_vs_oracle(gpu, n, levels, transpose):
    lo, hi = {"full": (1, n), "upper": (min(2, n), n), "one": (1, 1), "mid": (min(3, n), max(min(3, n), n - 1))}[levels]
    N = 1 << n
    K = 64 if n < 11 else 8
    rng = np.random.default_rng(n * 10 + transpose)
    z = np.exp(1j * rng.uniform(-np.pi, np.pi, K))
    X = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    lib = gpu.load()
    Xg = X.copy()
    lib.bfly_rows(n, K, gpu.ptr(z), gpu.ptr(Xg), lo, hi, transpose, 1)
    ref = (orw.bflyT if transpose else orw.bfly)(n, z, X, lo, hi)
    assert relmax(Xg, ref) < 1e-13


@pytest.mark.parametrize("n,lo", [(4, 1), (6, 2), (8, 2), (10, 1)])
def test_cheb_rows_vs_oracle(gpu, n, lo):
    N = 1 << n
    K = 96
    rng = np.random.default_rng(n)
    z = np.exp(1j * 2 * np.pi * rng.integers(0, 2 * N, K) / (2 * N))
    deg = rng.integers(1, 40, K).astype(np.int32)
    lam = np.abs(rng.standard_normal(K)) * N * N * 0.3 + 1.0
    tau = 0.2 * N
    c, a = orw.cheb_coeffs(lam, tau)
    Y = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    Z = np.empty_like(Y)
    steps = ctypes.c_longlong(5)
    gpu.load().cheb_rows(n, K, gpu.ptr(z), gpu.ptr(deg), gpu.ptr(a), gpu.ptr(c), tau, gpu.ptr(Y), gpu.ptr(Z), lo, n, 1,
                         ctypes.byref(steps))
    ref, s = orw.cheb_filter(n, z, Y, deg, a, c, tau, lo, n)
    assert relmax(Z, ref) < 1e-12
    assert steps.value == 5 + s


@pytest.mark.parametrize("n", [3, 6, 10])
@pytest.mark.parametrize("lo", [1, 2])
def test_cert_rowsum_vs_oracle(gpu, n, lo):
    oms = 2 * np.pi * np.fft.fftfreq(2 << n)
    out = np.empty(oms.size)
    gpu.load().cert_rowsum(n, oms.size, gpu.ptr(oms), lo, gpu.ptr(out), 1)
    ref = np.array([orw.cert_rowsum(n, float(o), lo) for o in oms])
    assert relmax(out, ref) < 1e-13


@pytest.mark.parametrize("fixture", ["rows_n4_cap-1", "rows_n4_cap8", "rows_n5_cap8"])
@pytest.mark.parametrize("p", [1, 2])
def test_engine_vs_reference_fixture(gpu, fixture, p):
    """HipEngine.run on the reference RowEngine's random rows, every op list (fixtures from fastops.RowEngine)."""
    d = golden(fixture)
    n = int(d["n"])
    cap = int(d["dense_cap"])
    E = pkg("engine")
    eng = E.HipEngine(n, d["z"], d["deg"], d["lam_max"], float(d["tau"]), copies=2,
                      dense_cap=None if cap < 0 else cap, p_levels=(p, None))
    assert eng.n_dense == int(d[f"n_dense_p{p}"])
    for name, ops in OPS(n).items():
        eng.count = [0]
        Z = eng.run(d["Y"], ops)
        assert relmax(Z, d[f"{name}_p{p}"]) < 1e-12, name
        assert eng.count[0] == int(d[f"{name}_p{p}_count"]), name


def test_engine_fp32_close(gpu):
    d = golden("rows_n5_cap8")
    n = int(d["n"])
    E = pkg("engine")
    eng = E.HipEngine(n, d["z"], d["deg"], d["lam_max"], float(d["tau"]), dense_cap=None, precision="fp32")
    Z = eng.run(d["Y"], OPS(n)["gh"])
    eng64 = E.HipEngine(n, d["z"], d["deg"], d["lam_max"], float(d["tau"]), dense_cap=None)
    Z64 = eng64.run(d["Y"], OPS(n)["gh"])
    assert relmax(Z, Z64) < 1e-4


@pytest.mark.parametrize("kind", ["hybrid", "lp"])
def test_gpu_dense_setup_matches_host(gpu, kind):
    """Dense tables and pass matrices from rocSOLVER / rocBLAS equal the host eigh construction (fastops.py:28-39)."""
    E = pkg("engine")
    G = pkg("gpusetup")
    n, N = 5, 32
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(2 * N))
    ks = np.array([0, 1, 2, 62, 63], dtype=np.int32)
    tau = 0.2 * N
    lo = 2 if kind == "hybrid" else 1
    F, Wop, Wback, w = G.dense_setup(n, z, ks, tau, kind, keep_F=True)
    Fh = E.dense_tables(n, z, tau, ks, lo, n)
    Wo, Wb = E.pass_matrices(n, z, ks, Fh, kind)
    # one table per conjugate pair: 62 = L - 2 and 63 = L - 1 take the conjugates of 2's and 1's
    assert list(w) == [0, 1, 2, -1, -1]
    own = np.flatnonzero(w >= 0)
    assert relmax(G.to_row_major(F), Fh[own]) < 1e-12
    assert relmax(G.to_row_major(Wop), Wo[own]) < 1e-12
    assert relmax(G.to_row_major(Wback), Wb[own]) < 1e-12
    # what the plan relies on, from the host's own construction of every frequency: the partner's matrices are the
    # conjugates of the first's
    for i, j in ((1, 4), (2, 3)):
        for T in (Fh, Wo, Wb):
            assert relmax(np.conj(T[i]), T[j]) < 1e-12
    # without keep_F no F is returned, and the tables are the same
    F2, Wop2, Wback2, w2 = G.dense_setup(n, z, ks, tau, kind)
    assert F2 is None and list(w2) == list(w)
    assert relmax(G.to_row_major(Wop2), Wo[own]) < 1e-12 and relmax(G.to_row_major(Wback2), Wb[own]) < 1e-12


@pytest.mark.parametrize("kind", ["hybrid", "lp"])
def test_one_table_per_conjugate_pair(gpu, monkeypatch, kind):
    """Plans on one pass-matrix table per conjugate pair (the default) agree with plans handed every frequency's own
    table, and hold about half the table bytes."""
    G = pkg("gpusetup")
    I = pkg("inverses")
    ph = pkg("phantom")
    R = pkg("radon")
    n, N = 7, 128
    x = ph.ellipses(3, N)
    b = np.stack([R.adrt(x), 0.5 * R.adrt(ph.ellipses(4, N))])
    cls = I.FastHybridHip if kind == "hybrid" else I.FastLpHip
    outer = (0.16, 0.96, 12) if kind == "hybrid" else (1.0, 4.0, 12)
    out, nbytes, slots = [], [], G.table_slots
    try:
        for every in (False, True):
            G.clear_tables()
            if every:
                monkeypatch.setattr(G, "table_slots", lambda z, ks: np.arange(len(ks), dtype=np.int32))
            H = cls(n, dense_cap=10, outer=outer, max_batch=2)
            assert H.eng.n_dense > 0
            out.append(H.apply_batch(b))
            nbytes.append(H.table_bytes)
            ks, z = np.array(H.eng.dense_k), np.asarray(H.z)
            del H
    finally:
        G.clear_tables()
    assert relmax(out[0], out[1]) < 1e-13
    per = 2 * N * N * 16  # W_op and W_back of one frequency
    assert nbytes[1] == per * ks.size
    assert nbytes[0] == per * int(np.sum(slots(z, ks) >= 0)) < 0.6 * nbytes[1]


@pytest.mark.parametrize("kind", ["hybrid", "lp"])
@pytest.mark.parametrize("on_device", [0, 1])
@pytest.mark.parametrize("compact", [False, True])
@pytest.mark.parametrize("max_batch", [2, 20])
def test_plan_copies_host_or_device_tables(gpu, kind, on_device, compact, max_batch):
    """adrt_plan_create copying its pass matrices, host row-major (dense_on_device 0) or device column-major (1),
    every frequency's own (dense_w NULL) or one per conjugate pair (dense_w), built on the host by eigh
    (fastops.py:28-39), gives the results of the package's plan on its borrowed GPU-built tables.  At max_batch 2
    conjugate frequencies share GEMM units, at 20 (40 rows per frequency) each goes alone."""
    E = pkg("engine")
    G = pkg("gpusetup")
    I = pkg("inverses")
    L = pkg("lib")
    ph = pkg("phantom")
    R = pkg("radon")
    n, N = 6, 64
    if kind == "hybrid":
        H = I.FastHybridHip(n, dense_cap=8, outer=(0.16, 0.96, 30), max_batch=max_batch)
    else:
        H = I.FastLpHip(n, dense_cap=8, outer=(1.0, 4.0, 25), max_batch=max_batch)
    B = np.stack([R.adrt(ph.ellipses(s, N)) for s in range(3)])
    ref = H.apply_batch(B)
    ks = np.asarray(H.eng.dense_k)
    Fh = E.dense_tables(n, H.z, H.tau, ks, 2 if kind == "hybrid" else 1, n)
    Wo, Wb = E.pass_matrices(n, H.z, ks, Fh, kind)
    w = G.table_slots(H.z, ks) if compact else None
    if compact:
        assert 0 < int(np.sum(w >= 0)) < ks.size
        Wo, Wb = Wo[w >= 0], Wb[w >= 0]
    if on_device:
        tabs = [L.DeviceArray.from_host(np.ascontiguousarray(T.transpose(0, 2, 1))) for T in (Wo, Wb)]
    else:
        tabs = [np.ascontiguousarray(T) for T in (Wo, Wb)]
    d = L.PlanDesc.from_buffer_copy(H._desc)
    d.dense_F = None  # not read by a plan with pass matrices
    d.dense_Wop, d.dense_Wback = (L.ptr(t) for t in tabs)
    d.dense_on_device = on_device
    d.dense_w = L.ptr(w)
    h = L.P()
    L.check(L.load().adrt_plan_create(ctypes.byref(h), ctypes.byref(d)), "plan_create")
    tabs = None  # the plan holds copies
    saved = H._plan
    try:
        H._plan = h.value
        out = H.apply_batch(B)
    finally:
        H._plan = saved
        L.load().adrt_plan_destroy(h.value)
    assert relmax(out, ref) < 1e-12
