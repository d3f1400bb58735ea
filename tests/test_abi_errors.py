"""Argument errors of the C ABI, on the CPU: every entry point validates its arguments before any HIP call and
returns -1 with a message in adrt_last_error() (the reference's C functions do not validate: bad arguments there are
undefined behaviour, cpp/oblique.cpp:56-132)."""

import ctypes

import numpy as np
import pytest

from conftest import pkg


@pytest.fixture(scope="module")
def L():
    return pkg("lib")


def err(L):
    m = L.load().adrt_last_error()
    return m.decode() if m else ""


def desc(L, **kw):
    d = L.PlanDesc()
    z = np.ones(16, dtype=np.complex128)
    deg = np.ones(16, dtype=np.int32)
    ac = np.ones(16)
    d.kind, d.n, d.precision, d.precond, d.g1_deg, d.outer_deg = L.KIND_HYBRID, 3, 0, 1, 14, 10
    d.lo, d.hi, d.tau = 1.0, 2.0, 1.6
    d.z, d.deg, d.a, d.c = L.ptr(z), L.ptr(deg), L.ptr(ac), L.ptr(ac)
    d.max_batch = 1
    for k, v in kw.items():
        setattr(d, k, v)
    return d, (z, deg, ac)


@pytest.mark.parametrize("field,value,word", [("kind", 7, "kind"), ("n", 12, "n"), ("max_batch", 0, "max_batch"),
                                               ("precision", 5, "precision"), ("g1_deg", 0, "g1_deg"),
                                               ("outer_deg", 0, "outer_deg"), ("hi", -1.0, "hi")])
def test_plan_create_rejects(L, field, value, word):
    d, keep = desc(L, **{field: value})
    h = L.P()
    assert L.load().adrt_plan_create(ctypes.byref(h), ctypes.byref(d)) == -1
    assert h.value is None and word in err(L)


def test_normal_plan_rejects_dense_rows(L):
    d, keep = desc(L, kind=L.KIND_NORMAL, n_dense=2)
    h = L.P()
    assert L.load().adrt_plan_create(ctypes.byref(h), ctypes.byref(d)) == -1
    assert "dense" in err(L)


def test_row_kernels_reject_bad_levels(L):
    lib = L.load()
    z = np.ones(4, dtype=np.complex128)
    X = np.zeros((4, 8), dtype=np.complex128)
    assert lib.adrt_bfly_rows(3, 4, L.ptr(z), L.ptr(X), 2, 1, 0) == -1 and "m_lo" in err(L)
    assert lib.adrt_bfly_rows(3, 4, L.ptr(z), L.ptr(X), 1, 4, 0) == -1
    assert lib.adrt_bfly_rows(3, 4, L.ptr(z), L.ptr(X), 1, 3, 2) == -1 and "transpose" in err(L)
    deg = np.ones(4, dtype=np.int32)
    ac = np.ones(4)
    assert lib.adrt_cheb_rows(3, 4, L.ptr(z), L.ptr(deg), L.ptr(ac), L.ptr(ac), 1.0, L.ptr(X), L.ptr(X), 0, 3,
                              None) == -1
    assert lib.adrt_bfly_rows(12, 4, L.ptr(z), L.ptr(X), 1, 3, 0) == -1 and "n" in err(L)


def test_memory_and_plan_calls_reject(L):
    lib = L.load()
    buf = np.zeros(4)
    assert lib.adrt_memcpy(L.ptr(buf), L.ptr(buf), 32, 9, None) == -1 and "kind" in err(L)
    assert lib.adrt_host_malloc(None, 16) == -1
    assert lib.adrt_plan_set_outer(None, 1.0, 2.0, 3) == -1 and "null plan" in err(L)
    assert lib.adrt_plan_apply(None, None, None, 1, None) == -1
    h = L.P()
    assert lib.adrt_rows_plan_create(ctypes.byref(h), 3, 16, None, None, None, None, 1.0, 0, None, None, 0) == -1


def test_dense_config_validates(L):
    """Classes 0 / 1 (FP64 / FP32 passes) and 2 / 3 (the same over one sinogram, following 0 / 1 until set)."""
    lib = L.load()
    assert lib.adrt_dense_config_set(12, 0, 1) == -1 and "n" in err(L)
    assert lib.adrt_dense_config_set(8, 4, 1) == -1 and "class" in err(L)
    assert lib.adrt_dense_config_set(8, -1, 1) == -1 and "class" in err(L)
    assert lib.adrt_dense_config_set(8, 1, 9) == -1 and "mode" in err(L)
    assert lib.adrt_dense_config_set(8, 1, -2) == -1
    old = lib.adrt_dense_config_get(8, 1)
    try:
        for m in (-1, 0, 1, 2, 4, 8):
            assert lib.adrt_dense_config_set(8, 1, m) == 0
            assert lib.adrt_dense_config_get(8, 1) == m
            assert lib.adrt_dense_config_get(8, 3) == m  # one sinogram: as FP32 until set
        assert lib.adrt_dense_config_set(8, 3, 4) == 0
        assert lib.adrt_dense_config_get(8, 3) == 4 and lib.adrt_dense_config_get(8, 1) == 8
        assert lib.adrt_dense_config_set(8, 3, 0) == 0
        assert lib.adrt_dense_config_get(8, 3) == 8
    finally:
        lib.adrt_dense_config_set(8, 1, old)
        lib.adrt_dense_config_set(8, 3, 0)


def test_rows_config_classes(L):
    """Class 2 (the FP64 refinement passes of FP32 plans) follows class 0 until set; classes above 2 are errors."""
    lib = L.load()
    assert lib.adrt_rows_config_set(8, 3, 2) == -1 and "class" in err(L)
    old0, old2 = lib.adrt_rows_config_get(8, 0), lib.adrt_rows_config_get(8, 2)
    try:
        assert lib.adrt_rows_config_set(8, 2, 0) == 0
        assert lib.adrt_rows_config_set(8, 0, 3) == 0
        assert lib.adrt_rows_config_get(8, 2) == 3
        assert lib.adrt_rows_config_set(8, 2, 4) == 0
        assert lib.adrt_rows_config_get(8, 2) == 4 and lib.adrt_rows_config_get(8, 0) == 3
    finally:
        lib.adrt_rows_config_set(8, 0, old0)
        lib.adrt_rows_config_set(8, 2, old2 if old2 != old0 else 0)


def test_ctx_calls_reject(L):
    """Context calls validate their arguments before touching the device."""
    lib = L.load()
    assert lib.adrt_ctx_create(None, 0) == -1 and "null" in err(L)
    assert lib.adrt_ctx_destroy(None) == 0
    assert lib.adrt_ctx_stream(None) is None
    used = ctypes.c_ulonglong()
    assert lib.adrt_ctx_usage(None, ctypes.byref(used), None) == -1 and "null context" in err(L)
    h = L.P()
    assert lib.adrt_ctx_malloc(None, ctypes.byref(h), 16) == -1 and "null" in err(L)
    assert lib.adrt_ctx_free(None, None) == -1 and "null context" in err(L)
    d, keep = desc(L)
    assert lib.adrt_plan_create_in(None, ctypes.byref(h), ctypes.byref(d)) == -1 and "null context" in err(L)


def test_own_queue_stream_rejects_null(L):
    assert L.load().adrt_stream_create_own_queue(None) == -1 and "null" in err(L)


def test_time_rows_ring_rejects(L):
    lib = L.load()
    assert lib.adrt_plan_time_rows_ring(None, None, None, 0) == -1 and "null plan" in err(L)


@pytest.mark.parametrize("m_lo", [0, -1, -40])
def test_cert_rowsum_rejects_low_levels(L, m_lo):
    """Level m of the row-sum recursion reads r[2a +- 1 + A] for |a| <= 2^(n-m) - 1: inside its slot only for m >= 1.
    Below that the call is an argument error before any device work (the reference's loop would read out of bounds)."""
    om = np.linspace(-3.0, 3.0, 8)
    out = np.zeros(8)
    assert L.load().adrt_cert_rowsum(3, 8, L.ptr(om), m_lo, L.ptr(out)) == -1 and "m_lo" in err(L)
    assert np.all(out == 0)


def test_timing_and_refine_queries_reject_null(L):
    lib = L.load()
    n = L.LL()
    assert lib.adrt_plan_time_rows_count(None, ctypes.byref(n)) == -1 and "null" in err(L)
    assert lib.adrt_plan_refine_steps(None) == 0


def test_own_queue_cap_declared(L):
    """The cap on own-queue streams is part of the header, and no stream is alive before one is made."""
    import os
    import re

    from conftest import REPO

    hdr = open(os.path.join(REPO, "include", "adrt_hip.h")).read()
    assert int(re.search(r"#define ADRT_MAX_OWN_QUEUES (\d+)", hdr).group(1)) == 4
    assert L.load().adrt_own_queue_count() == 0


def test_noalloc_forward_rejects_before_device_work(L):
    """adrt_forward_noalloc checks its arguments and its per-sinogram scratch need before touching the device: N =
    4096 needs more than the static scratch in every dtype; S = 0 is a no-op; bad n / dtype / S are errors."""
    lib = L.load()
    assert lib.adrt_noalloc_scratch_bytes() == 256 << 20
    for dtype in (0, 1, 2):
        assert lib.adrt_forward_noalloc(12, dtype, None, None, 1, None) == -1 and "scratch" in err(L)
    assert lib.adrt_forward_noalloc(11, 2, None, None, 0, None) == 0
    assert lib.adrt_forward_noalloc(0, 0, None, None, 1, None) == -1 and "n" in err(L)
    assert lib.adrt_forward_noalloc(5, 3, None, None, 1, None) == -1 and "dtype" in err(L)
    assert lib.adrt_forward_noalloc(5, 0, None, None, -1, None) == -1 and "S" in err(L)
    assert lib.adrt_forward_noalloc(5, 0, None, None, 1, None) == -1 and "null" in err(L)


def _dense_desc(L, n, ks, w, with_w=True):
    N = 1 << n
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(2 * N))
    d, keep = desc(L, n=n, z=None)
    deg, ac = np.ones(2 * N, dtype=np.int32), np.ones(2 * N)
    d.z, d.deg, d.a, d.c = L.ptr(z), L.ptr(deg), L.ptr(ac), L.ptr(ac)
    keep = (keep, deg, ac)
    kk = np.asarray(ks, dtype=np.int32)
    ww = None if w is None else np.asarray(w, dtype=np.int32)
    d.n_dense, d.dense_k, d.dense_w = len(ks), L.ptr(kk), L.ptr(ww)
    if with_w:
        d.dense_Wop = d.dense_Wback = 4096  # never dereferenced: the descriptor is rejected first
    return d, (keep, z, kk, ww)


@pytest.mark.parametrize("n,ks,w,with_w,word", [
    (4, [1, 31, 5], [0, -1, -1], True, "no conjugate partner"),  # 5's partner 27 is not dense
    (4, [1, 31], [0, 2], True, "dense_w out of range"),
    (4, [1, 31], [-1, -1], True, "no conjugate partner"),
    (3, [1, 15], [0, -1], True, "n >= 4"),
    (4, [1, 31], [0, -1], False, "pass matrices"),
    (4, [1, 1], None, True, "twice"),
])
def test_plan_create_rejects_dense_tables(L, n, ks, w, with_w, word):
    d, keep = _dense_desc(L, n, ks, w, with_w)
    h = L.P()
    assert L.load().adrt_plan_create(ctypes.byref(h), ctypes.byref(d)) == -1
    assert h.value is None and word in err(L)


def test_table_slots_pairs_conjugate_frequencies():
    """One pass-matrix table per conjugate pair (k, L - k), the first in the dense list's order; self-conjugate
    frequencies (0 and N) and frequencies whose partner is not dense keep their own."""
    G = pkg("gpusetup")
    N = 32
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(2 * N))
    assert list(G.table_slots(z, [0, 1, 2, 62, 63])) == [0, 1, 2, -1, -1]
    assert list(G.table_slots(z, [63, 32, 1, 5])) == [0, 1, -1, 2]
    # twiddles that are not exact conjugates are not paired
    z2 = z.copy()
    z2[63] = np.conj(z[1]) * (1 + 1e-16j)
    z2[63] = complex(z2[63].real, np.nextafter(z2[63].imag, 0.0))
    assert list(G.table_slots(z2, [1, 63])) == [0, 1]


def test_plan_bytes_counts_the_tables_given(L):
    """adrt_plan_bytes (host arithmetic only) charges copied tables: F per dense frequency when given, and the two
    pass matrices per table, max(dense_w) + 1 of them with dense_w, else one per dense frequency."""
    n, N = 4, 16
    lib = L.load()
    per = N * N * 16
    base, keep = _dense_desc(L, n, [], None, with_w=False)
    b0 = lib.adrt_plan_bytes(ctypes.byref(base))
    cases = [([1, 31, 5], [0, -1, 1], False, 2 * 2), ([1, 31, 5], None, False, 3 * 2), ([1, 31, 5], None, True, 3 * 3),
             ([1, 31, 5], [0, -1, 1], True, 3 + 2 * 2)]
    for ks, w, with_f, tables in cases:
        d, keep = _dense_desc(L, n, ks, w)
        d.dense_F = 4096 if with_f else None
        assert lib.adrt_plan_bytes(ctypes.byref(d)) - b0 == tables * per, (ks, w, with_f)
    d, keep = _dense_desc(L, n, [1, 31, 5], [0, -1, 1])
    d.dense_on_device = 2  # borrowed: not the plan's bytes
    assert lib.adrt_plan_bytes(ctypes.byref(d)) == b0
