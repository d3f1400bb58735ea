"""Every row-kernel variant (generic LDS kernel, register kernel with 4 / 8 / 16 values per lane) against the oracle,
FP64 and FP32, for every op list the inverses use, with and without dense rows."""

import numpy as np
import pytest

from conftest import pkg, relmax
from oracle import rows as orw

pytestmark = pytest.mark.gpu


def op_lists(n):
    return {
        "gh": [("B", 1, 1), ("P", 2, n), ("B", 2, n), ("BT", 2, n), ("BT", 1, 1)],
        "hyb_back": [("BT", 2, n), ("P", 2, n), ("BT", 1, 1)],
        "lp_frame": [("P", 1, n), ("B", 1, n), ("BT", 1, n)],
        "lp_back": [("BT", 1, n), ("P", 1, n)],
        "B_2n": [("B", 2, n)],
        "BT_1n": [("BT", 1, n)],
        "B_11": [("B", 1, 1)],
    }


_CACHE = {}


def setup(n, seed):
    """Frequencies of the full L = 2N grid, bounds above lambda_max (N^2 bounds every level Gram), random degrees,
    and 96 random rows (rows r use frequency r mod L)."""
    N = 1 << n
    L = 2 * N
    rng = np.random.default_rng(seed)
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(L))
    tau = 0.2 * N
    lam = float(N) ** 2 * rng.uniform(1.0, 2.0, L)
    lam = np.maximum(lam, lam[(-np.arange(L)) % L])
    deg = rng.integers(2, 24, L)
    deg = np.maximum(deg, deg[(-np.arange(L)) % L])
    R = min(96, 2 * L)
    Y = rng.standard_normal((R, N)) + 1j * rng.standard_normal((R, N))
    return z, deg, lam, tau, Y


def reference(n, cap, eng, name, ops, Y, z, deg, lam, tau):
    key = (n, cap, name)
    if key not in _CACHE:
        ref_eng = orw.Rows(n, z, deg, lam, tau, dense_cap=cap, p_levels=(2, None))
        if cap is not None:  # the GPU-built dense tables (eigh is not part of this check)
            for i, k in enumerate(eng.dense_k):
                ref_eng.dense[int(k)] = eng.dense_F[i]
        _CACHE[key] = ref_eng.run(Y, ops)
    return _CACHE[key]


@pytest.mark.parametrize("n", [5, 6, 8, 9, 11])
@pytest.mark.parametrize("loge", [-1, 2, 3, 4])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_variant(gpu, n, loge, prec):
    E = pkg("engine")
    lib = gpu.load()
    p = gpu.PRECISION[prec]
    old = lib.adrt_rows_config_get(n, p)
    gpu.check(lib.adrt_rows_config_set(n, p, loge), "config")
    try:
        z, deg, lam, tau, Y = setup(n, n)
        cap = 18 if n <= 8 else None
        eng = E.HipEngine(n, z, deg, lam, tau, dense_cap=cap, p_levels=(2, None), precision=prec)
        tol = 1e-12 if prec == "fp64" else 2e-5
        for name, ops in op_lists(n).items():
            ref = reference(n, cap, eng, name, ops, Y, z, deg, lam, tau)
            got = eng.run(Y, ops)
            assert relmax(got, ref) < tol, (name, relmax(got, ref))
    finally:
        lib.adrt_rows_config_set(n, p, old)
