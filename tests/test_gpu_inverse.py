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
