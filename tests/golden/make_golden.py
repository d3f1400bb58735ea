"""Writes the golden fixtures under tests/golden/ from the reference Python modules (run in the build container only).

It imports fastops / fasthybrid / fastspec / adrtcore from /root/reference (they need only numpy and scipy) for
what no closed form or invariant pins: RowEngine.run on random rows (fastops.py:156-167), the numpy power
iteration bounds (fastops.py:266-272, 309-316) and the end-to-end FastLp / FastHybrid outputs at fixed setup
scalars (fastops.py:261-338).  The fixtures hold inputs and outputs only.

    python tests/golden/make_golden.py            # writes tests/golden/*.npz
"""

import importlib
import os
import sys

import numpy as np

sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("ADRT_REFERENCE", "/root/reference")
sys.path.insert(0, REF)
sys.path.insert(0, REPO)

import adrtcore as ac  # noqa: E402  (reference)
import fastops as fo  # noqa: E402  (reference)

phantom = importlib.import_module("adrt-pinv_amd.phantom")


def ops_for(n):
    return {
        "B_1n": [("B", 1, n)],
        "BT_1n": [("BT", 1, n)],
        "B_2n": [("B", 2, n)],
        "BT_2n": [("BT", 2, n)],
        "B_11": [("B", 1, 1)],
        "BT_11": [("BT", 1, 1)],
        "P_1n": [("P", 1, n)],
        "P_2n": [("P", 2, n)],
        "gh": [("B", 1, 1), ("P", 2, n), ("B", 2, n), ("BT", 2, n), ("BT", 1, 1)],
        "hyb_back": [("BT", 2, n), ("P", 2, n), ("BT", 1, 1)],
        "lp_frame": [("P", 1, n), ("B", 1, n), ("BT", 1, n)],
        "lp_back": [("BT", 1, n), ("P", 1, n)],
    }


def rows_fixture(n, dense_cap, seed):
    """RowEngine.run on random Y for every op list, with the degrees of a FastHybrid-like setup."""
    N = 1 << n
    L = 2 * N
    rng = np.random.default_rng(seed)
    z = np.exp(1j * 2 * np.pi * np.fft.fftfreq(L))
    tau = 0.2 * N
    lam = fo._sym(np.abs(rng.standard_normal(L)) * N * N * 0.5 + 1.0)
    kap = 1.0 + lam / tau
    rate = (np.sqrt(kap) - 1) / (np.sqrt(kap) + 1)
    deg = np.maximum(2, np.ceil(np.log(0.05) / np.log(np.maximum(rate, 1e-12)))).astype(int)
    Y = rng.standard_normal((2 * L, N)) + 1j * rng.standard_normal((2 * L, N))
    out = {"n": n, "tau": tau, "z": z, "deg": deg, "lam_max": lam, "Y": Y, "dense_cap": -1 if dense_cap is None else dense_cap}
    for name, ops in ops_for(n).items():
        for p_levels in ((1, None), (2, None)):
            eng = fo.RowEngine(n, z, deg, lam, tau, threads=1, copies=2, dense_cap=dense_cap, p_levels=p_levels)
            Z = eng.run(Y, ops)
            out[f"{name}_p{p_levels[0]}"] = Z
            out[f"{name}_p{p_levels[0]}_count"] = eng.count[0]
            out[f"n_dense_p{p_levels[0]}"] = eng.n_dense
    return out


def inverse_fixture(n, dense_cap, x, b):
    """FastHybrid and FastLp setup scalars and outputs (fastops defaults, numpy engine, threads=1)."""
    out = {"n": n, "x": x, "b": b, "dense_cap": -1 if dense_cap is None else dense_cap}
    H = fo.FastHybrid(n, threads=1, dense_cap=dense_cap)
    out.update(
        h_lam_max=H.lam_max, h_deg=H.deg, h_lo=H.lo, h_hi=H.hi, h_D=H.outer_deg, h_n_dense=H.eng.n_dense,
        h_gh=H.Gh(x), h_back=H.backproject(b), h_apply=H.apply(b), h_P=H.P(x),
    )
    F = fo.FastLp(n, threads=1, dense_cap=dense_cap)
    out.update(
        l_lam_max=F.lam_max, l_deg=F.deg, l_lo=F.lo, l_hi=F.hi, l_D=F.outer_deg, l_n_dense=F.eng.n_dense,
        l_frame=F.frame_window(x), l_back=F.backproject(b), l_apply=F.apply(b),
    )
    return out


def main():
    os.makedirs(HERE, exist_ok=True)
    for n, cap in ((4, None), (4, 8), (5, 8)):
        d = rows_fixture(n, cap, seed=10 + n)
        np.savez_compressed(os.path.join(HERE, f"rows_n{n}_cap{d['dense_cap']}.npz"), **d)
        print("rows", n, cap, flush=True)
    for n, cap in ((3, None), (4, None), (4, 8)):
        N = 1 << n
        x = np.random.default_rng(100 + n).standard_normal((N, N))
        d = inverse_fixture(n, cap, x, ac.adrt(n, x))
        np.savez_compressed(os.path.join(HERE, f"inverse_n{n}_cap{d['dense_cap']}.npz"), **d)
        print("inverse", n, cap, d["h_D"], d["l_D"], flush=True)
    # config 1: N = 64 ellipse phantom (seed 0), reference defaults
    n = 6
    x = phantom.ellipses(0, 1 << n)
    d = inverse_fixture(n, None, x, ac.adrt(n, x))
    np.savez_compressed(os.path.join(HERE, "inverse_n6_phantom.npz"), **d)
    print("inverse phantom n=6", d["h_D"], d["l_D"], flush=True)


if __name__ == "__main__":
    main()
