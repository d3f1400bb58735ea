"""Autotuner of the row kernel's mapping per (N, precision).

The register row kernel holds a row of N values in G = N / E lanes with E = 2^e values per lane, so a 64-lane
workgroup carries 64 / G rows (one row spans G / 64 waves when G > 64), and the butterfly bits split into ceil(n / e)
register passes with one LDS exchange between passes.  Larger E means fewer passes and more instruction-level
parallelism but more registers (lower occupancy).  The tuner times the plan's own row pass (events on the plan's
stream) for each candidate (e = 2, 3, 4 and the generic LDS kernel) on the rows of a real pass, keeps the fastest and
records the choice in the library (adrt_rows_config_set) and in a JSON file next to the plan cache.

With the mapping fixed it then places the dense low-frequency GEMM of the pass (plans with dense rows): as MFMA waves
inside the row kernel's launch, at most 1, 2 or 4 per SIMD, or as a launch of its own after the rows
(adrt_dense_config_set).  Few dense waves leave the vector ALUs to the rows; more hide the GEMM's load latency better
when the dense work is large (N >= 1024, where a third of the frequencies can be dense).  Placements within 1% of
the fastest count as a tie, won by two dense waves per SIMD, then one, then four: with two each worker takes half
the tiles it would take with one, so fewer of its frequency's rows have left L2 when it reads them again (at N=256,
64 sinograms: 1.30 against 1.58 times the algorithmic HBM bytes, at the same time).

An FP32 plan with FP64 refinement passes (refine >= 1) runs those passes at the FP32 plan's batch depth, where the
best FP64 mapping can differ from a plain FP64 plan's: the tuner times them too, on the rows an apply's last (FP64)
pass leaves, and installs the choice as class 2 of adrt_rows_config_set (measured at N=256 with 64 sinograms: 16
values per lane, 3% off the whole apply against the default 4).
"""

import json
import os

from . import lib as L
from . import plancache

CANDIDATES = (-1, 2, 3, 4)
DENSE_CANDIDATES = (-1, 1, 2, 4)
DENSE_TIE = 0.01


DENSE_TIE_ORDER = (2, 1, 4, 8)


def pick_dense(times):
    """The dense placement of ``times`` ({mode: ms}): the fastest, or among those within DENSE_TIE of it the first of
    DENSE_TIE_ORDER (a launch of its own, -1, only when it is the fastest outright).  Two waves per SIMD give each
    worker half the tiles of one, so fewer of a frequency's rows are re-read from HBM; four take wave slots from the
    rows that these sizes' dense work does not need."""
    best = min(times, key=times.get)
    if best == -1:
        return best
    near = [m for m, t in times.items() if m > 0 and t <= times[best] * (1 + DENSE_TIE)]
    return min(near, key=lambda m: DENSE_TIE_ORDER.index(m) if m in DENSE_TIE_ORDER else len(DENSE_TIE_ORDER))


class TuneResult(tuple):
    """(loge, {loge: ms}); the dense placement chosen after it is in .dense, its timings in .dense_times; for an FP32
    plan with FP64 refinement passes, their row mapping in .f64 (None otherwise), its timings in .f64_times; the dense
    placement of one-sinogram passes in .dense1 (None when not tuned), its timings in .dense1_times."""

    def __new__(cls, loge, times, dense=0, dense_times=None, f64=None, f64_times=None):
        r = super().__new__(cls, (loge, times))
        r.dense = dense
        r.dense_times = dense_times or {}
        r.f64 = f64
        r.f64_times = f64_times or {}
        r.dense1, r.dense1_times = None, {}
        return r


def _table_path():
    d = plancache.cache_dir()
    return os.path.join(d, "rows_tuning.json") if d else None


def load_table():
    p = _table_path()
    if p and os.path.exists(p):
        try:
            with open(p) as f:
                out = {}
                for k, v in json.load(f).items():
                    loge, dense = (int(v), 0) if isinstance(v, int) else (int(v[0]), int(v[1]))
                    out[tuple(map(int, k.split(",")))] = (loge, dense)
                return out
        except (OSError, ValueError):
            return {}
    return {}


def save_table(table):
    p = _table_path()
    if not p:
        return
    tmp = p + f".{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump({f"{n},{prec}": [int(v[0]), int(v[1])] for (n, prec), v in table.items()}, f)
    os.replace(tmp, p)


def apply_table(table=None):
    """Push a tuning table {(n, precision_code): (loge, dense)} into the library."""
    lib = L.load()
    for (n, prec), (loge, dense) in (table if table is not None else load_table()).items():
        lib.adrt_rows_config_set(n, prec, loge)
        if prec < 2:  # class 2 (FP64 refinement passes of FP32 plans) shares the FP64 dense placement
            lib.adrt_dense_config_set(n, prec, dense)


def _time_pass(lib, plan, S, st, e0, e1, reps, which=1):
    L.check(lib.adrt_plan_rows_pass(plan._plan, which, S, st.handle), "rows_pass")  # warm-up
    e0.record(st)
    for _ in range(reps):
        L.check(lib.adrt_plan_rows_pass(plan._plan, which, S, st.handle), "rows_pass")
    e1.record(st)
    return e0.elapsed_ms(e1) / reps


def tune(plan, S, stream=None, reps=3, candidates=CANDIDATES, dense_candidates=DENSE_CANDIDATES, persist=True):
    """Time the row pass of ``plan`` (a FastLpHip / FastHybridHip with S <= max_batch sinograms) for every candidate
    mapping, then (plans with dense rows) every dense placement at the chosen mapping, and with S > 1 again for passes
    over one sinogram (dense class precision + 2); install them and return TuneResult(best_loge, {loge: ms}) with
    .dense, .dense_times, .dense1 and .dense1_times."""
    lib = L.load()
    prec = L.CONFIG_PRECISION[plan.precision]
    n = plan.n
    st = stream or L.Stream()
    dx = L.DeviceArray((S, plan.N, plan.N))
    dy = L.DeviceArray((S, plan.N, plan.N))
    dx.zero(st)
    plan.operator_device(dx, dy, S, st)  # a real pass leaves the plan's rows in place
    e0, e1 = L.Event(), L.Event()
    times = {}
    old = lib.adrt_rows_config_get(n, prec)
    old_dense = lib.adrt_dense_config_get(n, prec)
    L.check(lib.adrt_dense_config_set(n, prec, 0), "dense_config_set")
    for c in candidates:
        if c > 0 and (1 << c) > plan.N:
            continue
        L.check(lib.adrt_rows_config_set(n, prec, c), "rows_config_set")
        times[c] = _time_pass(lib, plan, S, st, e0, e1, reps)
    best = min(times, key=times.get) if times else old
    lib.adrt_rows_config_set(n, prec, best)
    dense, dense_times = old_dense, {}
    if getattr(plan.eng, "n_dense", 0) > 0 and best != -1:  # the generic kernel has no dense waves
        for m in dense_candidates:
            L.check(lib.adrt_dense_config_set(n, prec, m), "dense_config_set")
            dense_times[m] = _time_pass(lib, plan, S, st, e0, e1, reps)
        dense = pick_dense(dense_times)
    lib.adrt_dense_config_set(n, prec, dense)
    # passes over one sinogram (dense class prec + 2), where the W stream dominates: N=1024 at one sinogram takes
    # 299 us per row pass with two dense waves per SIMD against 337 us with one, while with more sinograms one wins
    dense1, dense1_times = None, {}
    if getattr(plan.eng, "n_dense", 0) > 0 and best != -1 and S > 1:
        plan.operator_device(dx, dy, 1, st)
        for m in dense_candidates:
            L.check(lib.adrt_dense_config_set(n, prec + 2, m), "dense_config_set")
            dense1_times[m] = _time_pass(lib, plan, 1, st, e0, e1, reps)
        dense1 = pick_dense(dense1_times)
        lib.adrt_dense_config_set(n, prec + 2, dense1)
        plan.operator_device(dx, dy, S, st)
    # the FP64 refinement passes of an FP32 plan (class 2): an apply leaves the rows of its last, FP64 pass
    f64, f64_times = None, {}
    if plan.precision != "fp64" and getattr(plan, "refine", 0) > 0:
        db = L.DeviceArray((S, 4, 2 * plan.N - 1, plan.N))
        db.zero(st)
        plan.apply_device(db, dy, S, st, count=False)
        for c in candidates:
            if c > 0 and (1 << c) > plan.N:
                continue
            L.check(lib.adrt_rows_config_set(n, 2, c), "rows_config_set")
            f64_times[c] = _time_pass(lib, plan, S, st, e0, e1, reps, which=3)
        f64 = min(f64_times, key=f64_times.get)
        lib.adrt_rows_config_set(n, 2, f64)
    if persist:
        table = load_table()
        table[(n, prec)] = (best, dense)
        if f64 is not None:
            table[(n, 2)] = (f64, 0)
        save_table(table)
    res = TuneResult(best, times, dense, dense_times, f64, f64_times)
    res.dense1, res.dense1_times = dense1, dense1_times
    return res


REFINE_CHOICES = (1, 2, 3, 4, 6, 8, 12)


def refine_for(n, dense_cap, lam_max, outer, target=1e-6, probes=2, seed=8000, cache=True, **kw):
    """The smallest number of FP64 refinement steps (outer steps at the end of an FP32 apply that run in FP64) for
    which the FP32 hybrid plan stays within ``target`` of the FP64 plan on ``probes`` ellipse phantoms with 1% noise
    (seeds ``seed``+i), at most the outer degree minus one.  Plans whose non-dense rows include ill-conditioned
    low frequencies (a large dense cap at small N) carry more FP32 rounding into the iterate than one FP64 step
    removes: measured at N=128, dense cap 32 (no dense rows), 1 step leaves 1.2e-5 and 6 steps 5e-7; caps of 16 and
    every size up to 1024 need 1 or 2.  The choice is cached beside the plan's setup scalars."""
    import numpy as np

    from . import phantom as ph
    from . import radon as R
    from .inverses import FastHybridHip

    D = int(outer[2])
    params = {"dense_cap": dense_cap, "target": target, "probes": probes, "seed": seed, "setup_outer": list(outer),
              "lam": float(np.sum(lam_max)), **kw}
    hit = plancache.load("refine", n, **params) if cache else None
    if hit is not None:
        return int(hit["outer"][2])
    N = 1 << n
    B = np.stack([ph.add_noise(R.adrt(ph.ellipses(seed + i, N)), 0.01, seed + 100 + i) for i in range(probes)])
    H64 = FastHybridHip(n, lam_max=lam_max, outer=outer, dense_cap=dense_cap, max_batch=probes, **kw)
    x64 = H64.apply_batch(B)
    H64.close()
    best = D - 1
    for r in REFINE_CHOICES:
        if r >= D - 1:
            break
        H = FastHybridHip(n, lam_max=lam_max, outer=outer, dense_cap=dense_cap, max_batch=probes, precision="fp32",
                          refine=r, **kw)
        x = H.apply_batch(B)
        H.close()
        err = max(float(np.linalg.norm(x[i] - x64[i]) / np.linalg.norm(x64[i])) for i in range(probes))
        if err <= target:
            best = r
            break
    if cache:
        plancache.store("refine", n, np.zeros(1), (0.0, 0.0, best), **params)
    return best

