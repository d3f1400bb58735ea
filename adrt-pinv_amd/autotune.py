"""Autotuner of the row kernel's mapping per (N, precision).

The register row kernel holds a row of N values in G = N / E lanes with E = 2^e values per lane, so a 64-lane
workgroup carries 64 / G rows (one row spans G / 64 waves when G > 64), and the butterfly bits split into ceil(n / e)
register passes with one LDS exchange between passes.  Larger E means fewer passes and more instruction-level
parallelism but more registers (lower occupancy).  The tuner times the plan's own row pass (events on the plan's
stream) for each candidate (e = 2, 3, 4 and the generic LDS kernel) on the rows of a real pass, keeps the fastest and
records the choice in the library (adrt_rows_config_set) and in a JSON file next to the plan cache.
"""

import json
import os

from . import lib as L
from . import plancache

CANDIDATES = (-1, 2, 3, 4)


def _table_path():
    d = plancache.cache_dir()
    return os.path.join(d, "rows_tuning.json") if d else None


def load_table():
    p = _table_path()
    if p and os.path.exists(p):
        try:
            with open(p) as f:
                return {tuple(map(int, k.split(","))): int(v) for k, v in json.load(f).items()}
        except (OSError, ValueError):
            return {}
    return {}


def save_table(table):
    p = _table_path()
    if not p:
        return
    tmp = p + f".{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump({f"{n},{prec}": v for (n, prec), v in table.items()}, f)
    os.replace(tmp, p)


def apply_table(table=None):
    """Push a tuning table {(n, precision_code): loge} into the library."""
    lib = L.load()
    for (n, prec), v in (table if table is not None else load_table()).items():
        lib.adrt_rows_config_set(n, prec, v)


def tune(plan, S, stream=None, reps=3, candidates=CANDIDATES, persist=True):
    """Time the row pass of ``plan`` (a FastLpHip / FastHybridHip with S <= max_batch sinograms) for every candidate
    mapping; install and return (best_loge, {loge: ms})."""
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
    for c in candidates:
        if c > 0 and (1 << c) > plan.N:
            continue
        L.check(lib.adrt_rows_config_set(n, prec, c), "rows_config_set")
        L.check(lib.adrt_plan_rows_pass(plan._plan, 1, S, st.handle), "rows_pass")  # warm-up
        e0.record(st)
        for _ in range(reps):
            L.check(lib.adrt_plan_rows_pass(plan._plan, 1, S, st.handle), "rows_pass")
        e1.record(st)
        times[c] = e0.elapsed_ms(e1) / reps
    best = min(times, key=times.get) if times else old
    lib.adrt_rows_config_set(n, prec, best)
    if persist:
        table = load_table()
        table[(n, prec)] = best
        save_table(table)
    return best, times
