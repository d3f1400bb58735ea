"""Benchmark: hybrid ADRT inverses per second at N = 256 on one MI355X, plus mixed-N batching and serving.

    python bench.py [--steps K] [--warmup W] [--batch B] [--n 8] [--dense-cap 16] [--no-cpu] [--no-serve]
    python bench.py --rows-only [--mapping LOGE,DENSE]   # short workload for rocprofv3 --pmc runs of the row kernel
    python bench.py --no-large       # without the N=2048, batch-16 FP32 vs FP64 leg

A step is one batched hybrid inverse (FP32 row engine with TwoSum-compensated Chebyshev accumulation; FP64 dense
table GEMM, preconditioner, outer iteration and the last refinement steps) of B noisy ellipse-phantom sinograms that
are resident in HBM, replayed from one hipGraph.  ``value`` is inverses/s (B * K / elapsed).  Beside it the line
reports the plain FP64 GPU path (one sinogram per call, no graph, no batching), the relative error of the FP32 result
against the FP64 one, the roofline of the dominant kernel (the register row kernel of the G_h pass, timed with HIP
the kernel's own dispatch timestamps inside whole applies), the reference C++ row engine timed on the host cores, a 512-sinogram
uneven batch in one graph launch and a Poisson serving trace over mixed N with p50 / p99 latency, served twice: with
batches in slots, and with N = 512 and 1024 served by continuous batching at outer-step granularity.
"""

import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import importlib  # noqa: E402

PEAK = {"fp32": 157.3e12, "fp32_nonpacked": 78.6e12, "fp64": 78.6e12, "hbm": 8.0e12}

def pmc_summary(N, batch):
    """The committed rocprofv3 --pmc summary of the row kernel for this workload (N, sinograms per launch)."""
    return os.path.join(REPO, "profiles", f"rows_reg_kernel_n{N}_b{batch}_pmc.txt")


def pkg(sub):
    return importlib.import_module(f"adrt-pinv_amd.{sub}")


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def dense_tables_read(H):
    """Dense pass matrices a pass reads: frequencies k and L - k with exactly conjugate twiddles share k's matrix (the
    plan applies its conjugate to the partner), so a pair reads one."""
    ks = {int(k) for k in H.eng.dense_k}
    z = np.asarray(H.eng.z)
    Lr = len(z)
    pairs = sum(1 for k in ks if (Lr - k) % Lr in ks and (Lr - k) % Lr > k and z[(Lr - k) % Lr] == np.conj(z[k]))
    return len(ks) - pairs


def dense_work(H, S):
    """FP64 MFMA FLOPs and HBM bytes of the dense tiles in one row-kernel launch over S sinograms: per dense frequency
    the 2S rows times the N x N complex pass matrix (8 N^2 flops per row), each pass matrix read once (complex128; a
    conjugate pair of frequencies reads one) and the rows read and written once (complex64)."""
    N, nd, SP = H.N, H.eng.n_dense, 2 * S
    return 8.0 * nd * SP * N * N, dense_tables_read(H) * N * N * 16.0 + 2.0 * nd * SP * N * 8


def rows_kernel_work(H, S, ops):
    """Algorithmic FLOPs (vector ALU) and HBM bytes of one row-kernel launch over S sinograms.  The kernel takes the
    rows of the non-dense frequencies (both packed pairs of every sinogram); per row: 8N flops per B / BT level and
    (d - 1)(16 N U + 18 N) + 2N per Chebyshev (P) op over U levels; it reads and writes each row once (complex64).
    The bytes include the dense tiles of the same launch (dense_work); their FP64 MFMA flops are counted apart."""
    N, Lr = H.N, H.L
    dense = np.zeros(Lr, dtype=bool)
    dense[H.eng.dense_k] = True
    rows_per_k = 2 * S
    lev = sum(hi - lo + 1 for o, lo, hi in ops if o in ("B", "BT"))
    d = H.deg.astype(float)[~dense]
    f = rows_per_k * (~dense).sum() * 8.0 * N * lev
    for o, lo, hi in ops:
        if o == "P":
            U = hi - lo + 1
            f += rows_per_k * ((np.maximum(d - 1, 0)) * (16 * N * U + 18 * N) + 2 * N).sum()
    nbytes = 2.0 * rows_per_k * (~dense).sum() * N * 8 + dense_work(H, S)[1]
    return f, nbytes


def pmc_traffic(N, batch):
    """(HBM bytes per launch of the row kernel, "loge,dense" mapping they were counted at) from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known size), or (None, None)."""
    try:
        txt = open(pmc_summary(N, batch)).read()
    except OSError:
        return None, None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt, flags=re.M)
    mp = re.search(r"--mapping (\S+)", txt)
    return (float(m.group(1)) if m else None), (mp.group(1) if mp else None)


def cpu_baseline(H, cap, b, threads, budget_s=20.0):
    """The reference's C++ row engine (cpp/oblique.cpp built into oracle/_ref) under the oracle's NumPy host, same
    setup scalars, one sinogram; tmin of up to three runs within a time budget (exp_compiled_tables.py:20-27)."""
    from oracle import inverses as oiv
    from oracle import refcpp

    if not refcpp.available():
        return None

    class RefHybrid(oiv.Hybrid):
        engine_cls = refcpp.RefRows

    R = RefHybrid(H.n, lam_max=H.lam_max, dense_cap=cap, threads=threads, outer=(H.lo, H.hi, H.outer_deg))
    ts = []
    t_all = time.time()
    x = None
    while len(ts) < 3 and (not ts or time.time() - t_all + ts[-1] < budget_s):
        t0 = time.time()
        x = R.apply(b)
        ts.append(time.time() - t0)
    return min(ts), len(ts), x


def make_sinos(n, count, seed0, noise_seed0, level=0.01):
    ph = pkg("phantom")
    R = pkg("radon")
    N = 1 << n
    X = np.stack([ph.ellipses(seed0 + i, N) for i in range(count)])
    B0 = R.adrt(X)
    return np.stack([ph.add_noise(B0[i], level, noise_seed0 + i) for i in range(count)])


def time_rows_kernel(L, H, d_b, d_x, S, st, reps=3):
    """Average duration of the FP32 row kernel of the G_h pass where it runs: every row-kernel launch of ``reps`` whole
    applies stamps its own start and end (adrt_plan_time_rows_ring: the dispatch's timestamps), and the average is
    over the FP32 G_h launches, i.e. all but the first (the FP64 backprojection) and the last ``refine`` (the FP64
    refinement steps) of each apply."""
    lib = L.load()
    n_launch = H.outer_deg  # one backprojection pass and outer_deg - 1 G_h passes
    refine = lib.adrt_plan_refine_steps(H._plan)  # FP64 passes at the end: min(refine, outer_deg - 1)
    evs = [(L.Event(), L.Event()) for _ in range(n_launch)]
    b0 = (L.P * n_launch)(*[a.handle for a, _ in evs])
    b1 = (L.P * n_launch)(*[z.handle for _, z in evs])
    H.apply_device(d_b, d_x, S, st, count=False)  # warm-up
    st.sync()
    ts = []
    try:
        for _ in range(reps):
            L.check(lib.adrt_plan_time_rows_ring(H._plan, b0, b1, n_launch), "plan_time_rows_ring")
            H.apply_device(d_b, d_x, S, st, count=False)
            st.sync()
            stamped = L.LL()
            L.check(lib.adrt_plan_time_rows_count(H._plan, ctypes.byref(stamped)), "plan_time_rows_count")
            # one row-kernel launch per pass, or the slice below would average the wrong launches
            if stamped.value != n_launch:
                raise RuntimeError(f"{stamped.value} row-kernel launches stamped in one apply, expected {n_launch}")
            t = [a.elapsed_ms(z) for a, z in evs]
            ts += t[1:n_launch - refine]
    finally:
        L.check(lib.adrt_plan_time_rows_ring(H._plan, None, None, 0), "plan_time_rows_ring")
    return float(np.mean(ts)), ts


def mixed_batch_leg(count=512, dense_cap=16):
    """Uneven batch: N drawn from {64, 128, 256, 512} with weights .4/.3/.2/.1 (default_rng(42)), phantoms 1000+i,
    noise 2000+i at 1%, FP32 hybrid, one graph launch, tuned row mapping per N."""
    B = pkg("batch")
    ns = B.draw_sizes(count)
    ph = pkg("phantom")
    R = pkg("radon")
    sinos = [ph.add_noise(R.adrt(ph.ellipses(1000 + i, 1 << n)), 0.01, 2000 + i) for i, n in enumerate(ns)]
    t0 = time.time()
    mb = B.MixedBatch(ns, precision="fp32", refine=1, dense_cap=dense_cap)
    setup_s = time.time() - t0
    mb.set_inputs(sinos)
    mb.capture()
    mb.launch()
    mb.main.sync()
    L = pkg("lib")
    e0, e1 = L.Event(), L.Event()
    reps = 3
    e0.record(mb.main)
    for _ in range(reps):
        mb.launch()
    e1.record(mb.main)
    ms = e0.elapsed_ms(e1) / reps
    counts = {int(1 << n): ns.count(n) for n in sorted(set(ns))}
    tuned = {int(1 << n): int(v[0]) for n, v in mb.tuning.items()}
    dense = {int(1 << n): int(v.dense) for n, v in mb.tuning.items()}
    # accuracy of the last replay: every sinogram against its FP64 plan (same setup), one batched FP64 call per N
    x = mb.outputs()
    mb.close()
    rel = fp64_relerr(ns, sinos, x, dense_cap)
    # the same batch with the explicit (cross-quadrant) level in FP64, as the workload's description has it (the row
    # mappings tuned above stay installed)
    mbx = B.MixedBatch(ns, precision="fp32", refine=1, dense_cap=dense_cap, tune=False, l1_fp64=True)
    mbx.set_inputs(sinos)
    mbx.capture()
    mbx.launch()
    mbx.main.sync()
    e0.record(mbx.main)
    for _ in range(reps):
        mbx.launch()
    e1.record(mbx.main)
    msx = e0.elapsed_ms(e1) / reps
    relx = fp64_relerr(ns, sinos, mbx.outputs(), dense_cap)
    mbx.close()
    log(f"mixed batch of {count}: {count * 1000.0 / ms:.0f} inverses/s (relerr max {np.max(rel):.2e}); explicit level "
        f"in fp64 {count * 1000.0 / msx:.0f} inverses/s (relerr max {np.max(relx):.2e})")
    return {"sinograms": count, "sizes": counts, "ms_per_launch": ms, "inverses_per_s": count * 1000.0 / ms,
            "setup_s": setup_s, "row_mapping_loge": tuned, "dense_placement": dense, "dense_cap": dense_cap,
            "relerr_fp32_vs_fp64_max": float(np.max(rel)), "relerr_fp32_vs_fp64_median": float(np.median(rel)),
            "l1_fp64": {"ms_per_launch": msx, "inverses_per_s": count * 1000.0 / msx,
                        "relerr_max": float(np.max(relx)), "relerr_median": float(np.median(relx))}}


def fp64_relerr(ns, sinos, xs, dense_cap, caps=None):
    """Relative error of each result xs[i] against the FP64 hybrid inverse of sinos[i] with the same setup scalars
    (N = 2^ns[i], dense cap caps[i] or dense_cap), run as one batched FP64 call per (N, cap)."""
    I = pkg("inverses")
    C = pkg("plancache")
    caps = caps or [dense_cap] * len(ns)
    groups = {}
    for i, key in enumerate(zip(ns, caps)):
        groups.setdefault(key, []).append(i)
    rel = np.empty(len(ns))
    for (n, cap), idx in sorted(groups.items()):
        lam, outer = C.hybrid_setup(n, dense_cap=cap)
        H = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=cap, max_batch=min(len(idx), 64))
        ref = H.apply_batch(np.stack([sinos[i] for i in idx]))
        H.close()
        for j, i in enumerate(idx):
            rel[i] = np.linalg.norm(xs[i] - ref[j]) / np.linalg.norm(ref[j])
    return rel


SERVE_BATCH = {7: 16, 8: 16, 9: 8, 10: 8}  # sinograms per launch by log2 N
SERVE_CONTINUOUS = "9,10"  # log2 N of the sizes of the second serving leg, served at outer-step granularity
SERVE_LANES = "9:8,10:8"


def serving_leg(rates, duration, sizes=(7, 8, 9, 10), weights=(0.4, 0.3, 0.2, 0.1), caps=(16, 32), slots_per_key=2,
                max_wait_s=0.25, continuous=(), lanes=None, tick_depth=None):
    """Poisson arrivals (default_rng(7)) of FP32 hybrid requests over mixed N and p at each rate; p50 / p99 latency
    and throughput of the served requests (upload and download inside the latency), how many were shed (not admitted
    into device memory within max_wait_s: the server at overload), and the relative error of every served result
    against the FP64 inverse of its sinogram with the same setup.  The sizes in ``continuous`` run one StepSlot per
    key (requests join a running batch between outer steps) with ``lanes[n]`` lanes instead of batches in slots."""
    S = pkg("serve")
    C = pkg("plancache")
    keys = [(n, c) for n in sizes for c in caps]
    mb = dict(SERVE_BATCH)
    mb.update({n: v for n, v in (lanes or {}).items() if n in continuous})
    t0 = time.time()
    setups = {k: C.hybrid_setup(k[0], dense_cap=k[1]) for k in keys}
    # FP64 refinement steps per (N, p): the fewest that keep the FP32 result within 1e-6 of FP64 on probe sinograms
    refine = {k: pkg("autotune").refine_for(k[0], k[1], *setups[k]) for k in keys}
    srv = S.Server(keys, slots_per_key=slots_per_key, max_batch=mb, setups=setups, max_wait_s=max_wait_s,
                   refine=refine, continuous=set(continuous), tick_depth=tick_depth)
    setup_s = time.time() - t0
    # the requests' sinograms arrive in page-locked host memory, as from a network or IPC receive buffer, so an
    # upload reads them where they are
    L = pkg("lib")
    pool = {}
    for n in sizes:
        pool[n] = []
        for b in make_sinos(n, 4, 3000, 4000):
            pb = L.PinnedArray(b.shape)
            pb.a[...] = b
            pool[n].append(pb)
    # FP64 results of every (sinogram, cap) the trace can draw: the served results are checked against them
    ref = {}
    I = pkg("inverses")
    for (n, c), (lam, outer) in setups.items():
        H = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=c, max_batch=len(pool[n]))
        for j, x in enumerate(H.apply_batch(np.stack([pb.a for pb in pool[n]]))):
            ref[(n, c, j)] = x
        H.close()
    # warm every slot once (first launches load code objects)
    warm = [(0.0, S.Request(-1 - i, s.n, s.cap, pool[s.n][0].a, 0.0, b_ptr=pool[s.n][0].ptr))
            for i, s in enumerate(srv.slots)]
    srv.run_trace(warm)
    out = []
    for rate in rates:
        srv.done.clear()
        srv.shed.clear()
        srv.host_s = dict.fromkeys(srv.host_s, 0)
        trace = S.poisson_trace(rate, duration, list(sizes), list(weights), list(caps), pool)
        t_run = time.perf_counter()
        done = srv.run_trace(trace)
        t_run = time.perf_counter() - t_run
        host = {k: (v / t_run if k != "ticks" else v) for k, v in srv.host_s.items()}
        rel, relk = [], {}
        for r in done:
            j = next(i for i, pb in enumerate(pool[r.n]) if pb.a is r.b)
            x64 = ref[(r.n, r.cap, j)]
            rel.append(float(np.linalg.norm(r.x - x64) / np.linalg.norm(x64)))
            key = f"{1 << r.n},{r.cap}"
            relk[key] = max(relk.get(key, 0.0), rel[-1])
        st = S.latency_stats(done, t_begin=min(r.t_arrive for _, r in trace), by_size=True) if done else {"count": 0}
        out.append({"rate_per_s": rate, "duration_s": duration, "offered": len(trace), **st, "shed": len(srv.shed),
                    "relerr_fp32_vs_fp64_max": max(rel) if rel else None,
                    "relerr_fp32_vs_fp64_median": float(np.median(rel)) if rel else None,
                    "relerr_fp32_vs_fp64_max_by_key": relk, "host_loop_fraction": host})
        log(f"serving at {rate:.0f}/s: p50 {st.get('p50_ms', 0):.1f} ms p99 {st.get('p99_ms', 0):.1f} ms, served "
            f"{st.get('throughput_per_s', 0):.0f}/s, shed {len(srv.shed)} of {len(trace)}, relerr max "
            f"{max(rel) if rel else float('nan'):.2e} ({', '.join(f'{k}: {v:.1e}' for k, v in sorted(relk.items()))})"
            f"; host loop: poll {host['poll']:.0%} launch {host['launch']:.0%} tick {host['tick']:.0%} "
            f"({host['ticks']} ticks)")
        for nn, e in st.get("by_size", {}).items():
            log(f"   N={nn}: {e['count']} served, p50 {e['p50_ms']:.1f} p99 {e['p99_ms']:.1f} ms; mean wait "
                f"{e.get('wait_ms', 0):.1f} ms, in slot {e.get('slot_ms', 0):.1f} ms")
    res = {"mix": "N in {128,256,512,1024} w .4/.3/.2/.1, p in {16,32}", "slots": len(srv.slots),
           "max_batch": {str(1 << n): b for n, b in mb.items()},
           "continuous": [1 << n for n in sorted(continuous)],
           "tick_depth": {str(1 << s.n): s.depth for s in srv.slots if getattr(s, "continuous", False)},
           "max_wait_s": max_wait_s, "fp64_refine_steps": {f"{1 << n},{c}": r for (n, c), r in refine.items()},
           "device_GiB": srv.device_bytes / 2**30, "setup_s": setup_s, "rates": out}
    srv.close()
    for pbs in pool.values():
        for pb in pbs:
            pb.free()
    return res


def large_leg(n=11, batch=16, dense_cap=16):
    """N=2048 hybrid, batch 16: ellipse phantoms (seeds 5000+i) + 0.5% noise (seeds 6000+i); the FP32 plan (TwoSum,
    FP64 dense rows / outer / refinement) against the FP64 plan on the same batch, both one device call."""
    L = pkg("lib")
    I = pkg("inverses")
    N = 1 << n
    t0 = time.time()
    H64 = I.FastHybridHip(n, dense_cap=dense_cap, precision="fp64", max_batch=batch)
    H32 = I.FastHybridHip(n, lam_max=H64.lam_max, dense_cap=dense_cap, precision="fp32", max_batch=batch,
                          outer=(H64.lo, H64.hi, H64.outer_deg), refine=1)
    pkg("autotune").tune(H32, batch, persist=False)
    setup_s = time.time() - t0
    Bn = make_sinos(n, batch, 5000, 6000, level=0.005)
    st = L.Stream()
    d_b = L.DeviceArray.from_host(Bn, st)
    d_x = L.DeviceArray((batch, N, N))
    e0, e1 = L.Event(), L.Event()
    out, x = {}, {}
    # the precision of the workload's description: FP32 bulk with the explicit (cross-quadrant) level in FP64
    H32x = I.FastHybridHip(n, lam_max=H64.lam_max, dense_cap=dense_cap, precision="fp32", max_batch=batch,
                           outer=(H64.lo, H64.hi, H64.outer_deg), refine=1, l1_fp64=True)
    for name, H in (("fp64", H64), ("fp32", H32), ("fp32_l1_fp64", H32x)):
        H.apply_device(d_b, d_x, batch, st, count=False)
        st.sync()
        e0.record(st)
        H.apply_device(d_b, d_x, batch, st, count=False)
        e1.record(st)
        ms = e0.elapsed_ms(e1)
        x[name] = d_x.download(st)
        out[f"{name}_inverses_per_s"] = batch * 1000.0 / ms
    rel = [float(np.linalg.norm(x["fp32"][i] - x["fp64"][i]) / np.linalg.norm(x["fp64"][i])) for i in range(batch)]
    relx = [float(np.linalg.norm(x["fp32_l1_fp64"][i] - x["fp64"][i]) / np.linalg.norm(x["fp64"][i]))
            for i in range(batch)]
    res = {"workload": f"hybrid inverse N={N}, batch {batch}, dense_cap={dense_cap}, 0.5% noise", **out,
           "relerr_fp32_vs_fp64_max": max(rel), "relerr_fp32_vs_fp64_median": float(np.median(rel)),
           "relerr_fp32_l1_fp64_vs_fp64_max": max(relx),
           "outer_deg": H64.outer_deg, "n_dense": H64.eng.n_dense, "table_GiB": H64.table_bytes / 2**30,
           "setup_s": setup_s}
    H32x.close()
    H32.close()
    H64.close()
    log(f"large N={N}: fp32 {out['fp32_inverses_per_s']:.2f}/s, fp64 {out['fp64_inverses_per_s']:.2f}/s, "
        f"relerr max {max(rel):.2e}; fp32 with the explicit level in fp64 {out['fp32_l1_fp64_inverses_per_s']:.2f}/s, "
        f"relerr max {max(relx):.2e}")
    return res


def rows_only(a):
    """Short workload for hardware-counter runs: K launches of the row kernel of the G_h pass at the headline
    configuration, then one device-to-device copy of a known size (the counter calibration)."""
    L = pkg("lib")
    I = pkg("inverses")
    n, N, B = a.n, 1 << a.n, a.batch
    cap = a.dense_cap if a.dense_cap >= 0 else None
    H64 = I.FastHybridHip(n, dense_cap=cap, precision="fp64", max_batch=1)
    H32 = I.FastHybridHip(n, lam_max=H64.lam_max, dense_cap=cap, precision="fp32", max_batch=B,
                          outer=(H64.lo, H64.hi, H64.outer_deg), refine=a.refine)
    if a.mapping:
        # counter runs serialise and slow the kernels, so the tuner's timings there are not the real ones: take the
        # mapping the timed run chose
        loge, dense = (int(v) for v in a.mapping.split(","))
        L.check(L.load().adrt_rows_config_set(n, 1, loge), "rows_config_set")
        L.check(L.load().adrt_dense_config_set(n, 1, dense), "dense_config_set")
    else:
        pkg("autotune").tune(H32, B, persist=False)
    st = L.Stream()
    d_x = L.DeviceArray.from_host(np.random.default_rng(0).standard_normal((B, N, N)), st)
    d_y = L.DeviceArray((B, N, N))
    H32.operator_device(d_x, d_y, B, st)
    for _ in range(5):
        L.check(L.load().adrt_plan_rows_pass(H32._plan, 1, B, st.handle), "rows_pass")
    copy_bytes = 1 << 28
    src = L.DeviceArray((copy_bytes // 8,))
    dst = L.DeviceArray((copy_bytes // 8,))
    src.zero(st)
    L.check(L.load().adrt_memcpy(dst.ptr, src.ptr, copy_bytes, 3, st.handle), "d2d")
    st.sync()
    fl, nb = rows_kernel_work(H32, B, H32.ops_op())
    print(json.dumps({"rows_kernel_launches": 5, "flops_per_launch": fl, "alg_bytes_per_launch": nb,
                      "calibration_copy_bytes": copy_bytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    # sinograms per launch of the batched FP32 path.  Measured on one box (inverses/s at N=256): 32: 3355, 64: 3544,
    # 128: 3668, 256: 3655 -- 128 is where the per-launch costs stop shrinking
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--dense-cap", type=int, default=16)
    ap.add_argument("--refine", type=int, default=1,
                    help="FP64 outer steps at the end of the FP32 apply (1: the final residual and update in FP64)")
    ap.add_argument("--plain-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-ablations", action="store_true",
                    help="skip the precision ablations (no TwoSum, G_1 / explicit level in double)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rates", default="200,500,1000,2000,5000",
                    help="Poisson arrival rates (req/s), comma separated")
    ap.add_argument("--serve-duration", type=float, default=4.0,
                    help="seconds of Poisson arrivals per rate (4 s at 200 req/s: about 800 requests behind the p99)")
    ap.add_argument("--serve-max-wait", type=float, default=0.25,
                    help="a request not admitted into device memory within this many seconds is shed")
    ap.add_argument("--serve-continuous", default=SERVE_CONTINUOUS,
                    help="a second serving leg, with these sizes (log2 N, comma separated) served by continuous "
                         "batching at outer-step granularity (requests join a running batch between outer steps); "
                         "empty: no second leg")
    ap.add_argument("--serve-lanes", default=SERVE_LANES,
                    help="lanes of the continuous sizes as n:lanes, comma separated")
    ap.add_argument("--serve-batch", default=None,
                    help="sinograms per launch of the batched sizes as n:batch, comma separated (default: "
                         + ",".join(f"{k}:{v}" for k, v in SERVE_BATCH.items()) + ")")
    ap.add_argument("--serve-tick-depth", type=int, default=None,
                    help="ticks kept queued per continuous slot (default 16)")
    ap.add_argument("--rows-only", action="store_true")
    ap.add_argument("--mapping", default=None,
                    help="with --rows-only: 'loge,dense' (row mapping and dense placement) instead of autotuning")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the N=2048 batch-16 FP32 vs FP64 leg (about 45 s of setup and 50 GiB of dense tables)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the sinogram batch would shard by "
              "sinogram across GPUs with no exchange, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    L = pkg("lib")
    I = pkg("inverses")
    if L.load().adrt_device_count() < 1:
        raise SystemExit("no GPU visible")
    if a.rows_only:
        rows_only(a)
        return
    n, N, B = a.n, 1 << a.n, a.batch
    cap = a.dense_cap if a.dense_cap >= 0 else None

    t0 = time.time()
    H64 = I.FastHybridHip(n, dense_cap=cap, precision="fp64", max_batch=1)
    setup_s = time.time() - t0
    outer = (H64.lo, H64.hi, H64.outer_deg)
    H32 = I.FastHybridHip(n, lam_max=H64.lam_max, dense_cap=cap, precision="fp32", max_batch=B, outer=outer,
                          refine=a.refine)
    log(f"setup {setup_s:.1f}s: D={H64.outer_deg} sum(deg)/N={H64.deg.sum() / N:.1f} dense rows={H64.eng.n_dense}")
    # the row kernel's mapping for this (N, FP32) chosen by the project's autotuner on the plan's own row pass
    tuned = pkg("autotune").tune(H32, B, persist=False)
    loge, tune_ms = tuned
    log(f"autotuned row mapping: loge={loge} ({', '.join(f'{k}: {v * 1e3:.0f} us' for k, v in tune_ms.items())}); "
        f"dense placement {tuned.dense} ({', '.join(f'{k}: {v * 1e3:.0f} us' for k, v in tuned.dense_times.items())}); "
        f"FP64 refinement passes loge={tuned.f64} "
        f"({', '.join(f'{k}: {v * 1e3:.0f} us' for k, v in tuned.f64_times.items())}); one sinogram: dense placement "
        f"{tuned.dense1} ({', '.join(f'{k}: {v * 1e3:.0f} us' for k, v in tuned.dense1_times.items())})")

    # inputs: ellipse phantoms (seeds 1000+i) through the GPU forward ADRT, plus 1% noise (seeds 2000+i)
    Bn = make_sinos(n, B, 1000, 2000)
    st = L.Stream()
    d_b = L.DeviceArray.from_host(Bn, st)
    d_x32 = L.DeviceArray((B, N, N))
    d_x64 = L.DeviceArray((1, N, N))

    # plain FP64 path: one sinogram per call, no graph, no batching
    x64 = np.empty((B, N, N))
    for i in range(B):
        H64.apply_device(d_b.ptr + i * Bn[0].nbytes, d_x64, 1, st, count=False)
        st.sync()
        x64[i] = d_x64.download(st)[0]
    e0, e1 = L.Event(), L.Event()
    e0.record(st)
    for r in range(a.plain_reps):
        H64.apply_device(d_b.ptr + (r % B) * Bn[0].nbytes, d_x64, 1, st, count=False)
    e1.record(st)
    plain_ms = e0.elapsed_ms(e1) / a.plain_reps
    plain_ips = 1000.0 / plain_ms
    log(f"plain fp64 {plain_ms:.2f} ms per inverse")

    def batched(H):
        with L.Graph(st) as g:
            H.apply_device(d_b, d_x32, B, st, count=False)
        for _ in range(a.warmup):
            g.launch(st)
        st.sync()
        t_wall = time.time()
        e0.record(st)
        for _ in range(a.steps):
            g.launch(st)
        e1.record(st)
        ms = e0.elapsed_ms(e1) / a.steps
        wall = (time.time() - t_wall) * 1000.0 / a.steps
        # the result of the last timed replay (the graph recomputes it from the resident inputs every time)
        x = d_x32.download(st)
        rel = np.linalg.norm((x - x64).reshape(B, -1), axis=1) / np.linalg.norm(x64.reshape(B, -1), axis=1)
        return ms, wall, rel

    step_ms, wall_ms, rel = batched(H32)
    ips = B * 1000.0 / step_ms
    log(f"batched fp32 {step_ms:.2f} ms per step of {B}: {ips:.0f} inverses/s, relerr max {rel.max():.2e}")

    # precision ablations on the same batch (each its own plan and graph), skipped with --no-ablations
    abl = {}
    variants = (("fp32_plain", dict(precision="fp32_plain")),  # without the TwoSum compensation
                ("fp32_g1_fp64", dict(g1_fp64=True)),  # the G_1 inner iteration in double in the FP32 steps too
                # the explicit level 1 with the FFTs and combine around it in double, and that together with G_1 in
                # double (FP64 for the explicit cross-quadrant level, the table GEMM and the residual)
                ("fp32_l1_fp64", dict(l1_fp64=True)),
                ("fp32_l1_g1_fp64", dict(l1_fp64=True, g1_fp64=True)))
    for key, kw in (() if a.no_ablations else variants):
        kw = {"precision": "fp32", **kw}
        Hx = I.FastHybridHip(n, lam_max=H64.lam_max, dense_cap=cap, max_batch=B, outer=outer, refine=a.refine, **kw)
        ms_x, _, rel_x = batched(Hx)
        Hx.close()
        abl[key] = {"inverses_per_s": B * 1000.0 / ms_x, "relerr_max": float(rel_x.max()),
                    "relerr_median": float(np.median(rel_x))}
        log(f"{key}: {abl[key]['inverses_per_s']:.0f} inverses/s, relerr max {rel_x.max():.2e}")

    # dominant kernel: the register row kernel of the G_h pass, timed by its own dispatch inside whole applies
    ops = H32.ops_op()
    d_tmp = L.DeviceArray((B, N, N))
    rows_ms, rows_ts = time_rows_kernel(L, H32, d_b, d_tmp, B, st)
    fl, nb = rows_kernel_work(H32, B, ops)
    dfl = dense_work(H32, B)[0]
    achieved = fl / (rows_ms * 1e-3)
    traffic, traffic_map = pmc_traffic(H32.N, B)  # counted at the mapping the profile names (row loge, dense placement)
    roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK["fp32"], "traffic": traffic, "traffic_mapping": traffic_map,
            "frac_nonpacked": achieved / PEAK["fp32_nonpacked"],
            "kernel": f"rows_reg_kernel<float, {n}, e, comp> (G_h pass, with the dense MFMA tiles)",
            "kernel_ms": rows_ms, "kernel_ms_median": float(np.median(rows_ts)), "kernel_ms_min": float(np.min(rows_ts)),
            "flops_per_launch": fl, "alg_bytes_per_launch": nb,
            "units_per_launch": f"{2 * B} rows x {H32.L - H32.eng.n_dense} non-dense frequencies (+ {2 * B} rows x "
            f"{H32.eng.n_dense} dense frequencies as FP64 MFMA tiles)",
            "dense_fp64_mfma_flops_per_launch": dfl, "dense_fp64_mfma_tflops": dfl / (rows_ms * 1e-3) / 1e12}
    log(f"row kernel {rows_ms * 1e3:.1f} us: {achieved / 1e12:.1f} TF/s")

    mixed = None
    if not a.no_mixed:
        mixed = mixed_batch_leg()
        log(f"mixed batch: {mixed['inverses_per_s']:.0f} inverses/s")
    serving = serving_cont = None
    if not a.no_serve:
        rates = [float(r) for r in a.serve_rates.split(",")]
        if a.serve_batch:
            SERVE_BATCH.update({int(k): int(v) for k, v in (e.split(":") for e in a.serve_batch.split(",") if e)})
        serving = serving_leg(rates, a.serve_duration, max_wait_s=a.serve_max_wait)
        cont = [int(v) for v in a.serve_continuous.split(",") if v.strip()]
        if cont:
            lanes = {int(k): int(v) for k, v in (e.split(":") for e in a.serve_lanes.split(",") if e.strip())}
            log(f"serving again with N in {[1 << n for n in cont]} at outer-step granularity")
            serving_cont = serving_leg(rates, a.serve_duration, max_wait_s=a.serve_max_wait, continuous=cont,
                                       lanes=lanes, tick_depth=a.serve_tick_depth)

    large = None if a.no_large else large_leg()

    cpu = None
    if not a.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        res = cpu_baseline(H64, cap, Bn[0], threads, a.cpu_budget)
        if res is not None:
            tmin, reps, xr = res
            from oracle import refcpp

            cpu = {"value": 1.0 / tmin, "unit": "inverses/s", "cores": threads, "kind": "reference",
                   "cpu_model": refcpp.cpu_info()[0], "build": f"g++ -O3 -march={refcpp.MARCH} -fopenmp",
                   "sample": f"1 hybrid inverse at N={N} (dense_cap={cap}, D={H64.outer_deg}), tmin of {reps}; "
                             f"reference cpp/oblique.cpp row kernels (OpenMP, {threads} threads) under the oracle's "
                             "NumPy host (FFTs, packing, outer loop)",
                   "relerr_vs_gpu_fp64": float(np.linalg.norm(xr - x64[0]) / np.linalg.norm(x64[0]))}
            log(f"cpu baseline {tmin:.2f} s per inverse")
            # one thread, as the paper times the engine (README.md:65-66): a single run
            res1 = cpu_baseline(H64, cap, Bn[0], 1, budget_s=0.0)
            if res1 is not None:
                cpu["one_thread"] = {"value": 1.0 / res1[0], "unit": "inverses/s", "cores": 1,
                                     "sample": f"1 hybrid inverse at N={N}, one run"}
                log(f"cpu baseline, one thread: {res1[0]:.2f} s per inverse")

    line = {
        "metric": "hybrid inverses/s at N=256, FP32 vs FP64 (relerr<=1e-5); p99 latency mixed-N",
        "value": ips,
        "unit": "inverses/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 offset FFTs, rows (explicit level 1 included; TwoSum-compensated Chebyshev) and G1 inner "
                 "iteration; f64 dense table GEMM, outer iteration and residual, backprojection and last refine pass",
        "data": "synthetic: ellipse phantoms (seeds 1000+i) + 1% Gaussian noise (seeds 2000+i)",
        "config": {"workload": f"hybrid inverse N={N}, batch {B}, dense_cap={cap}, eps_in=0.1, tau=0.2N, g1_deg=14",
                   "n": n, "batch": B, "dense_cap": cap, "outer_deg": H64.outer_deg, "fp64_refine_steps": a.refine,
                   "sum_deg_over_N": float(H64.deg.sum() / N), "n_dense": H64.eng.n_dense,
                   "row_mapping_loge": int(loge), "dense_placement": int(tuned.dense),
                   "fp64_refine_row_mapping_loge": tuned.f64},
        "plain_fp64_inverses_per_s": plain_ips,
        "speedup_vs_plain_fp64": ips / plain_ips,
        "relerr_fp32_vs_fp64_max": float(rel.max()),
        "relerr_fp32_vs_fp64_median": float(np.median(rel)),
        **{k: abl.get(k) for k, _ in variants},
        "wall_ms_per_step": wall_ms,
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed_batch": mixed,
        "serving": serving,
        "serving_continuous": serving_cont,
        "large": large,
    }
    if cpu:
        line["speedup_vs_cpu"] = ips / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
