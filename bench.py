"""Benchmark: hybrid ADRT inverses per second at N = 256 on one MI355X.

    python bench.py [--steps K] [--warmup W] [--batch B] [--n 8] [--dense-cap 16] [--no-cpu]

A step is one batched hybrid inverse (FP32 row engine, FP64 dense table / preconditioner / outer iteration) of B
noisy ellipse-phantom sinograms that are resident in HBM, replayed from one hipGraph.  ``value`` is inverses/s
(B * K / elapsed).  Beside it the line reports the plain FP64 GPU path (one sinogram per call, no graph, no batching),
the relative error of the FP32 result against the FP64 one, the roofline of the dominant kernel (the fused row kernel,
timed with HIP events on its own stream) and the reference C++ row engine timed on the host cores.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import importlib  # noqa: E402

PEAK = {"fp32": 157.3e12, "fp64": 78.6e12, "hbm": 8.0e12}


def pkg(sub):
    return importlib.import_module(f"adrt-pinv_amd.{sub}")


def rows_flops(H, S, ops):
    """Algorithmic FLOPs of one row-kernel launch over S sinograms (R = 4N rows each):
    8N per butterfly level-row, (d-1)(16 N U + 18 N) + 2N per Chebyshev row, 8 N^2 per dense row."""
    n, N, Lr = H.n, H.N, H.L
    dense = np.zeros(Lr, dtype=bool)
    dense[H.eng.dense_k] = True
    lev = sum(hi - lo + 1 for o, lo, hi in ops if o in ("B", "BT"))
    f = 0.0
    for o, lo, hi in ops:
        if o != "P":
            continue
        U = hi - lo + 1
        d = H.deg.astype(float)
        cheb = np.where(dense, 0.0, (np.maximum(d - 1, 0)) * (16 * N * U + 18 * N) + 2 * N)
        f += cheb.sum() + dense.sum() * 8.0 * N * N
    f += Lr * 8.0 * N * lev
    return 2 * S * f  # two packed pairs per frequency


def rows_bytes(H, S, elem):
    return 2.0 * (2 * H.L * S * H.N * elem) + H.eng.n_dense * H.N * H.N * 16.0


def cpu_baseline(H, cap, b, threads, budget_s=20.0):
    """The reference's C++ row engine (cpp/oblique.cpp built into oracle/_ref) under the oracle's NumPy host, same
    setup scalars, one sinogram; tmin protocol of exp_compiled_tables.py:20-27 within a time budget."""
    from oracle import inverses as oiv
    from oracle import refcpp

    if not refcpp.available():
        return None

    class RefHybrid(oiv.Hybrid):
        engine_cls = refcpp.RefRows

    R = RefHybrid(H.n, lam_max=H.lam_max, dense_cap=cap, threads=threads, outer=(H.lo, H.hi, H.outer_deg))
    ts = []
    t_all = time.time()
    while time.time() - t_all < budget_s or len(ts) < 1:
        t0 = time.time()
        x = R.apply(b)
        ts.append(time.time() - t0)
        if len(ts) >= 3:
            break
    return min(ts), len(ts), x


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--dense-cap", type=int, default=16)
    ap.add_argument("--refine", type=int, default=2, help="FP64 refinement steps at the end of the FP32 apply")
    ap.add_argument("--plain-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the sinogram batch would shard by "
              "sinogram across GPUs with no exchange, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    L = pkg("lib")
    I = pkg("inverses")
    ph = pkg("phantom")
    R = pkg("radon")
    L.load()
    if L.load().adrt_device_count() < 1:
        raise SystemExit("no GPU visible")
    n, N, B = a.n, 1 << a.n, a.batch
    cap = a.dense_cap if a.dense_cap >= 0 else None

    t0 = time.time()
    H64 = I.FastHybridHip(n, dense_cap=cap, precision="fp64", max_batch=1)
    setup_s = time.time() - t0
    H32 = I.FastHybridHip(n, lam_max=H64.lam_max, dense_cap=cap, precision="fp32", max_batch=B,
                          outer=(H64.lo, H64.hi, H64.outer_deg), refine=a.refine)
    print(f"# setup {setup_s:.1f}s: D={H64.outer_deg} sum(deg)/N={H64.deg.sum() / N:.1f} dense rows={H64.eng.n_dense}",
          file=sys.stderr, flush=True)

    # inputs: ellipse phantoms (seeds 1000+i) through the GPU forward ADRT, plus 1% noise (seeds 2000+i)
    X = np.stack([ph.ellipses(1000 + i, N) for i in range(B)])
    B0 = R.adrt(X)
    Bn = np.stack([ph.add_noise(B0[i], 0.01, 2000 + i) for i in range(B)])
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

    # batched FP32 path in one hipGraph
    with L.Graph(st) as g:
        H32.apply_device(d_b, d_x32, B, st, count=False)
    for _ in range(a.warmup):
        g.launch(st)
    st.sync()
    x32 = d_x32.download(st)
    rel = np.linalg.norm((x32 - x64).reshape(B, -1), axis=1) / np.linalg.norm(x64.reshape(B, -1), axis=1)
    t_wall = time.time()
    e0.record(st)
    for _ in range(a.steps):
        g.launch(st)
    e1.record(st)
    step_ms = e0.elapsed_ms(e1) / a.steps
    wall_ms = (time.time() - t_wall) * 1000.0 / a.steps
    ips = B * 1000.0 / step_ms

    # dominant kernel: the fused row kernel of the G_h pass, on the same stream, events around K launches
    ops = H32.ops_op()
    d_tmp = L.DeviceArray((B, N, N))
    H32.operator_device(d_x32, d_tmp, B, st)  # fills the plan's rows with a real G_h pass
    K = 10
    e0.record(st)
    for _ in range(K):
        L.check(L.load().adrt_plan_rows_pass(H32._plan, 1, B, st.handle), "rows_pass")
    e1.record(st)
    rows_ms = e0.elapsed_ms(e1) / K
    fl = rows_flops(H32, B, ops)
    achieved = fl / (rows_ms * 1e-3)
    roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK["fp32"], "traffic": None, "kernel": "rows_kernel<float,float>",
            "kernel_ms": rows_ms, "flops_per_launch": fl, "bytes_per_launch": rows_bytes(H32, B, 8)}

    cpu = None
    if not a.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        res = cpu_baseline(H64, cap, Bn[0], threads, a.cpu_budget)
        if res is not None:
            tmin, reps, xr = res
            cpu = {"value": 1.0 / tmin, "unit": "inverses/s", "cores": threads, "kind": "reference",
                   "sample": f"1 hybrid inverse at N={N} (dense_cap={cap}), tmin of {reps} after none; reference "
                             f"cpp/oblique.cpp kernels (OpenMP, {threads} threads) under the oracle's NumPy host",
                   "relerr_vs_gpu_fp64": float(np.linalg.norm(xr - x64[0]) / np.linalg.norm(x64[0]))}

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
        "dtype": "fp32 rows + fp64 dense/precond/outer",
        "data": "synthetic: ellipse phantoms (seeds 1000+i) + 1% Gaussian noise (seeds 2000+i)",
        "config": {"workload": f"hybrid inverse N={N}, batch {B}, dense_cap={cap}, eps_in=0.1, tau=0.2N, g1_deg=14",
                   "n": n, "batch": B, "dense_cap": cap, "outer_deg": H64.outer_deg, "fp64_refine_steps": a.refine,
                   "sum_deg_over_N": float(H64.deg.sum() / N), "n_dense": H64.eng.n_dense},
        "plain_fp64_inverses_per_s": plain_ips,
        "speedup_vs_plain_fp64": ips / plain_ips,
        "relerr_fp32_vs_fp64_max": float(rel.max()),
        "relerr_fp32_vs_fp64_median": float(np.median(rel)),
        "wall_ms_per_step": wall_ms,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = ips / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
