"""Mixed-N batches in one graph, the autotuner, and the serving loop on the GPU."""

import numpy as np
import pytest

from conftest import pkg, relmax

pytestmark = pytest.mark.gpu


def _sinos(ns, seed0=300, noise=0.01):
    ph = pkg("phantom")
    R = pkg("radon")
    return [ph.add_noise(R.adrt(ph.ellipses(seed0 + i, 1 << n)), noise, seed0 + 1000 + i) for i, n in enumerate(ns)]


def test_mixed_batch_graph(gpu):
    B = pkg("batch")
    I = pkg("inverses")
    C = pkg("plancache")
    ns = [6, 7, 6, 8, 7, 6, 8]
    b = _sinos(ns)
    mb = B.MixedBatch(ns, precision="fp32", refine=2, dense_cap=12)
    mb.set_inputs(b)
    mb.run()
    x1 = mb.outputs()
    mb.run()  # a second replay of the same graph gives the same result
    x2 = mb.outputs()
    for i, n in enumerate(ns):
        assert np.array_equal(x1[i], x2[i])
        lam, outer = C.hybrid_setup(n, dense_cap=12)
        H = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=12)
        ref = H.apply(b[i])
        assert np.linalg.norm(x1[i] - ref) / np.linalg.norm(ref) < 1e-5, (i, n)
        # the same FP32 plan on this sinogram alone gives the same result as in the batch
        alone = mb.plans[n].apply(b[i])
        assert relmax(x1[i], alone) < 1e-12
    for n, (best, times) in mb.tuning.items():
        assert best in times and times[best] == min(times.values())


def test_server_results_same_alone_and_in_mix(gpu):
    S = pkg("serve")
    C = pkg("plancache")
    keys = [(6, 8), (7, 8), (7, 16)]
    setups = {k: C.hybrid_setup(k[0], dense_cap=k[1]) for k in keys}
    srv = S.Server(keys, slots_per_key=2, max_batch=3, setups=setups)
    ns = [6, 7, 7, 6, 7, 6, 7, 7, 6, 7]
    b = _sinos(ns, seed0=500)
    reqs = [S.Request(i, n, 8 if i % 2 else 16 if n == 7 else 8, b[i], 0.0) for i, n in enumerate(ns)]
    trace = [(0.0005 * i, r) for i, r in enumerate(reqs)]
    bad = S.Request(99, 9, 8, np.zeros((4, 1023, 512)), 0.0)
    done = srv.run_trace(sorted(trace + [(0.001, bad)], key=lambda t: t[0]))
    assert bad.error and bad not in done
    assert len(done) == len(reqs) and all(r.error is None for r in reqs)
    # alone: each request through a fresh single-request call of the same slot type
    for r in reqs:
        slot = next(s for s in srv.slots if (s.n, s.cap) == (r.n, r.cap))
        alone = slot.plan.apply(r.b)
        assert relmax(r.x, alone) < 1e-12, r.rid
    st = S.latency_stats(done)
    assert st["count"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_context_budget_and_stream(gpu):
    """A plan created in a context charges its bytes to the budget, fails with -3 beyond it (allocating nothing), keeps
    the context alive (-4 on destroy) and gives the same bits as a plan outside any context, also on the context's
    stream."""
    L = pkg("lib")
    I = pkg("inverses")
    C = pkg("plancache")
    n = 6
    lam, outer = C.hybrid_setup(n, dense_cap=8)
    kw = dict(lam_max=lam, outer=outer, dense_cap=8, precision="fp32", refine=1, max_batch=2)
    b = np.stack(_sinos([n, n], seed0=700))
    ref = I.FastHybridHip(n, **kw)
    want = ref.apply_batch(b)
    tiny = L.Context(budget=4096)
    with pytest.raises(L.HipError, match=r"\(-3\)"):
        I.FastHybridHip(n, ctx=tiny, **kw)
    assert tiny.usage() == (0, 4096)
    tiny.close()
    N = 1 << n
    staging = 8 * 2 * (2 * N * N + 4 * (2 * N - 1) * N)  # two image arrays and one sinogram array, max_batch 2
    ctx = L.Context(budget=ref.plan_bytes + staging + 4096)
    H = I.FastHybridHip(n, ctx=ctx, **kw)
    used, budget = ctx.usage()
    assert used == H.plan_bytes + H._dimg.nbytes + H._dimg2.nbytes + H._dsino.nbytes and used <= budget
    with pytest.raises(L.HipError, match=r"\(-4\)"):
        ctx.close()
    with pytest.raises(L.HipError, match=r"\(-3\)"):  # a second plan does not fit in what is left
        I.FastHybridHip(n, ctx=ctx, **kw)
    assert ctx.usage()[0] == used
    assert np.array_equal(H.apply_batch(b), want)
    assert np.array_equal(H.apply_batch(b, stream=ctx.stream), want)
    H.close()
    ref.close()
    assert ctx.usage()[0] == 0
    ctx.close()


def test_independent_contexts_run_concurrently(gpu):
    """Two host threads, each with its own context, plan and stream, apply at the same time and get the bits of the
    same plans run one after the other."""
    import threading

    L = pkg("lib")
    I = pkg("inverses")
    C = pkg("plancache")
    n = 6
    lam, outer = C.hybrid_setup(n, dense_cap=8)
    kw = dict(lam_max=lam, outer=outer, dense_cap=8, precision="fp32", refine=1, max_batch=2)
    bs = [np.stack(_sinos([n, n], seed0=800 + 10 * t)) for t in range(2)]
    ctxs = [L.Context() for _ in range(2)]
    plans = [I.FastHybridHip(n, ctx=c, **kw) for c in ctxs]
    want = [plans[t].apply_batch(bs[t], stream=ctxs[t].stream) for t in range(2)]
    got = [None, None]
    errs = []

    def work(t):
        try:
            for _ in range(3):
                got[t] = plans[t].apply_batch(bs[t], stream=ctxs[t].stream)
        except Exception as e:  # reported below, in the main thread
            errs.append(e)

    th = [threading.Thread(target=work, args=(t,)) for t in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=120)
    assert not errs and all(not x.is_alive() for x in th)
    for t in range(2):
        assert np.array_equal(got[t], want[t])
    for p in plans:
        p.close()
    for c in ctxs:
        c.close()
