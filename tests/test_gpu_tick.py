"""Continuous batching at outer-step granularity on the GPU (adrt_plan_tick, serve.StepSlot).

Sinograms join a running batch between outer steps and leave after their last one.  Each result must be, bit for
bit, the plan's whole apply of that sinogram alone: the tick runs the apply's arithmetic step for step, per lane.
Covered: the FP32 hybrid plan with its FP64 refinement step (three kinds of step in one tick), the FP64 hybrid plan,
the lorr preconditioner and a G_1 degree above the fused kernel's halo (the unfused update, per lane), the explicit
level and G_1 in FP64 inside FP32 steps, FP32 without TwoSum or refinement, a plan whose max_batch gives its FP32
passes the paired inverse FFT (N = 256, 64), L_p, and the server with continuous keys at N = 512 and 1024 against FP64.
"""

import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _sino(n, seed, noise_seed, level=0.01):
    ph = pkg("phantom")
    R = pkg("radon")
    return ph.add_noise(R.adrt(ph.ellipses(seed, 1 << n)), level, noise_seed)


def run_staggered(plan, B, starts):
    """Sinograms B[i] joining at tick starts[i] (as lanes free up) through a StepSlot on ``plan``; returns the
    results in order, each checked by the slot's own bookkeeping."""
    S = pkg("serve")
    slot = S.StepSlot(plan.n, 0, dev=S.StepLanes(plan.n, 0, None, None, None, None, None, 4, plan=plan))
    reqs = [S.Request(i, plan.n, 0, B[i], 0.0) for i in range(len(B))]
    waiting = list(range(len(B)))
    t = 0
    while len([r for r in reqs if r.x is not None]) < len(reqs):
        new = [i for i in waiting if starts[i] <= t][: slot.free_lanes()]
        for i in new:
            waiting.remove(i)
        if slot.lanes or new:
            slot.tick([reqs[i] for i in new])
        slot.streams.compute.sync()
        slot.poll()
        t += 1
        assert t < 10 * plan.outer_deg * len(B)
    slot.close()  # the plan stays the caller's
    return [r.x for r in reqs]


@pytest.mark.parametrize("n,kw", [
    (8, dict(precision="fp32", refine=1, dense_cap=16)),
    (6, dict(precision="fp32", refine=2, dense_cap=8)),
    (6, dict(precision="fp64")),
    (6, dict(precision="fp64", precond="lorr")),
    (6, dict(precision="fp32", refine=1, g1_deg=20)),
    (7, dict(precision="fp32", refine=1, dense_cap=8, l1_fp64=True, g1_fp64=True)),
    (7, dict(precision="fp32_plain", refine=0, dense_cap=8)),
    (8, dict(precision="fp32", refine=1, dense_cap=16, max_batch=64)),
])
def test_tick_lanes_same_bits_as_apply(gpu, n, kw):
    """Seven sinograms on four lanes, joining at ticks 0, 0, 2, 3, 3 and later as lanes free: lanes at the plan's
    own steps, at FP64 refinement steps and at step 0 share ticks, lanes leave and move down, and every result is
    the plan's apply of that sinogram alone, bit for bit."""
    I = pkg("inverses")
    P = I.FastHybridHip(n, **{"max_batch": 4, **kw})
    D = P.outer_deg
    B = [_sino(n, 300 + i, 400 + i) for i in range(7)]
    starts = [0, 0, 2, 3, 3, D - 1, D]
    xs = run_staggered(P, B, starts)
    for i in range(len(B)):
        assert np.array_equal(xs[i], P.apply(B[i])), i
    P.close()


def test_tick_lp_same_bits_as_apply(gpu):
    """L_p (no preconditioner: x_0 = a y and the plain Chebyshev update, per lane) through the same lanes."""
    I = pkg("inverses")
    n = 6
    P = I.FastLpHip(n, max_batch=3, precision="fp32", refine=1, dense_cap=8)
    B = [_sino(n, 500 + i, 600 + i) for i in range(5)]
    xs = run_staggered(P, B, [0, 1, 1, 4, 5])
    for i in range(len(B)):
        assert np.array_equal(xs[i], P.apply(B[i])), i
    P.close()


def test_tick_rejects_bad_lane_orders(gpu):
    """adrt_plan_tick checks its lanes before any launch: steps in range, oldest first, sinograms for step 0."""
    L = pkg("lib")
    I = pkg("inverses")
    P = I.FastHybridHip(5, max_batch=3)
    D = P.outer_deg
    N = 32
    d_b = L.DeviceArray((3, 4, 2 * N - 1, N))
    xa, xb = L.DeviceArray((3, N, N)), L.DeviceArray((3, N, N))
    for ks, b in (([1, 2], d_b), ([D, 0], d_b), ([-1], d_b), ([2, 0], None), ([1, 1, 1, 1], d_b)):
        with pytest.raises(L.HipError):
            P.tick_device(ks, b, xa, xb)
    with pytest.raises(L.HipError):
        P.tick_device([1], d_b, xa, xa)
    with pytest.raises(L.HipError):
        P.lane_move_device(0, 3, xa, xb)
    P.close()


def test_server_continuous_keys_n512_n1024(gpu):
    """The server with N = 512 and 1024 as continuous keys (requests join between outer steps) beside batched
    N = 128: every result equal, bit for bit, to its slot's plan applied to it alone, and within 1e-5 of the FP64
    inverse with the same setup.  Every other request arrives in page-locked memory (Request.b_ptr) and uploads from
    there without staging, in both kinds of slot."""
    L = pkg("lib")
    S = pkg("serve")
    C = pkg("plancache")
    I = pkg("inverses")
    A = pkg("autotune")
    keys = [(7, 16), (9, 16), (10, 16), (10, 32)]
    setups = {k: C.hybrid_setup(k[0], dense_cap=k[1]) for k in keys}
    refine = {k: A.refine_for(k[0], k[1], *setups[k]) for k in keys}
    srv = S.Server(keys, slots_per_key=1, max_batch={7: 4, 9: 4, 10: 4}, setups=setups, refine=refine,
                   continuous={9, 10})
    assert sum(1 for s in srv.slots if getattr(s, "continuous", False)) == 3
    reqs, pinned = [], []
    for i in range(14):
        n, p = keys[i % len(keys)]
        b = _sino(n, 700 + i, 800 + i)
        if i % 2:
            pb = L.PinnedArray(b.shape)
            pb.a[...] = b
            pinned.append(pb)
            reqs.append(S.Request(i, n, p, pb.a, 0.0, b_ptr=pb.ptr))
        else:
            reqs.append(S.Request(i, n, p, b, 0.0))
    trace = [(0.004 * i, r) for i, r in enumerate(reqs)]
    done = srv.run_trace(trace)
    assert len(done) == len(reqs) and all(r.error is None for r in reqs)
    for r in reqs:
        slot = next(s for s in srv.slots if (s.n, s.cap) == (r.n, r.cap))
        assert np.array_equal(r.x, slot.plan.apply(r.b)), r.rid
    for n, p in keys:
        r = next(q for q in reqs if (q.n, q.cap) == (n, p))
        lam, outer = setups[(n, p)]
        H64 = I.FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=p)
        assert relerr(r.x, H64.apply(r.b)) < 1e-5, (n, p)
        H64.close()
    srv.close()
    for pb in pinned:
        pb.free()
