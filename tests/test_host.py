"""CPU tests of the host logic: setup arithmetic, level-row counts, phantoms, plan cache, the serving scheduler
(with a host stand-in for the GPU slots) and its failure handling.  No GPU call."""

import numpy as np
import pytest

from conftest import golden, other_package, pkg
from oracle import inverses as oiv
from oracle import layout as olay
from oracle import radon as orad


def test_degrees_and_sym_match_oracle():
    S = pkg("spectral")
    d = golden("inverse_n6_phantom")
    n = 6
    lam, deg = S.degrees(d["h_lam_max"], 0.2 * (1 << n), 0.1)
    assert np.array_equal(deg, d["h_deg"])
    lam2, deg2 = oiv.degrees(d["h_lam_max"], 0.2 * (1 << n), 0.1)
    assert np.array_equal(lam, lam2) and np.array_equal(deg, deg2)
    oms = 2 * np.pi * np.fft.fftfreq(2 << n)
    assert np.array_equal(S.explicit_lam_max(n, oms), oiv.explicit_bound(n, oms))


def test_g1_host_stencil_matches_oracle():
    S = pkg("spectral")
    x = np.random.default_rng(0).standard_normal((16, 16))
    assert np.array_equal(S.g1_apply(x), oiv.G1_apply(x))
    assert np.array_equal(S.pi_apply(x), oiv.Pi_apply(x))


@pytest.mark.parametrize("fixture", ["rows_n4_cap-1", "rows_n4_cap8", "rows_n5_cap8"])
def test_level_step_counts_match_reference(fixture):
    """The analytic level-row count equals the reference engine's count (fastops.py:137-154)."""
    E = pkg("engine")
    d = golden(fixture)
    n, N = int(d["n"]), 1 << int(d["n"])
    cap = int(d["dense_cap"])
    L = len(d["z"])
    ops = {"gh": [("B", 1, 1), ("P", 2, n), ("B", 2, n), ("BT", 2, n), ("BT", 1, 1)], "P_1n": [("P", 1, n)],
           "B_1n": [("B", 1, n)], "lp_back": [("BT", 1, n), ("P", 1, n)]}
    for p in (1, 2):
        dense = np.nonzero(d["deg"] > cap)[0] if cap >= 0 else []
        for name, o in ops.items():
            assert E.level_steps(o, 2 * L, L, d["deg"], dense, n, N) == int(d[f"{name}_p{p}_count"]), name


def test_phantom_and_noise():
    ph = pkg("phantom")
    a, b = ph.ellipses(3, 64), ph.ellipses(3, 64)
    assert np.array_equal(a, b) and not np.array_equal(a, ph.ellipses(4, 64))
    b0 = np.random.default_rng(0).standard_normal((4, 127, 64))
    bn = ph.add_noise(b0, 0.01, 5)
    assert abs(np.linalg.norm(bn - b0) / np.linalg.norm(b0) - 0.01) < 1e-12
    x = ph.int_image(1, 32)
    assert x.dtype == np.int32 and x.min() >= 0 and x.max() <= 255


def test_plan_cache_roundtrip(tmp_path, monkeypatch):
    monkeypatch.setenv("ADRT_PLAN_CACHE", str(tmp_path))
    C = pkg("plancache")
    assert C.load("hybrid", 5, dense_cap=8) is None
    C.store("hybrid", 5, np.arange(3.0), (0.1, 0.9, 17), dense_cap=8)
    hit = C.load("hybrid", 5, dense_cap=8)
    assert np.array_equal(hit["lam_max"], np.arange(3.0)) and hit["outer"] == (0.1, 0.9, 17)
    assert C.load("hybrid", 5, dense_cap=16) is None


def test_plan_cache_key_resolves_defaults(tmp_path, monkeypatch):
    """The key holds every setup keyword of the class with its default filled in: a default given explicitly hits
    the same entry, a changed default (tau_fac, g1_deg, safety ...) misses, and keywords that do not change the setup
    scalars (precision, max_batch, refine) are left out.  Unknown keywords are an error."""
    monkeypatch.setenv("ADRT_PLAN_CACHE", str(tmp_path))
    C = pkg("plancache")
    I = pkg("inverses")
    p0 = C.resolved(I.FastHybridHip, dense_cap=8)
    assert p0 == C.resolved(I.FastHybridHip, dense_cap=8, tau_fac=0.2, g1_deg=14, precision="fp32", max_batch=9)
    for k in ("tau_fac", "eps_in", "eps", "g1_deg", "safety", "power_iters", "precond", "lam_bound", "dense_cap"):
        assert k in p0, k
    for k in ("precision", "max_batch", "refine", "ctx"):
        assert k not in p0, k
    C.store("hybrid", 5, np.ones(3), (0.1, 0.9, 17), **p0)
    assert C.load("hybrid", 5, **p0) is not None
    assert C.load("hybrid", 5, **C.resolved(I.FastHybridHip, dense_cap=8, tau_fac=0.3)) is None
    assert C.load("hybrid", 5, **C.resolved(I.FastHybridHip, dense_cap=8, g1_deg=12)) is None
    with pytest.raises(TypeError):
        C.resolved(I.FastHybridHip, no_such_keyword=1)


def test_plan_cache_holds_refine_choices(tmp_path, monkeypatch):
    """The refinement-step choice is cached under the parameters refine_for keys it by (a list-valued interval and a
    float among them)."""
    monkeypatch.setenv("ADRT_PLAN_CACHE", str(tmp_path))
    C = pkg("plancache")
    params = {"dense_cap": 32, "target": 1e-6, "probes": 2, "seed": 8000, "setup_outer": [0.16, 0.96, 39],
              "lam": 1234.5}
    assert C.load("refine", 7, **params) is None
    C.store("refine", 7, np.zeros(1), (0.0, 0.0, 6), **params)
    assert C.load("refine", 7, **params)["outer"][2] == 6
    assert C.load("refine", 7, **{**params, "target": 1e-7}) is None


def test_plan_cache_ignores_stale_entries(tmp_path, monkeypatch):
    """An entry under the right file name whose stored parameters differ (written by a build of another VERSION, or
    by the old format without parameters) is not reused."""
    monkeypatch.setenv("ADRT_PLAN_CACHE", str(tmp_path))
    C = pkg("plancache")
    p = {"dense_cap": 8}
    path = tmp_path / (C.key("hybrid", 5, **p) + ".npz")
    np.savez(path, lam_max=np.ones(3), lo=0.1, hi=0.9, D=17)  # old format: no parameters
    assert C.load("hybrid", 5, **p) is None
    blob = C._blob("hybrid", 5, p).replace(f'"version": {C.VERSION}', f'"version": {C.VERSION - 1}')
    assert blob != C._blob("hybrid", 5, p)
    np.savez(path, lam_max=np.ones(3), lo=0.1, hi=0.9, D=17, params=np.array(blob))
    assert C.load("hybrid", 5, **p) is None
    C.store("hybrid", 5, np.ones(3), (0.1, 0.9, 17), **p)
    assert C.load("hybrid", 5, **p)["outer"] == (0.1, 0.9, 17)


class HostSlot:
    """Stand-in for a GPU slot: 'computes' x = sum of the sinogram per request after a fixed number of polls."""

    def __init__(self, n, cap, max_batch, latency_polls=2):
        self.n, self.cap, self.max_batch = n, cap, max_batch
        self.busy, self.polls, self.latency = [], 0, latency_polls
        self.batches = []

    def device_bytes(self):
        return 1000

    def launch(self, reqs):
        self.busy, self.polls = reqs, 0
        self.batches.append([r.rid for r in reqs])

    def poll(self):
        if not self.busy:
            return []
        self.polls += 1
        if self.polls < self.latency:
            return []
        done, self.busy = self.busy, []
        for r in done:
            r.x = np.full((1, 1), float(np.sum(r.b)))
        return done

    def close(self):
        pass


def host_server(keys, slots_per_key=2, max_batch=3, mem_budget=10**9, **kw):
    S = pkg("serve")
    made = []

    def factory(n, cap, mb):
        s = HostSlot(n, cap, mb)
        made.append(s)
        return s

    srv = S.Server(keys, slots_per_key=slots_per_key, max_batch=max_batch, mem_budget=mem_budget,
                   slot_factory=factory, **kw)
    return srv, made


def test_server_continuous_batching_and_rejection():
    S = pkg("serve")
    srv, made = host_server([(7, 16), (8, 16)])
    rng = np.random.default_rng(0)
    reqs = [S.Request(i, 7 if i % 3 else 8, 16, rng.standard_normal(5), 0.0) for i in range(20)]
    bad = S.Request(99, 11, 16, np.ones(3), 0.0)  # a size the server does not hold
    for r in reqs[:10]:
        srv.submit(r)
    assert not srv.submit(bad) and bad.error
    srv.step()
    for r in reqs[10:]:  # arrivals while the first batches are in flight
        srv.submit(r)
    for _ in range(50):
        srv.step()
        if srv.idle():
            break
    assert srv.idle() and len(srv.done) == 20
    for r in reqs:
        assert r.x[0, 0] == np.sum(r.b) and r.error is None
    assert all(len(b) <= 3 for s in made for b in s.batches)
    assert sum(len(s.batches) for s in made) < 20  # requests were batched


def test_server_memory_budget_limits_slots():
    srv, made = host_server([(7, 16), (8, 32)], slots_per_key=4, max_batch=2, mem_budget=3000)
    assert len(srv.slots) == 3 and srv.device_bytes <= 3000
    assert srv.served[(7, 16)] and srv.served[(8, 32)]


def test_server_grows_slots_while_memory_holds_them():
    """Requests waiting with every slot of their key busy get another slot while the budget holds one (up to
    max_slots_per_key); with the budget spent they wait for a slot to free instead, and all complete."""
    S = pkg("serve")
    for budget, want in ((10**9, 3), (2000, 2)):
        srv, made = host_server([(7, 16)], slots_per_key=1, max_batch=2, mem_budget=budget, max_slots_per_key=3)
        reqs = [S.Request(i, 7, 16, np.full(2, float(i)), 0.0) for i in range(12)]
        for r in reqs:
            srv.submit(r)
        for _ in range(100):
            srv.step()
            if srv.idle():
                break
        assert srv.idle() and len(srv.done) == 12 and not srv.shed
        assert len(srv.slots) == want and srv.grown == want - 1 and srv.device_bytes <= budget
        assert [r.rid for s in made for b in s.batches for r in reqs if r.rid in b] == \
            [rid for s in made for b in s.batches for rid in b]
        for r in reqs:
            assert r.x[0, 0] == np.sum(r.b)


def test_server_sheds_requests_that_wait_too_long():
    """Under overload a request not started within max_wait_s is shed with an error; started requests always
    complete."""
    S = pkg("serve")
    import time

    srv, made = host_server([(8, 16)], slots_per_key=1, max_batch=2, max_slots_per_key=1, max_wait_s=30.0)
    fresh = [S.Request(i, 8, 16, np.ones(2), time.perf_counter()) for i in range(2)]
    for r in fresh:
        srv.submit(r)
    srv.step()  # both start on the one slot
    assert all(r.t_start is not None for r in fresh)
    old = [S.Request(10 + i, 8, 16, np.ones(2), time.perf_counter() - 60.0) for i in range(4)]  # waited a minute
    new = [S.Request(20 + i, 8, 16, np.ones(2), time.perf_counter()) for i in range(2)]
    for r in old + new:
        srv.submit(r)
    for _ in range(50):
        srv.step()
        if srv.idle():
            break
    assert srv.idle()
    assert sorted(r.rid for r in srv.shed) == [10, 11, 12, 13]
    assert all(r.error.startswith("shed") and r.x is None for r in srv.shed)
    assert sorted(r.rid for r in srv.done) == [0, 1, 20, 21]
    assert all(r.error is None and r.x is not None for r in srv.done)


class HostLanes:
    """Host stand-in for StepLanes with adrt_plan_tick's contract on a toy outer iteration: lane i at step k keeps
    its right-hand side y, reads x_(k-1) from the current buffer and x_(k-2) from the next one, writes x_k to the
    next one; ticks complete after `latency` polls of their event."""

    def __init__(self, D, max_batch, m=3, latency=2, depth=8):
        self.outer_deg, self.max_batch, self.latency, self.depth = D, max_batch, latency, depth
        self.b = np.zeros((max_batch, m))
        self.y = np.full((max_batch, m), np.nan)
        self.x = [np.full((max_batch, m), np.nan), np.full((max_batch, m), np.nan)]
        self.staged = {}
        self.polls = {}
        self.lane_counts = []
        self.queued = []

    @staticmethod
    def step(k, y, xk, xp):
        return 2.0 * y if k == 0 else 0.5 * xk + (0.25 * xp if k >= 2 else 0.0) + y + k

    @classmethod
    def apply(cls, b, D):
        xs = [cls.step(0, b, None, None)]
        for k in range(1, D):
            xs.append(cls.step(k, b, xs[-1], xs[-2] if k >= 2 else None))
        return xs[-1]

    def device_bytes(self):
        return 1000

    def shared_bytes(self):
        return 0

    def upload(self, tick, lane, b, b_ptr=None):
        self.b[lane] = b

    def tick(self, ks, cur):
        assert all(a >= b for a, b in zip(ks, ks[1:])) and 0 <= min(ks) and max(ks) < self.outer_deg
        self.lane_counts.append(len(ks))
        xc, xn = self.x[cur], self.x[1 - cur]
        for i, k in enumerate(ks):
            if k == 0:
                self.y[i] = self.b[i]
            xn[i] = self.step(k, self.y[i], xc[i], xn[i])

    def download(self, tick, buf, lane):
        h = (tick, lane)
        self.staged[h] = self.x[buf][lane].copy()
        return h

    def move(self, src, dst):
        for a in (self.y, self.x[0], self.x[1]):
            a[dst] = a[src]
            a[src] = np.nan

    def record(self, tick):
        self.polls[tick] = 0
        self.queued.append(tick)
        assert sum(1 for t in self.queued if self.polls[t] < self.latency) <= self.depth

    def done(self, tick):
        self.polls[tick] += 1
        return self.polls[tick] >= self.latency

    def result(self, h):
        return self.staged.pop(h)

    def close(self):
        pass


def test_step_slot_lanes_join_and_leave_between_steps():
    """StepSlot (continuous batching at outer-step granularity) on the host stand-in: requests that arrive while
    others are mid-iteration join at the next tick, lanes stay ordered oldest first, finished lanes leave from the
    front while the others move down, and every result equals the iteration run alone."""
    S = pkg("serve")
    D, lanes = 7, 4
    made = []

    def factory(n, cap, mb):
        s = S.StepSlot(n, cap, max_batch=mb, dev=HostLanes(D, mb))
        made.append(s)
        return s

    srv = S.Server([(10, 16)], slots_per_key=2, max_batch=lanes, mem_budget=10**9, slot_factory=factory,
                   continuous={10})
    assert len(srv.slots) == 1  # one slot of lanes per continuous key
    rng = np.random.default_rng(3)
    reqs = [S.Request(i, 10, 16, rng.standard_normal(3), 0.0) for i in range(11)]
    arrive = {0: [0, 1], 3: [2], 4: [3, 4, 5], 9: [6, 7, 8, 9, 10]}
    for t in range(200):
        for i in arrive.get(t, []):
            srv.submit(reqs[i])
        srv.step()
        if t > max(arrive) and srv.idle():
            break
    assert srv.idle() and len(srv.done) == len(reqs) and not srv.shed
    for r in reqs:
        np.testing.assert_array_equal(r.x, HostLanes.apply(r.b, D))
    dev = made[0].dev
    assert max(dev.lane_counts) == lanes and min(dev.lane_counts) >= 1
    # requests 2..5 arrived while 0 and 1 were mid-iteration and shared their ticks
    assert reqs[2].t_start < reqs[0].t_done and reqs[5].t_start < reqs[1].t_done


def test_step_slot_rejects_overfull_ticks():
    S = pkg("serve")
    s = S.StepSlot(10, 16, max_batch=2, dev=HostLanes(9, 2, depth=3))
    assert s.depth == 3
    rs = [S.Request(i, 10, 16, np.ones(3), 0.0) for i in range(3)]
    with pytest.raises(ValueError):
        s.tick(rs)
    s.tick(rs[:2])
    s.tick()
    s.tick()
    with pytest.raises(RuntimeError):  # depth ticks already queued
        s.tick()
    # the queue depth stays below the outer degree, so a lane never starts and finishes within the queued ticks
    assert S.StepSlot(10, 16, max_batch=2, dev=HostLanes(3, 2, depth=8)).depth == 2


def test_poisson_trace_and_stats():
    S = pkg("serve")
    pool = {7: [np.zeros(2)], 8: [np.ones(2)]}
    tr = S.poisson_trace(1000.0, 0.5, [7, 8], [0.5, 0.5], [16, 32], pool, seed=7)
    gaps = np.diff([0.0] + [t for t, _ in tr])
    assert 350 < len(tr) < 650 and abs(gaps.mean() - 1e-3) < 3e-4
    for i, (_, r) in enumerate(tr):
        r.t_arrive, r.t_done = 0.0, 0.001 * (i + 1)
    st = S.latency_stats([r for _, r in tr])
    assert st["count"] == len(tr) and st["p99_ms"] >= st["p50_ms"]


def test_gram_and_lstsq_cheb_validation():
    """The argument checks of adrt.core.gram / lstsq_cheb (patch:629-639, 712-714) run before any device call."""
    lsq = pkg("lsq")
    for shape in [(4, 8), (6, 6), (8,)]:
        with pytest.raises(ValueError):
            lsq.gram(np.zeros(shape))
    with pytest.raises(TypeError):
        lsq.gram(np.zeros((8, 8), dtype=np.int64))
    with pytest.raises(ValueError):
        lsq.lstsq_cheb(np.zeros((4, 15, 8)), lam_min=10.0, lam_max=5.0)
    with pytest.raises(ValueError):
        lsq.lstsq_cheb(np.zeros((4, 8, 8)), lam_min=1.0, lam_max=5.0)


@pytest.mark.parametrize("n", [1, 2, 4, 6])
def test_layout_maps_pair_like_the_reference(n):
    """Exact integer fingerprints give the pairs of the reference's rounded-fingerprint grouping
    (skimage_adrt_pkg.py:17-35), duplicates included (both pair a group's entries in ascending index order)."""
    LY = pkg("layout")
    core = lambda x: orad.adrt(n, x)
    pk = other_package(n, core)
    ci, pi = LY.layout_maps(n, pk, core)
    rci, rpi = olay.layout_maps(n, core, pk)
    assert len(ci) == len(rci)
    assert set(zip(ci.tolist(), pi.tolist())) == set(zip(rci.tolist(), rpi.tolist()))


def test_layout_maps_reject_foreign_transforms():
    LY = pkg("layout")
    n, N = 3, 8
    core = lambda x: orad.adrt(n, x)
    with pytest.raises(ValueError, match="same lines"):
        LY.layout_maps(n, lambda x: 2.0 * core(x), core)
    with pytest.raises(ValueError, match="entries"):
        LY.layout_maps(n, lambda x: core(x)[:, :N], core)
    with pytest.raises(ImportError, match="pkg_adrt"):  # the adrt package is not installed here
        LY.layout_maps(n, None, core)


def test_dense_placement_tie_break():
    """Placements within 1% of the fastest tie, won by 2, then 1, then 4 waves per SIMD; a launch of its own only
    outright."""
    A = pkg("autotune")
    assert A.pick_dense({-1: 0.304, 1: 0.277, 2: 0.277, 4: 0.281}) == 2
    assert A.pick_dense({-1: 0.318, 1: 0.296, 2: 0.298, 4: 0.297}) == 2
    assert A.pick_dense({1: 0.300, 2: 0.290, 4: 0.280}) == 4
    assert A.pick_dense({1: 0.2801, 2: 0.290, 4: 0.280}) == 1
    assert A.pick_dense({-1: 0.270, 1: 0.272, 2: 0.2725}) == -1
    assert A.pick_dense({-1: 0.300, 1: 0.250, 2: 0.260, 4: 0.300}) == 1
