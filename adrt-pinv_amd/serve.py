"""Serving loop: continuous batching of hybrid inversions of mixed N with device-memory admission.

Requests (a sinogram, its N and its dense cap p) arrive at any time.  The server keeps device plans resident for every
(N, p) it serves, several instances ("slots") each, with their twiddle / degree / dense tables in HBM.  A slot owns a
stream, input / output buffers for up to ``max_batch`` sinograms and a completion event.  The loop

  * polls the slots' events; a finished slot hands its results back and becomes free again,
  * admits waiting requests of a free slot's (N, p) as one batch (up to max_batch) and launches the whole inverse on
    the slot's stream (upload from page-locked memory, one hipGraph replay of the inverse for that batch size,
    download into page-locked memory, completion event),

so requests that arrive while others are on the GPU start as soon as a slot of their size frees up, and the batches of
different sizes overlap on the device.  Slots are created at start-up while their device memory fits the budget;
requests for a size the server does not hold are rejected with an error status, the others are unaffected.

The slot factory is pluggable so the scheduling logic can be exercised without a GPU (tests use a host stand-in).
"""

import time
from collections import deque

import numpy as np


class Request:
    __slots__ = ("rid", "n", "cap", "b", "t_arrive", "t_start", "t_done", "x", "error")

    def __init__(self, rid, n, cap, b, t_arrive):
        self.rid, self.n, self.cap, self.b, self.t_arrive = rid, n, cap, b, t_arrive
        self.t_start = self.t_done = None
        self.x = None
        self.error = None


class GpuSlot:
    """One resident device plan plus buffers, stream and event.  The sinograms go up from and the images come back into
    page-locked host buffers on the slot's stream; the inverse itself is replayed from a hipGraph captured once per
    batch size, so a launch costs three enqueues and the completion is one event."""

    def __init__(self, n, cap, precision="fp32", refine=1, max_batch=4, setup=None, stream=None, **kw):
        from . import lib as L
        from .inverses import FastHybridHip

        self.L = L
        self.n, self.cap, self.max_batch = n, cap, max_batch
        N = 1 << n
        lam, outer = setup
        self.plan = FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=cap, precision=precision,
                                  max_batch=max_batch, refine=refine, **kw)
        self.stream = stream or L.Stream()  # a stream shared with other slots runs their launches in turn
        self.d_b = L.DeviceArray((max_batch, 4, 2 * N - 1, N))
        self.d_x = L.DeviceArray((max_batch, N, N))
        self.event = L.Event()
        self.host_b = L.PinnedArray((max_batch, 4, 2 * N - 1, N))
        self.host_x = L.PinnedArray((max_batch, N, N))
        self.graphs = {}
        self.busy = []

    @staticmethod
    def bytes_needed(n, cap, max_batch, plan_bytes=None):
        N = 1 << n
        return (plan_bytes or 0) + 8 * max_batch * (4 * (2 * N - 1) * N + N * N)

    def device_bytes(self):
        """Device bytes of this slot alone; the dense tables are shared by the slots of one (N, p)."""
        return self.plan.plan_bytes + self.d_b.nbytes + self.d_x.nbytes

    def shared_bytes(self):
        return getattr(self.plan, "table_bytes", 0)

    def prepare(self):
        """Capture the graphs of every batch size up front (the first request of a size then pays no capture)."""
        for S in range(1, self.max_batch + 1):
            self.graph(S)

    def graph(self, S):
        g = self.graphs.get(S)
        if g is None:
            with self.L.Graph(self.stream) as g:
                self.plan.apply_device(self.d_b, self.d_x, S, self.stream, count=False)
            self.graphs[S] = g
        return g

    def launch(self, reqs):
        S = len(reqs)
        for j, r in enumerate(reqs):
            self.host_b.a[j] = r.b
        L, lib = self.L, self.L.load()
        g = self.graph(S)
        L.check(lib.adrt_memcpy(self.d_b.ptr, self.host_b.ptr, S * self.host_b.a[0].nbytes, 1, self.stream.handle),
                "h2d")
        g.launch(self.stream)
        L.check(lib.adrt_memcpy(self.host_x.ptr, self.d_x.ptr, S * self.host_x.a[0].nbytes, 4, self.stream.handle),
                "d2h")
        self.event.record(self.stream)
        self.busy = reqs

    def poll(self):
        if not self.busy or not self.event.done():
            return []
        done, self.busy = self.busy, []
        for j, r in enumerate(done):
            r.x = self.host_x.a[j].copy()
        return done

    def close(self):
        self.graphs.clear()
        self.plan.close()
        self.host_b.free()
        self.host_x.free()


class Server:
    def __init__(self, keys, slots_per_key=2, max_batch=4, mem_budget=None, slot_factory=None, setups=None,
                 precision="fp32", refine=1, tune=True, stream_per_size=True, streams_per_size=1, **kw):
        """keys: iterable of (n, cap).  max_batch: sinograms per launch, one number or {n: number} (small N batch
        deeper at little latency cost).  mem_budget: device bytes the slots may hold (default: 60% of free memory).
        stream_per_size: the slots of one N share one stream on a hardware queue of its own, so a long batch of one
        size never holds up another size's launches (ordinary streams share a few hardware queues, and a small-N
        batch could wait behind an N=1024 one); the slots of one size then run their batches in turn, on
        ``streams_per_size`` such streams (dealt to the slots of the size in turn)."""
        self.keys = list(keys)
        self.pending = {k: deque() for k in self.keys}
        self.slots = []
        self.rejected = []
        self.done = []
        self._streams = {}
        made = {}

        def gpu_slot(n, cap, mb):
            st = None
            if stream_per_size:
                from . import lib as L

                j = made.get(n, 0) % max(1, streams_per_size)
                made[n] = made.get(n, 0) + 1
                if (n, j) not in self._streams:
                    self._streams[(n, j)] = L.Stream(own_queue=True)
                st = self._streams[(n, j)]
            return GpuSlot(n, cap, precision, refine, mb, setups[(n, cap)], stream=st, **kw)

        if stream_per_size and slot_factory is None:
            from . import lib as L

            want = len({n for n, _ in self.keys}) * max(1, streams_per_size)
            if want > L.MAX_OWN_QUEUES:
                raise ValueError(f"{want} own-queue streams asked for (sizes x streams_per_size); the library allows "
                                 f"{L.MAX_OWN_QUEUES} at once: serve fewer sizes or set stream_per_size=False")
        factory = slot_factory or gpu_slot
        if mem_budget is None and slot_factory is None:
            from . import lib as L

            free = L.ctypes.c_ulonglong()
            total = L.ctypes.c_ulonglong()
            L.check(L.load().adrt_device_mem_info(L.ctypes.byref(free), L.ctypes.byref(total)), "mem_info")
            mem_budget = int(0.6 * free.value)
        self.mem_budget = mem_budget
        used = 0
        for rnd in range(slots_per_key):
            for n, cap in self.keys:
                s = factory(n, cap, max_batch.get(n, 4) if isinstance(max_batch, dict) else max_batch)
                b = s.device_bytes() + (getattr(s, "shared_bytes", lambda: 0)() if rnd == 0 else 0)
                if mem_budget is not None and used + b > mem_budget and rnd > 0:
                    s.close()  # admission by memory: no further instances of this key
                    continue
                used += b
                self.slots.append(s)
        self.device_bytes = used
        self.served = {k: any((s.n, s.cap) == k for s in self.slots) for k in self.keys}
        if tune and slot_factory is None:
            # the row kernel's mapping per N, timed on the first slot of that N at its batch depth (before the graphs
            # are captured, which bake the mapping in)
            from . import autotune

            for n in sorted({s.n for s in self.slots}):
                s = next(x for x in self.slots if x.n == n)
                autotune.tune(s.plan, s.max_batch, s.stream, persist=False)
        for s in self.slots:
            if hasattr(s, "prepare"):
                s.prepare()

    def submit(self, req):
        k = (req.n, req.cap)
        if k not in self.pending or not self.served.get(k, False):
            req.error = f"no resident plan for N=2^{req.n}, p={req.cap}"
            req.t_done = time.perf_counter()
            self.rejected.append(req)
            return False
        self.pending[k].append(req)
        return True

    def step(self):
        now = time.perf_counter()
        for s in self.slots:
            for r in s.poll():
                r.t_done = now
                self.done.append(r)
        for s in self.slots:
            if s.busy:
                continue
            q = self.pending[(s.n, s.cap)]
            if q:
                batch = [q.popleft() for _ in range(min(s.max_batch, len(q)))]
                t = time.perf_counter()
                for r in batch:
                    r.t_start = t
                s.launch(batch)

    def idle(self):
        return all(not s.busy for s in self.slots) and all(not q for q in self.pending.values())

    def run_trace(self, trace, timeout=600.0):
        """trace: list of (t_offset_s, Request-without-time) sorted by offset.  Arrivals are released by the wall clock;
        returns the completed requests."""
        t0 = time.perf_counter()
        i = 0
        while True:
            now = time.perf_counter() - t0
            while i < len(trace) and trace[i][0] <= now:
                off, r = trace[i]
                r.t_arrive = t0 + off
                self.submit(r)
                i += 1
            self.step()
            if i == len(trace) and self.idle():
                break
            if now > timeout:
                raise TimeoutError("serving trace did not drain")
        return self.done

    def close(self):
        for s in self.slots:
            s.close()
        for st in self._streams.values():
            st.sync()
            st.close()
        self._streams.clear()


def poisson_trace(rate, duration, sizes, weights, caps, pool, seed=7):
    """Arrivals with exponential gaps (default_rng(seed)); N from ``sizes`` with ``weights``, p uniform over ``caps``;
    the sinogram of a request is drawn from pool[n] in turn."""
    rng = np.random.default_rng(seed)
    t, out, rid = 0.0, [], 0
    used = {n: 0 for n in sizes}
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        n = int(rng.choice(sizes, p=weights))
        cap = int(caps[rng.integers(len(caps))])
        b = pool[n][used[n] % len(pool[n])]
        used[n] += 1
        out.append((t, Request(rid, n, cap, b, None)))
        rid += 1
    return out


def latency_stats(done, t_begin=None):
    lat = np.array([r.t_done - r.t_arrive for r in done])
    if lat.size == 0:
        return {"count": 0}
    t_first = min(r.t_arrive for r in done) if t_begin is None else t_begin
    t_last = max(r.t_done for r in done)
    return {
        "count": int(lat.size),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "throughput_per_s": float(lat.size / max(t_last - t_first, 1e-9)),
    }
