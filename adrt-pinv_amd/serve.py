"""Serving loop: continuous batching of hybrid inversions of mixed N, admitted by device memory.

Requests (a sinogram, its N and its dense cap p) arrive at any time.  The server keeps, per (N, p) it serves, the dense
/ twiddle / degree tables resident and runs its batches in **slots**: a device plan with scratch, input and output
buffers for ``max_batch`` sinograms, page-locked staging and a hipGraph of the whole inverse per batch size.  Slots
are the unit of device memory: a (N, p) gets ``slots_per_key`` of them at start while the device-memory budget holds
them, and (with ``max_slots_per_key`` above that) another one whenever its requests wait with every slot busy and the
budget still holds one.  A request that has waited ``max_wait_s`` without being started is shed with an error (the
server's answer to overload); started requests always complete.  The loop

  * polls the slots' completion events; a finished batch hands its images back and frees its slot,
  * sheds requests that waited too long,
  * launches, on every idle slot, the waiting requests of its (N, p) as one batch: sinograms into the slot's
    page-locked staging, then on the size's compute stream one upload, the captured inverse and one download,
  * grows the slots of keys whose requests still wait, as memory allows, and launches on them at once,

so requests that arrive while others are on the GPU start as soon as a slot of their size frees up or fits in memory,
and the batches of different sizes overlap on the device (a size's compute stream has a hardware queue of its own,
adrt_stream_create_own_queue, so a small-N batch never waits behind an N=1024 batch's kernels on one queue).
Requests for a size the server does not hold are rejected with an error; the others are unaffected.

Sizes named in ``continuous`` are served at outer-step granularity instead: one StepSlot per (N, p) whose lanes hold
sinograms at different outer steps (adrt_plan_tick), so a request joins a running batch at the next outer step rather
than waiting for the batch to finish.

The slot factory is pluggable so the scheduling logic can be exercised without a GPU (tests use a host stand-in).
"""

import time
from collections import deque

import numpy as np


class Request:
    """One inversion: sinogram ``b`` of size N = 2^n with dense cap ``cap``.  ``b_ptr``: the host address of ``b``
    when it already sits in page-locked memory (lib.PinnedArray); its upload then reads it there, with no staging
    copy, and the caller keeps it unchanged until the request is done."""

    __slots__ = ("rid", "n", "cap", "b", "b_ptr", "t_arrive", "t_start", "t_done", "x", "error")

    def __init__(self, rid, n, cap, b, t_arrive, b_ptr=None):
        self.rid, self.n, self.cap, self.b, self.t_arrive = rid, n, cap, b, t_arrive
        self.b_ptr = b_ptr
        self.t_start = self.t_done = None
        self.x = None
        self.error = None


class SizeStreams:
    """The streams of one N: the inverses and downloads on a stream with a hardware queue of its own (or an ordinary
    one), and the uploads on the same stream or, with separate_upload, on an ordinary stream that never waits on
    anything (the compute stream then waits for each batch's upload)."""

    def __init__(self, own_queue=True, separate_upload=False):
        from . import lib as L

        self.compute = L.Stream(own_queue=own_queue)
        self.up = L.Stream() if separate_upload else self.compute

    def close(self):
        for s in {id(self.up): self.up, id(self.compute): self.compute}.values():
            s.sync()
            s.close()


class GpuSlot:
    """One resident device plan with its buffers, page-locked staging and events.  A batch: sinograms into the host
    staging, one upload (upload stream), then on the compute stream, once the upload is in, the inverse replayed from
    the graph captured for that batch size and one download."""

    def __init__(self, n, cap, precision="fp32", refine=1, max_batch=4, setup=None, streams=None, **kw):
        from . import lib as L
        from .inverses import FastHybridHip

        self.L = L
        self.n, self.cap, self.max_batch = n, cap, max_batch
        N = 1 << n
        lam, outer = setup
        self.plan = FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=cap, precision=precision,
                                  max_batch=max_batch, refine=refine, **kw)
        self._own_streams = streams is None
        self.streams = streams or SizeStreams(own_queue=False)
        self.d_b = L.DeviceArray((max_batch, 4, 2 * N - 1, N))
        self.d_x = L.DeviceArray((max_batch, N, N))
        self.host_b = L.PinnedArray((max_batch, 4, 2 * N - 1, N))
        self.host_x = L.PinnedArray((max_batch, N, N))
        self.ev_up, self.ev_done = L.Event(), L.Event()
        self.graphs = {}
        self.busy = []

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
            st = self.streams.compute
            with self.L.Graph(st) as g:
                self.plan.apply_device(self.d_b, self.d_x, S, st, count=False)
            self.graphs[S] = g
        return g

    def launch(self, reqs):
        S = len(reqs)
        L, lib = self.L, self.L.load()
        ss = self.streams
        row = self.host_b.nbytes // self.max_batch
        staged = []
        for j, r in enumerate(reqs):
            if r.b_ptr is not None:  # page-locked already: uploaded from where it is
                L.check(lib.adrt_memcpy(self.d_b.ptr + j * row, r.b_ptr, row, 1, ss.up.handle), "h2d")
            else:
                self.host_b.a[j] = r.b
                staged.append(j)
        g = self.graph(S)
        sx = S * (self.host_x.nbytes // self.max_batch)
        for j in staged:
            L.check(lib.adrt_memcpy(self.d_b.ptr + j * row, self.host_b.ptr + j * row, row, 1, ss.up.handle), "h2d")
        if ss.up is not ss.compute:
            self.ev_up.record(ss.up)
            ss.compute.wait_event(self.ev_up)
        g.launch(ss.compute)
        L.check(lib.adrt_memcpy(self.host_x.ptr, self.d_x.ptr, sx, 4, ss.compute.handle), "d2h")
        self.ev_done.record(ss.compute)
        self.busy = reqs

    def poll(self):
        if not self.busy or not self.ev_done.done():
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
        if self._own_streams:
            self.streams.close()


class StepLanes:
    """The device side of a StepSlot: a resident plan whose ``max_batch`` lanes hold sinograms at different outer
    steps, two iterate buffers and the lanes' sinograms on the device; page-locked rings of 2 max_batch sinograms
    (made on the first upload that needs staging) and max_batch images for the uploads and downloads; an event per
    queued tick.  A sinogram entry is reused once the tick that uploaded from it has completed (the slot waits for
    that tick in the rare case it has not).  The plan is a hybrid plan made
    from ``setup``, or any plan (hybrid or L_p) given as ``plan``."""

    def __init__(self, n, cap, precision, refine, max_batch, setup, streams, depth, plan=None, **kw):
        from . import lib as L
        from .inverses import FastHybridHip

        self.L = L
        N = 1 << n
        self.own_plan = plan is None
        if plan is None:
            lam, outer = setup
            plan = FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=cap, precision=precision,
                                 max_batch=max_batch, refine=refine, **kw)
        self.plan = plan
        max_batch = plan.max_batch
        self.outer_deg = self.plan.outer_deg
        self.max_batch = max_batch
        self.depth = depth
        self.own_streams = streams is None
        self.streams = streams or SizeStreams(own_queue=False)
        self.d_b = L.DeviceArray((max_batch, 4, 2 * N - 1, N))
        self.d_x = [L.DeviceArray((max_batch, N, N)), L.DeviceArray((max_batch, N, N))]
        # at most 2 max_batch sinograms join within the queued ticks (the lanes, and as many again after lanes
        # finish) and at most max_batch images await reading (a lane cannot start and finish within them)
        self.nb = 2 * max_batch
        self.host_b = None
        self.host_x = L.PinnedArray((max_batch, N, N))
        self.b_row = 4 * (2 * N - 1) * N * 8
        self.x_row = self.host_x.nbytes // max_batch
        self.b_owner = [-1] * self.nb  # tick that last uploaded from each ring entry
        self.x_unread = [False] * max_batch  # image entries downloaded into and not yet read
        self.b_next = self.x_next = 0
        self.events = [L.Event() for _ in range(depth)]
        self.recorded = {}  # tick -> event index, for the queued ticks
        self.completed = -1

    def device_bytes(self):
        return self.plan.plan_bytes + self.d_b.nbytes + 2 * self.d_x[0].nbytes

    def shared_bytes(self):
        return getattr(self.plan, "table_bytes", 0)

    def _free(self, owner):
        while owner > self.completed and not self.done(owner):
            pass

    def upload(self, tick, lane, b, b_ptr=None):
        L, st = self.L, self.streams.compute
        if b_ptr is not None:  # page-locked already: uploaded from where it is
            L.check(L.load().adrt_memcpy(self.d_b.ptr + lane * self.b_row, b_ptr, self.b_row, 1, st.handle), "h2d")
            return
        if self.host_b is None:
            self.host_b = L.PinnedArray((self.nb, self.b_row // 8))
        e = self.b_next
        self.b_next = (e + 1) % self.nb
        self._free(self.b_owner[e])
        self.host_b.a[e] = np.asarray(b, dtype=np.float64).reshape(-1)
        self.b_owner[e] = tick
        L.check(L.load().adrt_memcpy(self.d_b.ptr + lane * self.b_row, self.host_b.ptr + e * self.b_row, self.b_row,
                                     1, st.handle), "h2d")

    def tick(self, ks, cur):
        self.plan.tick_device(ks, self.d_b, self.d_x[cur], self.d_x[1 - cur], self.streams.compute)

    def download(self, tick, buf, lane):
        """Queues the download of a lane's image; returns the ring entry that will hold it."""
        L = self.L
        e = self.x_next
        self.x_next = (e + 1) % self.max_batch
        if self.x_unread[e]:  # cannot happen while outer_deg > depth: at most max_batch lanes finish per window
            raise RuntimeError("image ring full: a finished lane's image has not been read")
        self.x_unread[e] = True
        L.check(L.load().adrt_memcpy(self.host_x.ptr + e * self.x_row, self.d_x[buf].ptr + lane * self.x_row,
                                     self.x_row, 4, self.streams.compute.handle), "d2h")
        return e

    def move(self, src, dst):
        self.plan.lane_move_device(src, dst, self.d_x[0], self.d_x[1], self.streams.compute)

    def record(self, tick):
        i = tick % self.depth
        self.events[i].record(self.streams.compute)
        self.recorded[tick] = i

    def done(self, tick):
        if tick <= self.completed:
            return True
        if not self.events[self.recorded[tick]].done():
            return False
        for t in [t for t in self.recorded if t <= tick]:
            del self.recorded[t]
        self.completed = tick
        return True

    def result(self, e):
        x = self.host_x.a[e].copy()
        self.x_unread[e] = False
        return x

    def close(self):
        self.streams.compute.sync()
        if self.own_plan:
            self.plan.close()
        if self.host_b is not None:
            self.host_b.free()
        self.host_x.free()
        if self.own_streams:
            self.streams.close()


class StepSlot:
    """Continuous batching at outer-step granularity for one (N, p): ``max_batch`` lanes hold sinograms at different
    outer steps, and every tick (adrt_plan_tick) advances each lane by one step.  The lanes at FP64 refinement
    steps, those at the plan's own steps and those that start share one pass per kind: the row launch, and with it
    the stream of the dense tables, serves all of them.  A request joins at the next tick once a lane is free,
    instead of waiting for a whole batch to finish, and leaves after its last step; its result is the same bits as
    the plan's apply of that sinogram alone.  Up to ``depth`` ticks are kept queued on the size's compute stream
    (default 16, at most outer_deg - 1: about 10 ms of device work at N = 1024), so the device does not wait for the
    host between ticks, nor while the host loop stages other slots' sinograms; a new request joins behind them.

    A tick, in stream order: the new sinograms upload into the lanes after the running ones; the tick; the lanes
    that took their last step (the oldest, so a prefix of the lanes) download; the running lanes move down over them
    (adrt_plan_lane_move); an event.  ``dev`` is the device side (StepLanes; tests use a host stand-in)."""

    continuous = True

    def __init__(self, n, cap, precision="fp32", refine=1, max_batch=8, setup=None, streams=None, dev=None,
                 depth=None, **kw):
        self.n, self.cap = n, cap
        depth = depth or 16
        self.dev = dev or StepLanes(n, cap, precision, refine, max_batch, setup, streams, depth, **kw)
        self.max_batch = self.dev.max_batch
        self.plan = getattr(self.dev, "plan", None)
        self.streams = getattr(self.dev, "streams", None)
        self.D = self.dev.outer_deg
        # a lane cannot start and finish within the queued ticks, so at most max_batch images await reading
        self.depth = max(1, min(depth, self.dev.depth, self.D - 1))
        self.lanes = []     # [request, step of its next tick], oldest first: device lane i holds lanes[i]
        self.cur = 0        # iterate buffer holding the lanes' current iterates
        self.inflight = []  # (tick, [(request, ring entry of its image)]) of the queued ticks, oldest first
        self.ticks = 0

    @property
    def busy(self):
        return bool(self.lanes) or bool(self.inflight)

    def free_lanes(self):
        return self.max_batch - len(self.lanes)

    def device_bytes(self):
        return self.dev.device_bytes()

    def shared_bytes(self):
        return self.dev.shared_bytes()

    def can_tick(self):
        return len(self.inflight) < self.depth

    def tick(self, new_reqs=()):
        """Queue one tick: the lanes advance one step, ``new_reqs`` (at most free_lanes()) join at step 0."""
        new_reqs = list(new_reqs)
        if len(new_reqs) > self.free_lanes():
            raise ValueError("more new sinograms than free lanes")
        if not self.can_tick():
            raise RuntimeError("tick queue full")
        if not self.lanes and not new_reqs:
            return
        t = self.ticks
        self.ticks += 1
        for r in new_reqs:
            self.dev.upload(t, len(self.lanes), r.b, r.b_ptr)
            self.lanes.append([r, 0])
        ks = [k for _, k in self.lanes]
        self.dev.tick(ks, self.cur)
        nxt = 1 - self.cur
        nf = sum(1 for k in ks if k == self.D - 1)
        finished = [(self.lanes[i][0], self.dev.download(t, nxt, i)) for i in range(nf)]
        if nf:
            for i in range(nf, len(ks)):
                self.dev.move(i, i - nf)
        self.dev.record(t)
        self.lanes = [[r, k + 1] for r, k in self.lanes[nf:]]
        self.cur = nxt
        self.inflight.append((t, finished))

    def poll(self):
        """The requests whose last tick has completed."""
        done = []
        while self.inflight and self.dev.done(self.inflight[0][0]):
            _, finished = self.inflight.pop(0)
            for r, e in finished:
                r.x = self.dev.result(e)
                done.append(r)
        return done

    def close(self):
        self.dev.close()


class Server:
    def __init__(self, keys, slots_per_key=2, max_batch=4, mem_budget=None, slot_factory=None, setups=None,
                 precision="fp32", refine=1, tune=True, stream_per_size=True, max_slots_per_key=None,
                 max_wait_s=None, separate_upload=False, continuous=(), tick_depth=None, **kw):
        """keys: iterable of (n, cap).  max_batch: sinograms per launch, one number or {n: number} (small N batch
        deeper at little latency cost).  mem_budget: device bytes the slots may hold (default: 60% of free memory;
        a (N, p)'s tables are counted with its first slot).  slots_per_key slots per key are made at start while they
        fit, more on demand up to max_slots_per_key (default: slots_per_key, no growth) while the budget holds them.
        refine: FP64 refinement steps of the FP32 plans, one number or {(n, cap): number}.
        stream_per_size: a size's inverses run on a stream with a hardware queue of its own.  separate_upload: a
        size's uploads go on an ordinary stream of their own instead of the compute stream.
        max_wait_s: a request not started within that many seconds is shed with an error (None: requests wait).
        continuous: the sizes n whose keys run one StepSlot (requests join a running batch between outer steps)
        instead of batches in slots; max_batch is then its number of lanes, tick_depth its queued ticks (None: the
        StepSlot default).

        Measured on one MI355X at 200 req/s over N in {128 .. 1024}, p in {16, 32}: uploads on the compute stream
        and two slots per key give p50 / p99 10 / 120 ms; a separate upload stream 38 / 145 ms (HIP spreads
        ordinary streams over a few shared hardware queues, so the sizes' uploads and waits hold each other up);
        growing to three slots per key 54 / 185 ms (the slots of a size share its compute stream, so more slots only
        queue more batches there)."""
        self.keys = list(keys)
        self.max_wait_s = max_wait_s
        self.max_slots = max_slots_per_key or slots_per_key
        self.pending = {k: deque() for k in self.keys}
        self.slots = []
        self.rejected = []
        self.shed = []
        self.done = []
        self.grown = 0
        # host seconds of the loop: polling, launching batches, queueing ticks (and the ticks queued)
        self.host_s = {"poll": 0.0, "launch": 0.0, "tick": 0.0, "ticks": 0}
        self._streams = {}
        sizes = sorted({n for n, _ in self.keys})
        self._mb = (lambda n: max_batch.get(n, 4)) if isinstance(max_batch, dict) else (lambda n: max_batch)
        self.continuous = set(continuous)
        if stream_per_size and slot_factory is None:
            from . import lib as L

            if len(sizes) > L.MAX_OWN_QUEUES:
                raise ValueError(f"{len(sizes)} own-queue streams asked for (one per size); the library allows "
                                 f"{L.MAX_OWN_QUEUES} at once: serve fewer sizes or set stream_per_size=False")

        def gpu_slot(n, cap, mb):
            if n not in self._streams:
                self._streams[n] = SizeStreams(own_queue=stream_per_size, separate_upload=separate_upload)
            rf = refine.get((n, cap), 1) if isinstance(refine, dict) else refine
            if n in self.continuous:
                return StepSlot(n, cap, precision, rf, mb, setups[(n, cap)], streams=self._streams[n],
                                depth=tick_depth, **kw)
            return GpuSlot(n, cap, precision, rf, mb, setups[(n, cap)], streams=self._streams[n], **kw)

        self._factory = slot_factory or gpu_slot
        if mem_budget is None and slot_factory is None:
            from . import lib as L

            free = L.ctypes.c_ulonglong()
            total = L.ctypes.c_ulonglong()
            L.check(L.load().adrt_device_mem_info(L.ctypes.byref(free), L.ctypes.byref(total)), "mem_info")
            mem_budget = int(0.6 * free.value)
        self.mem_budget = mem_budget
        self.device_bytes = 0
        self._slot_bytes = {}
        for rnd in range(slots_per_key):
            for k in self.keys:
                if rnd > 0 and k[0] in self.continuous:
                    continue  # one slot of lanes per continuous key
                if rnd > 0 and not self._fits(k):
                    continue  # admission by memory: no further instances of this key
                self._add_slot(k, prepare=False)
        self.served = {k: any((s.n, s.cap) == k for s in self.slots) for k in self.keys}
        if tune and slot_factory is None:
            # the row kernel's mapping per N, timed on the first slot of that N at its batch depth (before the graphs
            # are captured, which bake the mapping in)
            from . import autotune

            for n in sizes:
                s = next(x for x in self.slots if x.n == n)
                autotune.tune(s.plan, s.max_batch, s.streams.compute, persist=False)
        for s in self.slots:
            if hasattr(s, "prepare"):
                s.prepare()

    def _fits(self, key):
        """Would one more slot of ``key`` fit in the budget (its bytes as measured on the key's first slot)?"""
        if self.mem_budget is None:
            return True
        b = self._slot_bytes.get(key)
        return b is not None and self.device_bytes + b <= self.mem_budget

    def _add_slot(self, key, prepare=True):
        n, cap = key
        s = self._factory(n, cap, self._mb(n))
        first = key not in self._slot_bytes
        self._slot_bytes.setdefault(key, s.device_bytes())
        self.device_bytes += s.device_bytes() + (getattr(s, "shared_bytes", lambda: 0)() if first else 0)
        self.slots.append(s)
        if prepare and hasattr(s, "prepare"):
            s.prepare()
        return s

    def submit(self, req):
        k = (req.n, req.cap)
        if k not in self.pending or not self.served.get(k, False):
            req.error = f"no resident plan for N=2^{req.n}, p={req.cap}"
            req.t_done = time.perf_counter()
            self.rejected.append(req)
            return False
        self.pending[k].append(req)
        return True

    def _shed_old(self, now):
        if self.max_wait_s is None:
            return
        for q in self.pending.values():
            while q and now - q[0].t_arrive > self.max_wait_s:
                r = q.popleft()
                r.error = f"shed: not started within {self.max_wait_s} s"
                r.t_done = now
                self.shed.append(r)

    def _start(self, slot, q):
        batch = [q.popleft() for _ in range(min(slot.max_batch, len(q)))]
        t = time.perf_counter()
        for r in batch:
            r.t_start = t
        slot.launch(batch)

    def _tick(self, slot, q):
        new = [q.popleft() for _ in range(min(slot.free_lanes(), len(q)))]
        t = time.perf_counter()
        for r in new:
            r.t_start = t
        slot.tick(new)

    def step(self):
        now = time.perf_counter()
        for s in self.slots:
            for r in s.poll():
                r.t_done = now
                self.done.append(r)
        self._shed_old(now)
        t1 = time.perf_counter()
        ticks = 0
        for s in self.slots:
            q = self.pending[(s.n, s.cap)]
            if getattr(s, "continuous", False):
                t = time.perf_counter()
                while s.can_tick() and (s.lanes or q):
                    self._tick(s, q)
                    ticks += 1
                self.host_s["tick"] += time.perf_counter() - t
            elif q and not s.busy:
                t = time.perf_counter()
                self._start(s, q)
                self.host_s["launch"] += time.perf_counter() - t
        self.host_s["poll"] += t1 - now
        self.host_s["ticks"] += ticks
        # requests still waiting with every slot of their key busy: one more slot, where memory holds it
        for k, q in self.pending.items():
            if k[0] in self.continuous:
                continue
            if q and sum(1 for s in self.slots if (s.n, s.cap) == k) < self.max_slots and self._fits(k):
                self.grown += 1
                self._start(self._add_slot(k), q)

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
            st.close()
        self._streams.clear()


def poisson_trace(rate, duration, sizes, weights, caps, pool, seed=7):
    """Arrivals with exponential gaps (default_rng(seed)); N from ``sizes`` with ``weights``, p uniform over ``caps``;
    the sinogram of a request is drawn from pool[n] in turn (an array, or a lib.PinnedArray: the request then
    carries its page-locked address)."""
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
        if hasattr(b, "ptr"):
            out.append((t, Request(rid, n, cap, b.a, None, b_ptr=b.ptr)))
        else:
            out.append((t, Request(rid, n, cap, b, None)))
        rid += 1
    return out


def latency_stats(done, t_begin=None, by_size=False):
    """p50 / p99 / mean latency (arrival to result) and throughput of the completed requests; with ``by_size`` also
    per N, with the mean time spent waiting for a slot (arrival to launch) and in the slot (launch to result)."""
    lat = np.array([r.t_done - r.t_arrive for r in done])
    if lat.size == 0:
        return {"count": 0}
    t_first = min(r.t_arrive for r in done) if t_begin is None else t_begin
    t_last = max(r.t_done for r in done)
    out = {
        "count": int(lat.size),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "throughput_per_s": float(lat.size / max(t_last - t_first, 1e-9)),
    }
    if by_size:
        per = {}
        for n in sorted({r.n for r in done}):
            rs = [r for r in done if r.n == n]
            lt = np.array([r.t_done - r.t_arrive for r in rs]) * 1e3
            e = {"count": len(rs), "p50_ms": float(np.percentile(lt, 50)), "p99_ms": float(np.percentile(lt, 99))}
            if all(r.t_start is not None for r in rs):
                e["wait_ms"] = float(np.mean([r.t_start - r.t_arrive for r in rs]) * 1e3)
                e["slot_ms"] = float(np.mean([r.t_done - r.t_start for r in rs]) * 1e3)
            per[str(1 << n)] = e
        out["by_size"] = per
    return out
