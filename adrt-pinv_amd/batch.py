"""Batches of sinograms of uneven N in one hipGraph launch.

The sinograms are grouped by N.  Every group gets its own device plan (max_batch = group size, FP32 row engine with
FP64 dense table / preconditioner / outer iteration and FP64 refinement steps by default), its own HIP stream and its
inputs and outputs resident in HBM.  One graph is captured from a main stream that forks into the group streams and
joins them again, so a single ``hipGraphLaunch`` runs the whole batch with the groups' kernels overlapping on the
device.  Setup scalars come from the plan cache; the row kernel's mapping per N comes from the autotuner.
"""

import numpy as np

from . import autotune as AT
from . import lib as L
from . import plancache
from .inverses import FastHybridHip, FastLpHip


class MixedBatch:
    def __init__(self, ns, kind="hybrid", precision="fp32", refine=1, dense_cap=None, tune=True, **kw):
        self.ns = [int(n) for n in ns]
        self.kind = kind
        self.groups = {}
        for i, n in enumerate(self.ns):
            self.groups.setdefault(n, []).append(i)
        self.main = L.Stream()
        self.plans, self.streams, self.d_b, self.d_x = {}, {}, {}, {}
        self.tuning = {}
        for n, idx in sorted(self.groups.items()):
            S = len(idx)
            N = 1 << n
            if kind == "hybrid":
                lam, outer = plancache.hybrid_setup(n, dense_cap=dense_cap, **kw)
                P = FastHybridHip(n, lam_max=lam, outer=outer, dense_cap=dense_cap, precision=precision, max_batch=S,
                                  refine=refine, **kw)
            else:
                probe = FastLpHip(n, dense_cap=dense_cap, **kw)
                P = FastLpHip(n, lam_max=probe.lam_max, outer=(probe.lo, probe.hi, probe.outer_deg),
                              dense_cap=dense_cap, precision=precision, max_batch=S, refine=refine, **kw)
                probe.close()
            self.plans[n] = P
            self.streams[n] = L.Stream()
            self.d_b[n] = L.DeviceArray((S, 4, 2 * N - 1, N))
            self.d_x[n] = L.DeviceArray((S, N, N))
            if tune and n >= 5:
                self.tuning[n] = AT.tune(P, S, self.streams[n])
        self.graph = None

    def set_inputs(self, sinos):
        """Upload the batch's sinograms (list in the order of ``ns``)."""
        for n, idx in self.groups.items():
            self.d_b[n].upload(np.stack([np.asarray(sinos[i], dtype=float) for i in idx]), self.streams[n])

    def capture(self):
        """Capture the whole batch (every group's apply) into one graph."""
        with L.Graph(self.main) as g:
            for n, idx in sorted(self.groups.items()):
                s = self.streams[n]
                s.wait(self.main)
                self.plans[n].apply_device(self.d_b[n], self.d_x[n], len(idx), s, count=False)
                self.main.wait(s)
        self.graph = g
        return g

    def launch(self):
        if self.graph is None:
            self.capture()
        self.graph.launch(self.main)

    def run(self):
        self.launch()
        self.main.sync()

    def outputs(self):
        out = [None] * len(self.ns)
        for n, idx in self.groups.items():
            x = self.d_x[n].download()
            for j, i in enumerate(idx):
                out[i] = x[j]
        return out

    def close(self):
        for P in self.plans.values():
            P.close()


def draw_sizes(count, choices=(6, 7, 8, 9), weights=(0.4, 0.3, 0.2, 0.1), seed=42):
    """N = 2^n drawn with numpy default_rng(seed).choice, as in the uneven-batch workload."""
    return [int(v) for v in np.random.default_rng(seed).choice(choices, size=count, p=weights)]
