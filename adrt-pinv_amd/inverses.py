"""FastLpHip / FastHybridHip: the reference's public inverse classes on the GPU.

Same constructor keywords, attributes and methods as fastcpp.FastLpCpp / FastHybridCpp over fastops.FastLp /
FastHybrid (fastcpp.py:137-169, fastops.py:261-338): ``lam_max``, ``deg``, ``lo``, ``hi``, ``outer_deg``, ``eng``
(with ``n_dense`` and ``count``), ``apply(b)``, ``backproject(b)``, ``Gh(x)`` / ``frame_window(x)`` and ``P(y)``.

Unlike the reference, the whole packed pass (strips, offset FFT, row ops, inverse FFT, quadrant sum) and the outer
Chebyshev iteration run on the GPU through a device plan; only the setup (bounds, dense tables, the outer interval's
eigensolve) touches the host.  Extra keywords:
  precision  'fp64' (default, the plain path), 'fp32' (float FFTs / butterflies / Chebyshev steps with the Chebyshev
             iterate accumulated by TwoSum compensation; dense table, preconditioner and outer iteration stay in
             double) or 'fp32_plain' ('fp32' without the compensation)
  max_batch  sinograms per device call (``apply_batch``)
  outer      (lo, hi, D) to skip the eigensolve (e.g. to reuse the reference's setup scalars)
  refine     with precision 'fp32': the backprojection and the last ``refine`` outer steps run in FP64
  g1_fp64    (hybrid) with precision 'fp32': the G_1 preconditioner's inner iteration runs in double in the FP32
             outer steps too (default: in float there)
  l1_fp64    (hybrid) with precision 'fp32': the explicit level 1 and the offset FFTs and combine around it run in
             double in the FP32 outer steps (default: in float); the rows' levels 2..n and Chebyshev filter stay float
  ctx        a lib.Context whose device-memory budget the plan is charged to
"""

import ctypes

import numpy as np

from . import engine as E
from . import gpusetup as G
from . import lib as L
from . import spectral


class _PlanBase:
    """Shared part of the GPU plans.  ctx: an optional lib.Context; the plan's device bytes (adrt_plan_bytes) and its
    staging arrays are then charged to that context's budget, and creation fails with HipError (-3) beyond it."""

    KIND = None

    def _setup_common(self, n, tau_fac, eps_in, threads):
        self.n = n
        N = self.N = 1 << n
        self.L = 2 * N
        self.tau = tau_fac * N
        self.eps_in = eps_in
        self.threads = threads
        self.oms = 2 * np.pi * np.fft.fftfreq(self.L)
        self.z = np.exp(1j * self.oms)

    def _degrees(self, lam_max, dense_cap, p_levels):
        self.lam_max, self.deg = spectral.degrees(lam_max, self.tau, self.eps_in)
        self.eng = E.HipEngine(self.n, self.z, self.deg, self.lam_max, self.tau, self.threads, copies=2,
                               dense_cap=dense_cap, p_levels=p_levels, precision=self.precision)

    def _make_plan(self, lo=1.0, hi=2.0, D=1):
        e = self.eng
        self._keep = [np.ascontiguousarray(e.z), np.ascontiguousarray(e.deg, dtype=np.int32), e.a, e.c]
        Fd = Wop = Wback = w = None
        self.table_bytes = 0
        if e.n_dense:
            # dense tables and pass matrices built on the GPU (rocSOLVER Cholesky + rocBLAS GEMMs), column-major on
            # the device, shared by every plan of this size (the plan borrows them; _keep holds them alive)
            kind = "hybrid" if self.KIND == L.KIND_HYBRID else "lp"
            Fd, Wop, Wback, w = G.dense_tables(self.n, self.z, e.dense_k, self.tau, kind)
            self.table_bytes = G.table_bytes((Fd, Wop, Wback))
            self._keep += [np.ascontiguousarray(e.dense_k, dtype=np.int32), Fd, Wop, Wback, w]
        d = L.PlanDesc()
        d.kind = self.KIND
        d.n = self.n
        d.precision = L.PRECISION[self.precision]
        # fastops.FastHybrid.P applies lorr_inv for every precond other than 'g1' (fastops.py:322-323)
        pc = getattr(self, "precond", None)
        d.precond = L.PRECOND["none"] if pc is None else L.PRECOND["g1" if pc == "g1" else "lorr"]
        d.g1_deg = int(getattr(self, "g1_deg", 1))
        d.outer_deg = int(D)
        d.lo, d.hi, d.tau = float(lo), float(hi), float(self.tau)
        d.lam_max = None
        d.z, d.deg, d.a, d.c = (L.ptr(a) for a in self._keep[:4])
        d.n_dense = e.n_dense
        d.dense_k = L.ptr(self._keep[4]) if e.n_dense else None
        d.dense_F = L.ptr(Fd)
        d.max_batch = int(self.max_batch)
        d.dense_Wop = L.ptr(Wop)
        d.dense_Wback = L.ptr(Wback)
        d.refine = int(self.refine)
        d.dense_on_device = 2
        d.dense_w = L.ptr(w)
        d.g1_fp64 = int(bool(getattr(self, "g1_fp64", False)))
        d.l1_fp64 = int(bool(getattr(self, "l1_fp64", False)))
        self._desc = d  # the plan's descriptor, kept current by set_outer (its arrays are held in _keep)
        self.plan_bytes = int(L.load().adrt_plan_bytes(ctypes.byref(d)))
        h = L.P()
        ctx = getattr(self, "ctx", None)
        if ctx is None:
            L.check(L.load().adrt_plan_create(ctypes.byref(h), ctypes.byref(d)), "plan_create")
        else:
            L.check(L.load().adrt_plan_create_in(ctx.handle, ctypes.byref(h), ctypes.byref(d)), "plan_create_in")
        self._plan = h.value
        N = self.N
        self._dimg = L.DeviceArray((self.max_batch, N, N), np.float64, ctx=ctx)
        self._dimg2 = L.DeviceArray((self.max_batch, N, N), np.float64, ctx=ctx)
        self._dsino = L.DeviceArray((self.max_batch, 4, 2 * N - 1, N), np.float64, ctx=ctx)

    def set_outer(self, lo, hi, D):
        self.lo, self.hi, self.outer_deg = float(lo), float(hi), int(D)
        L.check(L.load().adrt_plan_set_outer(self._plan, self.lo, self.hi, self.outer_deg), "plan_set_outer")
        self._desc.lo, self._desc.hi, self._desc.outer_deg = self.lo, self.hi, self.outer_deg

    # ---- per-pass level-row counts (added to eng.count like the reference's engine does)
    def _count(self, ops, S=1):
        self.eng.count[0] += self.eng.steps(ops, 2 * self.L * S)

    # ---- device entry points (device pointers in, device pointers out)
    def backproject_device(self, d_b, d_y, S=1, stream=None):
        L.check(L.load().adrt_plan_backproject(self._plan, L.ptr(d_b), L.ptr(d_y), S, stream.handle if stream else None),
                "plan_backproject")
        self._count(self.ops_back(), S)

    def operator_device(self, d_x, d_y, S=1, stream=None):
        L.check(L.load().adrt_plan_operator(self._plan, L.ptr(d_x), L.ptr(d_y), S, stream.handle if stream else None),
                "plan_operator")
        self._count(self.ops_op(), S)

    def precond_device(self, d_x, d_y, S=1, stream=None):
        L.check(L.load().adrt_plan_precond(self._plan, L.ptr(d_x), L.ptr(d_y), S, stream.handle if stream else None),
                "plan_precond")

    def apply_device(self, d_b, d_x, S=1, stream=None, count=True):
        L.check(L.load().adrt_plan_apply(self._plan, L.ptr(d_b), L.ptr(d_x), S, stream.handle if stream else None),
                "plan_apply")
        if count:
            self._count(self.ops_back(), S)
            for _ in range(self.outer_deg - 1):
                self._count(self.ops_op(), S)

    def tick_device(self, ks, d_b, d_xcur, d_xnext, stream=None):
        """One outer step of every lane of a continuous batch (adrt_plan_tick): lane i at step ks[i], oldest first."""
        S = len(ks)
        arr = (L.I * max(S, 1))(*[int(k) for k in ks])
        L.check(L.load().adrt_plan_tick(self._plan, arr, S, L.ptr(d_b) if d_b is not None else None, L.ptr(d_xcur),
                                        L.ptr(d_xnext), stream.handle if stream else None), "plan_tick")

    def lane_move_device(self, src, dst, d_xa, d_xb, stream=None):
        L.check(L.load().adrt_plan_lane_move(self._plan, int(src), int(dst), L.ptr(d_xa), L.ptr(d_xb),
                                             stream.handle if stream else None), "plan_lane_move")

    # ---- host (numpy) entry points, one sinogram or image at a time like the reference
    def _img_call(self, fn, x):
        x = np.ascontiguousarray(x, dtype=float).reshape(1, self.N, self.N)
        N = self.N
        di = L.DeviceArray((1, N, N))
        do = L.DeviceArray((1, N, N))
        di.upload(x)
        fn(di, do, 1)
        out = do.download()[0]
        di.free()
        do.free()
        return out

    def _sino_call(self, fn, b):
        N = self.N
        b = np.ascontiguousarray(b, dtype=float).reshape(1, 4, 2 * N - 1, N)
        db = L.DeviceArray(b.shape)
        do = L.DeviceArray((1, N, N))
        db.upload(b)
        fn(db, do, 1)
        out = do.download()[0]
        db.free()
        do.free()
        return out

    def backproject(self, b):
        return self._sino_call(self.backproject_device, b)

    def P(self, y):
        return self._img_call(self.precond_device, y)

    def apply(self, b):
        return self._sino_call(self.apply_device, b)

    def apply_batch(self, B, stream=None):
        """B: (S, 4, 2N-1, N) -> (S, N, N); chunks of max_batch sinograms per device call."""
        N = self.N
        B = np.ascontiguousarray(B, dtype=float).reshape(-1, 4, 2 * N - 1, N)
        out = np.empty((B.shape[0], N, N))
        for s0 in range(0, B.shape[0], self.max_batch):
            S = min(self.max_batch, B.shape[0] - s0)
            src = np.ascontiguousarray(B[s0 : s0 + S])
            dst = out[s0 : s0 + S]  # leading-axis slice of a C-contiguous array: contiguous
            L.check(L.load().adrt_memcpy(self._dsino.ptr, src.ctypes.data, src.nbytes, 1, stream.handle if stream else None),
                    "h2d")
            self.apply_device(self._dsino, self._dimg, S, stream)
            L.check(L.load().adrt_memcpy(dst.ctypes.data, self._dimg.ptr, dst.nbytes, 2, stream.handle if stream else None),
                    "d2h")
        return out

    def close(self):
        if getattr(self, "_plan", None):
            L.load().adrt_plan_destroy(self._plan)
            self._plan = None
            self._keep = []  # releases this plan's hold on the shared dense tables
        for a in ("_dimg", "_dimg2", "_dsino"):
            if getattr(self, a, None) is not None:
                getattr(self, a).free()
        if getattr(self, "eng", None):
            self.eng.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FastLpHip(_PlanBase):
    """L_p on the GPU (fastops.FastLp, fastops.py:261-288; fastcpp.FastLpCpp, fastcpp.py:137-153)."""

    KIND = L.KIND_LP

    def __init__(self, n, lam_bound="power", tau_fac=0.2, eps_in=0.1, eps=1e-10, threads=8, lam_max=None,
                 power_iters=40, safety=1.5, dense_cap=None, precision="fp64", max_batch=1, outer=None, refine=0,
                 ctx=None):
        self.ctx = ctx
        self.precision = precision
        self.max_batch = max_batch
        self.refine = refine
        self._setup_common(n, tau_fac, eps_in, threads)
        if lam_max is None:
            lam_max = {
                "power": lambda: E.power_lam_max(n, self.z, 1, power_iters, safety),
                "explicit": lambda: E.explicit_lam_max(n, self.oms),
                "cert": lambda: E.cert_lam_max(n, self.oms, 1),
            }[lam_bound]()
        self._degrees(lam_max, dense_cap, (1, None))
        self._make_plan()
        if outer is None:
            outer = spectral.outer_interval(self.N, self.frame_window, eps=eps)
        self.set_outer(*outer)
        self.eng.count = [0]

    def ops_back(self):
        n = self.n
        return [("BT", 1, n), ("P", 1, n)]

    def ops_op(self):
        n = self.n
        return [("P", 1, n), ("B", 1, n), ("BT", 1, n)]

    def frame_window(self, x):
        return self._img_call(self.operator_device, x)


class FastHybridHip(_PlanBase):
    """Hybrid inverse on the GPU (fastops.FastHybrid, fastops.py:291-338; fastcpp.FastHybridCpp, fastcpp.py:156-169)."""

    KIND = L.KIND_HYBRID

    def __init__(self, n, lam_bound="power", tau_fac=0.2, eps_in=0.1, eps=1e-10, threads=8, lam_max=None, g1_deg=14,
                 power_iters=40, safety=1.5, dense_cap=None, precond="g1", precision="fp64", max_batch=1, outer=None,
                 refine=0, ctx=None, g1_fp64=False, l1_fp64=False):
        self.ctx = ctx
        self.g1_fp64 = g1_fp64
        self.l1_fp64 = l1_fp64
        self.precision = precision
        self.max_batch = max_batch
        self.refine = refine
        self.g1_deg = g1_deg
        self.precond = precond
        self._setup_common(n, tau_fac, eps_in, threads)
        if lam_max is None:
            lam_max = {
                "power": lambda: E.power_lam_max(n, self.z, 2, power_iters, safety),
                "cert": lambda: E.cert_lam_max(n, self.oms, 2),
            }[lam_bound]()
        self._degrees(lam_max, dense_cap, (2, None))
        self._make_plan()
        if outer is None:
            Pinv = spectral.g1_apply if precond == "g1" else (lambda y: 4.0 * y + 8.0 * spectral.pi_apply(y))
            outer = spectral.outer_interval(self.N, self.Gh, Pinv=Pinv, P=self.P, eps=eps)
        self.set_outer(*outer)
        self.eng.count = [0]

    def ops_back(self):
        n = self.n
        return [("BT", 2, n), ("P", 2, n), ("BT", 1, 1)]

    def ops_op(self):
        n = self.n
        return [("B", 1, 1), ("P", 2, n), ("B", 2, n), ("BT", 2, n), ("BT", 1, 1)]

    def Gh(self, x):
        return self._img_call(self.operator_device, x)
