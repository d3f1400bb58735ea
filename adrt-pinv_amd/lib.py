"""ctypes binding of libadrt_hip.so (declared in include/adrt_hip.h) and a small device-buffer type.

The library is loaded from this package directory; if it is missing the import fails with the build command to run
(there is no CPU fallback).  ``DeviceArray`` owns a hipMalloc'ed buffer with a numpy shape and dtype; PyTorch-ROCm
tensors can be used instead anywhere a device pointer is taken (pass ``tensor.data_ptr()``).
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ADRT_HIP_LIB", os.path.join(HERE, "libadrt_hip.so"))

P = ctypes.c_void_p
I = ctypes.c_int
LL = ctypes.c_longlong
ULL = ctypes.c_ulonglong
D = ctypes.c_double

# symbol -> (restype, argtypes); every entry point of include/adrt_hip.h
SIGNATURES = {
    "bfly_rows": (None, [I, I, P, P, I, I, I, I]),
    "cheb_rows": (None, [I, I, P, P, P, P, D, P, P, I, I, I, P]),
    "cert_rowsum": (None, [I, I, P, I, P, I]),
    "omp_max_threads": (I, []),
    "adrt_bfly_rows": (I, [I, I, P, P, I, I, I]),
    "adrt_cheb_rows": (I, [I, I, P, P, P, P, D, P, P, I, I, P]),
    "adrt_cert_rowsum": (I, [I, I, P, I, P]),
    "adrt_last_error": (ctypes.c_char_p, []),
    "adrt_device_count": (I, []),
    "adrt_malloc": (I, [ctypes.POINTER(P), ULL]),
    "adrt_free": (I, [P]),
    "adrt_memcpy": (I, [P, P, ULL, I, P]),
    "adrt_memset": (I, [P, I, ULL, P]),
    "adrt_host_malloc": (I, [ctypes.POINTER(P), ULL]),
    "adrt_host_free": (I, [P]),
    "adrt_stream_create": (I, [ctypes.POINTER(P)]),
    "adrt_stream_create_own_queue": (I, [ctypes.POINTER(P)]),
    "adrt_own_queue_count": (I, []),
    "adrt_stream_destroy": (I, [P]),
    "adrt_stream_sync": (I, [P]),
    "adrt_stream_wait": (I, [P, P]),
    "adrt_stream_wait_event": (I, [P, P]),
    "adrt_device_sync": (I, []),
    "adrt_device_mem_info": (I, [ctypes.POINTER(ULL), ctypes.POINTER(ULL)]),
    "adrt_event_create": (I, [ctypes.POINTER(P)]),
    "adrt_event_destroy": (I, [P]),
    "adrt_event_record": (I, [P, P]),
    "adrt_event_elapsed_ms": (I, [P, P, ctypes.POINTER(ctypes.c_float)]),
    "adrt_event_query": (I, [P]),
    "adrt_rows_config_set": (I, [I, I, I]),
    "adrt_rows_config_get": (I, [I, I]),
    "adrt_dense_config_set": (I, [I, I, I]),
    "adrt_dense_config_get": (I, [I, I]),
    "adrt_rows_plan_create": (I, [ctypes.POINTER(P), I, I, P, P, P, P, D, I, P, P, I]),
    "adrt_rows_plan_destroy": (I, [P]),
    "adrt_rows_run": (I, [P, P, P, LL, P, I, P]),
    "adrt_plan_bytes": (ULL, [P]),
    "adrt_plan_create": (I, [ctypes.POINTER(P), P]),
    "adrt_plan_destroy": (I, [P]),
    "adrt_ctx_create": (I, [ctypes.POINTER(P), ULL]),
    "adrt_ctx_destroy": (I, [P]),
    "adrt_ctx_stream": (P, [P]),
    "adrt_ctx_usage": (I, [P, ctypes.POINTER(ULL), ctypes.POINTER(ULL)]),
    "adrt_ctx_malloc": (I, [P, ctypes.POINTER(P), ULL]),
    "adrt_ctx_free": (I, [P, P]),
    "adrt_plan_create_in": (I, [P, ctypes.POINTER(P), P]),
    "adrt_plan_set_outer": (I, [P, D, D, I]),
    "adrt_plan_backproject": (I, [P, P, P, I, P]),
    "adrt_plan_operator": (I, [P, P, P, I, P]),
    "adrt_plan_precond": (I, [P, P, P, I, P]),
    "adrt_plan_rows_pass": (I, [P, I, I, P]),
    "adrt_plan_time_rows": (I, [P, P, P]),
    "adrt_plan_time_rows_ring": (I, [P, ctypes.POINTER(P), ctypes.POINTER(P), I]),
    "adrt_plan_time_rows_count": (I, [P, ctypes.POINTER(LL)]),
    "adrt_plan_refine_steps": (I, [P]),
    "adrt_cgls_work_bytes": (ULL, [I, I, I]),
    "adrt_cgls": (I, [I, P, P, I, D, I, I, P, P, P, I, P]),
    "adrt_cgls_cb": (I, [I, P, P, I, D, I, I, P, P, P, I, P, P, P]),
    "adrt_plan_apply": (I, [P, P, P, I, P]),
    "adrt_plan_tick": (I, [P, ctypes.POINTER(I), I, P, P, P, P]),
    "adrt_plan_lane_move": (I, [P, I, I, P, P, P]),
    "adrt_graph_begin": (I, [P]),
    "adrt_graph_end": (I, [P, ctypes.POINTER(P)]),
    "adrt_graph_launch": (I, [P, P]),
    "adrt_graph_destroy": (I, [P]),
    "adrt_work_bytes_fwd": (ULL, [I, I, I]),
    "adrt_work_bytes_T": (ULL, [I, I, I]),
    "adrt_forward": (I, [I, I, P, P, P, I, P]),
    "adrt_forward_noalloc": (I, [I, I, P, P, I, P]),
    "adrt_noalloc_scratch_bytes": (ULL, []),
    "adrt_layout_permute": (I, [I, P, P, P, I, P]),
    "adrt_transpose": (I, [I, I, P, P, P, I, P]),
}


# adrt_cgls_callback: (it, d_x, gnorm, ran, S, user)
CGLS_CALLBACK = ctypes.CFUNCTYPE(None, I, P, ctypes.POINTER(D), ctypes.POINTER(I), I, P)


class PlanDesc(ctypes.Structure):
    """adrt_plan_desc of include/adrt_hip.h."""

    _fields_ = [
        ("kind", I), ("n", I), ("precision", I), ("precond", I), ("g1_deg", I), ("outer_deg", I),
        ("lo", D), ("hi", D), ("tau", D),
        ("lam_max", P), ("z", P), ("deg", P), ("a", P), ("c", P),
        ("n_dense", I), ("dense_k", P), ("dense_F", P), ("max_batch", I),
        ("dense_Wop", P), ("dense_Wback", P), ("refine", I), ("dense_on_device", I), ("g1_fp64", I),
        ("l1_fp64", I), ("dense_w", P),
    ]


KIND_LP, KIND_HYBRID, KIND_NORMAL = 0, 1, 2
MAX_OWN_QUEUES = 4  # ADRT_MAX_OWN_QUEUES: own-queue streams alive at once
PRECOND = {"none": 0, "g1": 1, "lorr": 2}
PRECISION = {"fp64": 0, "fp32": 1, "fp32_plain": 2}
# the row-kernel configuration table is kept per compute type: both FP32 variants share one entry
CONFIG_PRECISION = {"fp64": 0, "fp32": 1, "fp32_plain": 1}
OPS = {"B": 0, "BT": 1, "P": 2}
DTYPES = {np.dtype(np.int32): 0, np.dtype(np.float32): 1, np.dtype(np.float64): 2}


class HipError(RuntimeError):
    pass


_lib = None


def load():
    """Load libadrt_hip.so once; raises ImportError (with the build command) if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python adrt-pinv_amd/build.py` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status, what=""):
    if status != 0:
        msg = load().adrt_last_error()
        raise HipError(f"{what} failed ({status}): {msg.decode() if msg else ''}")


def ptr(a):
    """Raw pointer of a numpy array, a DeviceArray, a torch tensor or an int."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    raise TypeError(type(a))


def gpu_available():
    try:
        return load().adrt_device_count() > 0
    except (ImportError, OSError):
        return False


class Stream:
    """A HIP stream owned by Python: non-blocking, or with own_queue on a hardware queue of its own
    (adrt_stream_create_own_queue), so its work never waits behind another stream's (it synchronises with the null
    stream)."""

    def __init__(self, own_queue=False):
        h = P()
        if own_queue:
            check(load().adrt_stream_create_own_queue(ctypes.byref(h)), "stream_create_own_queue")
        else:
            check(load().adrt_stream_create(ctypes.byref(h)), "stream_create")
        self.handle = h.value

    def sync(self):
        check(load().adrt_stream_sync(self.handle), "stream_sync")

    def wait(self, other):
        check(load().adrt_stream_wait(self.handle, other.handle if other else None), "stream_wait")

    def wait_event(self, ev):
        """Later work on this stream waits for the last record of ``ev`` (an Event)."""
        check(load().adrt_stream_wait_event(self.handle, ev.handle), "stream_wait_event")

    def close(self):
        if getattr(self, "handle", None):
            load().adrt_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """A stream and a device-memory budget (adrt_ctx_*): plans and device arrays created with ``ctx=`` charge their
    bytes to it, and a request beyond the budget raises HipError (status -3) without allocating.  budget 0: no
    limit."""

    def __init__(self, budget=0):
        h = P()
        check(load().adrt_ctx_create(ctypes.byref(h), int(budget)), "ctx_create")
        self.handle = h.value
        self.stream = _BorrowedStream(load().adrt_ctx_stream(self.handle))

    def usage(self):
        used, budget = ULL(), ULL()
        check(load().adrt_ctx_usage(self.handle, ctypes.byref(used), ctypes.byref(budget)), "ctx_usage")
        return used.value, budget.value

    def close(self):
        """Fails (HipError, status -4) while plans or arrays of the context are alive."""
        if getattr(self, "handle", None):
            check(load().adrt_ctx_destroy(self.handle), "ctx_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _BorrowedStream(Stream):
    """A stream owned by someone else (a context): usable like Stream, never destroyed from Python."""

    def __init__(self, handle):
        self.handle = handle

    def close(self):
        self.handle = None


class PinnedArray:
    """Page-locked host buffer (adrt_host_malloc) seen as a numpy array ``a``: host<->device copies from or into it
    run asynchronously on their stream."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        h = P()
        check(load().adrt_host_malloc(ctypes.byref(h), max(1, self.nbytes)), "host_malloc")
        self.ptr = h.value
        buf = (ctypes.c_char * max(1, self.nbytes)).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(self.shape, dtype=np.int64))).reshape(self.shape)

    def free(self):
        if getattr(self, "ptr", None):
            self.a = None
            load().adrt_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """hipMalloc'ed buffer with a numpy shape/dtype; with ``ctx`` its bytes are charged to that Context."""

    def __init__(self, shape, dtype=np.float64, ctx=None):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._ctx = ctx
        h = P()
        if ctx is None:
            check(load().adrt_malloc(ctypes.byref(h), max(1, self.nbytes)), "malloc")
        else:
            check(load().adrt_ctx_malloc(ctx.handle, ctypes.byref(h), max(1, self.nbytes)), "ctx_malloc")
        self.ptr = h.value

    @classmethod
    def from_host(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(load().adrt_memcpy(self.ptr, a.ctypes.data, self.nbytes, 1, stream.handle if stream else None), "h2d")
        if stream is None:
            check(load().adrt_stream_sync(None), "sync")
        else:
            stream.sync()

    def download(self, stream=None, out=None):
        out = np.empty(self.shape, self.dtype) if out is None else out
        check(load().adrt_memcpy(out.ctypes.data, self.ptr, self.nbytes, 2, stream.handle if stream else None), "d2h")
        return out

    def zero(self, stream=None):
        check(load().adrt_memset(self.ptr, 0, self.nbytes, stream.handle if stream else None), "memset")

    def free(self):
        if getattr(self, "ptr", None):
            if self._ctx is None:
                load().adrt_free(self.ptr)
            else:
                load().adrt_ctx_free(self._ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = P()
        check(load().adrt_event_create(ctypes.byref(h)), "event_create")
        self.handle = h.value

    def record(self, stream=None):
        check(load().adrt_event_record(self.handle, stream.handle if stream else None), "event_record")

    def done(self):
        st = load().adrt_event_query(self.handle)
        if st < 0:
            check(st, "event_query")
        return st == 1

    def elapsed_ms(self, stop):
        ms = ctypes.c_float()
        check(load().adrt_event_elapsed_ms(self.handle, stop.handle, ctypes.byref(ms)), "event_elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            load().adrt_event_destroy(self.handle)
        except Exception:
            pass


class Graph:
    """A hipGraph captured from everything enqueued on ``stream`` between begin() and end()."""

    def __init__(self, stream):
        self.stream = stream
        self.handle = None

    def __enter__(self):
        check(load().adrt_graph_begin(self.stream.handle), "graph_begin")
        return self

    def __exit__(self, et, ev, tb):
        h = P()
        st = load().adrt_graph_end(self.stream.handle, ctypes.byref(h))
        if et is None:
            check(st, "graph_end")
        self.handle = h.value
        return False

    def launch(self, stream=None):
        check(load().adrt_graph_launch(self.handle, (stream or self.stream).handle), "graph_launch")

    def __del__(self):
        try:
            if self.handle:
                load().adrt_graph_destroy(self.handle)
        except Exception:
            pass


def ops_array(ops):
    """[('B'|'BT'|'P', lo, hi), ...] -> int32 triples."""
    return np.array([[OPS[o], int(lo), int(hi)] for o, lo, hi in ops], dtype=np.int32).reshape(-1)
