"""Plan cache: the data-independent setup scalars of an inverse (bounds Lambda_k, outer interval lo / hi, outer degree D)
persisted per (kind, N, parameters) as .npz, so a server or a batch does not repeat the setup's power iteration and
eigensolve.  The dense tables are rebuilt on the GPU from these scalars (they are large and cheap to rebuild).

Cache directory: $ADRT_PLAN_CACHE, else ~/.cache/adrt-pinv_amd.  Setting ADRT_PLAN_CACHE=off disables it.
"""

import hashlib
import inspect
import json
import os

import numpy as np

# Bumped whenever the setup arithmetic (bounds, degrees, outer interval) or this file's format changes, so entries
# written by an older build are not reused.
VERSION = 2
# keywords that do not change the setup scalars of an FP64 plan
_NOT_SETUP = ("max_batch", "precision", "refine", "ctx", "outer", "lam_max", "g1_fp64", "l1_fp64")


def cache_dir():
    d = os.environ.get("ADRT_PLAN_CACHE")
    if d == "off":
        return None
    d = d or os.path.join(os.path.expanduser("~"), ".cache", "adrt-pinv_amd")
    try:
        os.makedirs(d, exist_ok=True)
    except OSError:
        return None
    return d


def resolved(cls, **kw):
    """Every setup keyword of ``cls`` (the given ones over the constructor's defaults), without those that do not
    change the setup scalars."""
    sig = inspect.signature(cls.__init__)
    out = {k: v.default for k, v in sig.parameters.items()
           if k not in ("self", "n") and v.default is not inspect.Parameter.empty}
    unknown = set(kw) - set(out)
    if unknown:
        raise TypeError(f"{cls.__name__} has no keyword {sorted(unknown)}")
    out.update(kw)
    return {k: out[k] for k in sorted(out) if k not in _NOT_SETUP}


def _blob(kind, n, params):
    return json.dumps({"version": VERSION, "kind": kind, "n": n, **params}, sort_keys=True, default=str)


def key(kind, n, **params):
    return f"{kind}_n{n}_" + hashlib.sha1(_blob(kind, n, params).encode()).hexdigest()[:16]


def load(kind, n, **params):
    """The cached setup of (kind, n, params) or None; an entry whose stored parameters or version differ from the
    request (a hash collision or a file of another build) is ignored."""
    d = cache_dir()
    if d is None:
        return None
    path = os.path.join(d, key(kind, n, **params) + ".npz")
    if not os.path.exists(path):
        return None
    try:
        z = np.load(path, allow_pickle=False)
        if "params" not in z.files or str(z["params"]) != _blob(kind, n, params):
            return None
        return {"lam_max": z["lam_max"], "outer": (float(z["lo"]), float(z["hi"]), int(z["D"]))}
    except (OSError, ValueError, KeyError):
        return None


def store(kind, n, lam_max, outer, **params):
    d = cache_dir()
    if d is None:
        return
    path = os.path.join(d, key(kind, n, **params) + ".npz")
    tmp = path + f".{os.getpid()}.tmp.npz"
    np.savez(tmp, lam_max=np.asarray(lam_max, dtype=float), lo=outer[0], hi=outer[1], D=outer[2],
             params=np.array(_blob(kind, n, params)))
    os.replace(tmp, path)


def hybrid_setup(n, **kw):
    """(lam_max, (lo, hi, D)) of FastHybridHip(n, **kw) in FP64, from the cache or computed and stored.  The key is
    every setup keyword of FastHybridHip with its default filled in, plus VERSION."""
    from .inverses import FastHybridHip

    params = resolved(FastHybridHip, **kw)
    hit = load("hybrid", n, **params)
    if hit is not None:
        return hit["lam_max"], hit["outer"]
    H = FastHybridHip(n, max_batch=1, precision="fp64", **params)
    out = (H.lo, H.hi, H.outer_deg)
    store("hybrid", n, H.lam_max, out, **params)
    lam = H.lam_max
    H.close()
    return lam, out
