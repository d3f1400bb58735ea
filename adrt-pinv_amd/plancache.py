"""Plan cache: the data-independent setup scalars of an inverse (bounds Lambda_k, outer interval lo / hi, outer degree D)
persisted per (kind, N, parameters) as .npz, so a server or a batch does not repeat the setup's power iteration and
eigensolve.  The dense tables are rebuilt on the GPU from these scalars (they are large and cheap to rebuild).

Cache directory: $ADRT_PLAN_CACHE, else ~/.cache/adrt-pinv_amd.  Setting ADRT_PLAN_CACHE=off disables it.
"""

import hashlib
import json
import os

import numpy as np


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


def key(kind, n, **params):
    blob = json.dumps({"kind": kind, "n": n, **{k: params[k] for k in sorted(params)}}, sort_keys=True, default=str)
    return f"{kind}_n{n}_" + hashlib.sha1(blob.encode()).hexdigest()[:16]


def load(kind, n, **params):
    d = cache_dir()
    if d is None:
        return None
    path = os.path.join(d, key(kind, n, **params) + ".npz")
    if not os.path.exists(path):
        return None
    z = np.load(path, allow_pickle=False)
    return {"lam_max": z["lam_max"], "outer": (float(z["lo"]), float(z["hi"]), int(z["D"]))}


def store(kind, n, lam_max, outer, **params):
    d = cache_dir()
    if d is None:
        return
    path = os.path.join(d, key(kind, n, **params) + ".npz")
    tmp = path + f".{os.getpid()}.tmp.npz"
    np.savez(tmp, lam_max=np.asarray(lam_max, dtype=float), lo=outer[0], hi=outer[1], D=outer[2])
    os.replace(tmp, path)


def hybrid_setup(n, **kw):
    """(lam_max, (lo, hi, D)) of FastHybridHip(n, **kw) in FP64, from the cache or computed and stored."""
    from .inverses import FastHybridHip

    params = {k: kw[k] for k in sorted(kw) if k not in ("max_batch", "precision", "refine")}
    hit = load("hybrid", n, **params)
    if hit is not None:
        return hit["lam_max"], hit["outer"]
    H = FastHybridHip(n, max_batch=1, precision="fp64", **params)
    out = (H.lo, H.hi, H.outer_deg)
    store("hybrid", n, H.lam_max, out, **params)
    lam = H.lam_max
    H.close()
    return lam, out
