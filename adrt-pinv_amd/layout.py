"""The adrt package's sinogram layout and this library's: layout_maps / pkg_to_core / core_to_pkg
(skimage_adrt_pkg.py:17-51), with the permutation applied on the GPU.

Both transforms sum the same digital lines of an N x N image into a (4, 2N-1, N) array, in different places.
``layout_maps`` finds where each line sits in each layout by fingerprints: three images, one triple of line sums per
entry.  The reference fingerprints with standard-normal images and matches values rounded to 8 decimals; here the
images hold random integers in [1, 2^20), so every line sum is an integer below 2^53 and exact in float64 in whatever
order either transform adds, and entries are matched by exact equality (a sort of the triples, no rounding and no
hashing).  All-zero entries (lines that miss the image) have no counterpart and map to 0.  Entries with equal
fingerprints (one line reached from two quadrants) are paired in ascending index order on both sides, as the
reference's grouping pairs them, so noisy data in either layout moves the same way; the pairs are the reference's,
listed in another order.

The package is not a dependency: ``pkg_adrt`` is its forward transform, ``adrt.adrt`` by default (imported on first
use; an ImportError names the keyword when the package is absent).  The core side is this library's GPU transform.
"""

import numpy as np

from . import lib as L
from . import radon


def _pkg_default():
    try:
        import adrt  # the adrt package (PyPI), optional
    except ImportError as e:
        raise ImportError("the adrt package is not installed: pass pkg_adrt=<its forward transform, (N, N) -> "
                          "(4, 2N-1, N)>") from e
    return adrt.adrt


def _entries(B):
    """Indices of the entries with a non-zero fingerprint, sorted by (fingerprint, index), and their fingerprints."""
    idx = np.flatnonzero(np.any(B != 0, axis=0))
    K = B[:, idx]
    order = np.lexsort((idx,) + tuple(K[i] for i in range(K.shape[0] - 1, -1, -1)))
    return idx[order], K[:, order]


def layout_maps(n, pkg_adrt=None, core_adrt=None):
    """(core_idx, pkg_idx) with core.ravel()[core_idx] == pkg.ravel()[pkg_idx] for every entry that a line of the
    image reaches (skimage_adrt_pkg.py:17-35).  core_adrt defaults to this library's GPU transform."""
    pkg_adrt = pkg_adrt or _pkg_default()
    core_adrt = core_adrt or radon.adrt
    N = 1 << n
    F = np.random.default_rng(7).integers(1, 1 << 20, size=(3, N, N)).astype(np.float64)
    fc = np.stack([np.asarray(core_adrt(f), dtype=np.float64).ravel() for f in F])
    fp = np.stack([np.asarray(pkg_adrt(f), dtype=np.float64).ravel() for f in F])
    if fc.shape != fp.shape:
        raise ValueError(f"layout_maps: the package's transform has {fp.shape[1]} entries, the core's {fc.shape[1]}")
    ci, kc = _entries(fc)
    pi, kp = _entries(fp)
    if kc.shape != kp.shape or not np.array_equal(kc, kp):
        raise ValueError("layout_maps: the two transforms do not sum the same lines (fingerprints differ)")
    return ci, pi


class LayoutMap:
    """Both directions of one size's permutation, resident on the device as int32 source maps (-1: no source)."""

    def __init__(self, n, pkg_adrt=None, core_adrt=None):
        self.n = n
        N = 1 << n
        self.M = M = 4 * (2 * N - 1) * N
        self.core_idx, self.pkg_idx = layout_maps(n, pkg_adrt, core_adrt)
        to_core = np.full(M, -1, np.int32)
        to_core[self.core_idx] = self.pkg_idx
        to_pkg = np.full(M, -1, np.int32)
        to_pkg[self.pkg_idx] = self.core_idx
        self._to_core = L.DeviceArray.from_host(to_core)
        self._to_pkg = L.DeviceArray.from_host(to_pkg)

    def _permute(self, dmap, d_in, d_out, S, stream):
        L.check(L.load().adrt_layout_permute(self.n, L.ptr(dmap), L.ptr(d_in), L.ptr(d_out), int(S),
                                              stream.handle if stream else None), "layout_permute")

    def pkg_to_core_device(self, d_bp, d_bc, S=1, stream=None):
        """S float64 sinograms in the package's layout -> this library's (device pointers)."""
        self._permute(self._to_core, d_bp, d_bc, S, stream)

    def core_to_pkg_device(self, d_bc, d_bp, S=1, stream=None):
        """S float64 sinograms in this library's layout -> the package's (device pointers)."""
        self._permute(self._to_pkg, d_bc, d_bp, S, stream)

    def _host(self, fn, b):
        N = 1 << self.n
        b = np.asarray(b, dtype=np.float64)
        single = b.ndim == 3
        bb = np.ascontiguousarray(b.reshape((1,) + b.shape) if single else b)
        if bb.shape[1:] != (4, 2 * N - 1, N):
            raise ValueError(f"expected (S,) 4 x {2 * N - 1} x {N} sinograms, got {b.shape}")
        S = bb.shape[0]
        di = L.DeviceArray.from_host(bb)
        do = L.DeviceArray(bb.shape, np.float64)
        fn(di, do, S)
        out = do.download()
        di.free()
        do.free()
        return out[0] if single else out

    def free(self):
        self._to_core.free()
        self._to_pkg.free()

    def pkg_to_core(self, bp):
        return self._host(self.pkg_to_core_device, bp)

    def core_to_pkg(self, bc):
        return self._host(self.core_to_pkg_device, bc)


_maps = {}  # (n, pkg_adrt) -> LayoutMap; each holds 2 * 4 (2N-1) N int32 on the device (268 MB at N = 2048)


def clear_maps():
    """Free the device maps that pkg_to_core / core_to_pkg keep per size."""
    for m in _maps.values():
        m.free()
    _maps.clear()


def _map(n, pkg_adrt):
    key = (n, pkg_adrt)
    if key not in _maps:
        _maps[key] = LayoutMap(n, pkg_adrt)
    return _maps[key]


def pkg_to_core(bp, n, pkg_adrt=None):
    """skimage_adrt_pkg.py:38-43: a sinogram (or a batch) in the package's layout -> this library's."""
    return _map(n, pkg_adrt or _pkg_default()).pkg_to_core(bp)


def core_to_pkg(bc, n, pkg_adrt=None):
    """skimage_adrt_pkg.py:46-51: a sinogram (or a batch) in this library's layout -> the package's."""
    return _map(n, pkg_adrt or _pkg_default()).core_to_pkg(bc)
