"""The adrt package <-> core sinogram layout maps, restated (oracle / test infrastructure only).

skimage_adrt_pkg.py:17-51: both transforms of three standard-normal images (default_rng(7)), entries keyed by their
triple of values rounded to 8 decimals, all-zero entries (lines that miss the image) dropped, groups of equal keys
paired in the order the entries appear (core order for the groups, ascending index inside a group on both sides).
``core_adrt`` and ``pkg_adrt`` are the two forward transforms (N, N) -> (4, 2N-1, N); the reference calls
adrtcore.adrt and the package's adrt.adrt, which is not installed here, so the tests pass known layouts.

Parity pins (tests/test_oracle.py): for a package whose layout is a known permutation of the core's (quadrants
reordered, slopes reversed, offsets rolled by a slope-dependent amount), the maps take fresh transforms of new
images from one layout to the other exactly, and every non-structural entry is paired once.
"""

from collections import defaultdict

import numpy as np


def layout_maps(n, core_adrt, pkg_adrt):
    """(core_idx, pkg_idx) with core.ravel()[core_idx] == pkg.ravel()[pkg_idx] (skimage_adrt_pkg.py:17-35)."""
    N = 1 << n
    F = np.random.default_rng(7).standard_normal((3, N, N))
    fc = np.stack([np.asarray(core_adrt(f), dtype=float).ravel() for f in F], 1)
    fp = np.stack([np.asarray(pkg_adrt(f), dtype=float).ravel() for f in F], 1)
    gp, gc = defaultdict(list), defaultdict(list)
    for i, v in enumerate(fp):
        if np.abs(v).max() > 1e-12:
            gp[tuple(np.round(v, 8))].append(i)
    for i, v in enumerate(fc):
        if np.abs(v).max() > 1e-12:
            gc[tuple(np.round(v, 8))].append(i)
    assert gp.keys() == gc.keys() and all(len(gp[k]) == len(gc[k]) for k in gp)
    return (np.array([i for k in gc for i in gc[k]]), np.array([i for k in gc for i in gp[k]]))


def pkg_to_core(bp, n, maps):
    """skimage_adrt_pkg.py:38-43: zeros, then the core entries taken from their package entries."""
    N = 1 << n
    ci, pi = maps
    out = np.zeros(4 * (2 * N - 1) * N)
    out[ci] = np.asarray(bp, dtype=float).ravel()[pi]
    return out.reshape(4, 2 * N - 1, N)


def core_to_pkg(bc, n, maps):
    """skimage_adrt_pkg.py:46-51: the inverse placement."""
    N = 1 << n
    ci, pi = maps
    out = np.zeros(4 * (2 * N - 1) * N)
    out[pi] = np.asarray(bc, dtype=float).ravel()[ci]
    return out.reshape(4, 2 * N - 1, N)
