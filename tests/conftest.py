"""Test configuration: the ``gpu`` marker, import helpers for the package (its directory name has a hyphen) and
the golden fixtures."""

import importlib
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def pkg(sub=None):
    name = "adrt-pinv_amd" + (f".{sub}" if sub else "")
    return importlib.import_module(name)


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="session")
def gpu():
    """The loaded HIP library; fails (does not skip) when no GPU is visible, so a GPU run cannot pass vacuously."""
    L = pkg("lib")
    lib = L.load()
    n = lib.adrt_device_count()
    assert n >= 1, "no GPU visible to the HIP runtime"
    return L


def other_layout(n):
    """Index map of a stand-in for another package's sinogram layout: pkg.ravel() = core.ravel()[idx].  Quadrants
    reordered (2, 0, 3, 1), slopes reversed in the new quadrants 1 and 3, and each slope column's offsets rolled by
    slope + 1 (so the lines that miss the image sit elsewhere).  The adrt package itself is not installed."""
    N = 1 << n
    idx = np.arange(4 * (2 * N - 1) * N).reshape(4, 2 * N - 1, N)[[2, 0, 3, 1]].copy()
    idx[[1, 3]] = idx[[1, 3]][:, :, ::-1]
    for s in range(N):
        idx[:, :, s] = np.roll(idx[:, :, s], s + 1, axis=0)
    return idx.ravel()


def other_package(n, core):
    """A forward transform in the other_layout(n) layout, built on the transform ``core`` (N, N) -> (4, 2N-1, N)."""
    N = 1 << n
    idx = other_layout(n)
    return lambda x: np.asarray(core(x)).ravel()[idx].reshape(4, 2 * N - 1, N)
