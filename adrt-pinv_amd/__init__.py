"""MI355X-native engine for the ADRT oblique inverses (L_p and the hybrid) of dsrim/adrt-pinv.

Python host code over a hand-written CDNA4 HIP shared library (``csrc/``, built into ``libadrt_hip.so``) reached
through a ctypes C ABI.  Submodules are imported lazily so that ``phantom`` works without the library.
"""

import importlib

__all__ = ["lib", "engine", "inverses", "radon", "batch", "serve", "phantom", "autotune"]


def __getattr__(name):
    if name in __all__:
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
