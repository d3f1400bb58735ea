"""CPU checks of the C-ABI library: it loads, exports every symbol include/adrt_hip.h declares, and the ctypes
table covers the header exactly.  No GPU call is made."""

import ctypes
import os
import re

from conftest import REPO, pkg

HEADER = os.path.join(REPO, "include", "adrt_hip.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


def test_header_parses():
    names = header_functions()
    assert {"bfly_rows", "cheb_rows", "cert_rowsum", "omp_max_threads", "adrt_plan_apply"} <= set(names)


def test_library_exports_every_declared_symbol():
    L = pkg("lib")
    lib = L.load()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    L = pkg("lib")
    assert sorted(L.SIGNATURES) == header_functions()


def test_reference_signatures():
    """The four reference symbols take the argument lists fastcpp.py:25-31 binds."""
    L = pkg("lib")
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    assert L.SIGNATURES["bfly_rows"][1] == [I, I, P, P, I, I, I, I]
    assert L.SIGNATURES["cheb_rows"][1] == [I, I, P, P, P, P, D, P, P, I, I, I, P]
    assert L.SIGNATURES["cert_rowsum"][1] == [I, I, P, I, P, I]
    assert L.SIGNATURES["omp_max_threads"] == (I, [])


def test_plan_desc_layout():
    """adrt_plan_desc field order and size as the C compiler lays it out
    (6 ints, 3 doubles, 5 pointers, int + pad, 2 pointers, int + pad, 2 pointers, 2 ints)."""
    L = pkg("lib")
    assert ctypes.sizeof(L.PlanDesc) == 6 * 4 + 3 * 8 + 5 * 8 + 8 + 2 * 8 + 8 + 2 * 8 + 8


def test_every_package_module_imports():
    """Every host module of the package imports without a GPU (syntax, imports, ctypes tables)."""
    import glob

    mods = sorted(os.path.basename(p)[:-3] for p in glob.glob(os.path.join(REPO, "adrt-pinv_amd", "*.py")))
    for m in mods:
        if m in ("__init__", "build"):
            continue
        pkg(m)
