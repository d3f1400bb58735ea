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
    src = re.sub(r"^typedef[^;]*;", "", src, flags=re.M | re.S)  # function-pointer types declare no symbol
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


def test_plan_desc_layout(tmp_path):
    """The ctypes adrt_plan_desc has every field of the header's struct at the offset the C compiler gives it, and
    the same size (gcc compiles the header and prints them)."""
    import subprocess

    L = pkg("lib")
    names = [f[0] for f in L.PlanDesc._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"adrt_hip.h\"\nint main(void) {\n"
                   '  printf("%zu\\n", sizeof(adrt_plan_desc));\n'
                   + "".join(f'  printf("%zu\\n", offsetof(adrt_plan_desc, {n}));\n' for n in names)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)])
    got = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    assert got[0] == ctypes.sizeof(L.PlanDesc)
    assert got[1:] == [getattr(L.PlanDesc, n).offset for n in names]


def test_every_package_module_imports():
    """Every host module of the package imports without a GPU (syntax, imports, ctypes tables)."""
    import glob

    mods = sorted(os.path.basename(p)[:-3] for p in glob.glob(os.path.join(REPO, "adrt-pinv_amd", "*.py")))
    for m in mods:
        if m in ("__init__", "build"):
            continue
        pkg(m)


C_CALLER = r"""
#include <stdio.h>
#include "adrt_hip.h"
int main(void) {
    int deg[128];
    double ac[128];
    adrt_plan_desc d = {0};
    adrt_ctx* ctx = 0;
    for (int i = 0; i < 128; ++i) { deg[i] = 1 + i % 5; ac[i] = 1.0; }
    d.kind = ADRT_KIND_HYBRID; d.n = 6; d.precision = ADRT_PREC_F32; d.precond = ADRT_PRECOND_G1;
    d.g1_deg = 14; d.outer_deg = 10; d.lo = 1.0; d.hi = 2.0; d.tau = 12.8;
    d.deg = deg; d.a = ac; d.c = ac; d.max_batch = 4;
    printf("%llu\n", adrt_plan_bytes(&d));
    int st = adrt_ctx_create(0, 0);
    printf("%d %s\n", st, adrt_last_error());
    printf("%d\n", adrt_ctx_destroy(ctx));
    return 0;
}
"""


def test_header_compiles_as_c_and_links(tmp_path):
    """A C99 caller of include/adrt_hip.h compiles with gcc, links the library and gets the same host-side answers as
    the ctypes binding (adrt_plan_bytes; argument errors before any device call)."""
    import subprocess

    import numpy as np

    L = pkg("lib")
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"), str(src), "-o",
                    str(exe), "-L", libdir, "-ladrt_hip", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    d = L.PlanDesc()
    deg = np.array([1 + i % 5 for i in range(128)], dtype=np.int32)
    ac = np.ones(128)
    d.kind, d.n, d.precision, d.precond = L.KIND_HYBRID, 6, L.PRECISION["fp32"], L.PRECOND["g1"]
    d.g1_deg, d.outer_deg, d.lo, d.hi, d.tau = 14, 10, 1.0, 2.0, 12.8
    d.deg, d.a, d.c, d.max_batch = L.ptr(deg), L.ptr(ac), L.ptr(ac), 4
    assert int(out[0]) == L.load().adrt_plan_bytes(ctypes.byref(d)) > 0
    assert out[1].startswith("-1 ") and "null" in out[1]
    assert out[2] == "0"
