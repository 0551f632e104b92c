"""The C ABI on a machine without a GPU: the library loads and exports every
entry point include/tealeaf_hip.h declares, the header compiles as C99 and
C++, the ctypes and Fortran bindings match its parameter counts and struct
layouts, argument checks that precede any device call hold, and the host
library exports its own entry points."""
import ctypes
import os
import re
import subprocess

import pytest

from tealeaf_ref_amd import _lib, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "tealeaf_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tl_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_exported():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_calls_fail_cleanly_without_init():
    L = _lib.lib()
    out = ctypes.c_double()
    assert L.tl_calc_2norm(1, ctypes.byref(out)) == 2  # TL_ERR_STATE
    assert b"tl_init" in L.tl_last_error()


def test_init_rejects_bad_sizes_before_any_device_call():
    """Size checks come before the first HIP call, so they hold here too."""
    L = _lib.lib()
    for args in ((0, 10, 1, 0), (10, -1, 1, 0), (10, 10, 0, 0), (10, 10, 33, 0)):
        assert L.tl_init(*args) == 3, args  # TL_ERR_ARG
        assert b"tl_init" in L.tl_last_error()
    n = ctypes.c_long()
    assert L.tl_profile_idle(0, ctypes.byref(n)) == 2  # TL_ERR_STATE


def test_header_is_plain_c_and_cxx(tmp_path):
    """include/tealeaf_hip.h compiles as C99 and as C++ on its own, and the
    C usage shown for callers compiles against it."""
    src = tmp_path / "use.c"
    src.write_text(
        '#include "tealeaf_hip.h"\n'
        "int demo(int nx, int ny, const double *rho, const double *e, double rx, double ry) {\n"
        "  int zb[4] = {1, 1, 1, 1};\n"
        "  long n; double ms, norm;\n"
        "  if (tl_init(nx, ny, 1, 0)) return 1;\n"
        "  tl_set_options(TL_CONDUCTIVITY, TL_PREC_NONE, 0);\n"
        "  tl_upload(TL_FIELD_DENSITY, rho, 1); tl_upload(TL_FIELD_ENERGY1, e, 1);\n"
        "  tl_common_init(rx, ry, zb); tl_calc_2norm(0, &norm);\n"
        "  tl_profile_enable(16); tl_profile_get(0, &n, &ms); tl_profile_idle(0, &n);\n"
        "  return tl_finalize();\n"
        "}\n")
    inc = os.path.join(ROOT, "include")
    for cmd in (["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", inc,
                 str(src)],
                ["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++",
                 os.path.join(inc, "tealeaf_hip.h")]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, (cmd, r.stderr)


def header_arity():
    """Parameter count of every function include/tealeaf_hip.h declares."""
    text = open(os.path.join(ROOT, "include", "tealeaf_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:int|const char \*)\s*(tl_\w+)\s*\(([^)]*)\)\s*;", text, re.M):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else len(args.split(","))
    return out


def test_bindings_match_the_header():
    """The Fortran interfaces (tl_hip_api.f90) and the ctypes table (_lib.py)
    bind every function with the header's parameter count."""
    want = header_arity()
    assert len(want) >= 35
    api = open(os.path.join(ROOT, "tealeaf_ref_amd", "host", "tl_hip_api.f90")).read()
    bound = re.findall(r"FUNCTION\s+(tl_\w+)\s*\(([^)]*)\)\s*BIND\(C,\s*name=\"(tl_\w+)\"\)", api)
    assert len(bound) >= 25
    for fname, args, cname in bound:
        assert fname == cname and cname in want, cname
        n = len([a for a in args.split(",") if a.strip()])
        assert n == want[cname], (cname, n, want[cname])
    L = _lib.lib()
    for name, n in want.items():
        if name == "tl_last_error":
            continue
        fn = getattr(L, name)
        assert fn.argtypes is not None and len(fn.argtypes) == n, (name, fn.argtypes, n)


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of tl_state, tl_loop_state and tl_mixed_state have
    the C compiler's size and field offsets."""
    structs = {"tl_state": _lib.TLState, "tl_loop_state": _lib.LoopState,
               "tl_mixed_state": _lib.MixedState}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tealeaf_hip.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.splitlines():
        cname, f, v = line.split()
        got[(cname, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


FLANG = "/opt/rocm/llvm/bin/amdflang"


def test_fortran_struct_layouts_match(tmp_path):
    """The BIND(C) derived types of tl_hip_api.f90 have the same size and
    component offsets as the C structs (checked above against ctypes)."""
    if not os.access(FLANG, os.X_OK):
        pytest.skip("amdflang not present")
    structs = {"tl_state": _lib.TLState, "tl_loop_state": _lib.LoopState,
               "tl_mixed_state": _lib.MixedState}
    body = ["PROGRAM layout", "  USE ISO_C_BINDING", "  USE tl_hip_api", "  IMPLICIT NONE"]
    for cname in structs:
        body.append(f"  TYPE({cname}), TARGET :: v_{cname}")
    body.append("  INTEGER(c_intptr_t) :: base")
    for cname, py in structs.items():
        v = f"v_{cname}"
        body.append(f"  base = TRANSFER(C_LOC({v}), base)")
        body.append(f"  PRINT '(A,1X,A,1X,I0)', '{cname}', 'size', C_SIZEOF({v})")
        for f, _ in py._fields_:
            body.append(f"  PRINT '(A,1X,A,1X,I0)', '{cname}', '{f}', "
                        f"TRANSFER(C_LOC({v}%{f}), base) - base")
    body.append("END PROGRAM layout")
    (tmp_path / "layout.f90").write_text("\n".join(body) + "\n")
    api = os.path.join(ROOT, "tealeaf_ref_amd", "host", "tl_hip_api.f90")
    r = subprocess.run([FLANG, "-c", api, "-o", "api.o"], cwd=tmp_path, capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([FLANG, "layout.f90", "api.o", "-o", "layout"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(tmp_path / "layout")], capture_output=True, text=True, check=True)
    got = {}
    for line in out.stdout.splitlines():
        cname, f, v = line.split()
        got[(cname, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_host_library_exports():
    L = host.host_lib()
    for n in ("tea_host_init", "tea_host_step", "tea_host_run", "tea_host_last",
              "tea_host_finalize", "tea_host_guarded_init", "tea_host_guarded_step",
              "tea_host_guarded_run", "tea_host_guarded_finalize", "tea_host_error"):
        assert hasattr(L, n)
    assert os.access(host.TEA_LEAF_BIN, os.X_OK)
