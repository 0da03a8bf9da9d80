"""The C ABI library loads and exports every entry point include/fsl.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import fsl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fsl.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fsl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    syms = declared_symbols()
    assert "fsl_create" in syms and "fsl_pcg_solve" in syms and "fsl_step" in syms
    assert len(syms) == len(fsl.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(fsl.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    assert set(declared_symbols()) == set(fsl.SIGNATURES)


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", fsl.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fsl_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared_symbols()) <= exported


def test_binding_loads_and_rejects_bad_dims():
    lib = fsl.load()
    h = ctypes.c_void_p()
    dims = (ctypes.c_int * 3)(12, 16, 16)  # not a multiple of 8
    assert lib.fsl_create(ctypes.byref(h), dims, None, 0, 0) == -1
    assert lib.fsl_destroy(None) == -1
    assert lib.fsl_last_error(None) == b"null context"


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of fsl_solve_params and fsl_solve_stats have the C
    header's field offsets and sizes (a C program compiled against
    include/fsl.h prints them)."""
    import ctypes as C
    import subprocess

    import fsl

    structs = {"fsl_solve_params": fsl.SolveParams, "fsl_solve_stats": fsl.SolveStats,
               "fsl_frame_stats": fsl.FrameStats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "fsl.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)
