"""The C-ABI library loads on a CPU-only machine and exports every function that
include/jja.h declares (no GPU compute is called here)."""
import ctypes
import os
import re

import pytest

from pyjjasim_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(jja_[a-z_0-9]+)\s*\(", src))
    return names


def test_library_exports_all_declared_symbols():
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libjja.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_native.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    assert set(_native.EXPORTS) <= declared


STRUCTS = {"jja_source": _native.Source, "jja_problem_run": _native.ProblemRun,
           "jja_problem_desc": _native.ProblemDesc, "jja_slice": _native.Slice,
           "jja_adjoint_run": _native.AdjointRun}


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirrors of the descriptor structs have the C compiler's layout."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "jja.h"', "int main(void){"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            name, field, val = line.split()
            got[(name, field)] = int(val)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
