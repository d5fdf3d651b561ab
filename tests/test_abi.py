"""The C ABI: the library loads, exports every declared entry point, and the
ctypes structures have the C layout (no compute calls: runs without a GPU)."""

import ctypes
import os
import re
import subprocess

import pytest

from betse_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "betse_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(bt_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_bound_functions():
    decl = declared_functions()
    assert decl == sorted(_native.SIGNATURES), (decl, sorted(_native.SIGNATURES))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    lib = _native.load()
    assert lib.bt_version() >= 100


def test_load_fails_loudly_when_missing(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _native.load(str(tmp_path / "nope.so"))


def test_struct_layout_matches_c(tmp_path):
    structs = {"p": ("bt_params", _native.BtParams), "b": ("bt_buffers", _native.BtBuffers),
               "c": ("bt_cluster", _native.BtCluster), "o": ("bt_obs", _native.BtObs),
               "r": ("bt_rate", _native.BtRate), "n": ("bt_network", _native.BtNetwork),
               "t": ("bt_tuning", _native.BtTuning)}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "betse_amd.h"', 'int main(void){']
    for tag, (cname, cls) in structs.items():
        src.append(f'printf("s {tag} %zu\\n", sizeof({cname}));')
        for f in (x[0] for x in cls._fields_):
            src.append(f'printf("{tag} {f} %zu\\n", offsetof({cname}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in out:
        if not line:
            continue
        a, b, c = line.split()
        if a == "s":
            assert int(c) == ctypes.sizeof(structs[b][1]), b
        else:
            assert getattr(structs[a][1], b).offset == int(c), (a, b)
