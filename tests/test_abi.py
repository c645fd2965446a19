"""The C-ABI library loads and exports every entry point include/vflip.h declares.

CPU-only: no compute calls are made (there may be no GPU).
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_functions():
    names = []
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"\b(vflip_[a-z0-9_]+)\s*\(", src):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "vflip_solver_solve" in names and "vflip_amgpcg_solve_host" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from vflip_amd._lib import LIB_PATH, SIGNATURES
    if not os.path.exists(LIB_PATH):
        pytest.fail("libvflip.so not built (python -m vflip_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"symbols declared but not exported: {missing}"
    # the Python binding table covers every declared symbol too
    unbound = [n for n in declared_functions() if n not in SIGNATURES]
    assert not unbound, f"declared but not bound in vflip_amd/_lib.py: {unbound}"


def test_error_path_without_compute():
    from vflip_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.vflip_solver_create(0, 4, 4, None, ctypes.byref(h))
    assert rc == -1
    assert b"grid dims" in L.vflip_last_error()
    assert L.vflip_version().startswith(b"vflip")
