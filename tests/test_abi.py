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


def test_mis_shaped_arrays_raise_before_any_launch():
    """Staggered (MAC) arrays are checked against their face counts before the C ABI is
    called: a wrongly shaped tensor raises ValueError instead of reaching a kernel."""
    import torch
    from vflip_amd import api
    ni, nj, nk = 6, 5, 4  # distinct dims: the three face counts differ
    u, v, w = torch.zeros((nk, nj, ni + 1)), torch.zeros((nk, nj + 1, ni)), torch.zeros((nk + 1, nj, ni))
    p = torch.zeros((nk, nj, ni))
    with pytest.raises(ValueError, match="vw has"):
        api.grad_apply(api.COARSE, 0.1, 1.0, p, u, w, w, u, v, w, u, v, w)
    with pytest.raises(ValueError, match="w has"):
        api.advect_rk2(0.1, 0.01, u, v, v, p, (0, 0, 0), torch.zeros_like(p))
    with pytest.raises(ValueError, match="out has"):
        api.advect_rk2(0.1, 0.01, u, v, w, p, (0, 0, 0), torch.zeros_like(u))
    with pytest.raises(ValueError, match="u has"):
        api.inflow_bc(torch.zeros((nk, nj, ni), dtype=torch.int8), 1.0, p, v, w)
    S = object.__new__(api.Solver)  # the checks run before the handle is used
    S.ni, S.nj, S.nk = ni, nj, nk
    S.n = ni * nj * nk
    S.h = None
    args = [u, v, w, u, v, w, u, v, w]
    bad = list(args)
    bad[7] = w  # vs with w's shape
    with pytest.raises(ValueError, match="vs has"):
        S.assemble(api.COARSE, 0.1, 1.0, *bad, p)
    with pytest.raises(ValueError, match="rhs has"):
        S.project(api.COARSE, 0.1, 1.0, *args, u, p)
    # right shapes on the host reach the device check
    with pytest.raises(ValueError, match="device tensor"):
        api.grad_apply(api.COARSE, 0.1, 1.0, p, u, v, w, u, v, w, u, v, w)
