"""The C-ABI library builds, loads without a GPU, and exports exactly what
include/nlsemagic_hip.h declares.  No compute calls here."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nlsemagic_hip.h")
LIB = os.path.join(ROOT, "nlsemagic_amd", "libnlsemagic_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*\**\s+\**([A-Za-z_][A-Za-z0-9_]*)\s*\(", text,
                       flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


def test_header_declares_drop_in_and_resident_api():
    names = declared_functions()
    for dim in (1, 2, 3):
        assert f"NLSE{dim}D_TAKE_STEPS_HIP" in names
        assert f"NLSE{dim}D_TAKE_STEPS_HIP_F" in names
    for n in ("nlse_create", "nlse_destroy", "nlse_step", "nlse_set_state", "nlse_get_state", "nlse_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built (python -m nlsemagic_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(LIB)
    for n in declared_functions():
        assert hasattr(lib, n)


def test_header_compiles_as_c():
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", HEADER],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_argument_validation_needs_no_gpu():
    """Bad arguments are rejected before any device is touched."""
    import nlsemagic_amd as nm
    L = nm.lib()
    h = ctypes.c_void_p()
    rc = L.nlse_create(ctypes.byref(h), 4, 8, 8, 1, 1, 2, 1, -1.0, 1.0, 0.01, 0.001)
    assert rc == 1 and b"dim" in L.nlse_last_error()
    rc = L.nlse_create(ctypes.byref(h), 2, 8, 2, 1, 1, 2, 1, -1.0, 1.0, 0.01, 0.001)
    assert rc == 1 and b"at least 3" in L.nlse_last_error()
    rc = L.nlse_create(ctypes.byref(h), 1, 3, 1, 1, 2, 2, 4, -1.0, 1.0, 0.01, 0.001)
    assert rc == 1 and b"BC=4" in L.nlse_last_error()


def test_missing_library_fails_loudly(monkeypatch):
    import nlsemagic_amd as nm
    monkeypatch.setattr(nm, "_lib", None)
    monkeypatch.setattr(nm, "LIB_PATH", "/nonexistent/libnlsemagic_hip.so")
    with pytest.raises(ImportError):
        nm.lib()


def test_vorticity_argument_checks_need_no_gpu():
    import nlsemagic_amd as nm
    import numpy as np
    L = nm.lib()
    a = np.zeros(2 * 3 * 3)
    p = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = L.NLSE3D_VORTICITY_HIP(p, p, 0.5, 1e-3, 2, 3, 3, p, p, p, p)
    assert rc == 1 and b">= 3 points" in L.nlse_last_error()
    rc = L.NLSE3D_VORTICITY_HIP(p, p, 0.5, 1e-3, 3, 3, 3, None, p, p, p)
    assert rc == 1 and b"null" in L.nlse_last_error()
    rc = L.nlse_vorticity(None, 0.5, 1e-3, p, p, p, p)
    assert rc == 1 and b"null solver" in L.nlse_last_error()
    with pytest.raises(ValueError):
        nm.rmc_vorticity3D(np.zeros((4, 4)), 0.5, 1e-3)
