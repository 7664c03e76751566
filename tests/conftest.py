"""Shared test setup: the `gpu` marker and helpers for loading fixtures."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_cases():
    """[(case dict, U, V, expected output)] from tests/golden/nlse_golden.*"""
    with open(os.path.join(GOLDEN_DIR, "nlse_golden.json")) as f:
        man = json.load(f)
    data = np.load(os.path.join(GOLDEN_DIR, "nlse_golden.npz"))
    out = []
    for c in man["cases"]:
        n = c["name"]
        out.append((c, data[f"{n}__U"], data[f"{n}__V"], data[f"{n}__out"]))
    return out


def bits_equal(a, b):
    """Bitwise equality of two complex128 arrays, component by component.

    NaN compares equal to NaN whatever its sign and payload (x86 and the GPU
    produce different default NaNs); every other value, infinities and signed
    zeros included, must match bit for bit.
    """
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    if a.shape != b.shape:
        return False
    fa, fb = a.view(np.float64), b.view(np.float64)
    na, nb = np.isnan(fa), np.isnan(fb)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(fa.view(np.uint64)[~na], fb.view(np.uint64)[~nb])


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
