"""nlse_diagnostics_noalloc (the allocation-free variant of nlse_diagnostics
for callers that cannot allocate) against nlse_diagnostics on the same state:
the same bits, on two calls in a row (the second on a smaller state), for
every precision, with no device allocation by the library during the calls.
Device buffers for the caller's state are made with the HIP runtime the
library itself uses."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import nlsemagic_amd as nm
    if nm.device_count() < 1:
        pytest.fail("no HIP device visible")
    return nm


@pytest.fixture(scope="module")
def rt(hip):
    r = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")  # already loaded by the library: the same runtime
    r.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    r.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    r.hipFree.argtypes = [C.c_void_p]
    r.hipDeviceSynchronize.argtypes = []
    return r


def _to_device(rt, a):
    a = np.ascontiguousarray(a)
    p = C.c_void_p()
    assert rt.hipMalloc(C.byref(p), a.nbytes) == 0
    assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    return p


def _state(n, seed):
    rng = np.random.default_rng(seed)
    U = rng.normal(size=n) + 1j * rng.normal(size=n)
    U[n // 3] = np.nan  # one non-finite point
    U[n // 2] = 1e300 + 0j if n > 10 else U[n // 2]
    return U


@pytest.mark.parametrize("prec", [2, 1, 3])
def test_noalloc_diagnostics_equal_the_solver_diagnostics(hip, rt, prec):
    print("CASE diagnostics_noalloc", prec, flush=True)
    sys.stdout.flush()
    L = hip.lib()
    out = C.c_void_p()
    assert rt.hipMalloc(C.byref(out), 24) == 0
    for call, n in enumerate((300_001, 4_099)):
        U = _state(n, 10 * prec + call)
        if prec == 2:
            hi, lo = U, None
        else:
            with np.errstate(over="ignore", invalid="ignore"):  # the 1e300 point becomes inf
                hi = U.astype(np.complex64)
                lo = (U - hi.astype(np.complex128)).astype(np.complex64) if prec == 3 else None
        dU = _to_device(rt, hi)
        dL = _to_device(rt, lo) if lo is not None else None
        before = hip.device_allocations()
        rc = L.nlse_diagnostics_noalloc(dU, dL, n, prec, out, None)
        assert rc == 0, hip.lib().nlse_last_error()
        assert rt.hipDeviceSynchronize() == 0
        assert hip.device_allocations() == before
        got = np.empty(3)
        assert rt.hipMemcpy(got.ctypes.data, out, 24, 2) == 0  # device to host
        sol = hip.Solver((n,), 1, prec, 1, -1.0, 1.0, 0.01, 0.001)
        sol.set_state(U if prec != 1 else hi.astype(np.complex128))
        exp = np.array(sol.diagnostics(), dtype=np.float64)
        sol.close()
        assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), (got, exp)
        rt.hipFree(dU)
        if dL is not None:
            rt.hipFree(dL)
    rt.hipFree(out)


def test_noalloc_rejects_bad_arguments(hip):
    L = hip.lib()
    assert L.nlse_diagnostics_noalloc(None, None, 10, 2, None, None) == 1
    assert L.nlse_diagnostics_noalloc(C.c_void_p(16), None, 0, 2, C.c_void_p(16), None) == 1
    assert L.nlse_diagnostics_noalloc(C.c_void_p(16), None, 1 << 31, 2, C.c_void_p(16), None) == 1
    assert L.nlse_diagnostics_noalloc(C.c_void_p(16), None, 10, 4, C.c_void_p(16), None) == 1
    assert L.nlse_diagnostics_noalloc(C.c_void_p(16), C.c_void_p(16), 10, 2, C.c_void_p(16), None) == 1
