"""Host <-> device copies of states: round trips through the pageable-copy
path (plain copies below 1 MB per plane, pinned 4 MB chunks from 1 MB up)
at sizes around its thresholds and chunk boundaries, alone and from
several threads at once.  A state uploaded and read back with no step taken
must come back bit for bit."""
from __future__ import annotations

import sys
import threading

import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import nlsemagic_amd as nm
    if nm.device_count() < 1:
        pytest.fail("no HIP device visible")
    return nm


def _state(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# points per plane: FP64 planes of 1 MB - 8 B, 1 MB, 4 MB + 8 B (a one-element
# second chunk), 12 MB + 24 B (the three chunk slots reused), and a 3D state
# of 16 MB per plane
@pytest.mark.parametrize("shape", [(131071,), (131072,), (524289,), (1572867,), (128, 128, 128)])
def test_round_trip_double(hip, shape):
    print("CASE round trip FP64", shape, flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(len(shape) * 1000 + shape[0])
    U = _state(rng, shape)
    V = rng.standard_normal(shape)
    sol = hip.Solver(shape, 1, 2, 1, -1.0, 1.0, 0.01, 0.001)
    try:
        for _ in range(2):
            sol.set_state(U, V)
            assert bits_equal(sol.state(), U)
            U = U[..., ::-1].copy() if len(shape) == 1 else np.ascontiguousarray(U.transpose(1, 2, 0))
    finally:
        sol.close()


@pytest.mark.parametrize("n", [262143, 262145, 1048577])
def test_round_trip_single(hip, n):
    """FP32 planes of 1 MB - 4 B, 1 MB + 4 B and 4 MB + 4 B: the state comes
    back as the float rounding of what went in."""
    print("CASE round trip FP32", n, flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(n)
    U = _state(rng, (n,))
    sol = hip.Solver((n,), 2, 1, 1, -1.0, 1.0, 0.01, 0.001)
    try:
        sol.set_state(U, np.zeros(n))
        exp = U.astype(np.complex64).astype(np.complex128)
        assert bits_equal(sol.state(), exp)
    finally:
        sol.close()


def test_drop_in_zero_steps_returns_input(hip):
    """The drop-in entry point with chunk_size 0 uploads and downloads only."""
    print("CASE drop-in zero steps", flush=True)
    rng = np.random.default_rng(7)
    U = _state(rng, (600, 700))
    V = rng.standard_normal((600, 700))
    got = hip.take_steps(U, V, -1.0, 1.0, 0.01, 1, 0, 0.001, precision=2, method=2)
    assert bits_equal(got, U)


def test_concurrent_round_trips(hip):
    """Three threads, each with its own solver, copying at once: every copy
    goes through the device's one set of pinned chunks in turn."""
    print("CASE concurrent round trips", flush=True)
    shapes = [(700, 900), (1024, 1024), (300, 5000)]
    errors = []

    def work(i, shape):
        try:
            rng = np.random.default_rng(100 + i)
            sol = hip.Solver(shape, 1, 2, 1, -1.0, 1.0, 0.01, 0.001)
            try:
                for _ in range(5):
                    U = _state(rng, shape)
                    sol.set_state(U, rng.standard_normal(shape))
                    if not bits_equal(sol.state(), U):
                        errors.append(f"thread {i} {shape}: state differs")
            finally:
                sol.close()
        except Exception as e:  # reported below, in the test's own thread
            errors.append(f"thread {i} {shape}: {e!r}")

    threads = [threading.Thread(target=work, args=(i, s)) for i, s in enumerate(shapes)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a copy thread did not finish"
    assert not errors, errors
