"""The device energy (nlse_energy) against the restatement
oracle/energy_oracle.py, its conservation on the GPU (CD, Dirichlet), and
that diagnostics, energy, kfull and error norms on a resident solver, and
drop-in calls alternating between shapes, allocate no device memory after
their first call."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle.energy_oracle import energy

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import nlsemagic_amd as nm
    if nm.device_count() < 1:
        pytest.fail("no HIP device visible")
    return nm


def _announce(*what):
    print("CASE", *what, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [2, 1, 3])
@pytest.mark.parametrize("shape", [(300,), (41, 70), (12, 17, 66), (1024, 1024)])
def test_energy_matches_oracle(hip, shape, prec):
    _announce("energy", shape, prec)
    rng = np.random.default_rng(len(shape) * 10 + prec)
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    s, a, h2 = -1.0, 0.7, 0.04
    sol = hip.Solver(shape, 2, prec, 2, s, a, h2, 0.01 * h2)
    sol.set_state(U, V)
    W = sol.state()  # the state as held (single precision: rounded)
    Vd = V if prec == 2 else V.astype(np.float32).astype(np.float64)
    got = sol.energy()
    sol.close()
    exp = energy(W, Vd, s, a, h2)
    scale = abs(exp[1]) + abs(exp[2])
    for g, e in zip(got, exp):
        assert abs(g - e) <= 1e-12 * max(scale, abs(e)), (got, exp)


@pytest.mark.parametrize("shape,h", [((400,), 0.2), ((96, 80), 0.25), ((24, 26, 22), 0.3)])
def test_energy_conserved_on_the_gpu(hip, shape, h):
    """CD, Dirichlet, FP64: the device's energy drift falls >= 12-fold when k
    halves over the same time (the semi-discrete flow conserves it)."""
    _announce("energy conservation", shape)
    grids = np.meshgrid(*[h * (np.arange(n) - (n - 1) / 2) for n in shape], indexing="ij")
    r2 = sum(g * g for g in grids)
    U = np.exp(-r2 / 4) * np.exp(1j * 0.5 * grids[0])
    V = 0.3 * np.tanh(r2 / 8)
    s, a, h2 = 1.0, 0.6, h * h
    drift = []
    for kf, n in ((0.05, 400), (0.025, 800)):
        sol = hip.Solver(shape, 1, 2, 1, s, a, h2, kf * h2 / a)
        sol.set_state(U, V)
        e0 = sol.energy()[0]
        sol.step(n)
        drift.append(abs(sol.energy()[0] / e0 - 1))
        sol.close()
    assert drift[0] < 1e-6 and drift[0] > 12 * drift[1], drift


def test_resident_calls_allocate_nothing(hip):
    """A chunk loop with per-chunk diagnostics, energy, kfull and error norms
    allocates device memory only when the solver is created."""
    _announce("allocations: resident")
    n = 256
    h = 0.1
    x = h * (np.arange(n) - n / 2)
    sol = hip.Solver((n,), 2, 2, 2, 1.0, 1.0, h * h, 0.05 * h * h)
    sol.set_exact(1, h, (x[0],), 0.0, a=1.0, s=1.0, OM=-1.0, c=0.5, x0=0.0)
    before = hip.device_allocations()
    for chunk in range(5):
        sol.step(10)
        sol.diagnostics()
        sol.energy()
        sol.kfull()
        sol.exact_error(1, h, (x[0],), 0.0, a=1.0, s=1.0, OM=-1.0, c=0.5, x0=0.0)
    assert hip.device_allocations() == before
    sol.close()


def test_dropin_alternating_shapes_allocate_once(hip):
    """Drop-in calls alternating between three shapes keep their solvers."""
    _announce("allocations: drop-in")
    rng = np.random.default_rng(3)
    probs = []
    for shape in [(50, 40), (33, 70), (20, 21)]:
        U = (1 + 0.2 * rng.uniform(-1, 1, shape)) + 0.1j * rng.uniform(-1, 1, shape)
        probs.append((U, np.zeros(shape)))
    for U, V in probs:
        hip.take_steps(U, V, -1.0, 1.0, 0.01, 2, 5, 0.001, precision=2, method=2)
    before = hip.device_allocations()
    first = [hip.take_steps(U, V, -1.0, 1.0, 0.01, 2, 5, 0.001, precision=2, method=2) for U, V in probs]
    for _ in range(3):
        for (U, V), f in zip(probs, first):
            assert np.array_equal(hip.take_steps(U, V, -1.0, 1.0, 0.01, 2, 5, 0.001, precision=2, method=2), f)
    assert hip.device_allocations() == before
