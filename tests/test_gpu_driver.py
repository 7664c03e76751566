"""nlse_kfull (the drivers' full RK4 stability bound, NLSE2D.m:427-477 and
its 1D/3D twins) on the GPU against the CPU restatement
oracle/driver_oracle.py, for every dimension, method, boundary and
precision; and the state is untouched by it."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import bits_equal
from oracle import driver_oracle as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import nlsemagic_amd as nm
    if nm.device_count() < 1:
        pytest.fail("no HIP device visible")
    return nm


def _state(dim, seed):
    rng = np.random.default_rng(seed)
    shape = {1: (300,), 2: (40, 70), 3: (12, 17, 66)}[dim]
    U = (0.9 + 0.3 * rng.uniform(-1, 1, shape)) + 1j * (0.3 * rng.uniform(-1, 1, shape))
    V = 0.5 * rng.uniform(-1, 1, shape)
    return U, V


@pytest.mark.parametrize("prec", [2, 1, 3])
@pytest.mark.parametrize("BC", [1, 2, 3])
@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_kfull_matches_oracle(hip, oracle, dim, method, BC, prec):
    print("CASE kfull", dim, method, BC, prec, flush=True)
    sys.stdout.flush()
    U, V = _state(dim, 100 * dim + 10 * method + BC)
    h, a, s = 0.2, 0.9, -1.0 if BC != 3 else 0.8
    sol = hip.Solver(U.shape, method, prec, BC, s, a, h * h, 0.1 * h * h)
    sol.set_state(U, V)
    got = sol.kfull()
    # the oracle sees the state as the solver holds it
    Us = sol.state()
    Vs = V if prec == 2 else V.astype(np.float32).astype(np.float64)
    sol.close()
    exp = D.kfull(Us, Vs, h, a, s, method, BC)
    # single precision: the MSD layer uses the RHS evaluated in single precision
    tol = 1e-12 if prec == 2 else 1e-5
    np.testing.assert_allclose(got, exp, rtol=tol)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_kfull_leaves_the_state_alone(hip, oracle, dim):
    """MSD evaluates the RHS into the stage buffers; the steps that follow
    equal the oracle's bit for bit."""
    U, V = _state(dim, 7 + dim)
    h2, k = 0.04, 0.004
    sol = hip.Solver(U.shape, 2, 2, 2, -1.0, 0.9, h2, k)
    sol.set_state(U, V)
    sol.step(1)
    kf = sol.kfull()
    assert np.all(np.isfinite(kf)) and kf[0] > 0
    sol.step(2)
    got = sol.state()
    sol.close()
    assert bits_equal(got, oracle.take_steps(U, V, -1.0, 0.9, h2, 2, 3, k, 2))


def test_step_with_the_kfull_step(hip, oracle):
    """The drivers' flow when they choose k: kfull of the initial state, then
    the steps with that k, bit-identical to the oracle run with the same k."""
    U, V = _state(2, 31)
    sol = hip.Solver(U.shape, 2, 2, 2, -1.0, 0.9, 0.04, 1e-3)
    sol.set_state(U, V)
    kf, kl, k = sol.kfull()
    assert k == pytest.approx(0.9 * min(kf, kl), rel=1e-15)
    sol.set_step(k)
    sol.step(5)
    got = sol.state()
    with pytest.raises(hip.HipError, match="positive"):
        sol.set_step(0.0)
    sol.close()
    assert bits_equal(got, oracle.take_steps(U, V, -1.0, 0.9, 0.04, 2, 5, k, 2))
