"""Batched 2D integrator and the compensated single-precision state.

- FP64 batch: every problem of a batch of uneven grids (own s, a, h, k, V)
  is bit-identical to the CPU oracle run on that problem alone, for every
  fusion depth / tile height / persistent grid / graph setting.
- Compensated single precision (precision 3): relative L2 <= 1e-6 against
  the FP64 oracle after hundreds of steps, and closer to it than plain
  single precision; the batched form equals the single-problem form bit for
  bit, problem by problem.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import bits_equal, rel_l2

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


def _problem(rng, shape, s):
    """A bright Gaussian (s > 0) or a vortex (s < 0) in a harmonic-type potential."""
    N, M = shape
    h = 0.1
    x = h * (np.arange(N) - N / 2)
    y = h * (np.arange(M) - M / 2)
    X, Y = np.meshgrid(x, y, indexing="ij")
    r2 = X * X + Y * Y
    w = rng.uniform(0.05, 0.2)
    V = -w * w * r2
    if s > 0:
        U = np.exp(-r2 / 2) + 0j
    else:
        U = np.tanh(np.sqrt(r2)) * np.exp(1j * np.arctan2(Y, X))
    U = U + 1e-3 * (rng.uniform(-1, 1, U.shape) + 1j * rng.uniform(-1, 1, U.shape))
    return U, V, h


SHAPES = [(17, 65), (64, 64), (3, 3), (100, 33), (31, 129), (9, 70), (48, 50)]


def _batch_inputs(seed, method):
    rng = np.random.default_rng(seed)
    probs = []
    for i, sh in enumerate(SHAPES):
        s = 1.0 if i % 2 else -1.0
        a = rng.uniform(0.5, 1.0)
        U, V, h = _problem(rng, sh, s)
        kfac = 0.8 * (0.75 if method == 2 else 1.0) / (2 * np.sqrt(2))
        probs.append(dict(shape=sh, s=s, a=a, h2=h * h, k=kfac * h * h / a, U=U, V=V))
    return probs


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("kind,fusion,tile,grid,graph", [(0, 1, 8, 64, 0), (0, 2, 8, 1024, 0), (0, 2, 16, 256, 4),
                                                         (0, 4, 32, 8, 0), (0, 4, 16, 2048, 3), (1, 1, 8, 64, 0),
                                                         (1, 2, 8, 1024, 0), (1, 2, 12, 256, 4), (1, 4, 16, 8, 0),
                                                         (1, 4, 8, 2048, 3), (2, 1, 3, 64, 0), (2, 2, 16, 1024, 0),
                                                         (2, 4, 32, 256, 4), (2, 4, 128, 8, 3), (2, 4, 5, 2048, 0)])
def test_fp64_batch_is_bitwise_each_problem_alone(hip, oracle, method, kind, fusion, tile, grid, graph):
    _announce("batch64", method, kind, fusion, tile, grid, graph)
    probs = _batch_inputs(10 + method, method)
    B = hip.Batch([p["shape"] for p in probs], [p["s"] for p in probs], [p["a"] for p in probs],
                  [p["h2"] for p in probs], [p["k"] for p in probs], method=method, precision=2, BC=1)
    B.set_kernel(kind, tile)
    B.configure(fusion, tile, grid, graph)
    assert B.config() == (fusion, kind, tile, grid)
    for i, p in enumerate(probs):
        B.set_state(i, p["U"], p["V"])
    B.step(5)
    B.step(2)
    for i, p in enumerate(probs):
        exp = oracle.take_steps(p["U"], p["V"], p["s"], p["a"], p["h2"], 1, 7, p["k"], method)
        assert bits_equal(B.state(i), exp), (i, p["shape"])
    B.close()


@pytest.mark.parametrize("dim,shape", [(1, (2048,)), (1, (6000,)), (2, (96, 130)), (3, (20, 24, 30))])
@pytest.mark.parametrize("method", [1, 2])
def test_compensated_single_precision_accuracy(hip, oracle, dim, shape, method):
    _announce("comp", dim, shape, method)
    rng = np.random.default_rng(dim + 10 * method)
    h = 0.1
    grids = np.meshgrid(*[h * (np.arange(n) - n / 2) for n in shape], indexing="ij")
    r = np.sqrt(sum(g * g for g in grids))
    U = np.tanh(r) + 0j
    if dim >= 2:
        U = U * np.exp(1j * np.arctan2(grids[-1], grids[-2]))
    U = U + 1e-3 * (rng.uniform(-1, 1, U.shape) + 1j * rng.uniform(-1, 1, U.shape))
    V = np.zeros(shape)
    klin = (1.0 if method == 1 else 0.75) * h * h / (np.sqrt(2) * dim)
    k = 0.8 * klin
    steps = 300
    BC = 2 if dim < 3 else 1  # the 3D case runs with Dirichlet boundaries, as the 3D workload does
    exp = oracle.take_steps(U, V, -1.0, 1.0, h * h, BC, steps, k, method)
    assert np.all(np.isfinite(exp))
    got = {}
    for prec in (1, 3):
        sol = hip.Solver(shape, method, prec, BC, -1.0, 1.0, h * h, k)
        sol.set_state(U, V)
        sol.step(steps)
        got[prec] = sol.state()
        sol.close()
    e3, e1 = rel_l2(got[3], exp), rel_l2(got[1], exp)
    mass = lambda W: np.sum(np.abs(W) ** 2)
    drift = abs(mass(got[3]) / mass(exp) - 1)
    print(f"rel L2 compensated {e3:.2e} plain {e1:.2e}; mass drift vs FP64 {drift:.1e}", flush=True)
    assert e3 <= 1e-6
    assert e3 < e1
    assert drift < 1e-7


@pytest.mark.parametrize("kind,fusion,tile", [(0, 2, 8), (0, 4, 16), (0, 1, 32), (1, 2, 8), (1, 4, 12), (1, 1, 16),
                                              (2, 4, 64), (2, 2, 16), (2, 1, 5), (2, 4, 7)])
def test_compensated_batch_equals_single_problem(hip, kind, fusion, tile):
    _announce("batch32c", kind, fusion, tile)
    probs = _batch_inputs(3, 2)
    B = hip.Batch([p["shape"] for p in probs], [p["s"] for p in probs], [p["a"] for p in probs],
                  [p["h2"] for p in probs], [p["k"] for p in probs], method=2, precision=3, BC=1)
    B.set_kernel(kind, tile)
    B.configure(fusion, tile, 512, 0)
    for i, p in enumerate(probs):
        B.set_state(i, p["U"], p["V"])
    B.step(9)
    for i, p in enumerate(probs):
        sol = hip.Solver(p["shape"], 2, 3, 1, p["s"], p["a"], p["h2"], p["k"])
        sol.set_state(p["U"], p["V"])
        sol.step(9)
        assert bits_equal(B.state(i), sol.state()), (i, p["shape"])
        sol.close()
    B.close()


def test_batch_autotune_keeps_state(hip, oracle):
    probs = _batch_inputs(5, 2)
    B = hip.Batch([p["shape"] for p in probs], [p["s"] for p in probs], [p["a"] for p in probs],
                  [p["h2"] for p in probs], [p["k"] for p in probs], method=2, precision=2, BC=1)
    for i, p in enumerate(probs):
        B.set_state(i, p["U"], p["V"])
    f, t, g, ms = B.autotune(2)
    assert f in (1, 2, 4) and t in (8, 12, 16, 32, 64) and g % 8 == 0 and ms > 0
    assert B.config()[0] == f and B.config()[2] == t
    B.step(3)
    for i, p in enumerate(probs):
        exp = oracle.take_steps(p["U"], p["V"], p["s"], p["a"], p["h2"], 1, 3, p["k"], 2)
        assert bits_equal(B.state(i), exp)
    B.close()


def test_solver_autotune_keeps_state(hip, oracle):
    rng = np.random.default_rng(9)
    U = 1 + 0.1 * rng.uniform(-1, 1, (70, 90)) + 0j
    V = np.zeros((70, 90))
    sol = hip.Solver((70, 90), 2, 2, 2, -1.0, 1.0, 0.01, 0.002)
    sol.set_state(U, V)
    f, t, ms = sol.autotune(2)
    assert f in (1, 2, 4) and t in (8, 10, 12, 16, 32, 64)
    sol.step(4)
    assert bits_equal(sol.state(), oracle.take_steps(U, V, -1.0, 1.0, 0.01, 2, 4, 0.002, 2))
    sol.close()


@pytest.mark.parametrize("prec", [2, 3])
def test_device_diagnostics(hip, prec):
    """FP64 on-device mass / max Re / non-finite count match numpy on the state."""
    rng = np.random.default_rng(12)
    shape = (130, 70)
    U = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    sol = hip.Solver(shape, 2, prec, 1, -1.0, 1.0, 0.01, 0.001)
    sol.set_state(U, np.zeros(shape))
    sol.step(3)
    W = sol.state()
    mass, mxr, bad = sol.diagnostics()
    assert bad == 0
    assert abs(mass / np.sum(np.abs(W) ** 2) - 1) < 1e-13
    assert mxr == np.max(W.real) or abs(mxr - np.max(W.real)) < 1e-15
    sol.close()
    probs = _batch_inputs(4, 2)
    B = hip.Batch([p["shape"] for p in probs], [p["s"] for p in probs], [p["a"] for p in probs],
                  [p["h2"] for p in probs], [p["k"] for p in probs], method=2, precision=prec, BC=1)
    for i, p in enumerate(probs):
        B.set_state(i, p["U"], p["V"])
    B.step(2)
    D = B.diagnostics()
    for i in range(len(probs)):
        W = B.state(i)
        assert abs(D[i, 0] / np.sum(np.abs(W) ** 2) - 1) < 1e-13
        assert D[i, 2] == 0
    B.close()


def test_blowup_is_counted(hip):
    """2SHOC at the 2D workload's k = 0.3 h^2 exceeds RK4's stability limit:
    the state overflows and the non-finite count reports it."""
    n = 128
    h = 0.1
    x = h * (np.arange(n) - n / 2)
    X, Y = np.meshgrid(x, x, indexing="ij")
    rng = np.random.default_rng(0)
    U = np.tanh(np.hypot(X, Y)) * np.exp(1j * np.arctan2(Y, X)) + 1e-3 * rng.uniform(-1, 1, X.shape)
    sol = hip.Solver((n, n), 2, 2, 2, -1.0, 1.0, h * h, 0.3 * h * h)
    sol.set_state(U, np.zeros((n, n)))
    sol.step(400)
    mass, mxr, bad = sol.diagnostics()
    assert bad > 0
    sol.close()


@pytest.mark.parametrize("BC", [2, 3, 0])
@pytest.mark.parametrize("kind,fusion,tile", [(0, 2, 16), (1, 2, 8), (1, 4, 12), (2, 2, 16), (2, 4, 32), (2, 1, 3)])
def test_fp64_batch_other_boundaries_bitwise(hip, oracle, BC, kind, fusion, tile):
    """MSD, Lap0 and the zero boundary through every batch kernel family:
    each problem bit-identical to the CPU oracle run alone."""
    _announce("batch64 BC", BC, kind, fusion, tile)
    probs = _batch_inputs(20 + BC, 2)
    B = hip.Batch([p["shape"] for p in probs], [p["s"] for p in probs], [p["a"] for p in probs],
                  [p["h2"] for p in probs], [p["k"] for p in probs], method=2, precision=2, BC=BC)
    B.set_kernel(kind, tile)
    B.configure(fusion, tile, 512, 0)
    for i, p in enumerate(probs):
        B.set_state(i, p["U"], p["V"])
    B.step(4)
    for i, p in enumerate(probs):
        exp = oracle.take_steps(p["U"], p["V"], p["s"], p["a"], p["h2"], BC, 4, p["k"], 2)
        assert bits_equal(B.state(i), exp), (i, p["shape"])
    B.close()
