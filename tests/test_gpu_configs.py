"""The benchmark configurations at their stated sizes on the GPU:

- 3D: 256^3, Dirichlet, the four-vortex-ring state (seed 0), CD and 2SHOC:
  FP64 bit for bit against the CPU oracle over 2 steps, FP32 and compensated
  FP32 within 1e-6 (relative L2) of it;
- batched: all 512 uneven 2D 2SHOC problems (sides 64-512, seed 1, harmonic
  traps) in one batch for 100 steps: FP64 bit for bit against the oracle on a
  sample that includes the largest and smallest problems and the strongest
  traps, compensated FP32 within 1e-6 of the FP64 oracle on the same sample.
The workload definitions are bench.py's."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import bench
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


@pytest.fixture(scope="module")
def rings():
    U, h = bench.vortex_rings_3d()
    return U, np.zeros(U.shape), h


@pytest.mark.parametrize("method", [1, 2])
def test_3d_256_four_rings(hip, oracle, rings, method):
    U, V, h = rings
    k = (0.8 if method == 1 else 0.8 * 0.75) * h * h / (3 * np.sqrt(2))
    _announce("3d 256^3 four rings", method)
    exp = oracle.take_steps(U, V, -1.0, 1.0, h * h, 1, 2, k, method)
    for prec in (2, 1, 3):
        sol = hip.Solver(U.shape, method, prec, 1, -1.0, 1.0, h * h, k)
        sol.set_state(U, V)
        sol.step(2)
        got = sol.state()
        sol.close()
        if prec == 2:
            assert bits_equal(got, exp)
        else:
            assert rel_l2(got, exp) < 1e-6, (prec, rel_l2(got, exp))


def test_batch_512_uneven_problems(hip, oracle):
    probs = bench.batched_problems()
    shapes = [(p["n"], p["n"]) for p in probs]
    states = [bench.batched_state(p) for p in probs]
    ns = np.array([p["n"] for p in probs])
    trap = np.array([p["w"] ** 2 * (p["n"] * bench.H / 2) ** 2 for p in probs])  # |V| at the edge midpoints
    sample = sorted(set([int(np.argmax(ns)), int(np.argmin(ns))] + [int(i) for i in np.argsort(trap)[-3:]] +
                        [int(i) for i in np.random.default_rng(0).choice(len(probs), 3, replace=False)]))
    steps = 100
    got = {}
    for prec in (2, 3):
        _announce("batch 512", prec)
        B = hip.Batch(shapes, [p["s"] for p in probs], [p["a"] for p in probs], [bench.H ** 2] * len(probs),
                      [p["k"] for p in probs], method=2, precision=prec, BC=1)
        for i, (U, V) in enumerate(states):
            B.set_state(i, U, V)
        B.step(steps)
        got[prec] = {i: B.state(i) for i in sample}
        B.close()
    for i in sample:
        p = probs[i]
        U, V = states[i]
        exp = oracle.take_steps(U, V, p["s"], p["a"], bench.H ** 2, 1, steps, p["k"], 2)
        assert bits_equal(got[2][i], exp), (i, p)
        assert rel_l2(got[3][i], exp) < 1e-6, (i, p, rel_l2(got[3][i], exp))
