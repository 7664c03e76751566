"""Pin the CPU oracle to the reference itself.

- Bitwise equality with the golden vectors written by the reference's own
  NLSE{1,2}D_TAKE_STEPS[_F] (tests/golden/make_golden.py).
- Where oracle/_ref is built (this container, and any box it travels to),
  bitwise equality with the compiled reference on fresh seeded inputs,
  including non-finite blow-ups and tiny grids.
"""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

from conftest import bits_equal, golden_cases


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0]["name"])
def test_oracle_matches_golden(oracle, case):
    c, U, V, expected = case
    assert hashlib.sha256(np.ascontiguousarray(expected).tobytes()).hexdigest() == c["sha256_out"]
    got = oracle.take_steps(U, V, c["s"], c["a"], c["h2"], c["BC"], c["chunk_size"], c["k"], c["method"],
                            precision=c["precision"])
    assert bits_equal(got, expected), c["name"]


def _ref_cases():
    rng = np.random.default_rng(7)
    out = []
    for dim, shapes, bcs in ((1, [(4,), (5,), (33,), (257,)], [1, 2, 3, 4, 0, 9]),
                             (2, [(3, 3), (4, 7), (19, 11), (65, 3)], [1, 2, 3, 0, 4])):
        for shape in shapes:
            for method in (1, 2):
                for prec in (2, 1):
                    for BC in bcs:
                        out.append((dim, shape, method, prec, BC, int(rng.integers(1 << 30))))
    return out


@pytest.mark.parametrize("dim,shape,method,prec,BC,seed", _ref_cases())
def test_oracle_matches_compiled_reference(oracle, dim, shape, method, prec, BC, seed):
    if oracle.ref(dim) is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(seed)
    U = (0.8 + 0.4 * rng.uniform(-1, 1, shape)) + 1j * (0.4 * rng.uniform(-1, 1, shape))
    V = rng.uniform(-1, 1, shape)
    args = (U, V, -1.0 if BC % 2 else 0.6, 0.7, 0.02, BC, 6, 0.003, method)
    a = oracle.take_steps(*args, precision=prec)
    b = oracle.take_steps(*args, precision=prec, use_ref=True)
    assert bits_equal(a, b)


def test_oracle_matches_reference_on_blowup(oracle):
    """An unstable time step: both must overflow to the same inf/NaN pattern."""
    if oracle.ref(2) is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(3)
    U = 1 + 0.01 * rng.uniform(-1, 1, (12, 10)) + 0j
    V = np.zeros((12, 10))
    args = (U, V, -1.0, 1.0, 0.01, 2, 400, 0.01, 2)
    a = oracle.take_steps(*args)
    b = oracle.take_steps(*args, use_ref=True)
    assert not np.all(np.isfinite(a))
    assert bits_equal(a, b)
