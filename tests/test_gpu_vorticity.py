"""3D vorticity on the GPU (rmc_vorticity3D.c:24-354) against the oracle.

Both evaluate in double precision in the reference's operand order and set
the gradient the reference leaves unset on edges and corners to 0, so the
results agree bit for bit at every point; the mathematical pins of the
oracle (and where the reference itself is defined) are in
test_oracle_math.py.  Single-precision solver states are widened to double
(hi + lo for the compensated state) before the same arithmetic.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu


def _state(shape, seed, real=False):
    rng = np.random.default_rng(seed)
    U = 1 + 0.3 * rng.standard_normal(shape)
    if not real:
        U = U + 0.3j * rng.standard_normal(shape)
    return U


def _same(got, exp):
    assert all(bits_equal(g, e) for g, e in zip(got, exp))


def _edges(shape):
    L, N, M = shape
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    nb = ((i == 0) | (i == L - 1)).astype(int) + ((j == 0) | (j == N - 1)) + ((k == 0) | (k == M - 1))
    return nb >= 2


@pytest.mark.parametrize("shape", [(3, 3, 3), (3, 4, 5), (5, 7, 9), (12, 17, 33), (9, 64, 65), (33, 20, 130),
                                   # tile edges: 16 rows, 60 columns, chunks of 16 planes
                                   (16, 16, 60), (17, 17, 61), (15, 33, 59), (32, 15, 121), (48, 31, 120)])
@pytest.mark.parametrize("veps", [1e-3, 0.0])
def test_drop_in_matches_oracle(oracle, shape, veps):
    import nlsemagic_amd as nm
    print("CASE drop-in", shape, veps, flush=True)
    U = _state(shape, sum(shape))
    lh2 = 1 / (2 * 0.17)
    got = nm.rmc_vorticity3D(U, lh2, veps)
    _same(got, oracle.vorticity3d(U, lh2, veps))
    for w in got:
        assert np.all(w[_edges(shape)] == 0)


def test_drop_in_real_input_and_zero_field(oracle):
    import nlsemagic_amd as nm
    U = _state((6, 7, 8), 3, real=True)
    _same(nm.rmc_vorticity3D(U, 2.0, 1e-3), oracle.vorticity3d(U + 0j, 2.0, 1e-3))
    # a real field has no phase gradient: zero velocity, zero vorticity
    assert all(np.all(w == 0) for w in nm.rmc_vorticity3D(U, 2.0, 1e-3))
    # U = 0 with veps = 0: 0 * (1/0) is NaN at every point, as in the reference
    Z = np.zeros((4, 4, 4), complex)
    _same(nm.rmc_vorticity3D(Z, 2.0, 0.0), oracle.vorticity3d(Z, 2.0, 0.0))


def test_drop_in_vortex_line_full_size(oracle):
    """A vortex line on a 160^3 grid: bitwise against the oracle, and the
    circulation of every slice is pi."""
    import nlsemagic_amd as nm
    n = 160
    h = 16.0 / n
    g = (np.arange(n) - (n - 1) / 2) * h
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    U = np.tanh(np.hypot(y - 0.3, z) / 0.6) * np.exp(1j * np.arctan2(z, y - 0.3)) * np.exp(0.2j * x)
    print("CASE vortex line", n, flush=True)
    got = nm.rmc_vorticity3D(U, 1 / (2 * h), 1e-3)
    _same(got, oracle.vorticity3d(U, 1 / (2 * h), 1e-3))
    circ = got[0][:, 2:-2, 2:-2].sum(axis=(1, 2)) * h * h
    np.testing.assert_allclose(circ, np.pi, rtol=2e-3)


@pytest.mark.parametrize("precision", [2, 1, 3])
@pytest.mark.parametrize("method", [1, 2])
def test_resident_solver_vorticity(oracle, precision, method):
    """After a few RK4 steps the resident state's vorticity equals the
    oracle's vorticity of the downloaded state (widened to double)."""
    import nlsemagic_amd as nm
    shape = (10, 13, 70)
    print("CASE resident", precision, method, flush=True)
    U = _state(shape, 11)
    S = nm.Solver(shape, method, precision, 1, -1.0, 1.0, 0.04, 0.004)
    S.set_state(U, 0.1 * np.ones(shape))
    S.step(3)
    state = S.state()
    lh2 = 1 / (2 * 0.2)
    got = S.vorticity(lh2, 1e-3)
    _same(got, oracle.vorticity3d(state, lh2, 1e-3))
    # the state is untouched and stepping goes on from it
    assert bits_equal(S.state(), state)
    S.step(2)
    ref = nm.Solver(shape, method, precision, 1, -1.0, 1.0, 0.04, 0.004)
    ref.set_state(U, 0.1 * np.ones(shape))
    ref.step(5)
    assert bits_equal(S.state(), ref.state())
    S.close()
    ref.close()


def test_resident_device_outputs_match_host_outputs():
    import nlsemagic_amd as nm
    shape = (8, 16, 64)
    n = int(np.prod(shape))
    S = nm.Solver(shape, 2, 2, 1, -1.0, 1.0, 0.04, 0.004)
    S.set_state(_state(shape, 4))
    host = S.vorticity(1.5, 1e-3)
    dev = [nm.DeviceArray(n) for _ in range(4)]
    S.vorticity_device(1.5, 1e-3, *[d.address for d in dev])
    S.sync()
    for d, hst in zip(dev, host):
        assert bits_equal(d.numpy().reshape(shape), hst)
    # outputs may be skipped
    w = nm.DeviceArray(n)
    S.vorticity_device(1.5, 1e-3, 0, 0, 0, w.address)
    S.sync()
    assert bits_equal(w.numpy().reshape(shape), host[3])
    for d in dev + [w]:
        d.free()
    S.close()


def test_vorticity_timed_gives_the_same_outputs():
    import nlsemagic_amd as nm
    shape = (6, 20, 70)
    n = int(np.prod(shape))
    S = nm.Solver(shape, 1, 3, 1, -1.0, 1.0, 0.04, 0.004)
    S.set_state(_state(shape, 9))
    host = S.vorticity(2.0, 1e-3)
    dev = [nm.DeviceArray(n) for _ in range(4)]
    ms = S.vorticity_timed(2.0, 1e-3, 3, *[d.address for d in dev])
    assert ms > 0
    for d, hst in zip(dev, host):
        assert bits_equal(d.numpy().reshape(shape), hst)
        d.free()
    with pytest.raises(nm.HipError, match="reps"):
        S.vorticity_timed(2.0, 1e-3, 0)
    S.close()


def test_vorticity_rejects_non_3d_solvers():
    import nlsemagic_amd as nm
    S = nm.Solver((16, 16), 1, 2, 1, -1.0, 1.0, 0.04, 0.004)
    with pytest.raises(nm.HipError, match="3D"):
        S.vorticity(1.0, 1e-3)
    S.close()
