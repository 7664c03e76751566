"""Closed-form fields, exact-solution error norms and vortex initial
conditions on the GPU against the numpy restatement (oracle/ic_oracle.py).

The device evaluates the same formulas in double precision; the remaining
differences are the last-bit rounding of the transcendental functions (GPU
math library vs glibc) and of the linear-interpolation formula, so double
results are compared to an absolute tolerance of 1e-12 on fields of order 1
(1e-11 for the ring, which interpolates twice).  Single-precision states are
compared with the oracle's double result to float rounding; the compensated
state (hi + lo) to 1e-12.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import ic_oracle as I

pytestmark = pytest.mark.gpu


def _solver(shape, precision=2, method=2, BC=1):
    import nlsemagic_amd as nm
    return nm.Solver(shape, method, precision, BC, -1.0, 1.0, 0.01, 0.001)


def _tol(precision, atol=1e-12):
    return {2: atol, 3: max(atol, 1e-12), 1: 2e-7}[precision]


EXACT_CASES = [
    (1, (1.0, 1.0, 0.5, 0.75, 0.0), (301,), (-15.0,), 0.1),
    (2, (1.0, -1.0, -1.0, 0.75, 0.5), (257,), (-12.8,), 0.1),
    (4, (1.0, -1.0, -1.0, 0.75, 0.0), (200,), (-10.0,), 0.1),
    (3, (1.0, 0, 0, 0, 0), (150,), (-7.45,), 0.1),
    (3, (0.8, 0, 0, 0, 0), (40, 33), (-4.0, -3.2), 0.2),
    (3, (1.0, 0, 0, 0, 0), (20, 24, 18), (-2.0, -2.3, -1.7), 0.2),
]


@pytest.mark.parametrize("precision", [2, 1, 3])
@pytest.mark.parametrize("case", range(len(EXACT_CASES)))
@pytest.mark.parametrize("t", [0.0, 0.7])
def test_set_exact_matches_oracle(case, precision, t):
    kind, par, shape, origin, h = EXACT_CASES[case]
    print("CASE set_exact", kind, shape, precision, t, flush=True)
    S = _solver(shape, precision)
    S.set_exact(kind, h, origin, t, *par)
    exp = I.exact(kind, par, shape, origin, h, t)
    got = S.state()
    assert np.max(np.abs(got - exp)) <= _tol(precision) * max(1.0, np.abs(exp).max())
    S.close()


def test_set_exact_linear_sets_the_potential(oracle):
    """kind 3 also writes V = |x|^2/a: one CD step from the linear ground
    state on the GPU equals the oracle's step from the same state and V."""
    shape, origin, h, par = (30, 28), (-2.9, -2.7), 0.2, (1.0, 0, 0, 0, 0)
    S = _solver(shape, 2, method=1)
    S.set_exact(3, h, origin, 0.0, *par)
    U0 = S.state()
    S.step(3)
    V = I.exact_potential(3, par, shape, origin, h)
    exp = oracle.take_steps(U0, V, -1.0, 1.0, 0.01, 1, 3, 0.001, 1)
    assert np.max(np.abs(S.state() - exp)) == 0
    S.close()


@pytest.mark.parametrize("precision", [2, 3, 1])
@pytest.mark.parametrize("case", [0, 1, 2, 4, 5])
def test_exact_error_matches_oracle(case, precision):
    kind, par, shape, origin, h = EXACT_CASES[case]
    print("CASE exact_error", kind, shape, precision, flush=True)
    a, s = par[0], par[1]
    import nlsemagic_amd as nm
    BC = 2 if kind in (2, 4) else 1
    S = nm.Solver(shape, 2, precision, BC, s, a, h * h, 0.1 * h * h)
    S.set_exact(kind, h, origin, 0.0, *par)
    S.step(20)
    t = 20 * 0.1 * h * h
    got = S.exact_error(kind, h, origin, t, *par)
    exp = I.exact_error(S.state(), kind, par, origin, h, t)
    np.testing.assert_allclose(got, exp, rtol=1e-9, atol=1e-15)
    assert max(got) > 0
    S.close()


VORTEX_2D = [  # charge, om, x0, y0, vx, vy, phase, a, s, amp, opt
    (1, -1.0, 0.3, -0.2, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0, 1),
    (1, -1.0, 0.3, -0.2, 0.4, -0.1, 0.5, 1.0, -1.0, 1.2, 0),   # direct tanh form
    (2, -1.0, -0.7, 0.4, 0.1, 0.2, 1.0, 0.8, -1.0, 1.0, 1),
    (3, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0, 0),     # tanh ansatz interpolated
    (1, 1 / 3, 0.2, 0.1, 0.3, 0.0, 0.2, 1.0, 1.0, 1.0, 1),     # bright: added
]


@pytest.mark.parametrize("precision", [2, 3, 1])
@pytest.mark.parametrize("case", range(len(VORTEX_2D)))
def test_vortex_2d_matches_oracle(case, precision):
    shape, h, origin = (52, 44), 0.25, (-6.5, -5.4)
    args = VORTEX_2D[case]
    print("CASE vortex2d", args, precision, flush=True)
    rng = np.random.default_rng(case)
    U = 1 + 0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    S = _solver(shape, precision)
    S.set_state(U)
    U_dev = S.state()
    charge, om, x0, y0, vx, vy, phase, a, s, amp, opt = args
    S.add_vortex_2d(charge, om, x0, y0, vx, vy, phase, h, a, s, amp, origin, opt)
    exp = I.add_vortex_2d(charge, om, x0, y0, vx, vy, phase, U_dev, h, a, s, amp, origin, opt)
    assert np.max(np.abs(S.state() - exp)) <= _tol(precision)
    S.close()


@pytest.mark.parametrize("precision", [2, 3])
@pytest.mark.parametrize("args", [(1, -1.0, 0.1, 0.2, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, -1.0),
                                  (2, -1.0, -0.4, 0.3, 0.0, 0.2, -0.1, 0.3, 0.4, 1.0, -1.0),
                                  (1, 1 / 3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0)])
def test_vortex_line_matches_oracle(args, precision):
    shape, h, origin = (10, 36, 40), 0.3, (-5.4, -6.0, -1.5)
    print("CASE line", args, precision, flush=True)
    m, OM, x0, y0, z0, vx, vy, vz, phase, a, s = args
    rng = np.random.default_rng(3)
    U = 1 + 0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    S = _solver(shape, precision)
    S.set_state(U)
    U_dev = S.state()
    S.add_vortex_line(m, OM, x0, y0, z0, vx, vy, vz, phase, h, a, s, origin)
    exp = I.add_vortex_line(m, OM, x0, y0, z0, vx, vy, vz, phase, U_dev, h, a, s, origin)
    assert np.max(np.abs(S.state() - exp)) <= _tol(precision)
    S.close()


RINGS = [  # m, OM, x0, y0, z0, thetax, thetay, thetaz, vx, vy, vz, phase, s, d, ea, eb
    (1, -1.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0, 0, 0.0, -1.0, 3.0, 1.0, 1.0),
    (1, -1.0, 0.4, -0.3, 0.2, 0, 0, 0, 0.0, 0.0, 0.2, 0.3, -1.0, 2.5, 1.2, 1.0),
    (2, -1.0, 0.0, 0.0, 0.0, math.pi / 2, 0, 0, 0, 0, 0, 0.0, -1.0, 3.0, 1.0, 1.0),
    (1, -1.0, 0.0, 0.0, 0.0, math.pi / 2, math.pi / 2, math.pi / 2, 0, 0, 0, 0.0, -1.0, 3.0, 1.0, 1.0),
    (1, -1.0, 0.2, 0.1, -0.3, 0.3, -0.2, 0.5, 0.1, 0.0, 0.0, 0.0, -1.0, 3.0, 1.0, 1.0),
    (1, 1 / 3, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0, 0, 0.0, 1.0, 3.0, 1.0, 1.0),
]


@pytest.mark.parametrize("precision", [2, 3])
@pytest.mark.parametrize("case", range(len(RINGS)))
def test_vortex_ring_matches_oracle(case, precision):
    n, h = 32, 0.35
    origin = (-(n - 1) * h / 2,) * 3
    shape = (n, n, n)
    m, OM, x0, y0, z0, tx, ty, tz, vx, vy, vz, phase, s, d, ea, eb = RINGS[case]
    print("CASE ring", RINGS[case], precision, flush=True)
    rng = np.random.default_rng(7)
    U = 1 + 0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    S = _solver(shape, precision)
    S.set_state(U)
    U_dev = S.state()
    S.add_vortex_ring(m, OM, x0, y0, z0, tx, ty, tz, vx, vy, vz, phase, h, 1.0, s, origin, d, 0, ea, eb)
    exp = I.add_vortex_ring(m, OM, x0, y0, z0, tx, ty, tz, vx, vy, vz, phase, U_dev, h, 1.0, s, origin, d, ea, eb)
    assert np.max(np.abs(S.state() - exp)) <= _tol(precision, 1e-11)
    S.close()


def test_ring_on_device_then_vorticity_circulation():
    """Generator and post-processor together on a 96^3 grid: a dark ring of
    charge 1 made on the GPU, its vorticity from the GPU, circulation -+pi
    through the half-planes of the x-z plane through the axis."""
    n, h = 96, 0.25
    origin = (-(n - 1) * h / 2,) * 3
    S = _solver((n, n, n), 2)
    S.set_state(np.ones((n, n, n), complex))
    S.add_vortex_ring(1, -1.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0, 0, 0.0, h, 1.0, -1.0, origin, 6.0)
    wx, wy, wz, W = S.vorticity(1 / (2 * h), 1e-3)
    k = n // 2
    X = origin[0] + np.arange(n) * h
    idx = np.arange(n)
    sl = wz[2:-2, :, k]
    plus = sl[:, (X > 0) & (idx < n - 2)].sum() * h * h
    minus = sl[:, (X < 0) & (idx > 1)].sum() * h * h
    np.testing.assert_allclose([plus, minus], [-np.pi, np.pi], rtol=5e-3)
    S.close()


def test_matlab_named_generators(oracle):
    import nlsemagic_amd as nm
    h = 0.25
    x = -5.0 + np.arange(40) * h
    y = -4.0 + np.arange(36) * h
    X, Y = np.meshgrid(x, y, indexing="ij")  # this package's (N, M) layout
    U = np.ones(X.shape, complex)
    got = nm.NLS_add_vortex(1, -1.0, 0.1, 0.0, 0.0, 0.0, 0.0, U, h, 1.0, -1.0, 1.0, X, Y, 1)
    exp = I.add_vortex_2d(1, -1.0, 0.1, 0.0, 0.0, 0.0, 0.0, U, h, 1.0, -1.0, 1.0, (x[0], y[0]), 1)
    assert np.max(np.abs(got - exp)) < 1e-12
    z = -1.0 + np.arange(8) * h
    Z3, X3, Y3 = np.meshgrid(z, x, y, indexing="ij")  # (L, N, M)
    U3 = np.ones(X3.shape, complex)
    got = nm.NLSE_add_vortex_line(1, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, U3, h, 1.0, -1.0, 1.0, X3, Y3, Z3)
    exp = I.add_vortex_line(1, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, U3, h, 1.0, -1.0, (x[0], y[0], z[0]))
    assert np.max(np.abs(got - exp)) < 1e-12


def test_generators_reject_bad_arguments():
    import nlsemagic_amd as nm
    S3 = _solver((8, 8, 8))
    with pytest.raises(nm.HipError, match="opt_2D"):
        S3.add_vortex_ring(1, -1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0.3, 1.0, -1.0, (-1, -1, -1), 1.0, 1)
    with pytest.raises(nm.HipError, match="2D state"):
        S3.add_vortex_2d(1, -1.0, 0, 0, 0, 0, 0, 0.3, 1.0, -1.0, 1.0, (-1, -1), 1)
    with pytest.raises(nm.HipError, match="1D"):
        S3.set_exact(1, 0.3, (-1, -1, -1), 0.0, 1.0, 1.0, 0.5, 0.0, 0.0)
    S2 = _solver((8, 9))
    with pytest.raises(nm.HipError, match="3D state"):
        S2.add_vortex_line(1, -1.0, 0, 0, 0, 0, 0, 0, 0, 0.3, 1.0, -1.0, (-1, -1, -1))
    S2.close()
    S3.close()


@pytest.mark.parametrize("precision", [2, 1, 3])
def test_fill_state(precision):
    S = _solver((5, 6, 7), precision)
    S.fill(0.3 - 0.7j)
    got = S.state()

    def split(x):  # the value a state of this precision holds (hi + lo for the compensated one)
        hi = np.float32(x)
        return float(hi) + (float(np.float32(x - float(hi))) if precision == 3 else 0.0)
    exp = complex(split(0.3), split(-0.7)) if precision != 2 else 0.3 - 0.7j
    assert np.all(got == exp)
    S.close()


def test_serving_default_ics_are_the_generators():
    """The serving workload's device-built states: 1D dark soliton, 2D
    vortex, 3D ring (oracle restatements of the same generator calls)."""
    from nlsemagic_amd import serving as SV
    h = 0.2
    for shape in [(101,), (30, 26), (24, 24, 24)]:
        S = _solver(shape, 2)
        SV.default_ic(shape, h)(S)
        got = S.state()
        S.close()
        origin = tuple(-(n - 1) * h / 2 for n in shape)
        one = np.ones(shape, complex)
        if len(shape) == 1:
            exp = I.exact(2, (1.0, -1.0, -1.0, 0.0, 0.0), shape, origin, h, 0.0)
        elif len(shape) == 2:
            exp = I.add_vortex_2d(1, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, one, h, 1.0, -1.0, 1.0, origin, 0)
        else:
            o3 = (origin[1], origin[2], origin[0])
            exp = I.add_vortex_ring(1, -1.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0, 0, 0.0, one, h, 1.0, -1.0, o3,
                                    0.25 * h * shape[2])
        print("CASE serving ic", shape, flush=True)
        assert np.max(np.abs(got - exp)) < 1e-11


@pytest.mark.parametrize("turns", list(range(8)))
def test_quarter_turn_rings_on_a_non_cubic_grid(turns):
    """Every combination of quarter turns (the reference's permute + flip
    path) on a grid with three different extents, against the oracle."""
    th = [math.pi / 2 if turns & (1 << b) else 0.0 for b in range(3)]
    shape, h, origin = (20, 24, 28), 0.3, (-3.45, -4.05, -2.85)
    print("CASE quarter turns", th, flush=True)
    S = _solver(shape, 2)
    S.fill(1.0)
    S.add_vortex_ring(1, -1.0, 0.2, -0.1, 0.3, *th, 0.0, 0.1, 0.0, 0.2, h, 1.0, -1.0, origin, 2.0)
    exp = I.add_vortex_ring(1, -1.0, 0.2, -0.1, 0.3, *th, 0.0, 0.1, 0.0, 0.2, np.ones(shape, complex), h, 1.0, -1.0,
                            origin, 2.0)
    assert np.max(np.abs(S.state() - exp)) <= 1e-11
    S.close()
