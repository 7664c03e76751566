"""Mathematical pins of the closed-form fields and vortex initial conditions
(oracle/ic_oracle.py), and the host-side radial profile solver of the
library (nlse_vortex_profile: host logic, no GPU).

- The FRS closed forms (NLSEmagic1D_FRS.m:535-545) are exact solutions of
  the NLSE: integrating the closed form at t = 0 with the CPU integrator
  leaves an error against the closed form at t that falls as h^2 (CD) and
  h^4 (2SHOC).  A wrong sign or factor in any phase term gives an O(1) error.
- The radial profile is the root of the discrete steady state
  (NLS_STEADY_F.m): two independent Newton solvers (complex-step banded
  Jacobian here, analytic tridiagonal Jacobian in the library) agree, the
  residual vanishes, the core behaves as r^|m| and the far field tends to
  sqrt(OM/s).
- The generated vortices carry their charge: winding number m of the phase
  (2D), circulation pi*m of the vorticity (3D line, each core of a ring).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import ic_oracle as I


def _frs_case(kind):
    # NLSEmagic1D_FRS.m:93-128: (a, s, OM, c, x0), BC, half-width
    return {1: ((1.0, 1.0, 0.5, 0.75, 0.0), 1, 30.0),
            2: ((1.0, -1.0, -1.0, 0.75, 0.0), 2, 30.0),
            4: ((1.0, -1.0, -1.0, 0.75, 0.0), 2, 30.0)}[kind]


def _evolve_error(oracle, kind, method, h, T=0.5):
    par, BC, W = _frs_case(kind)
    n = int(round(2 * W / h))
    origin = (-W,)
    U0 = I.exact(kind, par, (n,), origin, h, 0.0)
    k = 0.05 * h * h
    steps = int(round(T / k))
    k = T / steps
    a, s = par[0], par[1]
    U = oracle.take_steps(U0, np.zeros(n), s, a, h * h, BC, steps, k, method)
    return I.exact_error(U, kind, par, origin, h, T)[3]  # max |d Re|


@pytest.mark.parametrize("kind", [1, 2, 4])
@pytest.mark.parametrize("method,order", [(1, 2), (2, 4)])
def test_frs_closed_forms_are_solutions(oracle, kind, method, order):
    e1 = _evolve_error(oracle, kind, method, 0.4)
    e2 = _evolve_error(oracle, kind, method, 0.2)
    rate = math.log2(e1 / e2)
    assert abs(rate - order) < 0.35, (e1, e2, rate)


def test_closed_forms_at_zero_match_the_frs_initial_conditions():
    x = np.linspace(-5, 5, 41)
    h, origin = x[1] - x[0], (x[0],)
    a, s, OM, c, x0 = 1.0, 1.0, 0.5, 0.75, 0.0
    np.testing.assert_allclose(I.exact(1, (a, s, OM, c, x0), x.shape, origin, h, 0.0),
                               np.sqrt(2 * OM / s) / np.cosh(np.sqrt(OM / a) * x) * np.exp(1j * c / (2 * a) * x),
                               rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(I.exact(3, (a, 0, 0, 0, 0), x.shape, origin, h, 0.0), np.exp(-x ** 2 / (2 * a)),
                               rtol=1e-14)
    # the grey soliton's density dip: |U|^2 = (OM/s) (c^2/vc^2 + (1 - c^2/vc^2) tanh^2)
    a, s, OM, c = 1.0, -1.0, -1.0, 0.75
    vc = np.sqrt(-2 * a * OM)
    g = I.exact(4, (a, s, OM, c, 0.0), x.shape, origin, h, 0.0)
    assert abs(np.abs(g[20]) ** 2 - c ** 2 / vc ** 2) < 1e-14


def test_exact_error_norms():
    rng = np.random.default_rng(0)
    par = (1.0, 0.0, 0.0, 0.0, 0.0)
    E = I.exact(3, par, (6, 7), (-1.0, -1.5), 0.5, 0.3)
    d = 1e-3 * (rng.standard_normal(E.shape) + 1j * rng.standard_normal(E.shape))
    out = I.exact_error(E + d, 3, par, (-1.0, -1.5), 0.5, 0.3)
    np.testing.assert_allclose(out[0], np.sqrt(np.mean(d.real ** 2)), rtol=1e-10)
    np.testing.assert_allclose(out[4], np.max(np.abs(d.imag)), rtol=1e-10)
    assert I.exact_error(E, 3, par, (-1.0, -1.5), 0.5, 0.3) == (0.0,) * 6


@pytest.mark.parametrize("m,OM,s,kind,tol", [(1, -1.0, -1.0, 0, 1e-14), (2, -1.0, -1.0, 0, 1e-14),
                                             (3, -0.7, -1.0, 0, 1e-12), (1, 1 / 3, 1.0, 2, 1e-12),
                                             (1, -1.0, -1.0, 1, 1e-8)])
def test_radial_profile_library_matches_oracle(m, OM, s, kind, tol):
    import nlsemagic_amd as nm
    hr, n = 0.1, 500
    fo = I.profile(m, OM, 1.0, s, hr, n, kind, True, tol)
    fl = nm.vortex_profile(m, OM, 1.0, s, hr, n, kind, True, tol)
    np.testing.assert_allclose(fl, fo, rtol=0, atol=1e-12)
    assert np.max(np.abs(I.steady_residual(fl, m, OM, 1.0, s, hr))) < 1e-11
    if s < 0:
        r = np.arange(n) * hr
        # core ~ r^|m|, far field -> sqrt(OM/s + a m^2 / (s r^2))
        slope = math.log(fl[4] / fl[2]) / math.log(2)
        assert abs(slope - abs(m)) < 0.05
        far = np.sqrt(OM / s + m ** 2 / (s * r[-50] ** 2))
        assert abs(fl[-50] - far) < 1e-3


def test_radial_profile_without_newton_is_the_ansatz():
    import nlsemagic_amd as nm
    f = nm.vortex_profile(1, -1.0, 1.0, -1.0, 0.2, 40, 1, optimize=False)
    np.testing.assert_allclose(f, I.ansatz(1, -1.0, 1.0, -1.0, 0.2, 40, 1), rtol=1e-15, atol=1e-16)


def test_radial_profile_rejects_bad_arguments():
    import nlsemagic_amd as nm
    with pytest.raises(nm.HipError, match="n >= 3"):
        nm.vortex_profile(1, -1.0, 1.0, -1.0, 0.1, 2, 0)
    with pytest.raises(nm.HipError, match="ansatz"):
        nm.vortex_profile(1, -1.0, 1.0, -1.0, 0.1, 20, 5)
    with pytest.raises(nm.HipError, match="far-field"):
        nm.vortex_profile(1, 1.0, 1.0, -1.0, 0.1, 20, 0)  # OM/s < 0: no dark background


def _winding(V, i0, i1, j0, j1):
    loop = ([(i, j0) for i in range(i0, i1)] + [(i1, j) for j in range(j0, j1)] +
            [(i, j1) for i in range(i1, i0, -1)] + [(i0, j) for j in range(j1, j0, -1)])
    return sum(np.angle(V[b] / V[a]) for a, b in zip(loop, loop[1:] + loop[:1])) / (2 * np.pi)


@pytest.mark.parametrize("m,opt", [(1, 1), (2, 1), (1, 0), (-1, 1)])
def test_2d_vortex_winding_and_amplitude(m, opt):
    N, M, h, o = 48, 40, 0.25, (-6.0, -5.0)
    V = I.add_vortex_2d(m, -1.0, 0.3, -0.2, 0.0, 0.0, 0.0, np.ones((N, M), complex), h, 1.0, -1.0, 1.0, o, opt)
    assert abs(_winding(V, 10, 38, 8, 32) - m) < 1e-9
    # the product combine rescales to the new vortex's amplitude
    X, Y, _ = I.coords((N, M), o, h)
    R = np.hypot(X - 0.3, Y + 0.2)
    assert np.abs(V).max() <= 1.0 + 1e-15 and np.abs(V[R < 0.2]).max() < 0.15


def test_bright_vortex_is_added():
    N, M, h, o = 40, 40, 0.3, (-6.0, -6.0)
    U0 = 0.1 * np.ones((N, M), complex)
    V = I.add_vortex_2d(1, 1 / 3, 0.0, 0.0, 0.0, 0.0, 0.0, U0, h, 1.0, 1.0, 1.0, o, 1)
    Z = I.add_vortex_2d(1, 1 / 3, 0.0, 0.0, 0.0, 0.0, 0.0, 0 * U0, h, 1.0, 1.0, 1.0, o, 1)
    np.testing.assert_allclose(V - U0, Z, atol=1e-15)
    assert abs(_winding(Z, 5, 35, 5, 35) - 1) < 1e-9


@pytest.mark.parametrize("m", [1, 2])
def test_vortex_line_circulation(oracle, m):
    L, N, M, h, o = 8, 40, 40, 0.3, (-6.0, -6.0, -1.5)
    V = I.add_vortex_line(m, -1.0, 0.1, 0.2, 0.0, 0.0, 0.0, 0.0, 0.0, np.ones((L, N, M), complex), h, 1.0, -1.0, o)
    wx, wy, wz, W = oracle.vorticity3d(V, 1 / (2 * h), 1e-3)
    circ = wx[:, 2:-2, 2:-2].sum(axis=(1, 2)) * h * h
    np.testing.assert_allclose(circ, m * np.pi, rtol=5e-3)


def test_vortex_ring_circulation_and_quarter_turns(oracle):
    n, h, o = 40, 0.4, (-7.8, -7.8, -7.8)
    U = np.ones((n, n, n), complex)
    R = I.add_vortex_ring(1, -1.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0, 0, 0.0, U, h, 1.0, -1.0, o, 4.0)
    wx, wy, wz, W = oracle.vorticity3d(R, 1 / (2 * h), 1e-3)
    k = int(np.argmin(np.abs(o[1] + np.arange(n) * h)))
    X = o[0] + np.arange(n) * h
    idx = np.arange(n)
    sl = wz[2:-2, :, k]  # the x-z plane through the ring's axis (y = 0)
    plus = sl[:, (X > 0) & (idx < n - 2)].sum() * h * h
    minus = sl[:, (X < 0) & (idx > 1)].sum() * h * h
    np.testing.assert_allclose([plus, minus], [-np.pi, np.pi], rtol=5e-3)
    # a quarter turn about x moves the ring's axis from z to y: |U| permutes
    Q = I.add_vortex_ring(1, -1.0, 0.0, 0.0, 0.0, math.pi / 2, 0, 0, 0, 0, 0, 0.0, U, h, 1.0, -1.0, o, 4.0)
    np.testing.assert_allclose(np.abs(Q), np.abs(R).transpose(2, 1, 0), atol=1e-12)
