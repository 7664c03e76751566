"""Mathematical pins of the CPU oracle.

The 1D and 2D oracle is pinned bit-for-bit to the compiled reference
(test_oracle_golden.py).  The reference's 3D integrators cannot be built here
(they include MATLAB's mex.h), so the 3D oracle is pinned by what the
mathematics fixes; every check below also runs on 1D and 2D, where passing
shows the check itself is right:

- eigen-KAT: with Dirichlet boundaries, s = 0 and V = 0, a product of sine
  modes is an exact eigenvector of the discrete operator; one RK4 step
  multiplies it by R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24, z = i a k sigma.
  sigma is the symbol of the CD / 2SHOC stencil, and is itself checked to be
  a 2nd / 4th order approximation of -|theta|^2/h^2;
- convergence: the linear harmonic ground state exp(-r^2/2a), V = r^2/a,
  evolves as exp(-i d t) (NLSEmagic1D_FRS.m:540, NLSE2D.m:1011,
  NLSE3D.m:1272); the observed order is 2 (CD) and 4 (2SHOC);
- the dark soliton tanh(x/sqrt 2) e^{-it} (s = -1, MSD; NLSEmagic1D_FRS.m:538);
- a uniform field with Lap0 boundaries follows u' = i(s|u|^2 - v)u, which
  fixes the signs of the nonlinear and potential terms;
- boundary identities of one F evaluation (Dirichlet, Lap0, MSD);
- 3D CD on an L = 3 grid with zero Dirichlet end planes equals 2D CD with
  V + 2a/h^2 on the middle plane;
- one 3D F evaluation equals a restatement written with axis-shift operators.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from conftest import rel_l2


def R4(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def symbol(method, thetas, h2):
    c = [np.cos(t) for t in thetas]
    d = len(thetas)
    lam = (2 * sum(c) - 2 * d) / h2
    if method == 1:
        return lam
    if d == 1:
        return lam * (7 - c[0]) / 6
    if d == 2:
        return lam * (1 - (c[0] + c[1]) / 6) + (4 * c[0] * c[1] - 4) / (6 * h2)
    pairs = c[0] * c[1] + c[1] * c[2] + c[2] * c[0]
    return -lam * (2 * sum(c) - 10) / 12 + (4 * pairs - 12) / (6 * h2)


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("method,order", [(1, 2), (2, 4)])
def test_symbol_is_consistent_with_the_laplacian(d, method, order):
    """sigma(theta) h^2 = -|theta|^2 + O(theta^(order+2)), any direction."""
    rng = np.random.default_rng(d)
    direction = rng.uniform(0.3, 1.0, d)
    errs = []
    for eps in (0.08, 0.04):
        th = eps * direction
        errs.append(abs(symbol(method, th, 1.0) + np.sum(th ** 2)))
    observed = np.log2(errs[0] / errs[1])
    assert abs(observed - (order + 2)) < 0.1, observed


def test_2shoc_symbol_at_nyquist_is_four_thirds_of_cd():
    """The 3/4 factor of the reference's 2SHOC time-step rule (NLSEmagic2D.m:131-134)."""
    for d in (1, 2, 3):
        th = [np.pi] * d
        assert np.isclose(symbol(2, th, 1.0), 4 / 3 * symbol(1, th, 1.0))


def _shape(d):
    return {1: (23,), 2: (13, 17), 3: (9, 11, 14)}[d]  # (N,), (N, M), (L, N, M)


def _mode(shape, modes):
    axes = [np.sin(np.pi * m * np.arange(n) / (n - 1)) for n, m in zip(shape, modes)]
    out = axes[0]
    for a in axes[1:]:
        out = np.multiply.outer(out, a)
    return out


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("method", [1, 2])
def test_eigenmode_known_answer(oracle, d, method):
    shape = _shape(d)
    modes = {1: (3,), 2: (2, 5), 3: (1, 3, 4)}[d]
    thetas = [np.pi * m / (n - 1) for n, m in zip(shape, modes)]
    a, h2, k, steps = 0.7, 0.04, 0.005, 9
    sig = symbol(method, thetas, h2)
    U0 = (1 + 0.5j) * _mode(shape, modes)
    V = np.zeros(shape)
    got = oracle.take_steps(U0, V, 0.0, a, h2, 1, steps, k, method)
    exp = R4(1j * a * k * sig) ** steps * U0
    assert rel_l2(got, exp) < 1e-12


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("method", [1, 2])
def test_eigenmode_single_precision(oracle, d, method):
    shape = _shape(d)
    modes = {1: (3,), 2: (2, 5), 3: (1, 3, 4)}[d]
    thetas = [np.pi * m / (n - 1) for n, m in zip(shape, modes)]
    a, h2, k, steps = 0.7, 0.04, 0.005, 9
    U0 = (1 + 0.5j) * _mode(shape, modes)
    got = oracle.take_steps(U0, np.zeros(shape), 0.0, a, h2, 1, steps, k, method, precision=1)
    exp = R4(1j * a * k * symbol(method, thetas, h2)) ** steps * U0
    assert rel_l2(got, exp) < 1e-5


def _harmonic_error(oracle, d, method, n, a=0.8, L=7.0, T=0.25):
    h = 2 * L / (n - 1)
    x = -L + h * np.arange(n)
    grids = np.meshgrid(*([x] * d), indexing="ij")
    r2 = sum(g * g for g in grids)
    U0 = np.exp(-r2 / (2 * a)) + 0j
    V = r2 / a
    k = 0.05 * h * h / a
    steps = int(round(T / k))
    k = T / steps
    got = oracle.take_steps(U0, V, 0.0, a, h * h, 1, steps, k, method)
    exact = U0 * np.exp(-1j * d * T)
    return np.max(np.abs(got - exact))


@pytest.mark.parametrize("d,sizes", [(1, (61, 121)), (2, (41, 81)), (3, (25, 49))])
@pytest.mark.parametrize("method,order", [(1, 2), (2, 4)])
def test_harmonic_ground_state_convergence_order(oracle, d, sizes, method, order):
    e1 = _harmonic_error(oracle, d, method, sizes[0])
    e2 = _harmonic_error(oracle, d, method, sizes[1])
    h_ratio = (sizes[1] - 1) / (sizes[0] - 1)
    observed = np.log(e1 / e2) / np.log(h_ratio)
    assert abs(observed - order) < 0.35, (e1, e2, observed)


@pytest.mark.parametrize("method", [1, 2])
def test_dark_soliton_1d(oracle, method):
    n = 801
    L = 25.0
    h = 2 * L / (n - 1)
    x = -L + h * np.arange(n)
    U0 = np.tanh(x / np.sqrt(2)) + 0j
    k = 0.2 * h * h
    T = 1.0
    steps = int(round(T / k))
    k = T / steps
    got = oracle.take_steps(U0, np.zeros(n), -1.0, 1.0, h * h, 2, steps, k, method)
    err = np.max(np.abs(got - U0 * np.exp(-1j * T)))
    assert err < (2e-4 if method == 1 else 2e-6), err


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("precision", [2, 1])
def test_uniform_field_follows_the_local_ode(oracle, d, method, precision):
    shape = _shape(d)
    c0 = 0.6 - 0.8j
    s, v0, k, steps = -1.3, 0.4, 0.01, 50
    U0 = np.full(shape, c0)
    V = np.full(shape, v0)
    got = oracle.take_steps(U0, V, s, 0.9, 0.05, 3, steps, k, method, precision=precision)
    exact = c0 * np.exp(1j * (s * abs(c0) ** 2 - v0) * k * steps)
    # the discrete Laplacian of a constant is zero up to rounding of sums like c - 4c
    assert np.allclose(got, got.ravel()[0], rtol=0, atol=1e-12 if precision == 2 else 1e-5)
    # the classical RK4 tableau on the scalar ODE, written as a generic method
    f = lambda u: 1j * (s * abs(u) ** 2 - v0) * u
    rk = c0
    for _ in range(steps):
        k1 = f(rk)
        k2 = f(rk + k / 2 * k1)
        k3 = f(rk + k / 2 * k2)
        k4 = f(rk + k * k3)
        rk = rk + k / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert abs(got.ravel()[0] - rk) < (1e-12 if precision == 2 else 2e-6)
    assert abs(got.ravel()[0] - exact) < (1e-8 if precision == 2 else 1e-6)
    wrong = c0 * np.exp(1j * (-s * abs(c0) ** 2 + v0) * k * steps)
    assert abs(got.ravel()[0] - wrong) > 0.1


def _boundary_points(shape):
    for idx in itertools.product(*[range(n) for n in shape]):
        if any(i == 0 or i == n - 1 for i, n in zip(idx, shape)):
            yield idx


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("method", [1, 2])
def test_boundary_identities_of_one_evaluation(oracle, d, method):
    shape = _shape(d)
    rng = np.random.default_rng(d * 10 + method)
    U = 1 + 0.3 * rng.uniform(-1, 1, shape) + 0.3j * rng.uniform(-1, 1, shape)
    V = rng.uniform(-1, 1, shape)
    s, a, h2 = -0.7, 0.9, 0.05
    F1 = oracle.rhs(U, V, s, a, h2, 1, method)
    F2 = oracle.rhs(U, V, s, a, h2, 2, method)
    F3 = oracle.rhs(U, V, s, a, h2, 3, method)
    for b in _boundary_points(shape):
        q = tuple(min(max(i, 1), n - 2) for i, n in zip(b, shape))
        assert F1[b] == 0
        assert np.isclose(F3[b], 1j * (s * abs(U[b]) ** 2 - V[b]) * U[b], rtol=1e-12, atol=0)
        om = (F2[q] / U[q]).imag  # MSD: the partner's phase rate carried to the boundary
        assert np.isclose(F2[b], 1j * om * U[b], rtol=1e-13, atol=1e-15)


def test_3d_cd_reduces_to_2d_with_shifted_potential(oracle):
    rng = np.random.default_rng(11)
    N, M = 12, 15
    a, h2, s, k = 0.8, 0.05, -1.0, 0.004
    U2 = 1 + 0.3 * rng.uniform(-1, 1, (N, M)) + 0.3j * rng.uniform(-1, 1, (N, M))
    V2 = rng.uniform(-1, 1, (N, M))
    U3 = np.zeros((3, N, M), complex)
    U3[1] = U2
    V3 = np.zeros((3, N, M))
    V3[1] = V2
    got = oracle.take_steps(U3, V3, s, a, h2, 1, 6, k, 1)
    exp = oracle.take_steps(U2, V2 + 2 * a / h2, s, a, h2, 1, 6, k, 1)
    assert np.all(got[0] == 0) and np.all(got[2] == 0)
    assert rel_l2(got[1], exp) < 1e-13


def _shift(U, axis, off):
    out = np.zeros_like(U)
    n = U.shape[axis]
    src = [slice(None)] * U.ndim
    dst = [slice(None)] * U.ndim
    if off > 0:
        src[axis], dst[axis] = slice(off, n), slice(0, n - off)
    else:
        src[axis], dst[axis] = slice(0, n + off), slice(-off, n)
    out[tuple(dst)] = U[tuple(src)]
    return out


@pytest.mark.parametrize("method", [1, 2])
def test_3d_rhs_equals_shift_operator_form(oracle, method):
    """F = i[a L U + (s|U|^2 - V) U] inside, 0 on a Dirichlet boundary, with
    L = Lap_CD (CD) or L = D/3 - (h^2/12) Lap_CD(D) + (sum_{a<b} T_a T_b - 12)/(6h^2)
    (2SHOC), T_a = shift(+1) + shift(-1) along axis a and D = Lap_CD(U) inside,
    D = (V - s|U|^2) U / a on the boundary."""
    shape = (7, 9, 10)
    rng = np.random.default_rng(4)
    U = 1 + 0.3 * rng.uniform(-1, 1, shape) + 0.3j * rng.uniform(-1, 1, shape)
    V = rng.uniform(-1, 1, shape)
    s, a, h2 = -0.6, 0.85, 0.05
    inner = np.zeros(shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    T = [lambda X, ax=ax: _shift(X, ax, 1) + _shift(X, ax, -1) for ax in range(3)]
    lap = lambda X: (T[0](X) + T[1](X) + T[2](X) - 6 * X) / h2
    nl = (s * np.abs(U) ** 2 - V) * U
    if method == 1:
        Lu = lap(U)
    else:
        D = np.where(inner, lap(U), (V - s * np.abs(U) ** 2) * U / a)
        edges = T[0](T[1](U)) + T[1](T[2](U)) + T[2](T[0](U))
        Lu = D / 3 - h2 / 12 * lap(D) + (edges - 12 * U) / (6 * h2)
    exp = np.where(inner, 1j * (a * Lu + nl), 0)
    got = oracle.rhs(U, V, s, a, h2, 1, method)
    assert np.allclose(got, exp, rtol=1e-11, atol=1e-11 * np.max(np.abs(exp)))


def _msd_partner_d(D, U, V, s, a, partner, sign=1.0):
    """MSD boundary D (NLSE3D_TAKE_STEPS_2SHOC.c:100-190): from the partner's
    D and U, A = Re(D_q conj U_q)/|U_q|^2, D_b = (A + (n_q - n_b)/a) U_b with
    n = s|U|^2 - V; partner(b) gives the partner index of boundary point b
    (sign = -1 plants a sign mistake in n_q - n_b)."""
    out = D.copy()
    for b in _boundary_points(U.shape):
        q = partner(b, U.shape)
        A = (D[q] * np.conj(U[q])).real / abs(U[q]) ** 2
        nq, nb = s * abs(U[q]) ** 2 - V[q], s * abs(U[b]) ** 2 - V[b]
        out[b] = (A + sign * (nq - nb) / a) * U[b]
    return out


def _clamped(b, shape):
    return tuple(min(max(i, 1), n - 2) for i, n in zip(b, shape))


def _second_layer(b, shape):
    """A planted mistake: the partner two layers in."""
    return tuple(min(max(i, 2), n - 3) for i, n in zip(b, shape))


@pytest.mark.parametrize("BC", [2, 3])
def test_3d_2shoc_boundary_d_shift_operator_form(oracle, BC):
    """2SHOC with MSD (D on the boundary from the partner with every
    coordinate clamped into [1, n-2]) and Lap0 (D = 0 on the boundary): the
    interior F of the shift-operator form equals the oracle's, which pins the
    3D boundary D on every face (each interior point next to a face reads the
    face's D).  Edge and corner D are read by no point: the 2SHOC F takes D
    only from the six axis neighbours and the MSD boundary F only from
    interior partners, so their partner rule cannot change a result.
    Planted mistakes (a partner two layers in, a flipped n_q - n_b, a
    Dirichlet-style D on a Lap0 boundary) all fail."""
    shape = (6, 7, 8)
    rng = np.random.default_rng(8 + BC)
    U = 1 + 0.3 * rng.uniform(-1, 1, shape) + 0.3j * rng.uniform(-1, 1, shape)
    V = rng.uniform(-1, 1, shape)
    s, a, h2 = -0.6, 0.85, 0.05
    inner = np.zeros(shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    T = [lambda X, ax=ax: _shift(X, ax, 1) + _shift(X, ax, -1) for ax in range(3)]
    lap = lambda X: (T[0](X) + T[1](X) + T[2](X) - 6 * X) / h2
    nl = (s * np.abs(U) ** 2 - V) * U
    edges = T[0](T[1](U)) + T[1](T[2](U)) + T[2](T[0](U))

    def interior_f(D):
        Lu = D / 3 - h2 / 12 * lap(D) + (edges - 12 * U) / (6 * h2)
        return 1j * (a * Lu + nl)

    got = oracle.rhs(U, V, s, a, h2, BC, 2)
    Din = np.where(inner, lap(U), 0)
    if BC == 2:
        right = _msd_partner_d(Din, U, V, s, a, _clamped)
        wrongs = [_msd_partner_d(Din, U, V, s, a, _second_layer), _msd_partner_d(Din, U, V, s, a, _clamped, -1.0)]
    else:
        right = Din
        wrongs = [np.where(inner, Din, (V - s * np.abs(U) ** 2) * U / a)]
    exp = interior_f(right)
    tol = 1e-11 * np.max(np.abs(exp[inner]))
    assert np.max(np.abs(got[inner] - exp[inner])) < tol
    for wrong in wrongs:
        assert np.max(np.abs(got[inner] - interior_f(wrong)[inner])) > 1e3 * tol


def test_single_precision_tracks_double(oracle):
    rng = np.random.default_rng(2)
    for shape in ((64,), (20, 24), (8, 10, 12)):
        U = 1 + 0.1 * rng.uniform(-1, 1, shape) + 0.1j * rng.uniform(-1, 1, shape)
        V = 0.1 * rng.uniform(-1, 1, shape)
        for method in (1, 2):
            d = oracle.take_steps(U, V, -1.0, 1.0, 0.04, 2, 10, 0.004, method)
            f = oracle.take_steps(U, V, -1.0, 1.0, 0.04, 2, 10, 0.004, method, precision=1)
            assert rel_l2(f, d) < 1e-5


# ---- 3D vorticity (rmc_vorticity3D.c:24-354) ------------------------------
# The reference's vorticity MEX file needs MATLAB's mex.h and is not built
# here; its restatement (oracle/nlse_vorticity.c) is pinned by what the
# mathematics fixes: a plane wave carries no vorticity, a straight vortex line
# of winding +-1 carries circulation +-pi in the reference's normalisation
# (V = Im(conj U grad U) / (2|U|^2)), axis permutations permute the outputs,
# and the points where the reference's unset edge gradient leaks in are
# exactly oracle.vorticity_defined_mask.

def _grid3(shape, h):
    L, N, M = shape
    return np.meshgrid(*[(np.arange(n) - (n - 1) / 2) * h for n in (L, N, M)], indexing="ij")


def test_vorticity_of_a_plane_wave_vanishes(oracle):
    h = 0.3
    x, y, z = _grid3((12, 14, 16), h)
    U = np.exp(1j * (0.7 * x + 0.4 * y - 0.3 * z))
    wx, wy, wz, W = oracle.vorticity3d(U, 1 / (2 * h), 1e-3)
    mask = oracle.vorticity_defined_mask(U.shape)
    # the face rules keep V exact: Im(conj(U0)(U1 - U0)) = Im(conj(U0) U1) for |U0| = 1
    assert np.max(W[mask]) < 1e-13
    # velocity scale, for comparison: |k| / 2 ~ 0.4
    _, _, _, W2 = oracle.vorticity3d(U * np.exp(0.05j * x ** 2), 1 / (2 * h), 1e-3)
    assert np.max(W2[mask]) > 1e-3  # a curved phase front is not irrotational in this check's eyes


@pytest.mark.parametrize("sign", [1, -1])
@pytest.mark.parametrize("n", [40, 64])
def test_vortex_line_carries_circulation_pi(oracle, sign, n):
    h = 8.0 / n
    L = 8
    x, y, z = _grid3((L, n, n), h)
    r = np.hypot(y, z)
    U = np.tanh(r / 0.5) * np.exp(sign * 1j * np.arctan2(z, y))
    wx, wy, wz, W = oracle.vorticity3d(U, 1 / (2 * h), 1e-3)
    # over indices 2 .. n-3 of every slice, the x-faces included (whose wx
    # uses central y, z differences; the rows next to the edges read the
    # unset edge velocity and are left out)
    circ = wx[:, 2:-2, 2:-2].sum(axis=(1, 2)) * h * h
    np.testing.assert_allclose(circ, sign * np.pi, rtol=2e-3)
    # the line runs along the slowest axis: no other component where the
    # reference is defined
    mask = oracle.vorticity_defined_mask(U.shape)
    assert np.max(np.abs(wy[mask])) < 1e-12 and np.max(np.abs(wz[mask])) < 1e-12
    np.testing.assert_allclose(W[mask], np.abs(wx[mask]), rtol=1e-12, atol=1e-12)


def test_vorticity_axis_permutation(oracle):
    """U'(j, k, i) = U(i, j, k): (wx', wy', wz') = (wy, wz, wx) bit for bit
    where every difference is central (indices 2 .. n-3)."""
    rng = np.random.default_rng(5)
    U = 1 + 0.3 * rng.standard_normal((9, 10, 11)) + 0.3j * rng.standard_normal((9, 10, 11))
    lh2, veps = 1.7, 1e-3
    wx, wy, wz, W = oracle.vorticity3d(U, lh2, veps)
    P = np.ascontiguousarray(U.transpose(1, 2, 0))
    px, py, pz, PW = [np.ascontiguousarray(a.transpose(2, 0, 1)) for a in oracle.vorticity3d(P, lh2, veps)]
    core = (slice(2, -2),) * 3
    np.testing.assert_array_equal(px[core], wy[core])
    np.testing.assert_array_equal(py[core], wz[core])
    np.testing.assert_array_equal(pz[core], wx[core])
    np.testing.assert_allclose(PW[core], W[core], rtol=1e-15)


def test_vorticity_is_phase_and_scale_invariant(oracle):
    """V depends on U only through grad(phase) when veps = 0: a global phase
    and a global scale leave the vorticity unchanged up to rounding."""
    rng = np.random.default_rng(6)
    U = 1 + 0.3 * rng.standard_normal((7, 8, 9)) + 0.3j * rng.standard_normal((7, 8, 9))
    base = oracle.vorticity3d(U, 1.3, 0.0)
    for f in (np.exp(0.9j), 3.5, -2j):
        got = oracle.vorticity3d(U * f, 1.3, 0.0)
        for a, b in zip(got, base):
            np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 5, 6), (5, 6, 7), (9, 12, 8)])
def test_vorticity_defined_mask_is_where_unset_values_leak(oracle, shape):
    """Fill the gradient the reference leaves unset with NaN: the points that
    stay finite (edges and corners aside, which the reference never writes)
    are exactly the defined mask, and there the values do not depend on the
    fill."""
    rng = np.random.default_rng(7)
    U = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    zero = oracle.vorticity3d(U, 0.7, 1e-3)
    nan = oracle.vorticity3d(U, 0.7, 1e-3, edge_g=np.nan)
    L, N, M = shape
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    nb = ((i == 0) | (i == L - 1)).astype(int) + ((j == 0) | (j == N - 1)) + ((k == 0) | (k == M - 1))
    edge = nb >= 2
    mask = oracle.vorticity_defined_mask(shape)
    np.testing.assert_array_equal(np.isfinite(nan[3]) & ~edge, mask)
    for a, b in zip(zero, nan):
        np.testing.assert_array_equal(a[mask], b[mask])
        assert np.all(a[edge] == 0) and np.all(b[edge] == 0)
