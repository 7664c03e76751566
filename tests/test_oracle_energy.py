"""Pins of the energy restatement (oracle/energy_oracle.py), CPU only.

The reference has no energy diagnostic, so the formula is pinned by
mathematics: a plane wave's closed form, the identity between the energy's
gradient and the reference's CD right-hand side, and conservation by the
reference integrator (the C restatement, bitwise equal to the compiled
reference in 1D/2D) with Dirichlet boundaries.  Each check is shown to reject
planted mistakes in the formula.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.energy_oracle import energy


def _planted(kind):
    """The energy with one planted mistake."""
    def e(U, V, s, a, h2):
        hd = h2 ** (0.5 * U.ndim)
        g = sum(float(np.sum(np.abs(np.diff(U, axis=ax)) ** 2)) for ax in range(U.ndim))
        m2 = np.abs(U) ** 2
        if kind == "nl_sign":
            p = np.sum(V * m2 + 0.5 * s * m2 * m2)
        elif kind == "nl_half":
            p = np.sum(V * m2 - s * m2 * m2)
        elif kind == "v_sign":
            p = np.sum(-V * m2 - 0.5 * s * m2 * m2)
        else:  # "grad_a": the Laplacian coefficient dropped
            return hd * (1.0 / h2) * g + hd * float(np.sum(V * m2 - 0.5 * s * m2 * m2))
        return hd * (a / h2) * g + hd * float(p)
    return e


PLANTED = ["nl_sign", "nl_half", "v_sign", "grad_a"]


@pytest.mark.parametrize("shape", [(40,), (12, 9), (5, 6, 7)])
def test_plane_wave_closed_form(shape):
    """U = A exp(i kappa x) along the slowest axis: sum over pairs of
    |U_p - U_q|^2 = (n0 - 1) (points per slice) |A|^2 4 sin^2(kappa h / 2)."""
    h, a, s, A, kappa, v0 = 0.3, 0.8, -1.0, 1.3 + 0.4j, 0.9, 0.25
    x = h * np.arange(shape[0]).reshape((-1,) + (1,) * (len(shape) - 1))
    U = A * np.exp(1j * kappa * x) * np.ones(shape)
    V = np.full(shape, v0)
    E, G, P, M = energy(U, V, s, a, h * h)
    n = int(np.prod(shape))
    hd = h ** len(shape)
    pairs = (shape[0] - 1) * (n // shape[0])
    assert np.isclose(G, hd * a / h ** 2 * pairs * abs(A) ** 2 * 4 * np.sin(kappa * h / 2) ** 2, rtol=1e-13)
    assert np.isclose(P, hd * n * (v0 * abs(A) ** 2 - 0.5 * s * abs(A) ** 4), rtol=1e-13)
    assert np.isclose(M, hd * n * abs(A) ** 2, rtol=1e-13)
    assert E == G + P


def _grad_identity_error(oracle, efun, U, V, s, a, h2):
    """max over interior points of |F - (-i dE/dconj(U) / h^d)| / max |F|,
    dE/dconj(U) = (dE/dRe U + i dE/dIm U) / 2 by central differences."""
    F = oracle.rhs(U, V, s, a, h2, 1, 1)  # CD, Dirichlet
    hd = h2 ** (0.5 * U.ndim)
    eps = 1e-6
    worst = 0.0
    interior = tuple(slice(1, n - 1) for n in U.shape)
    idx = np.argwhere(np.ones(U.shape, bool)[interior]) + 1
    for p in map(tuple, idx):
        d = []
        for step in (eps, 1j * eps):
            Up, Um = U.copy(), U.copy()
            Up[p] += step
            Um[p] -= step
            d.append((efun(Up, V, s, a, h2)[0] if efun is energy else efun(Up, V, s, a, h2)) -
                     (efun(Um, V, s, a, h2)[0] if efun is energy else efun(Um, V, s, a, h2)))
        grad = (d[0] + 1j * d[1]) / (2 * eps) / 2
        worst = max(worst, abs(F[p] - (-1j) * grad / hd))
    return worst / np.max(np.abs(F))


@pytest.mark.parametrize("shape", [(9,), (6, 5), (4, 5, 4)])
def test_gradient_is_the_cd_right_hand_side(oracle, shape):
    """F = -i dE/dconj(U) / h^d at every interior point (the semi-discrete CD
    flow is Hamiltonian with this E); each planted mistake breaks it."""
    rng = np.random.default_rng(len(shape))
    U = (0.8 + 0.4 * rng.uniform(-1, 1, shape)) + 1j * 0.4 * rng.uniform(-1, 1, shape)
    V = 0.5 * rng.uniform(-1, 1, shape)
    s, a, h2 = -1.3, 0.7, 0.09
    assert _grad_identity_error(oracle, energy, U, V, s, a, h2) < 1e-6
    for kind in PLANTED:
        assert _grad_identity_error(oracle, _planted(kind), U, V, s, a, h2) > 1e-3, kind


@pytest.mark.parametrize("shape,h", [((80,), 0.2), ((22, 18), 0.25), ((10, 11, 9), 0.3)])
def test_conserved_by_the_reference_integrator(oracle, shape, h):
    """CD, Dirichlet, a smooth state: E drifts only by the RK4 error, which
    falls at least 12-fold when k halves over the same time (RK4 dissipates
    at O(k^4) or better); each planted variant drifts by far more and does
    not converge with k."""
    grids = np.meshgrid(*[h * (np.arange(n) - (n - 1) / 2) for n in shape], indexing="ij")
    r2 = sum(g * g for g in grids)
    U = np.exp(-r2 / 4) * np.exp(1j * 0.5 * grids[0])
    V = 0.3 * np.tanh(r2 / 8)
    s, a, h2 = 1.0, 0.6, h * h
    ends = [oracle.take_steps(U, V, s, a, h2, 1, n, kf * h2 / a, 1) for kf, n in ((0.05, 400), (0.025, 800))]
    drift = [abs(energy(W, V, s, a, h2)[0] / energy(U, V, s, a, h2)[0] - 1) for W in ends]
    assert drift[0] < 1e-6 and drift[0] > 12 * drift[1], drift
    for kind in PLANTED:
        e = _planted(kind)
        d = [abs(e(W, V, s, a, h2) / e(U, V, s, a, h2) - 1) for W in ends]
        assert d[1] > 100 * drift[0] and d[1] > 0.5 * d[0], (kind, d, drift)
