"""Pins of the drivers' full RK4 stability bound restated in
oracle/driver_oracle.py (NLSEmagic1D_FRS.m:233-275, NLSE2D.m:427-477,
NLSE3D.m:443-495), by mathematics rather than by retyping the tables:

- linear case (s = 0, V = 0): every x is 0, so kfull = (h^2/a) sqrt(8) /
  max|G|, and it must equal the linear bound klin; this fixes the largest G
  entry at the scheme's spectral radius, 4d for CD and 4/3 of it for 2SHOC;
- uniform field U = c, V = 0: x = s (h^2/a)|c|^2 everywhere, the Laplacian
  vanishes, so the MSD layer value (h^2/(i a)) F/U is x itself and Lap0's
  boundary maximum is x; with s = -1, Lmax = 4d(4/3 for 2SHOC) - x;
- MATLAB's complex max: largest magnitude, ties to the larger angle, NaN
  omitted.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import driver_oracle as D


def _shape(dim):
    return {1: (40,), 2: (12, 17), 3: (6, 7, 9)}[dim]


def _radius(dim, method):
    return 4.0 * dim * (1.0 if method == 1 else 4.0 / 3.0)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("BC", [1, 3])
def test_linear_case_gives_the_linear_bound(oracle, dim, method, BC):
    rng = np.random.default_rng(dim + 10 * method)
    shape = _shape(dim)
    U = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    kf, kl, k = D.kfull(U, np.zeros(shape), 0.1, 0.7, 0.0, method, BC)
    assert np.isclose(kf, kl, rtol=1e-13, atol=0)
    assert np.isclose(kl, (0.1 ** 2 / 0.7) * np.sqrt(8) / _radius(dim, method), rtol=1e-13)
    assert np.isclose(k, (0.6 if dim == 3 else 0.9) * kl, rtol=1e-15)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("BC", [1, 2, 3])
def test_uniform_field_closed_form(oracle, dim, method, BC):
    h, a, c = 0.2, 0.8, 1.3 - 0.4j
    shape = _shape(dim)
    U = np.full(shape, c)
    kf, kl, k = D.kfull(U, np.zeros(shape), h, a, -1.0, method, BC)
    x = -(h * h / a) * abs(c) ** 2
    lmax = _radius(dim, method) - x
    assert np.isclose(kf, (h * h / a) * np.sqrt(8) / lmax, rtol=1e-13)
    assert kf < kl
    assert np.isclose(k, (0.6 if dim == 3 else 0.9) * kf, rtol=1e-15)


def test_uniform_field_boundary_values():
    """The MSD layer value and Lap0's boundary value of a uniform field are x."""
    h, a, s, c = 0.2, 0.8, 1.0, 0.5 + 0.5j
    U = np.full((9, 11), c)
    x = s * (h * h / a) * abs(c) ** 2
    from oracle import oracle as O
    Ut = O.rhs(U, np.zeros(U.shape), s, a, h * h, 2, 2)
    _, layer = D._faces(U.shape)
    z = D.matlab_complex_max((h * h / (1j * a)) * Ut[layer] / U[layer])
    assert np.isclose(z.real, x, rtol=1e-13) and abs(z.imag) < 1e-13
    bnd, _ = D._faces(U.shape)
    assert bnd.sum() == 2 * 9 + 2 * 11 - 4 and layer.sum() == 2 * 7 + 2 * 9 - 4


def test_matlab_complex_max_order():
    assert D.matlab_complex_max(np.array([1, -2, 2j])) == -2          # tie in |z|: larger angle
    assert D.matlab_complex_max(np.array([1 + 0j, np.nan, -3j])) == -3j
    assert D.matlab_complex_max(np.array([0.5, 0.5j])) == 0.5j
