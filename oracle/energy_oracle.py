"""Discrete energy of an NLSE state (TEST INFRASTRUCTURE ONLY).

The reference reports the mass sum |U|^2 h^d (src/NLSEmagic1D/NLSEmagic1D_FRS.m:582-586)
and has no energy code, so this restatement is pinned by mathematics alone
(tests/test_oracle_energy.py): a closed form for plane waves, the identity
F = -i (dE/d conj U) / h^d between the energy and the reference's CD right-hand
side (F = i[a Lap_CD U + (s|U|^2 - V) U], NLSEmagic1D_CPU.c:27-30) at interior
points, and its conservation by the reference integrator with Dirichlet
boundaries.

  E = h^d [ a/h^2 sum_{neighbour pairs (p,q)} |U_p - U_q|^2
            + sum_p (V_p |U_p|^2 - (s/2) |U_p|^4) ]

Pairs are nearest neighbours along every axis, boundary points included.
Arrays use the reference layout (2D (N, M), 3D (L, N, M), last axis
contiguous).  Only tests use this module.
"""
from __future__ import annotations

import numpy as np


def energy(U, V, s, a, h2):
    """(E, gradient part, potential + nonlinear part, mass), all times h^d."""
    U = np.asarray(U, dtype=np.complex128)
    V = np.asarray(V, dtype=np.float64)
    hd = h2 ** (0.5 * U.ndim)
    g = 0.0
    for ax in range(U.ndim):
        d = np.diff(U, axis=ax)
        g += float(np.sum(d.real ** 2 + d.imag ** 2))
    m2 = U.real ** 2 + U.imag ** 2
    p = float(np.sum(V * m2 - 0.5 * s * m2 * m2))
    grad = hd * (a / h2) * g
    pot = hd * p
    return grad + pot, grad, pot, hd * float(np.sum(m2))
