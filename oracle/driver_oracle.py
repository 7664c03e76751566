"""CPU restatement of the reference drivers' full RK4 stability bound (test
infrastructure: only tests/ use it, as the checker of nlse_kfull).

Follows the `if(k==0)` blocks of src/NLSEmagic1D/NLSEmagic1D_FRS.m:233-275,
src/NLSEmagic2D/NLSEmagic2D_Full_Research_Scripts/NLSE2D.m:427-477 and
src/NLSEmagic3D/NLSE3D.m:443-495, step by step, in numpy:

    x      = (hmin^2/a) (s U conj(U) - V)                    every point
    Lmax   = max over G of max over points |x - G|
    Bmax   = 0                                                BC 1
           = real(max((hmin^2/(i a)) Ut ./ U)) on the first   BC 2 (MSD)
             interior layer, Ut the RHS (oracle.rhs), max by
             MATLAB's complex order (magnitude, then angle)
           = max x on the boundary                            BC 3 (Lap0)
    kfull  = (hmin^2/a) sqrt(8) / max(|Bmax|, Lmax)
    k      = c min(kfull, klin), c = 0.9 (1D, 2D), 0.6 (3D)

Pins (tests/test_oracle_driver.py): in the linear case (s = 0, V = 0) kfull
equals the linear bound klin in every dimension and method, which fixes the
largest entry of each G table (4d for CD, 4/3 of it for 2SHOC: the schemes'
spectral radii); a uniform field has closed-form Lmax and Bmax for every
BC.  The smallest G entries follow the reference text and are not pinned by
mathematics (parity unpinned for them).
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

# eigenvalue offsets of the stability analysis (NLSEmagic1D_FRS.m:236-242,
# NLSE2D.m:430-436, NLSE3D.m:447-453); 2SHOC tables in twelfths
G_TABLES = {
    (1, 1): [4, 3, 1, 0],
    (1, 2): [v / 12 for v in (64, 63, 46, 12, -3, -4)],
    (2, 1): [8, 7, 6, 2, 1, 0],
    (2, 2): [v / 12 for v in (128, 127, 126, 110, 109, 92, 24, 9, 8, -6, -7, -8)],
    (3, 1): [12, 11, 10, 9, 3, 2, 1, 0],
    (3, 2): [v / 12 for v in (192, 191, 190, 189, 174, 173, 172, 156, 155, 138, 36, 21, 20, 6, 5, 4,
                              -9, -10, -11, -12)],
}


def klin(dim, method, h, a):
    """NLSEmagic1D_FRS.m:238,241, NLSE2D.m:432,435, NLSE3D.m:449,452."""
    return (1.0 if method == 1 else 0.75) * h ** 2 / (dim * np.sqrt(2) * a)


def _faces(shape):
    """Boundary points (an index at 0 or n-1) and the first interior layer
    (every index in 1..n-2, one of them at 1 or n-2) as boolean masks."""
    idx = np.indices(shape)
    n = np.array(shape).reshape((-1,) + (1,) * len(shape))
    bnd = np.any((idx == 0) | (idx == n - 1), axis=0)
    inner = np.all((idx >= 1) & (idx <= n - 2), axis=0)
    near = np.any((idx == 1) | (idx == n - 2), axis=0)
    return bnd, inner & near


def matlab_complex_max(z):
    """max of a complex vector as MATLAB orders it: largest magnitude, ties
    to the larger angle; NaN entries are omitted."""
    z = z[~np.isnan(z)]
    if z.size == 0:
        return np.nan
    order = np.lexsort((np.angle(z), np.abs(z)))
    return z[order[-1]]


def kfull(U, V, h, a, s, method, BC):
    """(kfull, klin, k) for the state U, V (numpy arrays of one shape)."""
    U = np.asarray(U, dtype=complex)
    V = np.asarray(V, dtype=float)
    dim = U.ndim
    G = G_TABLES[(dim, method)]
    x = ((h ** 2 / a) * (s * U * np.conj(U) - V)).real
    lmax = max(np.max(np.abs(x - g)) for g in G)
    bnd, layer = _faces(U.shape)
    if BC == 2:
        Ut = O.rhs(U, V, s, a, h * h, BC, method)
        bmax = np.real(matlab_complex_max((h ** 2 / (1j * a)) * Ut[layer] / U[layer]))
    elif BC == 3:
        bmax = np.max(x[bnd])
    else:
        bmax = 0.0
    kf = (h ** 2 / a) * np.sqrt(8) / max(abs(bmax), lmax)
    kl = klin(dim, method, h, a)
    c = 0.6 if dim == 3 else 0.9
    return kf, kl, (c * kf if kf < kl else c * kl)
