"""Closed-form fields, exact-solution error norms and vortex initial
conditions: a numpy restatement of the reference's MATLAB generators.

TEST INFRASTRUCTURE ONLY: the checker for nlse_set_exact, nlse_exact_error,
nlse_add_vortex_{2d,line,ring} and nlse_vortex_profile.  Nothing in the
product imports it.

Followed, step by step:
  exact solutions       src/NLSEmagic1D/NLSEmagic1D_FRS.m:201-225 (ICs),
                        :535-545 (Ureal), :582-595 (error norms);
                        NLSEmagic2D_Full_Research_Scripts/NLSE2D.m:1008-1014,
                        src/NLSEmagic3D/NLSE3D.m:1270-1276 (linear case, RMS norms)
  radial profile        src/NLSEmagic3D/NLS_STEADY_F.m:10-55 (residual), solved
                        by Newton's method with a banded Jacobian obtained by
                        complex-step differentiation (the reference uses
                        Newton-Krylov, nsoli.m; both converge to the root of
                        the same discrete equations)
  2D vortex             NLSEmagic2D_Full_Research_Scripts/NLS_add_vortex.m:9-158
  3D vortex line        src/NLSEmagic3D/NLSE_add_vortex_line.m:10-125
  3D vortex ring        src/NLSEmagic3D/NLSE_add_vortex_ring.m:10-534 (opt_2D = 0)

Array layout is this package's: 1D (N,), 2D (N, M) with x along axis 0 and
y along axis 1, 3D (L, N, M) with z, x, y along the axes (MATLAB's M x N x L
arrays from meshgrid).  MATLAB's a:d:b has colon_count(a, d, b) elements
a + k d.
"""
from __future__ import annotations

import math

import numpy as np


def colon_count(a, d, b):
    if b < a:
        return 0
    q = (b - a) / d
    return int(math.floor(q + 1e-10 * max(1.0, abs(q)))) + 1


def coords(shape, origin, h):
    """(X, Y, Z) of the grid (None for absent axes), as the GPU kernels
    compute them: coordinate = origin + index * h."""
    o = tuple(origin) + (0.0,) * (3 - len(tuple(origin)))
    if len(shape) == 1:
        return o[0] + np.arange(shape[0]) * h, None, None
    if len(shape) == 2:
        N, M = shape
        X = np.broadcast_to((o[0] + np.arange(N) * h)[:, None], shape)
        Y = np.broadcast_to((o[1] + np.arange(M) * h)[None, :], shape)
        return X, Y, None
    L, N, M = shape
    Z = np.broadcast_to((o[2] + np.arange(L) * h)[:, None, None], shape)
    X = np.broadcast_to((o[0] + np.arange(N) * h)[None, :, None], shape)
    Y = np.broadcast_to((o[1] + np.arange(M) * h)[None, None, :], shape)
    return X, Y, Z


# ---- exact solutions ------------------------------------------------------

def exact(kind, par, shape, origin, h, t):
    a, s, OM, c, x0 = par
    X, Y, Z = coords(shape, origin, h)
    d = len(shape)
    if kind == 3:
        r2 = X ** 2
        if d >= 2:
            r2 = r2 + Y ** 2
        if d == 3:
            r2 = r2 + Z ** 2
        return np.exp(-r2 / (2 * a)) * np.exp(-1j * d * t)
    x = X
    xi = x - x0 - c * t
    if kind == 1:
        return np.sqrt((2 * OM) / s) / np.cosh(np.sqrt(OM / a) * xi) * np.exp(
            1j * ((c / (2 * a)) * x + (OM - c ** 2 / (4 * a)) * t))
    if kind == 2:
        return np.sqrt(OM / s) * np.tanh(np.sqrt(abs(OM / (2 * a))) * xi) * np.exp(
            1j * ((c / (2 * a)) * x + (OM - c ** 2 / (4 * a)) * t))
    if kind == 4:
        vc = np.sqrt(-OM * 2 * a)
        E = np.sqrt(a / -OM)
        q = np.sqrt(1 - c ** 2 / vc ** 2)
        return np.sqrt(OM / s) * (1j * (c / vc) + q * np.tanh((xi / (np.sqrt(2) * E)) * q)) * np.exp(1j * OM * t)
    raise ValueError(kind)


def exact_potential(kind, par, shape, origin, h):
    """V of the linear case: |x|^2 / a (NLSEmagic1D_FRS.m:212, NLSE3D.m:429)."""
    X, Y, Z = coords(shape, origin, h)
    r2 = X ** 2
    if len(shape) >= 2:
        r2 = r2 + Y ** 2
    if len(shape) == 3:
        r2 = r2 + Z ** 2
    return r2 / par[0]


def exact_error(U, kind, par, origin, h, t):
    """(RMS dRe, RMS dIm, RMS d|U|^2, max |dRe|, max |dIm|, max |d|U|^2|)."""
    Ur = exact(kind, par, U.shape, origin, h, t)
    dr = (Ur.real - U.real).ravel()
    di = (Ur.imag - U.imag).ravel()
    dm = (np.abs(Ur) ** 2 - np.abs(U) ** 2).ravel()
    n = dr.size
    rms = [math.sqrt(float(np.sum(v ** 2)) / n) for v in (dr, di, dm)]
    mx = [float(np.nanmax(np.abs(v))) for v in (dr, di, dm)]
    return tuple(rms + mx)


# ---- radial profile ---------------------------------------------------------

def steady_residual(f, m, OM, a, s, hr):
    """NLS_STEADY_F.m:20-55 at the interior points (f[0], f[-1] fixed)."""
    f = np.asarray(f)
    q = np.arange(1, len(f) - 1)
    r = q * hr
    rF, rB = r + hr / 2, r - hr / 2
    lap = (1 / r) * (1 / hr) * (rF * (1 / hr) * (f[q + 1] - f[q]) - rB * (1 / hr) * (f[q] - f[q - 1]))
    return -OM * f[q] - (a * m ** 2 / r ** 2) * f[q] + a * lap + s * f[q] ** 3


def ansatz(m, OM, a, s, hr, n, kind):
    r = np.arange(n) * hr
    with np.errstate(divide="ignore", invalid="ignore"):
        if kind == 2:
            B, Cc, R0 = np.sqrt((3 * OM) / s), np.sqrt((3 * OM) / (2 * a)), np.sqrt((2 * m ** 2 * a) / OM)
            f = B / np.cosh(Cc * (r - R0))
            rBC = 0.0
        else:
            if kind == 1:
                f = np.sqrt(OM / s) * np.tanh(np.sqrt(abs(OM / (2 * a))) * r)
            else:
                v = OM / s + a * m ** 2 / (s * r ** 2)
                f = np.where(v > 0, np.sqrt(np.abs(v)), 0.0)
            rBC = math.sqrt(OM / s + a * m ** 2 / (s * r[-1] ** 2))
    f = np.array(f, dtype=np.float64)
    f[0] = 0.0
    f[-1] = rBC
    return f


def profile(m, OM, a, s, hr, n, kind, optimize=True, tol=1e-14):
    f = ansatz(m, OM, a, s, hr, n, kind)
    if not optimize:
        return f
    F = steady_residual(f, m, OM, a, s, hr)
    f0 = np.linalg.norm(F)
    k = n - 2
    for _ in range(100):
        if np.linalg.norm(F) <= tol * (1 + f0):
            break
        # banded Jacobian by complex steps, three colours of columns
        band = np.zeros((3, k))
        eps = 1e-30
        for col in range(3):
            fc = f.astype(complex)
            idx = np.arange(1 + col, n - 1, 3)
            fc[idx] += 1j * eps
            dF = steady_residual(fc, m, OM, a, s, hr).imag / eps
            for j in idx - 1:  # interior column j (0-based), rows j-1, j, j+1
                for r_ in (j - 1, j, j + 1):
                    if 0 <= r_ < k:
                        band[1 + r_ - j, j] = dF[r_]
        from scipy.linalg import solve_banded
        f[1:-1] += solve_banded((1, 1), band, -F)
        F = steady_residual(f, m, OM, a, s, hr)
    if not np.linalg.norm(F) <= max(tol * (1 + f0), 1e-9 * (1 + f0)):
        raise RuntimeError("profile Newton did not converge")
    return f


# ---- vortex generators --------------------------------------------------------

def _combine_product(U, Uinit):
    ampU = np.nanmax(np.abs(U))
    ampUnew = np.nanmax(np.abs(Uinit))
    return (U / ampU) * (Uinit / ampUnew), ampUnew


def add_vortex_2d(charge, om, x0, y0, vx, vy, phase, U, h, a, s, amp, origin, opt):
    """NLS_add_vortex.m:9-158."""
    m, OM = charge, om
    X, Y, _ = coords(U.shape, origin, h)
    R = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2)
    Theta = np.arctan2(Y - y0, X - x0)
    rh = h
    rmax = R.max()
    rvec = np.arange(colon_count(0.0, rh, rmax + rh)) * rh
    if s < 0:
        OM = amp ** 2 * s
    if s < 0 and abs(m) == 1 and opt == 0:
        Uinit = (np.sqrt(OM / s) * np.tanh(np.sqrt(abs(OM / (2 * a))) * R) * np.exp(1j * (m * Theta + phase))
                 * np.exp(1j * vx * X) * np.exp(1j * vy * Y))
    else:
        kind = (1 if opt == 0 else 0) if s < 0 else 2
        f = profile(m, OM, a, s, rh, len(rvec), kind, optimize=(opt != 0), tol=1e-8)
        Finit = np.interp(R, rvec, f)
        Uinit = Finit * np.exp(1j * (m * Theta + phase)) * np.exp(1j * vx * X) * np.exp(1j * vy * Y)
    if s < 0:
        ampUinit = np.nanmax(np.abs(Uinit))
        Uinit = Uinit / ampUinit
        U = U / np.nanmax(np.abs(U))
        return (U * Uinit) * ampUinit
    return U + Uinit


def add_vortex_line(m, OM, x0, y0, z0, vx, vy, vz, phase, U, h, a, s, origin):
    """NLSE_add_vortex_line.m:10-125 (U shaped (L, N, M))."""
    X, Y, Z = coords(U.shape, origin, h)
    X2, Y2 = X[0], Y[0]
    R = np.sqrt((X2 - x0) ** 2 + (Y2 - y0) ** 2)
    Theta = np.arctan2(Y2 - y0, X2 - x0)
    rh = h
    rvec = np.arange(colon_count(0.0, rh, R.max() + rh)) * rh
    f = profile(m, OM, a, s, rh, len(rvec), 0 if s < 0 else 2, optimize=True, tol=1e-14)
    Uinit2D = np.interp(R, rvec, f) * np.exp(1j * (m * Theta + phase))
    Uinit3D = np.broadcast_to(Uinit2D[None], U.shape)
    if s < 0:
        P, ampUnew = _combine_product(U, Uinit3D)
        return ampUnew * P * np.exp(1j * (vx * X + vy * Y + vz * Z))
    return U + Uinit3D


def _mat3(A, B):
    return [[A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j] for j in range(3)] for i in range(3)]


def _rot(Rm, v):
    return [Rm[i][0] * v[0] + Rm[i][1] * v[1] + Rm[i][2] * v[2] for i in range(3)]


def add_vortex_ring(m, OM, x0, y0, z0, thetax, thetay, thetaz, vx, vy, vz, phase, U, h, a, s, origin, d,
                    ea=1.0, eb=1.0):
    """NLSE_add_vortex_ring.m:10-534 with opt_2D = 0 (U shaped (L, N, M))."""
    L, N, M = U.shape
    rh = h / 2
    cx, sx, cy, sy, cz, sz = (math.cos(thetax), math.sin(thetax), math.cos(thetay), math.sin(thetay),
                              math.cos(thetaz), math.sin(thetaz))
    Rx = [[1, 0, 0], [0, cx, -sx], [0, sx, cx]]
    Ry = [[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]]
    Rz = [[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]]
    Rm = _mat3(_mat3(Rx, Ry), Rz)
    x0R, y0R, z0R = _rot(Rm, [x0, y0, z0])
    xmin, ymin, zmin = origin
    xmax, ymax, zmax = xmin + (N - 1) * h, ymin + (M - 1) * h, zmin + (L - 1) * h
    corners = [_rot(Rm, [xx, yy, zz]) for xx in (xmin, xmax) for yy in (ymin, ymax) for zz in (zmin, zmax)]
    lo = [min(c[k] for c in corners) for k in range(3)]
    hi = [max(c[k] for c in corners) for k in range(3)]
    xvecR = lo[0] + np.arange(colon_count(lo[0], h, hi[0])) * h
    yvecR = lo[1] + np.arange(colon_count(lo[1], h, hi[1])) * h
    zvecR = lo[2] + np.arange(colon_count(lo[2], h, hi[2])) * h
    rxy_max = max(math.sqrt((x0R - xx) ** 2 + (y0R - yy) ** 2) for xx in (lo[0], hi[0]) for yy in (lo[1], hi[1]))
    rxy_res = int(math.ceil(rxy_max / h)) + 2
    rxy_vec = np.arange(rxy_res) * h
    RXY, Z2D = rxy_vec[:, None], zvecR[None, :]
    RXZ = np.sqrt((Z2D - z0R) ** 2 + (RXY - d) ** 2)
    Phi = np.arctan2(Z2D - z0R + 0 * RXY, RXY - d + 0 * Z2D)
    Phi2 = np.arctan2(Z2D - z0R + 0 * RXY, RXY + d + 0 * Z2D)
    rxz_res = int(math.ceil(2 * RXZ.max() / rh)) + 2
    rxz_vec = np.arange(rxz_res) * rh
    f = profile(m, OM, a, s, rh, rxz_res, 0 if s < 0 else 2, optimize=True, tol=1e-12)
    Uinit2D = np.interp(RXZ, rxz_vec, f) * np.exp(1j * ((m * Phi - m * Phi2) + phase))  # (rxy, z)
    # the ring on the rotated lattice, MATLAB order (y, x, z)
    YR, XR = yvecR[:, None], xvecR[None, :]
    RXYreal = np.sqrt(((XR - x0R) / ea) ** 2 + ((YR - y0R) / eb) ** 2)
    q = np.clip(np.floor(RXYreal / h).astype(int), 0, rxy_res - 2)
    r0 = q * h
    w = (RXYreal - r0) / ((q + 1) * h - r0)
    A, B = Uinit2D[q], Uinit2D[q + 1]  # (ny, nx, nz)
    U3R = A + w[..., None] * (B - A)
    U3R = U3R * np.exp(1j * (vx * XR[..., None] + vy * YR[..., None] + vz * zvecR[None, None, :]))
    q2 = math.pi / 2
    if all(t in (0.0, q2) for t in (thetax, thetay, thetaz)):
        if thetax == q2:
            U3R = np.flip(np.transpose(U3R, (2, 1, 0)), axis=2)
        if thetay == q2:
            U3R = np.flip(np.transpose(U3R, (0, 2, 1)), axis=1)
        if thetaz == q2:
            U3R = np.flip(np.transpose(U3R, (1, 0, 2)), axis=0)
        U3 = U3R
    else:
        sizeR = U3R.shape
        X, Y, Z = coords(U.shape, origin, h)
        Xm, Ym, Zm = [np.transpose(np.asarray(A_), (2, 1, 0)) for A_ in (X, Y, Z)]  # MATLAB order (y, x, z)
        xr = Rm[0][0] * Xm + Rm[0][1] * Ym + Rm[0][2] * Zm
        yr = Rm[1][0] * Xm + Rm[1][1] * Ym + Rm[1][2] * Zm
        zr = Rm[2][0] * Xm + Rm[2][1] * Ym + Rm[2][2] * Zm
        ir = np.clip(np.ceil((yr - lo[1]) / h + 1).astype(int), 1, sizeR[0]) - 1
        jr = np.clip(np.ceil((xr - lo[0]) / h + 1).astype(int), 1, sizeR[1]) - 1
        kr = np.clip(np.ceil((zr - lo[2]) / h + 1).astype(int), 1, sizeR[2]) - 1
        U3 = U3R[ir, jr, kr]
    Uinit3D = np.transpose(U3, (2, 1, 0))  # to (L, N, M)
    if Uinit3D.shape != U.shape:
        raise ValueError("ring lattice does not match the grid")
    if s < 0:
        P, ampUnew = _combine_product(U, Uinit3D)
        return ampUnew * P
    return U + Uinit3D
