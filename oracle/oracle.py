"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (this project's C restatement of the reference
CPU integrators) and, when present, ``oracle/_ref/libnlse_ref{1,2}d.so``
(the reference's own ``src/NLSEmagic1D/NLSEmagic1D_CPU.c`` and
``src/NLSEmagic2D/NLSEmagic2D_CPU.c`` compiled by ``oracle/Makefile``).
Only tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module; the product path never does.

Arrays follow the reference layout: split real/imaginary planes, 2D arrays of
shape (N, M) with M contiguous (MATLAB's M x N column-major), 3D arrays of
shape (L, N, M).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_lib = None
_ref = {}


def lib():
    """The restatement library (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        _lib = C.CDLL(path)
        for dim, extra in ((1, 1), (2, 2), (3, 3)):
            fd = getattr(_lib, f"oracle_nlse{dim}d")
            fd.restype = None
            fd.argtypes = [_D] * 5 + [C.c_double] * 3 + [C.c_int] * 2 + [C.c_double] + [C.c_int] * (extra + 1)
            ff = getattr(_lib, f"oracle_nlse{dim}d_f")
            ff.restype = None
            out = _F if dim == 1 else _D
            ff.argtypes = [_F] * 3 + [out] * 2 + [C.c_float] * 3 + [C.c_int] * 2 + [C.c_float] + [C.c_int] * (extra + 1)
            fr = getattr(_lib, f"oracle_rhs{dim}d")
            fr.restype = None
            fr.argtypes = [_D] * 5 + [C.c_double] * 3 + [C.c_int] * (extra + 2)
        _lib.oracle_sum_abs2.restype = C.c_double
        _lib.oracle_sum_abs2.argtypes = [_D, _D, C.c_long]
        _lib.oracle_vorticity3d.restype = None
        _lib.oracle_vorticity3d.argtypes = [_D, _D, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int] + [_D] * 4
        _lib.oracle_vorticity3d_fill.restype = None
        _lib.oracle_vorticity3d_fill.argtypes = ([_D, _D, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                                  C.c_double] + [_D] * 4)
    return _lib


def ref(dim):
    """The reference's own CPU integrator for dim 1 or 2, or None if unbuilt."""
    if dim not in _ref:
        L = _load(os.path.join(HERE, "_ref", f"libnlse_ref{dim}d.so"))
        if L is not None:
            extra = dim
            fd = getattr(L, f"NLSE{dim}D_TAKE_STEPS")
            fd.restype = None
            fd.argtypes = [_D] * 5 + [C.c_double] * 3 + [C.c_int] * 2 + [C.c_double] + [C.c_int] * (extra + 1)
            ff = getattr(L, f"NLSE{dim}D_TAKE_STEPS_F")
            ff.restype = None
            out = _F if dim == 1 else _D
            ff.argtypes = [_F] * 3 + [out] * 2 + [C.c_float] * 3 + [C.c_int] * 2 + [C.c_float] + [C.c_int] * (extra + 1)
        _ref[dim] = L
    return _ref[dim]


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _cplx(re_, im_):
    """complex128 array from two real planes (exact, no inf*1j artefacts)."""
    out = np.empty(re_.shape, np.complex128)
    out.real = re_
    out.imag = im_
    return out


def _dims(shape):
    """(N, M, L) integer tail of the reference C ABI for an array shape."""
    if len(shape) == 1:
        return (shape[0],)
    if len(shape) == 2:
        n, m = shape
        return (n, m)
    l, n, m = shape
    return (n, m, l)


def take_steps(U, V, s, a, h2, BC, chunk_size, k, method, precision=2, use_ref=False):
    """Run one chunk like the reference's ``NLSE{1,2,3}D_TAKE_STEPS[_F]``.

    U complex array (1D, (N,M) or (L,N,M)); V real of the same shape.
    precision 2 = double, 1 = single (inputs rounded to float first, as the
    gateway does: src/NLSEmagic2D/NLSEmagic2D_Take_Steps_CPU.c:48-63).
    Returns a new complex128 array (complex64 values widened for precision 1).
    ``use_ref`` calls the compiled reference instead of the restatement; the
    reference overwrites its input, so a private copy is passed.
    """
    U = np.asarray(U)
    dim = U.ndim
    shape = U.shape
    tail = _dims(shape)
    L = ref(dim) if use_ref else lib()
    if L is None:
        raise RuntimeError(f"reference library for {dim}D not built")
    name = (f"NLSE{dim}D_TAKE_STEPS" if use_ref else f"oracle_nlse{dim}d")
    if precision == 2:
        ur = np.ascontiguousarray(U.real, dtype=np.float64).copy()
        ui = np.ascontiguousarray(U.imag, dtype=np.float64).copy()
        v = np.ascontiguousarray(V, dtype=np.float64).copy()
        outr = np.empty(shape, np.float64)
        outi = np.empty(shape, np.float64)
        getattr(L, name)(_p(ur, C.c_double), _p(ui, C.c_double), _p(v, C.c_double),
                         _p(outr, C.c_double), _p(outi, C.c_double),
                         float(s), float(a), float(h2), int(BC), int(chunk_size), float(k),
                         *[int(t) for t in tail], int(method))
        return _cplx(outr, outi)
    ur = np.ascontiguousarray(U.real, dtype=np.float32).copy()
    ui = np.ascontiguousarray(U.imag, dtype=np.float32).copy()
    v = np.ascontiguousarray(V, dtype=np.float32).copy()
    odt, ot = (np.float32, C.c_float) if dim == 1 else (np.float64, C.c_double)
    outr = np.empty(shape, odt)
    outi = np.empty(shape, odt)
    f32 = lambda x: float(np.float32(x))
    getattr(L, name + "_F" if use_ref else name + "_f")(
        _p(ur, C.c_float), _p(ui, C.c_float), _p(v, C.c_float), _p(outr, ot), _p(outi, ot),
        f32(s), f32(a), f32(h2), int(BC), int(chunk_size), f32(k),
        *[int(t) for t in tail], int(method))
    return _cplx(outr, outi)


def rhs(U, V, s, a, h2, BC, method):
    """One evaluation of F(U) = U_t with its boundary values (double)."""
    U = np.asarray(U)
    dim = U.ndim
    ur = np.ascontiguousarray(U.real, dtype=np.float64)
    ui = np.ascontiguousarray(U.imag, dtype=np.float64)
    v = np.ascontiguousarray(V, dtype=np.float64)
    fr = np.empty(U.shape)
    fi = np.empty(U.shape)
    getattr(lib(), f"oracle_rhs{dim}d")(
        _p(ur, C.c_double), _p(ui, C.c_double), _p(v, C.c_double), _p(fr, C.c_double),
        _p(fi, C.c_double), float(s), float(a), float(h2), int(BC),
        *[int(t) for t in _dims(U.shape)], int(method))
    return _cplx(fr, fi)


def vorticity3d(U, lh2, veps, edge_g=None):
    """3D vorticity of a state (rmc_vorticity3D.c:24-354): returns
    (wx, wy, wz, W) as arrays of U's shape (L, N, M); lh2 = 1/(2h).  The
    gradient the reference leaves unset on edges and corners is 0, or
    edge_g when given."""
    U = np.asarray(U)
    assert U.ndim == 3
    L, N, M = U.shape
    ur = np.ascontiguousarray(U.real, dtype=np.float64)
    ui = np.ascontiguousarray(U.imag, dtype=np.float64)
    out = [np.empty(U.shape) for _ in range(4)]
    args = (_p(ur, C.c_double), _p(ui, C.c_double), float(lh2), float(veps), M, N, L)
    if edge_g is None:
        lib().oracle_vorticity3d(*args, *[_p(o, C.c_double) for o in out])
    else:
        lib().oracle_vorticity3d_fill(*args, float(edge_g), *[_p(o, C.c_double) for o in out])
    return tuple(out)


def vorticity_defined_mask(shape):
    """Points where the reference's vorticity is defined.  The reference
    leaves the gradient, and so the velocity, unset on edges and corners
    (rmc_vorticity3D.c:73-86,113-184); the curl at a point is defined when
    none of the velocities its stencil reads sits on an edge or corner:
    every interior point; on an x-face (i = 0 or L-1) the points with j and
    k at least 2 from the boundary (central y and z differences); on a y- or
    z-face the points whose central tangential index is at least 2 from the
    boundary and whose forward x-difference does not reach i = L-1
    (1 <= i <= L-3).  Edges and corners themselves are never defined."""
    L, N, M = shape
    i, j, k = np.meshgrid(np.arange(L), np.arange(N), np.arange(M), indexing="ij")
    bi, bj, bk = (i == 0) | (i == L - 1), (j == 0) | (j == N - 1), (k == 0) | (k == M - 1)
    nbound = bi.astype(int) + bj + bk
    far = lambda a, n: (a >= 2) & (a <= n - 3)
    xface = bi & (nbound == 1) & far(j, N) & far(k, M)
    yface = bj & (nbound == 1) & far(k, M) & (i <= L - 3)
    zface = bk & (nbound == 1) & far(j, N) & (i <= L - 3)
    return (nbound == 0) | xface | yface | zface
