"""nlsemagic_amd -- MI355X-native RK4 integrators for the NLSE (NLSEmagic's hot path).

Python mirror of the reference's MEX interface, calling the HIP library
``libnlsemagic_hip.so`` through its C ABI (``include/nlsemagic_hip.h``).

Array convention: a MATLAB ``M x N`` array (column-major, M contiguous) is
the numpy array of shape ``(N, M)`` in C order; ``M x N x L`` is ``(L, N, M)``.
1D arrays are plain vectors.  ``h2`` is h^2, as in the reference.

Reference entry points mirrored here:
  v1 MEX, 8 arguments ``U = NAME(U,V,s,a,h2,BC,chunk_size,k)``
    (src/NLSEmagic3D/README_B.txt:11-26; called at NLSEmagic2D.m:256-279,
    NLSEmagic3D.m:303-325): ``NLSE{1,2,3}D_TAKE_STEPS_{CD,2SHOC}_HIP_{D,F}``
    replace the ``..._CUDA_{D,F}`` files.
  v2 gateways, 10 arguments ``U = NAME(U,V,s,a,h2,BC,chunk_size,k,precision,method)``
    (src/NLSEmagic1D/NLSEmagic1D_Take_Steps_GPU.c:5-90,
    src/NLSEmagic2D/NLSEmagic2D_Take_Steps_GPU.c:5-78):
    ``NLSEmagic{1,2,3}D_Take_Steps_GPU``.
The result is a new complex128 array; the input is never modified.
There is no CPU fallback: without the library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnlsemagic_hip.so")

__all__ = [
    "lib", "Solver", "Batch", "Group", "HipError", "DeviceArray", "rmc_vorticity3D", "vortex_profile",
    "NLS_add_vortex", "NLSE_add_vortex_line", "NLSE_add_vortex_ring",
    "NLSEmagic1D_Take_Steps_GPU", "NLSEmagic2D_Take_Steps_GPU", "NLSEmagic3D_Take_Steps_GPU",
]

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_lib = None


class HipError(RuntimeError):
    """A nonzero status from the C ABI (message from nlse_last_error())."""


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m nlsemagic_amd.build`")
    L = C.CDLL(LIB_PATH)
    for dim in (1, 2, 3):
        nd = dim
        fd = getattr(L, f"NLSE{dim}D_TAKE_STEPS_HIP")
        fd.restype = C.c_int
        fd.argtypes = [_D] * 5 + [C.c_double] * 3 + [C.c_int] * 2 + [C.c_double] + [C.c_int] * (nd + 1)
        ff = getattr(L, f"NLSE{dim}D_TAKE_STEPS_HIP_F")
        ff.restype = C.c_int
        ff.argtypes = [_F] * 5 + [C.c_float] * 3 + [C.c_int] * 2 + [C.c_float] + [C.c_int] * (nd + 1)
    L.nlse_last_error.restype = C.c_char_p
    L.nlse_device_count.restype = C.c_int
    L.nlse_create.restype = C.c_int
    L.nlse_create.argtypes = [C.POINTER(C.c_void_p)] + [C.c_int] * 7 + [C.c_double] * 4
    L.nlse_destroy.restype = None
    L.nlse_destroy.argtypes = [C.c_void_p]
    L.nlse_set_state.restype = C.c_int
    L.nlse_set_state.argtypes = [C.c_void_p, _D, _D, _D]
    L.nlse_set_state_f.restype = C.c_int
    L.nlse_set_state_f.argtypes = [C.c_void_p, _F, _F, _F]
    L.nlse_get_state.restype = C.c_int
    L.nlse_get_state.argtypes = [C.c_void_p, _D, _D]
    L.nlse_get_state_f.restype = C.c_int
    L.nlse_get_state_f.argtypes = [C.c_void_p, _F, _F]
    L.nlse_step.restype = C.c_int
    L.nlse_step.argtypes = [C.c_void_p, C.c_int]
    L.nlse_step_timed.restype = C.c_int
    L.nlse_step_timed.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L.nlse_profile_launches.restype = C.c_int
    L.nlse_profile_launches.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.nlse_autotune.restype = C.c_int
    L.nlse_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_double)]
    L.nlse_set_tile.restype = C.c_int
    L.nlse_set_tile.argtypes = [C.c_void_p, C.c_int]
    L.nlse_set_fusion.restype = C.c_int
    L.nlse_set_fusion.argtypes = [C.c_void_p, C.c_int]
    L.nlse_set_kernel.restype = C.c_int
    L.nlse_set_kernel.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.nlse_get_config.restype = C.c_int
    L.nlse_get_config.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.nlse_sync.restype = C.c_int
    L.nlse_sync.argtypes = [C.c_void_p]
    L.nlse_query.restype = C.c_int
    L.nlse_query.argtypes = [C.c_void_p]
    L.nlse_set_graph.restype = C.c_int
    L.nlse_set_graph.argtypes = [C.c_void_p, C.c_int]
    _I = C.POINTER(C.c_int)
    L.nlse_batch_create.restype = C.c_int
    L.nlse_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, _I, _I, _D, _D, _D, _D, C.c_int, C.c_int,
                                    C.c_int]
    L.nlse_batch_destroy.restype = None
    L.nlse_batch_destroy.argtypes = [C.c_void_p]
    L.nlse_batch_points.restype = C.c_longlong
    L.nlse_batch_points.argtypes = [C.c_void_p]
    L.nlse_batch_set_state.restype = C.c_int
    L.nlse_batch_set_state.argtypes = [C.c_void_p, C.c_int, _D, _D, _D]
    L.nlse_batch_get_state.restype = C.c_int
    L.nlse_batch_get_state.argtypes = [C.c_void_p, C.c_int, _D, _D]
    L.nlse_batch_step.restype = C.c_int
    L.nlse_batch_step.argtypes = [C.c_void_p, C.c_int]
    L.nlse_batch_step_timed.restype = C.c_int
    L.nlse_batch_step_timed.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L.nlse_batch_sync.restype = C.c_int
    L.nlse_batch_sync.argtypes = [C.c_void_p]
    L.nlse_batch_configure.restype = C.c_int
    L.nlse_batch_configure.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    L.nlse_batch_set_kernel.restype = C.c_int
    L.nlse_batch_set_kernel.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.nlse_batch_get_config.restype = C.c_int
    L.nlse_batch_get_config.argtypes = [C.c_void_p, _I, _I, _I, _I]
    L.nlse_batch_autotune.restype = C.c_int
    L.nlse_batch_autotune.argtypes = [C.c_void_p, C.c_int, _I, _I, _I, C.POINTER(C.c_double)]
    L.nlse_kfull.restype = C.c_int
    L.nlse_kfull.argtypes = [C.c_void_p, _D]
    L.nlse_set_step.restype = C.c_int
    L.nlse_set_step.argtypes = [C.c_void_p, C.c_double]
    L.nlse_diagnostics.restype = C.c_int
    L.nlse_diagnostics.argtypes = [C.c_void_p, _D]
    L.nlse_energy.restype = C.c_int
    L.nlse_energy.argtypes = [C.c_void_p, _D]
    L.nlse_device_allocations.restype = C.c_longlong
    L.nlse_device_allocations.argtypes = []
    L.nlse_group_create.restype = C.c_int
    L.nlse_group_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int]
    L.nlse_group_step.restype = C.c_int
    L.nlse_group_step.argtypes = [C.c_void_p, C.c_int]
    L.nlse_group_done.restype = C.c_int
    L.nlse_group_done.argtypes = [C.c_void_p]
    L.nlse_group_sync.restype = C.c_int
    L.nlse_group_sync.argtypes = [C.c_void_p]
    L.nlse_wait.restype = C.c_int
    L.nlse_wait.argtypes = [C.c_void_p]
    L.nlse_group_destroy.restype = None
    L.nlse_group_destroy.argtypes = [C.c_void_p]
    L.nlse_diagnostics_noalloc.restype = C.c_int
    L.nlse_diagnostics_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p, C.c_void_p]
    L.nlse_batch_diagnostics.restype = C.c_int
    L.nlse_batch_diagnostics.argtypes = [C.c_void_p, _D]
    L.NLSE3D_VORTICITY_HIP.restype = C.c_int
    L.NLSE3D_VORTICITY_HIP.argtypes = [_D, _D, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int] + [_D] * 4
    L.nlse_vorticity.restype = C.c_int
    L.nlse_vorticity.argtypes = [C.c_void_p, C.c_double, C.c_double] + [_D] * 4
    L.nlse_vorticity_timed.restype = C.c_int
    L.nlse_vorticity_timed.argtypes = ([C.c_void_p, C.c_double, C.c_double] + [C.c_void_p] * 4 +
                                       [C.c_int, C.POINTER(C.c_double)])
    L.nlse_vorticity_device.restype = C.c_int
    L.nlse_vorticity_device.argtypes = [C.c_void_p, C.c_double, C.c_double] + [C.c_void_p] * 4
    L.nlse_set_exact.restype = C.c_int
    L.nlse_set_exact.argtypes = [C.c_void_p, C.c_int, _D] + [C.c_double] * 5
    L.nlse_fill_state.restype = C.c_int
    L.nlse_fill_state.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.nlse_exact_error.restype = C.c_int
    L.nlse_exact_error.argtypes = [C.c_void_p, C.c_int, _D] + [C.c_double] * 5 + [_D]
    L.nlse_add_vortex_2d.restype = C.c_int
    L.nlse_add_vortex_2d.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 12 + [C.c_int]
    L.nlse_add_vortex_line.restype = C.c_int
    L.nlse_add_vortex_line.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 14
    L.nlse_add_vortex_ring.restype = C.c_int
    L.nlse_add_vortex_ring.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 18 + [C.c_int, C.c_double, C.c_double]
    L.nlse_vortex_profile.restype = C.c_int
    L.nlse_vortex_profile.argtypes = ([C.c_int] + [C.c_double] * 4 + [C.c_int] * 3 + [C.c_double, _D])
    L.nlse_copy_probe.restype = C.c_int
    L.nlse_copy_probe.argtypes = [C.c_longlong, C.c_int, C.POINTER(C.c_double)]
    L.nlse_footprint_bytes.restype = C.c_longlong
    L.nlse_footprint_bytes.argtypes = [C.c_int] * 6
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise HipError(f"nlsemagic_hip error {rc}: {lib().nlse_last_error().decode()}")


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _tail(shape):
    """(N[, M[, L]]) integers of the C ABI for an array shape."""
    if len(shape) == 1:
        return (shape[0],)
    if len(shape) == 2:
        return (shape[0], shape[1])
    if len(shape) == 3:
        L, N, M = shape
        return (N, M, L)
    raise ValueError("U must be 1-, 2- or 3-dimensional")


def _cplx(re_, im_):
    out = np.empty(re_.shape, np.complex128)
    out.real = re_
    out.imag = im_
    return out


def take_steps(U, V, s, a, h2, BC, chunk_size, k, precision=2, method=1):
    """One chunk through NLSE{1,2,3}D_TAKE_STEPS_HIP[_F].

    precision 2: double throughout.  precision 1: U, V and the scalars are
    rounded to float and the integrator runs in single precision, as the v2
    gateway does (NLSEmagic1D_Take_Steps_GPU.c:50-79); the result is widened
    back to double.
    """
    U = np.asarray(U)
    V = np.asarray(V, dtype=np.float64)
    if V.shape != U.shape:
        raise ValueError(f"V shape {V.shape} differs from U shape {U.shape}")
    dim = U.ndim
    tail = _tail(U.shape)
    L = lib()
    if precision == 2:
        ur = np.ascontiguousarray(U.real, dtype=np.float64)
        ui = np.ascontiguousarray(U.imag if np.iscomplexobj(U) else np.zeros(U.shape), dtype=np.float64)
        v = np.ascontiguousarray(V)
        outr = np.empty(U.shape)
        outi = np.empty(U.shape)
        _check(getattr(L, f"NLSE{dim}D_TAKE_STEPS_HIP")(
            _p(ur, C.c_double), _p(ui, C.c_double), _p(v, C.c_double), _p(outr, C.c_double),
            _p(outi, C.c_double), float(s), float(a), float(h2), int(BC), int(chunk_size), float(k),
            *tail, int(method)))
        return _cplx(outr, outi)
    if precision != 1:
        raise ValueError("precision must be 1 (single) or 2 (double)")
    ur = np.ascontiguousarray(U.real, dtype=np.float32)
    ui = np.ascontiguousarray(U.imag if np.iscomplexobj(U) else np.zeros(U.shape), dtype=np.float32)
    v = np.ascontiguousarray(V, dtype=np.float32)
    outr = np.empty(U.shape, np.float32)
    outi = np.empty(U.shape, np.float32)
    f32 = lambda x: float(np.float32(x))
    _check(getattr(L, f"NLSE{dim}D_TAKE_STEPS_HIP_F")(
        _p(ur, C.c_float), _p(ui, C.c_float), _p(v, C.c_float), _p(outr, C.c_float), _p(outi, C.c_float),
        f32(s), f32(a), f32(h2), int(BC), int(chunk_size), f32(k), *tail, int(method)))
    return _cplx(outr.astype(np.float64), outi.astype(np.float64))


def _v2(dim):
    def gateway(U, V, s, a, h2, BC, chunk_size, k, precision, method):
        U = np.asarray(U)
        if U.ndim != dim:
            raise ValueError(f"expected a {dim}D array, got shape {U.shape}")
        return take_steps(U, V, s, a, h2, BC, chunk_size, k, precision=precision, method=method)

    gateway.__name__ = f"NLSEmagic{dim}D_Take_Steps_GPU"
    gateway.__doc__ = (f"v2 gateway U = NLSEmagic{dim}D_Take_Steps_GPU(U,V,s,a,h2,BC,chunk_size,k,precision,method)"
                       " (src/NLSEmagic1D/NLSEmagic1D_Take_Steps_GPU.c:5-90).")
    return gateway


def rmc_vorticity3D(U, lh2, veps):
    """``[wx, wy, wz, W] = rmc_vorticity3D(U, 1/(2h), veps)`` on the GPU
    (src/NLSEmagic3D/rmc_vorticity3D.c:24-354; called at NLSE3D.m:594,1109).

    U: 3D array of shape (L, N, M) (a MATLAB M x N x L array), real or
    complex.  Returns four float64 arrays of U's shape: the curl of the
    velocity Im(conj(U) grad U) / (2|U|^2 + veps) and its norm, evaluated in
    double precision with the reference's difference rules.  The reference
    leaves the gradient unset on edges and corners; here it is 0 there, so
    the result equals the reference's wherever its stencil avoids edges and
    corners, and edges and corners are 0.
    """
    U = np.asarray(U)
    if U.ndim != 3:
        raise ValueError(f"expected a 3D array, got shape {U.shape}")
    L_, N, M = U.shape
    ur = np.ascontiguousarray(U.real, dtype=np.float64)
    ui = np.ascontiguousarray(U.imag, dtype=np.float64) if np.iscomplexobj(U) else None
    out = [np.empty(U.shape) for _ in range(4)]
    _check(lib().NLSE3D_VORTICITY_HIP(_p(ur, C.c_double), _p(ui, C.c_double) if ui is not None else None,
                                      float(lh2), float(veps), M, N, L_, *[_p(o, C.c_double) for o in out]))
    return tuple(out)


NLSEmagic1D_Take_Steps_GPU = _v2(1)
NLSEmagic2D_Take_Steps_GPU = _v2(2)
NLSEmagic3D_Take_Steps_GPU = _v2(3)


def _v1(dim, method, precision):
    def mex(U, V, s, a, h2, BC, chunk_size, k):
        U = np.asarray(U)
        if U.ndim != dim:
            raise ValueError(f"expected a {dim}D array, got shape {U.shape}")
        return take_steps(U, V, s, a, h2, BC, chunk_size, k, precision=precision, method=method)

    meth = "CD" if method == 1 else "2SHOC"
    name = f"NLSE{dim}D_TAKE_STEPS_{meth}_HIP_{'D' if precision == 2 else 'F'}"
    mex.__name__ = name
    mex.__doc__ = (f"v1 MEX {name}(U,V,s,a,h2,BC,chunk_size,k); replaces "
                   f"NLSE{dim}D_TAKE_STEPS_{meth}_CUDA_{'D' if precision == 2 else 'F'}.")
    return name, mex


for _dim in (1, 2, 3):
    for _m in (1, 2):
        for _pr in (2, 1):
            _n, _f = _v1(_dim, _m, _pr)
            globals()[_n] = _f
            __all__.append(_n)


class Solver:
    """Device-resident RK4 integrator (the nlse_create/... C ABI).

    Keeps U, V and the RK4 work arrays in HBM across chunks; ``step`` is
    asynchronous, ``state`` downloads (and synchronises).
    """

    def __init__(self, shape, method, precision, BC, s, a, h2, k):
        dim = len(shape)
        tail = list(_tail(tuple(shape))) + [1, 1]
        N, M, L_ = tail[0], tail[1], tail[2]
        self.shape = tuple(shape)
        self.dim = dim
        self.precision = precision
        self._h = C.c_void_p()
        self._lib = lib()
        _check(self._lib.nlse_create(C.byref(self._h), dim, N, M, L_, int(method), int(precision), int(BC),
                                     float(s), float(a), float(h2), float(k)))

    def close(self):
        if self._h:
            self._lib.nlse_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, U, V=None):
        U = np.asarray(U)
        if U.shape != self.shape:
            raise ValueError("shape mismatch")
        ur = np.ascontiguousarray(U.real, dtype=np.float64)
        ui = np.ascontiguousarray(U.imag if np.iscomplexobj(U) else np.zeros(U.shape), dtype=np.float64)
        vp = None
        if V is not None:
            v = np.ascontiguousarray(V, dtype=np.float64)
            if v.shape != self.shape:
                raise ValueError("shape mismatch")
            vp = _p(v, C.c_double)
        _check(self._lib.nlse_set_state(self._h, _p(ur, C.c_double), _p(ui, C.c_double), vp))

    def state(self):
        outr = np.empty(self.shape)
        outi = np.empty(self.shape)
        _check(self._lib.nlse_get_state(self._h, _p(outr, C.c_double), _p(outi, C.c_double)))
        return _cplx(outr, outi)

    def step(self, steps):
        _check(self._lib.nlse_step(self._h, int(steps)))

    def step_timed(self, steps):
        ms = C.c_double()
        _check(self._lib.nlse_step_timed(self._h, int(steps), C.byref(ms)))
        return ms.value

    def diagnostics(self):
        """(sum |U|^2, max Re U, non-finite count), reduced on the GPU in FP64."""
        out = np.empty(3)
        _check(self._lib.nlse_diagnostics(self._h, _p(out, C.c_double)))
        return float(out[0]), float(out[1]), int(out[2])

    def energy(self):
        """(E, gradient part, potential + nonlinear part, mass) of the resident
        state, reduced on the GPU in FP64 (see nlse_energy): E = h^d [a/h^2
        sum over neighbour pairs |U_p - U_q|^2 + sum (V|U|^2 - (s/2)|U|^4)],
        the invariant of the semi-discrete CD flow with Dirichlet boundaries."""
        out = np.empty(4)
        _check(self._lib.nlse_energy(self._h, _p(out, C.c_double)))
        return tuple(float(x) for x in out)

    def kfull(self):
        """(kfull, klin, k): the reference drivers' full RK4 stability bound of
        the resident state, the linear bound, and the step they would take
        (NLSE2D.m:427-477; see nlse_kfull).  MSD evaluates the RHS once into
        the solver's stage buffers, which hold nothing between steps."""
        out = np.empty(3)
        _check(self._lib.nlse_kfull(self._h, _p(out, C.c_double)))
        return float(out[0]), float(out[1]), float(out[2])

    def set_step(self, k):
        """Change the time step k (e.g. to the k of :meth:`kfull`)."""
        _check(self._lib.nlse_set_step(self._h, float(k)))

    def vorticity(self, lh2, veps):
        """(wx, wy, wz, W) of the resident 3D state, as :func:`rmc_vorticity3D`
        (double precision whatever the solver's precision)."""
        out = [np.empty(self.shape) for _ in range(4)]
        _check(self._lib.nlse_vorticity(self._h, float(lh2), float(veps), *[_p(o, C.c_double) for o in out]))
        return tuple(out)

    def vorticity_device(self, lh2, veps, wx=0, wy=0, wz=0, W=0):
        """Enqueue the vorticity into device arrays given as raw addresses
        (n doubles each, e.g. DeviceArray.address; 0 to skip one) on the
        solver's stream."""
        _check(self._lib.nlse_vorticity_device(self._h, float(lh2), float(veps), *[C.c_void_p(int(x) or None)
                                                                                    for x in (wx, wy, wz, W)]))

    # ---- closed-form fields and vortex initial conditions (grid origin
    # (xmin, ymin, zmin) and spacing h; see include/nlsemagic_hip.h)
    def set_exact(self, kind, h, origin, t=0.0, a=1.0, s=0.0, OM=0.0, c=0.0, x0=0.0):
        """Write the exact solution `kind` (1 bright, 2 dark, 3 linear ground
        state, 4 grey; NLSEmagic1D_FRS.m:201-225) at time t into the state."""
        par = np.array([a, s, OM, c, x0], dtype=np.float64)
        o = tuple(origin) + (0.0,) * (3 - len(tuple(origin)))
        _check(self._lib.nlse_set_exact(self._h, int(kind), _p(par, C.c_double), float(t), *map(float, o), float(h)))

    def fill(self, value):
        """U = value at every point (V unchanged)."""
        value = complex(value)
        _check(self._lib.nlse_fill_state(self._h, value.real, value.imag))

    def exact_error(self, kind, h, origin, t, a=1.0, s=0.0, OM=0.0, c=0.0, x0=0.0):
        """(rms_re, rms_im, rms_mod2, max_re, max_im, max_mod2) of the state
        against the exact solution at time t (NLSE2D.m:1011-1013,
        NLSEmagic1D_FRS.m:591-593)."""
        par = np.array([a, s, OM, c, x0], dtype=np.float64)
        o = tuple(origin) + (0.0,) * (3 - len(tuple(origin)))
        out = np.empty(6)
        _check(self._lib.nlse_exact_error(self._h, int(kind), _p(par, C.c_double), float(t), *map(float, o), float(h),
                                          _p(out, C.c_double)))
        return tuple(float(x) for x in out)

    def add_vortex_2d(self, charge, om, x0, y0, vx, vy, phase, h, a, s, amp, origin, opt=1):
        """NLS_add_vortex (NLS_add_vortex.m:9-158) on the resident 2D state."""
        _check(self._lib.nlse_add_vortex_2d(self._h, int(charge), *map(float, (om, x0, y0, vx, vy, phase, h, a, s, amp,
                                                                             origin[0], origin[1])), int(opt)))

    def add_vortex_line(self, m, OM, x0, y0, z0, vx, vy, vz, phase, h, a, s, origin):
        """NLSE_add_vortex_line (NLSE_add_vortex_line.m:10-125) on the resident 3D state."""
        _check(self._lib.nlse_add_vortex_line(self._h, int(m), *map(float, (OM, x0, y0, z0, vx, vy, vz, phase, h, a, s,
                                                                           *origin))))

    def add_vortex_ring(self, m, OM, x0, y0, z0, thetax, thetay, thetaz, vx, vy, vz, phase, h, a, s, origin, d,
                        opt_2D=0, ea=1.0, eb=1.0):
        """NLSE_add_vortex_ring (NLSE_add_vortex_ring.m:10-534, opt_2D = 0) on
        the resident 3D state."""
        _check(self._lib.nlse_add_vortex_ring(self._h, int(m), *map(float, (OM, x0, y0, z0, thetax, thetay, thetaz, vx,
                                                                           vy, vz, phase, h, a, s, *origin, d)),
                                              int(opt_2D), float(ea), float(eb)))

    def vorticity_timed(self, lh2, veps, reps, wx=0, wy=0, wz=0, W=0):
        """Mean device ms of `reps` vorticity_device calls (HIP events on
        the solver's stream)."""
        ms = C.c_double()
        _check(self._lib.nlse_vorticity_timed(self._h, float(lh2), float(veps),
                                              *[C.c_void_p(int(x) or None) for x in (wx, wy, wz, W)],
                                              int(reps), C.byref(ms)))
        return ms.value

    def profile_launches(self, steps):
        """Mean device ms of each kernel launch of an RK4 step (HIP events)."""
        out = (C.c_double * 4)()
        n = C.c_int()
        _check(self._lib.nlse_profile_launches(self._h, int(steps), out, C.byref(n)))
        return list(out)[: n.value]

    def autotune(self, steps=5):
        """Pick the fastest kernel configuration (stages per launch, kernel
        kind, rows) for this problem.

        Returns (stages_per_launch, rows, ms_per_step); config() gives the
        kind as well.  The state is unchanged."""
        f, t, ms = C.c_int(), C.c_int(), C.c_double()
        _check(self._lib.nlse_autotune(self._h, int(steps), C.byref(f), C.byref(t), C.byref(ms)))
        return f.value, t.value, ms.value

    def set_tile(self, tile_rows):
        """2D: tile height of the stage kernels (8, 16 or 32 rows)."""
        _check(self._lib.nlse_set_tile(self._h, int(tile_rows)))

    def set_fusion(self, stages_per_launch):
        """RK4 stages per kernel launch: 2D 1, 2 or 4; 3D 1 or 2 (not MSD)."""
        _check(self._lib.nlse_set_fusion(self._h, int(stages_per_launch)))

    def set_kernel(self, kind, rows):
        """2D stage kernel family and its rows per tile: 0 LDS tiles (8/16/32),
        1 row-sweep tiles (8/12/16), 2 register strips (3..256-row segments)."""
        _check(self._lib.nlse_set_kernel(self._h, int(kind), int(rows)))

    def config(self):
        """(stages_per_launch, kernel_kind, rows) in use."""
        f, t, s = C.c_int(), C.c_int(), C.c_int()
        _check(self._lib.nlse_get_config(self._h, C.byref(f), C.byref(t), C.byref(s)))
        return f.value, t.value, s.value

    def sync(self):
        _check(self._lib.nlse_sync(self._h))

    def wait(self):
        """sync() with the calling thread asleep instead of spinning."""
        _check(self._lib.nlse_wait(self._h))

    def done(self):
        """True when every step enqueued so far has finished (non-blocking)."""
        rc = self._lib.nlse_query(self._h)
        if rc < 0:
            _check(-rc)
        return rc == 1

    def set_graph(self, steps_per_graph):
        _check(self._lib.nlse_set_graph(self._h, int(steps_per_graph)))


class Group:
    """Resident 2D solvers (one method and precision, any sizes) advanced
    together by shared launches (the nlse_group_* ABI); each member's result
    is bit for bit its result alone.  While the group is open its members
    must not be used on their own; close() waits and hands them back."""

    def __init__(self, solvers):
        self._lib = lib()
        self.members = list(solvers)
        arr = (C.c_void_p * len(self.members))(*[s._h.value for s in self.members])
        self._h = C.c_void_p()
        _check(self._lib.nlse_group_create(C.byref(self._h), arr, len(self.members)))

    def step(self, steps):
        """Enqueue `steps` RK4 steps of every member (returns at once)."""
        _check(self._lib.nlse_group_step(self._h, int(steps)))

    def done(self):
        rc = self._lib.nlse_group_done(self._h)
        if rc < 0:
            _check(-rc)
        return bool(rc)

    def sync(self):
        """Wait until the group's enqueued steps have finished."""
        _check(self._lib.nlse_group_sync(self._h))

    def close(self):
        if self._h:
            self._lib.nlse_group_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent 2D problems advanced together (the nlse_batch_* ABI).

    shapes: list of (N, M); s, a, h2, k: per-problem sequences (or scalars).
    precision 2 (double) or 3 (single, compensated).
    """

    def __init__(self, shapes, s, a, h2, k, method=2, precision=3, BC=1):
        n = len(shapes)
        per = lambda x: np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)).copy()
        self.shapes = [tuple(map(int, sh)) for sh in shapes]
        Ns = np.array([sh[0] for sh in self.shapes], dtype=np.int32)
        Ms = np.array([sh[1] for sh in self.shapes], dtype=np.int32)
        self._arrays = [per(s), per(a), per(h2), per(k)]
        self._h = C.c_void_p()
        self._lib = lib()
        _check(self._lib.nlse_batch_create(C.byref(self._h), n, _p(Ns, C.c_int), _p(Ms, C.c_int),
                                           *[_p(x, C.c_double) for x in self._arrays], int(method),
                                           int(precision), int(BC)))

    def close(self):
        if self._h:
            self._lib.nlse_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def points(self):
        return int(self._lib.nlse_batch_points(self._h))

    def set_state(self, p, U, V=None):
        U = np.asarray(U)
        if U.shape != self.shapes[p]:
            raise ValueError("shape mismatch")
        ur = np.ascontiguousarray(U.real, dtype=np.float64)
        ui = np.ascontiguousarray(U.imag if np.iscomplexobj(U) else np.zeros(U.shape), dtype=np.float64)
        vp = None
        if V is not None:
            v = np.ascontiguousarray(V, dtype=np.float64)
            vp = _p(v, C.c_double)
        _check(self._lib.nlse_batch_set_state(self._h, int(p), _p(ur, C.c_double), _p(ui, C.c_double), vp))

    def state(self, p):
        outr = np.empty(self.shapes[p])
        outi = np.empty(self.shapes[p])
        _check(self._lib.nlse_batch_get_state(self._h, int(p), _p(outr, C.c_double), _p(outi, C.c_double)))
        return _cplx(outr, outi)

    def step(self, steps):
        _check(self._lib.nlse_batch_step(self._h, int(steps)))

    def step_timed(self, steps):
        ms = C.c_double()
        _check(self._lib.nlse_batch_step_timed(self._h, int(steps), C.byref(ms)))
        return ms.value

    def sync(self):
        _check(self._lib.nlse_batch_sync(self._h))

    def diagnostics(self):
        """Per-problem (sum |U|^2, max Re U, non-finite count) as an (nprob, 3) array."""
        out = np.empty((len(self.shapes), 3))
        _check(self._lib.nlse_batch_diagnostics(self._h, _p(out, C.c_double)))
        return out

    def configure(self, stages_per_launch=2, tile_rows=8, grid=1024, steps_per_graph=0):
        _check(self._lib.nlse_batch_configure(self._h, int(stages_per_launch), int(tile_rows), int(grid),
                                              int(steps_per_graph)))

    def set_kernel(self, kind, rows):
        """Stage kernel family: 0 LDS tiles (8/16/32 rows), 1 row-sweep tiles
        (8/12/16 rows), 2 register strips (3..256-row segments)."""
        _check(self._lib.nlse_batch_set_kernel(self._h, int(kind), int(rows)))

    def config(self):
        """(stages_per_launch, kernel_kind, rows, grid) in use."""
        f, k, t, g = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(self._lib.nlse_batch_get_config(self._h, C.byref(f), C.byref(k), C.byref(t), C.byref(g)))
        return f.value, k.value, t.value, g.value

    def autotune(self, steps=3):
        """Pick (kernel family, stages per launch, tile rows, grid); returns
        (stages per launch, tile rows, grid, ms/step); config() gives the
        family."""
        f, t, g, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _check(self._lib.nlse_batch_autotune(self._h, int(steps), C.byref(f), C.byref(t), C.byref(g),
                                             C.byref(ms)))
        return f.value, t.value, g.value, ms.value


_hip_rt = None


def _hip():
    """The HIP runtime the library links (already loaded by lib()); a second
    runtime in the process (torch wheels bundle their own) would not see the
    GPU once this one has initialised it."""
    global _hip_rt
    if _hip_rt is None:
        lib()
        H = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
        H.hipMalloc.restype = C.c_int
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.restype = C.c_int
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.restype = C.c_int
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip_rt = H
    return _hip_rt


class DeviceArray:
    """A device buffer of `n` elements of `dtype` for the ``*_device`` entry
    points; ``ptr`` is its address, ``numpy()`` copies it to the host."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.ptr = C.c_void_p()
        rc = _hip().hipMalloc(C.byref(self.ptr), self.n * self.dtype.itemsize)
        if rc != 0:
            raise HipError(f"hipMalloc of {self.n * self.dtype.itemsize} bytes failed ({rc})")

    @property
    def address(self):
        return int(self.ptr.value or 0)

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        rc = _hip().hipMemcpy(out.ctypes.data, self.ptr, self.n * self.dtype.itemsize, 2)  # device to host
        if rc != 0:
            raise HipError(f"hipMemcpy failed ({rc})")
        return out

    def free(self):
        if self.ptr:
            _hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def vortex_profile(m, OM, a, s, hr, n, ansatz, optimize=True, tol=1e-14):
    """Radial vortex profile f(q*hr), q < n, of the discrete steady state
    (NLS_STEADY_F.m:10-55); ansatz 0 dark sqrt, 1 dark tanh, 2 bright sech.
    Host computation."""
    f = np.empty(int(n))
    _check(lib().nlse_vortex_profile(int(m), float(OM), float(a), float(s), float(hr), int(n), int(ansatz),
                                     int(bool(optimize)), float(tol), _p(f, C.c_double)))
    return f


def _with_state(U, fn, dim):
    """Run fn(solver) on a temporary resident FP64 state holding U; return the new U."""
    U = np.asarray(U)
    if U.ndim != dim:
        raise ValueError(f"expected a {dim}D array, got shape {U.shape}")
    S = Solver(U.shape, 1, 2, 1, -1.0, 1.0, 1.0, 1.0)
    try:
        S.set_state(U)
        fn(S)
        return S.state()
    finally:
        S.close()


def _axis_origin(X, Y, Z=None):
    if Z is None:
        return (float(np.asarray(X).flat[0]), float(np.asarray(Y).flat[0]))
    return (float(np.asarray(X).flat[0]), float(np.asarray(Y).flat[0]), float(np.asarray(Z).flat[0]))


def NLS_add_vortex(charge, om, x0, y0, vx, vy, phase, U, h, a, s, amp, X, Y, opt):
    """``U = NLS_add_vortex(charge,om,x0,y0,vx,vy,phase,U,h,a,s,amp,X,Y,opt)``
    on the GPU (NLSEmagic2D_Full_Research_Scripts/NLS_add_vortex.m:9-158).
    X, Y: meshgrid arrays in this package's layout (shape (N, M), x along
    the first axis, y along the second); only their origin is read, the
    grid is taken as uniform with spacing h."""
    o = _axis_origin(X, Y)
    return _with_state(U, lambda S: S.add_vortex_2d(charge, om, x0, y0, vx, vy, phase, h, a, s, amp, o, opt), 2)


def NLSE_add_vortex_line(m, OM, x0, y0, z0, vx, vy, vz, phase, U, h, a, s, amp, X, Y, Z):
    """``U = NLSE_add_vortex_line(m,OM,x0,y0,z0,vx,vy,vz,phase,U,h,a,s,amp,X,Y,Z)``
    on the GPU (NLSEmagic3D/NLSE_add_vortex_line.m:10-125); arrays of shape
    (L, N, M) with z, x, y along the axes."""
    o = _axis_origin(X, Y, Z)
    return _with_state(U, lambda S: S.add_vortex_line(m, OM, x0, y0, z0, vx, vy, vz, phase, h, a, s, o), 3)


def NLSE_add_vortex_ring(m, OM, x0, y0, z0, thetax, thetay, thetaz, vx, vy, vz, phase, U, h, a, s, X, Y, Z, d,
                         opt_2D, ea, eb):
    """``U = NLSE_add_vortex_ring(m,OM,x0,y0,z0,thetax,thetay,thetaz,vx,vy,vz,phase,U,h,a,s,X,Y,Z,d,opt_2D,ea,eb)``
    on the GPU (NLSEmagic3D/NLSE_add_vortex_ring.m:10-534) for opt_2D = 0."""
    o = _axis_origin(X, Y, Z)
    return _with_state(U, lambda S: S.add_vortex_ring(m, OM, x0, y0, z0, thetax, thetay, thetaz, vx, vy, vz, phase, h,
                                                      a, s, o, d, opt_2D, ea, eb), 3)


def copy_probe(nbytes, reps=10):
    """Mean ms of a plain device copy of nbytes (HBM counter calibration)."""
    ms = C.c_double()
    _check(lib().nlse_copy_probe(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def footprint_bytes(shape, method, precision):
    """Device bytes a Solver of this shape holds (admission control)."""
    t = list(_tail(tuple(shape))) + [1, 1]
    return int(lib().nlse_footprint_bytes(len(shape), t[0], t[1], t[2], int(method), int(precision)))


def device_count():
    return int(lib().nlse_device_count())


def device_allocations():
    """Device allocations the library has made since it was loaded (see
    nlse_device_allocations)."""
    return int(lib().nlse_device_allocations())
