"""ctypes binding of the C-ABI in include/rkdg_capi.h (librkdg.so, built in-tree).

This is the host-side mirror a Python caller uses; a C/C++ caller includes the
header directly. Device buffers are passed as raw pointers (torch tensors on
``cuda`` are accepted for convenience; torch is only used for device memory).
The library must be present: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import scheme as S

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librkdg.so")

RKDG_OK = 0
RKDG_E_ARG = -1
RKDG_E_UNPHYSICAL = -2
RKDG_E_HIP = -3
RKDG_E_UNSUPPORTED = -4
RKDG_E_NAN = -5


class MeshDesc(C.Structure):
    _fields_ = [
        ("geom", C.c_int), ("order", C.c_int), ("num_elements", C.c_int64),
        ("elem_coords", C.POINTER(C.c_double)), ("elem_verts", C.POINTER(C.c_int64)),
        ("num_bdr", C.c_int64), ("bdr_verts", C.POINTER(C.c_int64)),
        ("bdr_attr", C.POINTER(C.c_int32)),
    ]


class NcDesc(C.Structure):
    _fields_ = [
        ("count", C.c_int64), ("slave", C.POINTER(C.c_int64)), ("slave_face", C.POINTER(C.c_int32)),
        ("master", C.POINTER(C.c_int64)), ("master_face", C.POINTER(C.c_int32)), ("param", C.POINTER(C.c_double)),
    ]


class SchemeDesc(C.Structure):
    _fields_ = [
        ("riemann", C.c_int), ("wave_speeds", C.c_int), ("hllc_abs", C.c_int),
        ("gamma", C.c_double), ("gas_constant", C.c_double), ("covolume", C.c_double),
        ("num_bc", C.c_int), ("bc_attr", C.c_int * S.MAX_BC), ("bc_type", C.c_int * S.MAX_BC),
        ("bc_param", (C.c_double * 8) * S.MAX_BC),
        ("indicator", C.c_int), ("phi", C.c_int), ("venkat_k", C.c_double),
        ("limiter", C.c_int), ("last_hope", C.c_int), ("pow_nd", C.c_int), ("linearize", C.c_int),
    ]


class Info(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("num_equations", C.c_int), ("nodes_per_element", C.c_int),
        ("faces_per_element", C.c_int), ("num_elements", C.c_int64), ("num_dofs", C.c_int64),
        ("state_size", C.c_int64), ("num_interior_faces", C.c_int64),
        ("num_boundary_faces", C.c_int64), ("affine", C.c_int), ("two_colour", C.c_int),
    ]


# every symbol declared in include/rkdg_capi.h
EXPORTS = [
    "rkdg_create", "rkdg_destroy", "rkdg_last_error", "rkdg_get_info", "rkdg_set_stream",
    "rkdg_rhs", "rkdg_face_lambda", "rkdg_limit", "rkdg_step", "rkdg_means", "rkdg_hmin",
    "rkdg_node_coords", "rkdg_check_error", "rkdg_profile_enable", "rkdg_profile_read",
    "rkdg_create_ex", "rkdg_batch_create", "rkdg_batch_offsets", "rkdg_batch_init",
    "rkdg_batch_advance", "rkdg_batch_status", "rkdg_batch_set_options", "rkdg_tune_grid",
    "rkdg_set_bc_params", "rkdg_rhs_adjoint", "rkdg_limit_adjoint", "rkdg_objective",
    "rkdg_step_adjoint", "rkdg_forward_solve", "rkdg_adjoint_solve", "rkdg_step_noalloc",
    "rkdg_step_ssprk3", "rkdg_totals", "rkdg_primitives", "rkdg_interpolate",
    "rkdg_rhs_host", "rkdg_limit_host", "rkdg_step_host", "rkdg_riemann_eval", "rkdg_bc_right_state",
    "rkdg_error_elements", "rkdg_create_nc",
]

# include/rkdg_selftest.h (host-only)
SELFTEST_EXPORTS = [
    "rkdg_selftest_riemann", "rkdg_selftest_riemann_vjp", "rkdg_selftest_volume_flux",
    "rkdg_selftest_volume_flux_vjp", "rkdg_selftest_bc_state", "rkdg_selftest_bc_state_vjp",
    "rkdg_selftest_limiter_phi", "rkdg_selftest_pressure",
]

RKDG_CREATE_DENSE = 1
RKDG_CREATE_TWO_COLOUR = 2

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python rkdg_amd/build.py` "
                          "(there is no CPU fallback for the RKDG path)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.c_void_p
    L.rkdg_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(SchemeDesc), C.POINTER(C.c_void_p)]
    L.rkdg_destroy.argtypes = [P]
    L.rkdg_last_error.restype = C.c_char_p
    L.rkdg_get_info.argtypes = [P, C.POINTER(Info)]
    L.rkdg_set_stream.argtypes = [P, C.c_void_p]
    L.rkdg_rhs.argtypes = [P, dp, dp, C.POINTER(C.c_double)]
    L.rkdg_face_lambda.argtypes = [P, dp, C.POINTER(C.c_double)]
    L.rkdg_limit.argtypes = [P, dp, dp]
    L.rkdg_step.argtypes = [P, dp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double),
                            C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.rkdg_totals.argtypes = [P, dp, C.POINTER(C.c_double)]
    L.rkdg_primitives.argtypes = [P, dp, dp, dp, C.c_double]
    L.rkdg_interpolate.argtypes = [P, dp, C.c_int, C.POINTER(C.c_double), dp]
    L.rkdg_step_ssprk3.argtypes = [P, dp, C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.rkdg_step_noalloc.argtypes = [P, dp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), dp]
    L.rkdg_means.argtypes = [P, dp, dp]
    L.rkdg_hmin.argtypes = [P, C.POINTER(C.c_double)]
    L.rkdg_node_coords.argtypes = [P, C.POINTER(C.c_double)]
    L.rkdg_check_error.argtypes = [P, C.POINTER(C.c_int64)]
    L.rkdg_error_elements.argtypes = [P, C.POINTER(C.c_int64)]
    L.rkdg_profile_enable.argtypes = [P, C.c_int]
    L.rkdg_profile_read.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    pd = C.POINTER(C.c_double)
    L.rkdg_create_ex.argtypes = [C.POINTER(MeshDesc), C.POINTER(SchemeDesc), C.c_int, C.POINTER(C.c_void_p)]
    L.rkdg_create_nc.argtypes = [C.POINTER(MeshDesc), C.POINTER(SchemeDesc), C.c_int, C.POINTER(NcDesc),
                                 C.POINTER(C.c_void_p)]
    L.rkdg_batch_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(SchemeDesc), C.c_int, C.POINTER(C.c_void_p)]
    L.rkdg_batch_offsets.argtypes = [P, C.POINTER(C.c_int64)]
    L.rkdg_batch_init.argtypes = [P, dp, pd, pd, C.c_double, C.c_int64]
    L.rkdg_batch_advance.argtypes = [P, dp, C.c_int, pd, pd, pd, C.c_int64, C.POINTER(C.c_int)]
    L.rkdg_batch_status.argtypes = [P, pd, pd, C.POINTER(C.c_int64), C.POINTER(C.c_int)]
    L.rkdg_batch_set_options.argtypes = [P, C.c_int, C.c_int]
    L.rkdg_tune_grid.argtypes = [P, dp, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_int), pd]
    L.rkdg_set_bc_params.argtypes = [P, C.c_int, pd]
    L.rkdg_rhs_adjoint.argtypes = [P, dp, dp, dp, dp, dp]
    L.rkdg_limit_adjoint.argtypes = [P, dp, dp, dp]
    L.rkdg_objective.argtypes = [P, dp, C.c_int, C.c_double, dp, dp]
    L.rkdg_step_adjoint.argtypes = [P, dp, C.c_double, C.c_int, pd, pd, pd, dp, dp, pd]
    L.rkdg_forward_solve.argtypes = [P, dp, C.c_int64, pd, C.c_int, pd, pd, pd, C.c_int, pd]
    L.rkdg_adjoint_solve.argtypes = [P, dp, C.c_int64, pd, C.c_int, pd, pd, pd, C.c_int, C.c_int64,
                                     pd, dp, pd, dp]
    SD = C.POINTER(SchemeDesc)
    L.rkdg_rhs_host.argtypes = [P, pd, pd, pd]
    L.rkdg_limit_host.argtypes = [P, pd, C.POINTER(C.c_int8)]
    L.rkdg_step_host.argtypes = [P, pd, C.c_double, C.c_double, C.c_int, pd, pd, pd, pd]
    L.rkdg_riemann_eval.argtypes = [SD, C.c_int, pd, pd, pd, pd, pd]
    L.rkdg_bc_right_state.argtypes = [SD, C.c_int, C.c_int, pd, pd, pd]
    L.rkdg_selftest_riemann.restype = C.c_double
    L.rkdg_selftest_riemann.argtypes = [SD, pd, pd, pd, pd]
    L.rkdg_selftest_riemann_vjp.argtypes = [SD, pd, pd, pd, pd, pd, pd]
    L.rkdg_selftest_volume_flux.argtypes = [SD, pd, pd, pd]
    L.rkdg_selftest_volume_flux_vjp.argtypes = [SD, pd, pd, pd, pd]
    L.rkdg_selftest_bc_state.argtypes = [SD, C.c_int, pd, pd, pd]
    L.rkdg_selftest_bc_state_vjp.argtypes = [SD, C.c_int, pd, pd, pd, pd, pd]
    L.rkdg_selftest_limiter_phi.restype = C.c_double
    L.rkdg_selftest_limiter_phi.argtypes = [SD, C.c_double, C.c_double, C.c_double, pd]
    L.rkdg_selftest_pressure.restype = C.c_double
    L.rkdg_selftest_pressure.argtypes = [SD, pd, pd]
    _lib = L
    return L


class RKDGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rkdg error {code}: {msg}")
        self.code = code


class UnphysicalStateError(RKDGError):
    pass


class NonFiniteStateError(UnphysicalStateError):
    """A flux input held a NaN or an infinity (RKDG_E_NAN)."""


def _check(rc):
    if rc != RKDG_OK:
        msg = lib().rkdg_last_error().decode()
        if rc == RKDG_E_NAN:
            raise NonFiniteStateError(rc, msg)
        if rc == RKDG_E_UNPHYSICAL:
            raise UnphysicalStateError(rc, msg)
        raise RKDGError(rc, msg)


def scheme_desc(s: S.Scheme) -> SchemeDesc:
    d = SchemeDesc()
    d.riemann = s.riemann
    d.wave_speeds = s.wave_speeds
    d.hllc_abs = s.hllc_abs
    d.gamma = s.gamma
    d.gas_constant = s.gas_constant
    d.covolume = s.covolume
    d.num_bc = len(s.bcs)
    for i, bc in enumerate(s.bcs):
        d.bc_attr[i] = bc.attr
        d.bc_type[i] = bc.type
        for j, p in enumerate(bc.params):
            d.bc_param[i][j] = p
    d.indicator = s.indicator
    d.phi = s.phi
    d.venkat_k = s.venkat_k
    d.limiter = s.limiter
    d.last_hope = s.last_hope
    d.pow_nd = s.pow_nd
    d.linearize = s.linearize
    return d


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def check_buffer(t, n, name, dtype="float64", nullable=False):
    """Validates a device buffer before its pointer reaches the kernels: a torch
    tensor on the GPU, of the given dtype (float64 states, int8 flags),
    contiguous, with exactly n elements. Raises ValueError otherwise (a short,
    strided or host buffer would be read or written out of bounds on the
    device). Returns the tensor."""
    if t is None:
        if nullable:
            return None
        raise ValueError(f"{name}: a device buffer is required")
    import torch

    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor on the GPU, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name}: tensor is on {t.device}, the kernels need GPU memory")
    if t.dtype != getattr(torch, dtype):
        raise ValueError(f"{name}: dtype {t.dtype}, expected torch.{dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor is not contiguous")
    if t.numel() != n:
        raise ValueError(f"{name}: {t.numel()} elements, expected {n}")
    return t


def check_tableau(tableau):
    """(s, a, b, c) with 1 <= s <= 4, at least the s(s-1)/2 entries of a that
    the packed a[l++] rule reads (the reference's literal tableaus hold more,
    case_manager.cpp:829-884) and s weights b."""
    s, a, b, c = tableau
    if not 1 <= int(s) <= 4:
        raise ValueError(f"tableau: {s} stages, 1 to 4 supported")
    if len(a) < s * (s - 1) // 2:
        raise ValueError(f"tableau: {len(a)} a entries, the packed rule reads s(s-1)/2 = {s * (s - 1) // 2}")
    if len(b) != s:
        raise ValueError(f"tableau: {len(b)} b entries, expected {s}")
    return tableau


def _mesh_desc(mesh, order, md=None):
    """Fills a MeshDesc; returns it with the arrays it points into (keep them alive)."""
    keep = (np.ascontiguousarray(mesh.elem_coords, np.float64), np.ascontiguousarray(mesh.elem_verts, np.int64),
            np.ascontiguousarray(mesh.bdr_verts, np.int64), np.ascontiguousarray(mesh.bdr_attr, np.int32))
    md = md if md is not None else MeshDesc()
    md.geom = mesh.geom
    md.order = order
    md.num_elements = mesh.num_elements
    md.elem_coords = keep[0].ctypes.data_as(C.POINTER(C.c_double))
    md.elem_verts = keep[1].ctypes.data_as(C.POINTER(C.c_int64))
    md.num_bdr = len(mesh.bdr_attr)
    md.bdr_verts = keep[2].ctypes.data_as(C.POINTER(C.c_int64))
    md.bdr_attr = keep[3].ctypes.data_as(C.POINTER(C.c_int32))
    return md, keep


def _nc_desc(mesh):
    """NcDesc of a mesh's non-conforming faces (None when it has none), with
    the arrays it points into."""
    nc = getattr(mesh, "nc", None)
    if nc is None or len(nc) == 0:
        return None, ()
    keep = (np.ascontiguousarray(nc[:, 0], np.int64), np.ascontiguousarray(nc[:, 1], np.int32),
            np.ascontiguousarray(nc[:, 2], np.int64), np.ascontiguousarray(nc[:, 3], np.int32),
            np.ascontiguousarray(mesh.nc_param, np.float64))
    d = NcDesc()
    d.count = len(nc)
    d.slave = keep[0].ctypes.data_as(C.POINTER(C.c_int64))
    d.slave_face = keep[1].ctypes.data_as(C.POINTER(C.c_int32))
    d.master = keep[2].ctypes.data_as(C.POINTER(C.c_int64))
    d.master_face = keep[3].ctypes.data_as(C.POINTER(C.c_int32))
    d.param = keep[4].ctypes.data_as(C.POINTER(C.c_double))
    return d, keep


def _tableau_arrays(tableau):
    s, a, b, c = check_tableau(tableau)
    a = (C.c_double * max(1, len(a)))(*a) if a else (C.c_double * 1)(0.0)
    b = (C.c_double * len(b))(*b)
    c = (C.c_double * max(1, len(c)))(*c) if c else (C.c_double * 1)(0.0)
    return s, a, b, c


class Context:
    """One rkdg_ctx: a flattened mesh + scheme resident on the GPU. dense=True
    runs quadrilaterals on the dense bilinear kernels (RKDG_CREATE_DENSE);
    two_colour=True runs the right-hand side as a solver/receiver pair of
    kernels where the mesh allows it (RKDG_CREATE_TWO_COLOUR)."""

    def __init__(self, mesh, order: int, scheme: S.Scheme, dense: bool = False, _handle=None,
                 two_colour: bool = False):
        L = lib()
        if _handle is None:
            md, self._keep = _mesh_desc(mesh, order)
            self._sd = scheme_desc(scheme)
            h = C.c_void_p()
            flags = (RKDG_CREATE_DENSE if dense else 0) | (RKDG_CREATE_TWO_COLOUR if two_colour else 0)
            nd, nkeep = _nc_desc(mesh)
            if nd is None:
                _check(L.rkdg_create_ex(C.byref(md), C.byref(self._sd), flags, C.byref(h)))
            else:
                _check(L.rkdg_create_nc(C.byref(md), C.byref(self._sd), flags, C.byref(nd), C.byref(h)))
        else:
            h = _handle
        self.h = h
        info = Info()
        _check(L.rkdg_get_info(self.h, C.byref(info)))
        self.info = info
        self.mesh = mesh
        self.order = order
        self.scheme = scheme

    def close(self):
        if getattr(self, "h", None):
            lib().rkdg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def state_size(self):
        return self.info.state_size

    def set_stream(self, stream_ptr):
        _check(lib().rkdg_set_stream(self.h, C.c_void_p(stream_ptr)))

    def _st(self, t, name, nullable=False):
        return check_buffer(t, self.info.state_size, name, nullable=nullable)

    def rhs(self, x, y, want_lambda=True):
        self._st(x, "x")
        self._st(y, "y")
        lam = C.c_double(0.0)
        _check(lib().rkdg_rhs(self.h, _ptr(x), _ptr(y), C.byref(lam) if want_lambda else None))
        return lam.value

    def face_lambda(self, x):
        self._st(x, "x")
        lam = C.c_double(0.0)
        _check(lib().rkdg_face_lambda(self.h, _ptr(x), C.byref(lam)))
        return lam.value

    def limit(self, x, troubled=None):
        self._st(x, "x")
        check_buffer(troubled, self.info.num_elements, "troubled", dtype="int8", nullable=True)
        _check(lib().rkdg_limit(self.h, _ptr(x), _ptr(troubled)))

    def means(self, x, out):
        self._st(x, "x")
        check_buffer(out, self.info.num_equations * self.info.num_elements, "means")
        _check(lib().rkdg_means(self.h, _ptr(x), _ptr(out)))

    def step(self, x, dt, tableau, t=0.0, want_lambda=True):
        self._st(x, "x")
        s, a, b, c = _tableau_arrays(tableau)
        lam = C.c_double(0.0)
        _check(lib().rkdg_step(self.h, _ptr(x), t, dt, s, a, b, c,
                               C.byref(lam) if want_lambda else None))
        return lam.value

    def step_ssprk3(self, x, dt, t=0.0, want_lambda=True):
        """rkdg_step_ssprk3: one SSP-RK3 step in the Shu-Osher form, each stage
        value limited (the classical limited RKDG step)."""
        self._st(x, "x")
        lam = C.c_double(0.0)
        _check(lib().rkdg_step_ssprk3(self.h, _ptr(x), t, dt, C.byref(lam) if want_lambda else None))
        return lam.value

    def step_noalloc(self, x, dt, tableau, lam_dev=None, t=0.0):
        """rkdg_step_noalloc: the step with its stage vectors in the static
        scratch array; the wave speed goes to the device double lam_dev."""
        self._st(x, "x")
        check_buffer(lam_dev, 1, "lam_dev", nullable=True)
        s, a, b, c = _tableau_arrays(tableau)
        _check(lib().rkdg_step_noalloc(self.h, _ptr(x), t, dt, s, a, b, c, _ptr(lam_dev)))

    def totals(self, x):
        """Integral over the mesh of each conserved variable (rkdg_totals)."""
        self._st(x, "x")
        out = (C.c_double * self.info.num_equations)()
        _check(lib().rkdg_totals(self.h, _ptr(x), out))
        return np.array(out[:])

    def primitives(self, x, prim=None, acc=None, weight=1.0):
        """rho, u, p, T at the nodes into prim [(dim+3)*num_dofs] and/or acc += weight * them."""
        self._st(x, "x")
        n = (self.info.dim + 3) * self.info.num_dofs
        check_buffer(prim, n, "prim", nullable=True)
        check_buffer(acc, n, "acc", nullable=True)
        _check(lib().rkdg_primitives(self.h, _ptr(x), _ptr(prim), _ptr(acc), float(weight)))

    def interpolate(self, x, ref_pts, out):
        """The state at reference points of every element: out [neq*NE*npts] (rkdg_interpolate)."""
        self._st(x, "x")
        pts = np.ascontiguousarray(ref_pts, np.float64).reshape(-1, self.info.dim)
        check_buffer(out, self.info.num_equations * self.info.num_elements * len(pts), "out")
        _check(lib().rkdg_interpolate(self.h, _ptr(x), len(pts), pts.ctypes.data_as(C.POINTER(C.c_double)),
                                      _ptr(out)))

    def hmin(self):
        v = C.c_double(0.0)
        _check(lib().rkdg_hmin(self.h, C.byref(v)))
        return v.value

    def node_coords(self):
        out = np.zeros((self.info.num_elements, self.info.nodes_per_element, self.info.dim))
        _check(lib().rkdg_node_coords(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def profile(self, enable=True):
        _check(lib().rkdg_profile_enable(self.h, 1 if enable else 0))

    def profile_read(self):
        """(rhs_ms_total, rhs_launches, limit_ms_total, limit_launches)"""
        a, b = C.c_double(0), C.c_int64(0)
        c_, d = C.c_double(0), C.c_int64(0)
        _check(lib().rkdg_profile_read(self.h, C.byref(a), C.byref(b), C.byref(c_), C.byref(d)))
        return a.value, b.value, c_.value, d.value

    def tune_grid(self, x, candidates=(256, 512, 1024, 2048, 4096, 8192), reps=5):
        """Autotuner (rkdg_tune_grid): times the right-hand side of x for the
        persistent grid sizes 0 (one workgroup per work block) and candidates,
        keeps the fastest; returns (chosen cap, {cap: ms per launch})."""
        self._st(x, "x")
        cand = (C.c_int * len(candidates))(*candidates)
        ms = (C.c_double * (len(candidates) + 1))()
        best = C.c_int(0)
        _check(lib().rkdg_tune_grid(self.h, _ptr(x), cand, len(candidates), int(reps), C.byref(best), ms))
        return best.value, dict(zip((0,) + tuple(candidates), ms[:]))

    def set_grid_cap(self, cap):
        _check(lib().rkdg_batch_set_options(self.h, 1, int(cap)))

    # --- discrete adjoint (dense 2D contexts)
    def set_bc_params(self, slot, params):
        p = np.zeros(8)
        p[:len(params)] = params
        _check(lib().rkdg_set_bc_params(self.h, int(slot), p.ctypes.data_as(C.POINTER(C.c_double))))

    def rhs_adjoint(self, x, kbar, xbar, addend=None, pbar=None):
        for t, nm in ((x, "x"), (kbar, "kbar"), (xbar, "xbar")):
            self._st(t, nm)
        self._st(addend, "addend", nullable=True)
        check_buffer(pbar, S.MAX_BC * 8, "pbar", nullable=True)
        _check(lib().rkdg_rhs_adjoint(self.h, _ptr(x), _ptr(kbar), _ptr(xbar), _ptr(addend), _ptr(pbar)))

    def limit_adjoint(self, x, ubar, xbar):
        for t, nm in ((x, "x"), (ubar, "ubar"), (xbar, "xbar")):
            self._st(t, nm)
        _check(lib().rkdg_limit_adjoint(self.h, _ptr(x), _ptr(ubar), _ptr(xbar)))

    def objective(self, x, slot, coef, J=None, xbar=None):
        self._st(x, "x")
        check_buffer(J, 1, "J", nullable=True)
        self._st(xbar, "xbar", nullable=True)
        _check(lib().rkdg_objective(self.h, _ptr(x), int(slot), float(coef), _ptr(J), _ptr(xbar)))

    def step_adjoint(self, x, dt, tableau, ybar, xbar):
        for t, nm in ((x, "x"), (ybar, "ybar"), (xbar, "xbar")):
            self._st(t, nm)
        s, a, b, c = _tableau_arrays(tableau)
        pbar = np.zeros(S.MAX_BC * 8)
        pd = C.POINTER(C.c_double)
        _check(lib().rkdg_step_adjoint(self.h, _ptr(x), float(dt), s, a, b, c, _ptr(ybar), _ptr(xbar),
                                       pbar.ctypes.data_as(pd)))
        return pbar.reshape(S.MAX_BC, 8)

    def forward_solve(self, x, dts, tableau, obj_slot=-1):
        self._st(x, "x")
        s, a, b, c = _tableau_arrays(tableau)
        d = np.ascontiguousarray(dts, np.float64)
        J = C.c_double(0.0)
        pd = C.POINTER(C.c_double)
        _check(lib().rkdg_forward_solve(self.h, _ptr(x), len(d), d.ctypes.data_as(pd), s, a, b, c,
                                        int(obj_slot), C.byref(J)))
        return J.value

    def adjoint_solve(self, x0, dts, tableau, obj_slot, x0bar, every=0, xN=None):
        """Returns (J, bc_param_bar[MAX_BC, 8]); x0bar receives dJ/dx0."""
        self._st(x0, "x0")
        self._st(x0bar, "x0bar")
        self._st(xN, "xN", nullable=True)
        s, a, b, c = _tableau_arrays(tableau)
        d = np.ascontiguousarray(dts, np.float64)
        J = C.c_double(0.0)
        pbar = np.zeros(S.MAX_BC * 8)
        pd = C.POINTER(C.c_double)
        _check(lib().rkdg_adjoint_solve(self.h, _ptr(x0), len(d), d.ctypes.data_as(pd), s, a, b, c,
                                        int(obj_slot), int(every), C.byref(J), _ptr(x0bar),
                                        pbar.ctypes.data_as(pd), _ptr(xN)))
        return J.value, pbar.reshape(S.MAX_BC, 8)

    def check_error(self):
        info = (C.c_int64 * 2)()
        _check(lib().rkdg_check_error(self.h, info))


class BatchContext(Context):
    """A batched ensemble (rkdg_batch_*): instances that share element type,
    order and scheme (boundary parameters may differ) held as one union mesh.
    The state of the union is the byNODES layout of the union mesh; instance i
    owns elements offsets[i]:offsets[i+1] of every variable."""

    def __init__(self, meshes, order: int, schemes):
        L = lib()
        n = len(meshes)
        if len(schemes) != n or n < 1:
            raise ValueError("one scheme per mesh")
        mds = (MeshDesc * n)()
        self._keep = []
        for i, m in enumerate(meshes):
            _, k = _mesh_desc(m, order, mds[i])
            self._keep.append(k)
        sds = (SchemeDesc * n)()
        for i, s in enumerate(schemes):
            sds[i] = scheme_desc(s)
        self._sds = sds
        h = C.c_void_p()
        _check(L.rkdg_batch_create(mds, sds, n, C.byref(h)))
        super().__init__(None, order, schemes[0], _handle=h)
        self.count = n
        self.meshes = meshes
        off = (C.c_int64 * (n + 1))()
        _check(L.rkdg_batch_offsets(self.h, off))
        self.offsets = np.array(off[:], dtype=np.int64)

    def init(self, x, t_final, cfl, t0=0.0, max_steps=0):
        self._st(x, "x")
        tf = np.broadcast_to(np.asarray(t_final, np.float64), (self.count,)).copy()
        t0 = np.broadcast_to(np.asarray(t0, np.float64), (self.count,)).copy()
        pd = C.POINTER(C.c_double)
        _check(lib().rkdg_batch_init(self.h, _ptr(x), t0.ctypes.data_as(pd), tf.ctypes.data_as(pd),
                                     float(cfl), int(max_steps)))

    def advance(self, x, tableau, nsteps=1):
        """Returns the number of instances still running."""
        self._st(x, "x")
        s, a, b, c = _tableau_arrays(tableau)
        na = C.c_int(0)
        _check(lib().rkdg_batch_advance(self.h, _ptr(x), s, a, b, c, int(nsteps), C.byref(na)))
        return na.value

    def status(self):
        n = self.count
        t, dt = np.zeros(n), np.zeros(n)
        steps, act = np.zeros(n, np.int64), np.zeros(n, np.int32)
        pd = C.POINTER(C.c_double)
        _check(lib().rkdg_batch_status(self.h, t.ctypes.data_as(pd), dt.ctypes.data_as(pd),
                                       steps.ctypes.data_as(C.POINTER(C.c_int64)),
                                       act.ctypes.data_as(C.POINTER(C.c_int))))
        return t, dt, steps, act

    def set_options(self, use_graph=True, grid_cap=-1):
        _check(lib().rkdg_batch_set_options(self.h, 1 if use_graph else 0, int(grid_cap)))



class SelfTest:
    """Host evaluation of the 2D physics and its hand-written derivatives
    (include/rkdg_selftest.h); needs no GPU."""

    def __init__(self, scheme: S.Scheme):
        self.sd = scheme_desc(scheme)
        self.L = lib()

    @staticmethod
    def _a(v, n=4):
        return np.ascontiguousarray(np.asarray(v, np.float64).reshape(-1)[:n] if v is not None else np.zeros(n))

    def _p(self, a):
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def riemann(self, uL, uR, n):
        uL, uR, n = self._a(uL), self._a(uR), self._a(n, 2)
        f = np.zeros(4)
        lam = self.L.rkdg_selftest_riemann(C.byref(self.sd), self._p(uL), self._p(uR), self._p(n), self._p(f))
        return f, lam

    def riemann_vjp(self, uL, uR, n, Fbar):
        uL, uR, n, Fbar = self._a(uL), self._a(uR), self._a(n, 2), self._a(Fbar)
        a, b = np.zeros(4), np.zeros(4)
        self.L.rkdg_selftest_riemann_vjp(C.byref(self.sd), self._p(uL), self._p(uR), self._p(n), self._p(Fbar),
                                         self._p(a), self._p(b))
        return a, b

    def volume_flux(self, u, a):
        u, a = self._a(u), self._a(a, 2)
        G = np.zeros(4)
        self.L.rkdg_selftest_volume_flux(C.byref(self.sd), self._p(u), self._p(a), self._p(G))
        return G

    def volume_flux_vjp(self, u, a, Gbar):
        u, a, Gbar = self._a(u), self._a(a, 2), self._a(Gbar)
        ub = np.zeros(4)
        self.L.rkdg_selftest_volume_flux_vjp(C.byref(self.sd), self._p(u), self._p(a), self._p(Gbar), self._p(ub))
        return ub

    def bc_state(self, bi, u, n):
        u, n = self._a(u), self._a(n, 2)
        g = np.zeros(4)
        self.L.rkdg_selftest_bc_state(C.byref(self.sd), int(bi), self._p(u), self._p(n), self._p(g))
        return g

    def bc_state_vjp(self, bi, u, n, gbar):
        u, n, gbar = self._a(u), self._a(n, 2), self._a(gbar)
        ub, pb = np.zeros(4), np.zeros(8)
        self.L.rkdg_selftest_bc_state_vjp(C.byref(self.sd), int(bi), self._p(u), self._p(n), self._p(gbar),
                                          self._p(ub), self._p(pb))
        return ub, pb

    def limiter_phi(self, y, diff, eps2):
        d = np.zeros(3)
        v = self.L.rkdg_selftest_limiter_phi(C.byref(self.sd), float(y), float(diff), float(eps2), self._p(d))
        return v, d[1], d[2]

    def pressure(self, u):
        u = self._a(u)
        dp = np.zeros(4)
        p = self.L.rkdg_selftest_pressure(C.byref(self.sd), self._p(u), self._p(dp))
        return p, dp
