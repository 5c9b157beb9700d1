"""ctypes wrapper of the CPU oracle (oracle/refcpu.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATHS = {
    "double": os.path.join(_HERE, "_build", "librefcpu.so"),
    # the same restatement built with REF_REAL = long double (x87 80-bit, 64-bit
    # significand): how far a double evaluation lies from the operator
    "long double": os.path.join(_HERE, "_build", "librefcpu_ld.so"),
}
MAX_BC = 16


class RefScheme(C.Structure):
    _fields_ = [
        ("riemann", C.c_int), ("wave_speeds", C.c_int), ("hllc_abs", C.c_int),
        ("gamma", C.c_double), ("gas_constant", C.c_double), ("covolume", C.c_double),
        ("num_bc", C.c_int), ("bc_attr", C.c_int * MAX_BC), ("bc_type", C.c_int * MAX_BC),
        ("bc_param", (C.c_double * 8) * MAX_BC),
        ("indicator", C.c_int), ("phi", C.c_int), ("venkat_k", C.c_double),
        ("limiter", C.c_int), ("last_hope", C.c_int), ("pow_nd", C.c_int), ("linearize", C.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def lib(precision="double"):
    if precision not in _libs:
        path = _LIB_PATHS[precision]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int64)
        L.ref_create.restype = P
        L.ref_create.argtypes = [C.c_int, C.c_int, C.c_int64, dp, ip, C.c_int64, ip,
                                 C.POINTER(C.c_int), C.POINTER(RefScheme), C.c_char_p, C.c_int]
        L.ref_create_nc.restype = P
        L.ref_create_nc.argtypes = [C.c_int, C.c_int, C.c_int64, dp, ip, C.c_int64, ip,
                                    C.POINTER(C.c_int), C.POINTER(RefScheme), C.c_int64, ip,
                                    C.POINTER(C.c_int), ip, C.POINTER(C.c_int), dp, C.c_char_p, C.c_int]
        L.ref_destroy.argtypes = [P]
        L.ref_num_dofs.restype = C.c_int64
        L.ref_num_dofs.argtypes = [P]
        L.ref_nodes_per_element.argtypes = [P]
        L.ref_num_interior_faces.restype = C.c_int64
        L.ref_num_interior_faces.argtypes = [P]
        L.ref_node_coords.argtypes = [P, dp]
        L.ref_mult.argtypes = [P, dp, dp, dp, ip]
        L.ref_face_mult.argtypes = [P, dp, dp, dp, ip]
        L.ref_means.argtypes = [P, dp, dp]
        L.ref_limiter_update.argtypes = [P, dp, dp, dp]
        L.ref_step.argtypes = [P, dp, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp, ip]
        L.ref_hmin.restype = C.c_double
        L.ref_hmin.argtypes = [P]
        L.ref_element_mass_inverse.argtypes = [P, C.c_int64, dp]
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_riemann_eval.restype = C.c_double
        L.ref_riemann_eval.argtypes = [C.POINTER(RefScheme), C.c_int, dp, dp, dp, dp]
        L.ref_bc_state.argtypes = [C.POINTER(RefScheme), C.c_int, C.c_int, dp, dp, dp]
        L.ref_gl_rule.argtypes = [C.c_int, dp, dp]
        L.ref_shape.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        _libs[precision] = L
    return _libs[precision]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def ref_scheme(s) -> RefScheme:
    r = RefScheme()
    r.riemann = s.riemann
    r.wave_speeds = s.wave_speeds
    r.hllc_abs = s.hllc_abs
    r.gamma = s.gamma
    r.gas_constant = s.gas_constant
    r.covolume = s.covolume
    r.num_bc = len(s.bcs)
    for i, bc in enumerate(s.bcs):
        r.bc_attr[i] = bc.attr
        r.bc_type[i] = bc.type
        for j, p in enumerate(bc.params):
            r.bc_param[i][j] = p
    r.indicator = s.indicator
    r.phi = s.phi
    r.venkat_k = s.venkat_k
    r.limiter = s.limiter
    r.last_hope = s.last_hope
    r.pow_nd = s.pow_nd
    r.linearize = s.linearize
    return r


class RefSolver:
    """The oracle for one mesh + scheme: FE_Evolution::Mult, Limiter::update, Step."""

    def __init__(self, mesh, order, scheme, threads=None, precision="double"):
        """precision: "double" (the reference's arithmetic) or "long double"
        (the same operator evaluated wider, for measuring rounding errors)."""
        self.L = L = lib(precision)
        self.precision = precision
        if threads is not None:
            L.ref_set_threads(int(threads))
        self.mesh = mesh
        self.order = order
        self.scheme = scheme
        self._rs = ref_scheme(scheme)
        err = C.create_string_buffer(256)
        ba = np.ascontiguousarray(mesh.bdr_attr, dtype=np.int32)
        self._ba = ba
        nc = getattr(mesh, "nc", None)
        if nc is not None and len(nc):
            self._nc = [np.ascontiguousarray(nc[:, 0], dtype=np.int64), np.ascontiguousarray(nc[:, 1], dtype=np.int32),
                        np.ascontiguousarray(nc[:, 2], dtype=np.int64), np.ascontiguousarray(nc[:, 3], dtype=np.int32),
                        np.ascontiguousarray(mesh.nc_param, dtype=np.float64)]
            ci = C.POINTER(C.c_int)
            self.h = L.ref_create_nc(mesh.geom, order, mesh.num_elements, _dp(mesh.elem_coords),
                                     _ip(mesh.elem_verts), len(mesh.bdr_attr), _ip(mesh.bdr_verts),
                                     ba.ctypes.data_as(ci), C.byref(self._rs), len(nc), _ip(self._nc[0]),
                                     self._nc[1].ctypes.data_as(ci), _ip(self._nc[2]),
                                     self._nc[3].ctypes.data_as(ci), _dp(self._nc[4]), err, 256)
        else:
            self.h = L.ref_create(mesh.geom, order, mesh.num_elements, _dp(mesh.elem_coords),
                                  _ip(mesh.elem_verts), len(mesh.bdr_attr), _ip(mesh.bdr_verts),
                                  ba.ctypes.data_as(C.POINTER(C.c_int)), C.byref(self._rs), err, 256)
        if not self.h:
            raise RuntimeError("ref_create failed: " + err.value.decode())
        self.nd = L.ref_nodes_per_element(self.h)
        self.ndofs = L.ref_num_dofs(self.h)
        self.neq = mesh.dim + 2
        self.size = self.neq * self.ndofs

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def node_coords(self):
        out = np.zeros((self.mesh.num_elements, self.nd, self.mesh.dim))
        self.L.ref_node_coords(self.h, _dp(out))
        return out

    def mult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        m = np.zeros(1)
        bad = np.full(2, -1, np.int64)
        rc = self.L.ref_mult(self.h, _dp(x), _dp(y), _dp(m), _ip(bad))
        if rc:
            raise FloatingPointError(f"unphysical state at faces between elements {bad[0]} {bad[1]}")
        return y, float(m[0])

    def face_mult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        z = np.zeros_like(x)
        m = np.zeros(1)
        bad = np.full(2, -1, np.int64)
        rc = self.L.ref_face_mult(self.h, _dp(x), _dp(z), _dp(m), _ip(bad))
        if rc:
            raise FloatingPointError(f"unphysical state at faces between elements {bad[0]} {bad[1]}")
        return z, float(m[0])

    def means(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.neq * self.mesh.num_elements)
        self.L.ref_means(self.h, _dp(x), _dp(out))
        return out

    def limit(self, x):
        x = np.array(x, dtype=np.float64, copy=True)
        ind = np.zeros(self.neq * self.mesh.num_elements)
        mv = np.zeros(self.mesh.num_elements)
        self.L.ref_limiter_update(self.h, _dp(x), _dp(ind), _dp(mv))
        return x, ind, mv

    def step(self, x, dt, tableau, t=0.0):
        s, a, b, c = tableau
        x = np.array(x, dtype=np.float64, copy=True)
        a = np.ascontiguousarray(a + [0.0], dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        c = np.ascontiguousarray(c + [0.0], dtype=np.float64)
        m = np.zeros(1)
        bad = np.full(2, -1, np.int64)
        rc = self.L.ref_step(self.h, _dp(x), t, dt, s, _dp(a), _dp(b), _dp(c), _dp(m), _ip(bad))
        if rc:
            raise FloatingPointError(f"unphysical state at faces between elements {bad[0]} {bad[1]}")
        return x, float(m[0])

    def hmin(self):
        return self.L.ref_hmin(self.h)

    def patch_pressure(self, x, attr, gamma=1.4):
        """Integral of the interior-trace pressure over the boundary faces with
        attribute attr (2D): sum over faces and the p+1 Gauss-Legendre points of
        the edge of w_q |edge| p(u(x_q)). The objective of the adjoint checks."""
        m = self.mesh
        if m.dim != 2:
            raise ValueError("2D meshes only")
        if not hasattr(self, "_patch"):
            self._patch = {}
        if attr not in self._patch:
            tri = m.geom == 1
            edges = [(0, 1), (1, 2), (2, 0)] if tri else [(0, 1), (1, 2), (2, 3), (3, 0)]
            rv = np.array([[0, 0], [1, 0], [0, 1]] if tri else [[0, 0], [1, 0], [1, 1], [0, 1]], float)
            s, w = gl_rule(self.order + 1)
            where = {}
            for e in range(m.num_elements):
                for f, (a, b) in enumerate(edges):
                    where[tuple(sorted((int(m.elem_verts[e, a]), int(m.elem_verts[e, b]))))] = (e, f)
            rows = []
            for k in np.flatnonzero(m.bdr_attr == attr):
                e, f = where[tuple(sorted(int(v) for v in m.bdr_verts[k]))]
                a, b = edges[f]
                length = np.linalg.norm(m.elem_coords[e, b] - m.elem_coords[e, a])
                for q in range(len(s)):
                    xi = rv[a] + s[q] * (rv[b] - rv[a])
                    phi, _ = shape(m.geom, self.order, xi)
                    rows.append((e, w[q] * length, phi))
            self._patch[attr] = rows
        x = np.asarray(x, np.float64).reshape(self.neq, m.num_elements, self.nd)
        total = 0.0
        for e, wl, phi in self._patch[attr]:
            u = x[:, e, :] @ phi
            p = (gamma - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0])
            total += wl * p
        return total

    def mass_inverse(self, e):
        out = np.zeros((self.nd, self.nd))
        self.L.ref_element_mass_inverse(self.h, e, _dp(out))
        return out


def run(ref, x, t_final, cfl, tableau, t0=0.0, max_steps=None, callback=None):
    """The reference time loop (main.cpp:265, 467-553, 701, 733) on the oracle:
    limit the projected state, first dt = cfl hmin / lambda_faces / (2p+1)
    from the face wave speeds, then per step dt_real = min(dt, t_final - t),
    Step(x, dt_real), dt = cfl hmin / lambda_last_stage / (2p+1), and t += dt
    (the recomputed dt, quirk Q7), until t >= t_final - 1e-8 dt. Returns the
    final state and a dict with steps, t, dt, the list of dt_real and their
    sum (the time the state has actually been advanced to)."""
    x, _, _ = ref.limit(x)
    hmin = ref.hmin()
    p = ref.order
    _, lam = ref.face_mult(x)
    dt = cfl * hmin / lam / (2 * p + 1)
    t = t0
    dts = []
    steps = 0
    while True:
        dt_real = min(dt, t_final - t)
        x, lam = ref.step(x, dt_real, tableau, t)
        dts.append(dt_real)
        dt = cfl * hmin / lam / (2 * p + 1)
        t += dt
        steps += 1
        if callback is not None:
            callback(steps, t, dt, x)
        if t >= t_final - 1e-8 * dt or (max_steps is not None and steps >= max_steps):
            break
    return x, {"steps": steps, "t": t, "dt": dt, "dt_real": dts, "advanced": t0 + float(np.sum(dts)),
               "max_char_speed": lam}


def riemann_eval(scheme, dim, s1, s2, nor, precision="double"):
    rs = ref_scheme(scheme)
    s1 = np.ascontiguousarray(s1, np.float64)
    s2 = np.ascontiguousarray(s2, np.float64)
    nor = np.ascontiguousarray(nor, np.float64)
    f = np.zeros(dim + 2)
    lam = lib(precision).ref_riemann_eval(C.byref(rs), dim, _dp(s1), _dp(s2), _dp(nor), _dp(f))
    return f, lam


def bc_state(scheme, dim, bc_index, s1, nor, precision="double"):
    rs = ref_scheme(scheme)
    s1 = np.ascontiguousarray(s1, np.float64)
    nor = np.ascontiguousarray(nor, np.float64)
    s2 = np.zeros(dim + 2)
    lib(precision).ref_bc_state(C.byref(rs), dim, bc_index, _dp(s1), _dp(nor), _dp(s2))
    return s2


def gl_rule(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().ref_gl_rule(n, _dp(x), _dp(w))
    return x, w


def shape(geom, order, xi):
    from_dim = {0: 2, 1: 2, 2: 3}[geom]
    nd = {0: (order + 1) ** 2, 1: (order + 1) * (order + 2) // 2, 2: (order + 1) ** 3}[geom]
    xi = np.ascontiguousarray(xi, np.float64)
    phi = np.zeros(nd)
    dphi = np.zeros(nd * from_dim)
    lib().ref_shape(geom, order, _dp(xi), _dp(phi), _dp(dphi))
    return phi, dphi.reshape(nd, from_dim)
