"""Case definitions for the RKDG path (the host-side role of CaseManager,
src/case_manager/case_manager.cpp): mesh, scheme, initial state and time
settings for the workloads the path is benchmarked and tested on.

Initial conditions are projected by nodal interpolation at the element nodes,
as MFEM's ProjectCoefficient does for the L2 space (case_manager.cpp:515-516).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import mesh as M
from . import scheme as S


@dataclass
class Case:
    name: str
    mesh: M.Mesh
    order: int
    scheme: S.Scheme
    t_final: float
    cfl: float
    tableau: str = "ssprk3"
    params: dict = field(default_factory=dict)

    def initial_state(self, xyz):
        return self.params["ic"](xyz)


def _cons(rho, vel, p, gamma=1.4):
    E = p / (gamma - 1.0) + 0.5 * rho * sum(v * v for v in vel)
    return [rho] + [rho * v for v in vel] + [E]


def _flat(U):
    return np.ascontiguousarray(np.concatenate([np.asarray(u, float).ravel() for u in U]))


def sod2d(nx=200, ny=10, phi=S.PHI_BJ) -> Case:
    """Sod tube on [0,1]x[0,0.1], Q1, HLLC, BJ-family multiplier limiter, SSP-RK3, CFL 0.2."""
    m = M.structured_quad(nx, ny, 0.0, 1.0, 0.0, 0.1)
    bcs = [S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(2, S.BC_OUTLET), S.BC(4, S.BC_OUTLET)]
    sch = S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, phi=phi, limiter=S.LIM_MULTI, bcs=bcs)

    def ic(xyz):
        X = xyz[..., 0].ravel()
        left = X < 0.5
        rho = np.where(left, 1.0, 0.125)
        p = np.where(left, 1.0, 0.1)
        return _flat(_cons(rho, [0 * X, 0 * X], p))

    return Case("sod2d", m, 1, sch, 0.2, 0.2, params={"ic": ic})


def pulse3d(n=96, amp=1e-3, width=0.05) -> Case:
    """Periodic acoustic pulse on [0,1]^3, Q2 hex, Rusanov, Venkatakrishnan K=5, SSP-RK3."""
    m = M.structured_hex(n, n, n)
    sch = S.Scheme(riemann=S.RUSANOV, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0,
                   limiter=S.LIM_MULTI)
    gamma = 1.4

    def ic(xyz):
        r2 = ((xyz - 0.5) ** 2).sum(axis=-1).ravel()
        g = np.exp(-r2 / (width * width))
        rho = 1.0 + amp * g
        p = 1.0 / gamma + amp * g
        z = 0 * rho
        return _flat(_cons(rho, [z, z, z], p, gamma))

    return Case("pulse3d", m, 2, sch, 0.1, 0.2, params={"ic": ic})


def double_mach(nx=490, ny=123, seed=20240601, shock_width=0.0) -> Case:
    """Mach-10 shock at 60 degrees on [0,4]x[0,1] hitting a wall from x = 1/6:
    jittered Delaunay triangles, P2, HLLC with Einfeldt speeds, Michalak
    limiter, SSP-RK3, CFL 0.1; pre-shock (rho, p) = (1.4, 1). The bottom is the
    post-shock inflow for x < 1/6 (attribute 5, a supersonic inlet holding the
    post-shock state) and a slip wall beyond; the left side is the post-shock
    inflow, the right side an outlet. The classical problem moves the top
    boundary's state with the shock (x_s(t) = 1/6 + (1 + 20 t)/sqrt 3); that
    time-dependent boundary is no reference boundary type, so the top here is
    an outlet (zero-gradient) boundary, which matches the classical data until
    the shock's reflection reaches the top.

    shock_width = 0 interpolates the discontinuous data at the nodes. With it
    the scheme as configured (P2, this limiter) creates an unphysical face state
    within the first step at every resolution tried (the reference stops there,
    face_integrator.cpp:214-218). shock_width = w > 0 blends the primitive
    variables across the shock with 0.5 (1 + tanh(d / w)), d the distance to
    the shock line, which the scheme advances."""
    m = M.jittered_delaunay_tri(nx, ny, 0.0, 4.0, 0.0, 1.0, jitter=0.3, seed=seed)
    m = M.relabel_boundary(m, lambda attr, mid: 5 if attr == 1 and mid[0] < 1.0 / 6.0 else attr)
    post = (8.0, 8.25 * math.cos(math.pi / 6), -8.25 * math.sin(math.pi / 6), 116.5)
    inlet_state = (post[0], post[1], post[2], 0.0, post[3])
    bcs = [S.BC(4, S.BC_SUPERSONIC_INLET, inlet_state), S.BC(5, S.BC_SUPERSONIC_INLET, inlet_state),
           S.BC(1, S.BC_SLIP), S.BC(2, S.BC_OUTLET), S.BC(3, S.BC_OUTLET)]
    sch = S.Scheme(riemann=S.HLLC, wave_speeds=S.EINFELDT, indicator=S.IND_BJ, phi=S.PHI_MICHALAK,
                   limiter=S.LIM_MULTI, bcs=bcs)

    def ic(xyz):
        X, Y = xyz[..., 0].ravel(), xyz[..., 1].ravel()
        if shock_width > 0.0:
            # signed distance to the shock line x = 1/6 + y / sqrt 3 (positive behind it)
            d = (1.0 / 6.0 + Y / math.sqrt(3.0) - X) * (math.sqrt(3.0) / 2.0)
            s = 0.5 * (1.0 + np.tanh(d / shock_width))
        else:
            s = (X < 1.0 / 6.0 + Y / math.sqrt(3.0)).astype(float)
        rho = 1.4 + (post[0] - 1.4) * s
        u = post[1] * s
        v = post[2] * s
        p = 1.0 + (post[3] - 1.0) * s
        return _flat(_cons(rho, [u, v], p))

    return Case("double_mach", m, 2, sch, 0.2, 0.1, params={"ic": ic, "shock_width": shock_width})


def nozzle(num_elements, area_ratio, pressure_ratio, p_tot=1.0, T_tot=1.0) -> Case:
    """Converging-diverging nozzle: Q1 quads, HLL, BJ-family (Michalak) multiplier limiter,
    subsonicInletTotalPressure inlet, fixedPressureOutlet, slip walls."""
    m = M.nozzle_quad(num_elements, area_ratio)
    p_exit = p_tot / pressure_ratio
    bcs = [S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP),
           S.BC(4, S.BC_SUBSONIC_INLET_TOTAL_PRESSURE, (p_tot, T_tot)),
           S.BC(2, S.BC_FIXED_PRESSURE_OUTLET, (p_exit,))]
    sch = S.Scheme(riemann=S.HLL, indicator=S.IND_BJ, phi=S.PHI_MICHALAK, limiter=S.LIM_MULTI, bcs=bcs)

    def ic(xyz):
        X = xyz[..., 0].ravel()
        p = p_tot + (p_exit - p_tot) * X
        rho = p / T_tot
        return _flat(_cons(rho, [0.1 + 0 * X, 0 * X], p))

    return Case("nozzle", m, 1, sch, 1.0, 0.2,
                params={"ic": ic, "area_ratio": area_ratio, "pressure_ratio": pressure_ratio})


def forward_step(nx=240, ny=80) -> Case:
    """Mach-3 forward-facing step: [0,3]x[0,1] minus the step, Q1 quads, HLLC,
    Venkatakrishnan limiter, supersonic inlet (rho 1.4, u 3, p 1), outlet, slip walls."""
    m = M.forward_step_quad(nx, ny)
    bcs = [S.BC(4, S.BC_SUPERSONIC_INLET, (1.4, 3.0, 0.0, 0.0, 1.0)), S.BC(2, S.BC_OUTLET),
           S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(5, S.BC_SLIP), S.BC(6, S.BC_SLIP)]
    sch = S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0,
                   limiter=S.LIM_MULTI, bcs=bcs)

    def ic(xyz):
        X = xyz[..., 0].ravel()
        return _flat(_cons(1.4 + 0 * X, [3.0 + 0 * X, 0 * X], 1.0 + 0 * X))

    return Case("forward_step", m, 1, sch, 4.0, 0.1, params={"ic": ic})


def nozzle_ensemble(count=256, seed=12345, ne_min=2000, ne_max=40000):
    """The batched ensemble: element counts log-uniform in [ne_min, ne_max], throat/inlet
    area ratio U[0.3, 0.6], inlet-total to exit pressure ratio U[1.5, 8]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        ne = int(round(math.exp(rng.uniform(math.log(ne_min), math.log(ne_max)))))
        ar = rng.uniform(0.3, 0.6)
        pr = rng.uniform(1.5, 8.0)
        out.append((ne, ar, pr))
    return out
