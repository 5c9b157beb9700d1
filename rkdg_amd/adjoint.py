"""Discrete adjoint of the RKDG time loop, and its PyTorch checking harness.

The objective is the time-integrated pressure force on one boundary patch,
J = sum_{n=1..N} dt_{n-1} * integral over the patch of p(x_n), for the
explicit RK steps x_{n+1} = Step(x_n, dt_n) of FE_Evolution + the limited
Runge-Kutta solver (src/fe_evolution/fe_evolution.cpp:102-216,
src/ode_solver/rk_explicit_limited.cpp:30-82, src/limiter/limiter.cpp:22-94),
starting from the limited projected initial state (main.cpp:265). The dt
schedule is taken from one run of the reference time loop at the base
parameters (main.cpp:467-553) and then held fixed, so J is a function of the
initial state and the boundary parameters only.

The gradient comes from the C-ABI (rkdg_adjoint_solve: sqrt(N) checkpointing,
hand-written backward kernels); :class:`RKDGObjective` wraps it as a
``torch.autograd.Function`` for checking against finite differences only.
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as M
from . import scheme as S
from .capi import Context


class AdjointProblem:
    """A dense 2D context with an objective patch (boundary attribute obj_attr)."""

    def __init__(self, mesh, order: int, scheme: S.Scheme, obj_attr: int, tableau: str = "ssprk3"):
        self.ctx = Context(mesh, order, scheme, dense=True)
        self.scheme = scheme
        self.order = order
        self.tableau = S.TABLEAUS[tableau]
        slots = [i for i, bc in enumerate(scheme.bcs) if bc.attr == obj_attr]
        if not slots:
            raise ValueError(f"no boundary condition with attribute {obj_attr}")
        self.obj_slot = slots[0]
        self.size = self.ctx.state_size

    def bc_slot(self, attr: int) -> int:
        return [i for i, bc in enumerate(self.scheme.bcs) if bc.attr == attr][0]

    def set_bc(self, slot: int, params):
        self.ctx.set_bc_params(slot, params)

    def limit(self, x):
        y = x.clone()
        self.ctx.limit(y)
        return y

    def dt_schedule(self, x0, nsteps: int, cfl: float):
        """dt_n of the reference time loop (first dt from face speeds, then from
        the last stage's wave speed), from the limited initial state x0."""
        x = x0.clone()
        p = self.order
        hmin = self.ctx.hmin()
        dt = cfl * hmin / self.ctx.face_lambda(x) / (2 * p + 1)
        dts = []
        for _ in range(nsteps):
            dts.append(dt)
            lam = self.ctx.step(x, dt, self.tableau)
            dt = cfl * hmin / lam / (2 * p + 1)
        return np.array(dts)

    def objective(self, x0, dts):
        x = x0.clone()
        return self.ctx.forward_solve(x, dts, self.tableau, self.obj_slot)

    def gradient(self, x0, dts, every: int = 0):
        """(J, dJ/dx0, dJ/d bc_params[MAX_BC, 8])"""
        import torch

        x0bar = torch.empty_like(x0)
        J, pbar = self.ctx.adjoint_solve(x0, dts, self.tableau, self.obj_slot, x0bar, every=every)
        return J, x0bar, pbar


def _function_class():
    import torch

    class RKDGObjective(torch.autograd.Function):
        """J(x_ic, params) for a fixed dt schedule: x0 = Limit(x_ic), N steps,
        pressure objective. ``params`` are the 8 parameters of boundary slot
        ``slot``. Backward: rkdg_limit_adjoint o rkdg_adjoint_solve."""

        @staticmethod
        def forward(ctx, x_ic, params, prob: AdjointProblem, dts, slot: int, every: int = 0):
            prob.set_bc(slot, params.detach().cpu().numpy())
            x0 = prob.limit(x_ic.detach())
            J = prob.objective(x0, dts)
            ctx.save_for_backward(x_ic.detach(), params.detach())
            ctx.prob, ctx.dts, ctx.slot, ctx.every = prob, dts, slot, every
            return torch.tensor(J, dtype=torch.float64, device=x_ic.device)

        @staticmethod
        def backward(ctx, gJ):
            x_ic, params = ctx.saved_tensors
            prob = ctx.prob
            prob.set_bc(ctx.slot, params.cpu().numpy())
            x0 = prob.limit(x_ic)
            _, x0bar, pbar = prob.gradient(x0, ctx.dts, ctx.every)
            xbar = torch.empty_like(x_ic)
            prob.ctx.limit_adjoint(x_ic, x0bar, xbar)
            pb = torch.from_numpy(pbar[ctx.slot].copy()).to(params.device)
            return gJ * xbar, gJ * pb, None, None, None, None

    return RKDGObjective


_RKDGObjective = None


def rkdg_objective(x_ic, params, prob: AdjointProblem, dts, slot: int, every: int = 0):
    global _RKDGObjective
    if _RKDGObjective is None:
        _RKDGObjective = _function_class()
    return _RKDGObjective.apply(x_ic, params, prob, dts, slot, every)


# ---------------------------------------------------------------------------
# The Mach 3 forward-facing step with a seeded density perturbation

INLET_ATTR, OUTLET_ATTR, STEP_FRONT_ATTR = 4, 2, 5


def ffs_scheme(mach: float = 3.0, gamma: float = 1.4, rho: float = 1.4, p: float = 1.0, v_in: float = 0.0) -> S.Scheme:
    """HLLC, Venkatakrishnan (K = 5) multiplier limiter; supersonic inlet (rho,
    M c, v_in, 0, p), outlet, slip walls; the step front (attribute 5) is a wall."""
    c = math.sqrt(gamma * p / rho)
    bcs = [S.BC(INLET_ATTR, S.BC_SUPERSONIC_INLET, (rho, mach * c, v_in, 0.0, p)), S.BC(OUTLET_ATTR, S.BC_OUTLET),
           S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(STEP_FRONT_ATTR, S.BC_SLIP), S.BC(6, S.BC_SLIP)]
    return S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0, limiter=S.LIM_MULTI,
                    bcs=bcs, gamma=gamma)


class FFSProblem:
    """Forward-facing step on [0,3]x[0,1] minus [0.6,3]x[0,0.2] (nx, ny multiples
    of 5 so the step lies on grid lines), P1 on the dense quadrilateral kernels.
    Parameters: inflow Mach number and a nodal density perturbation field; the
    initial state is the inflow state with rho(1 + delta) (velocity and pressure
    unchanged).

    With v_in = 0 (the reference case) every transverse face starts with zero
    normal velocity on both sides, i.e. exactly at HLLC's star-side switch
    (S* = 0), where the flux is continuous but not differentiable; the adjoint
    differentiates the side the forward took. v_in != 0 moves the flow off it."""

    def __init__(self, nx=240, ny=80, mach=3.0, gamma=1.4, rho=1.4, p=1.0, v_in=0.0):
        if nx % 5 or ny % 5:
            raise ValueError("nx and ny must be multiples of 5")
        self.mesh = M.forward_step_quad(nx, ny)
        self.gamma, self.rho, self.p, self.mach, self.v_in = gamma, rho, p, mach, v_in
        self.c = math.sqrt(gamma * p / rho)
        self.prob = AdjointProblem(self.mesh, 1, ffs_scheme(mach, gamma, rho, p, v_in), STEP_FRONT_ATTR)
        self.inlet_slot = self.prob.bc_slot(INLET_ATTR)
        self.nodes = self.prob.ctx.info.num_dofs

    def perturbation(self, seed=1234, amp=1e-3):
        return amp * np.random.default_rng(seed).uniform(-1.0, 1.0, self.nodes)

    def initial_state(self, mach, delta):
        """torch: x_ic(M, delta) in the byNODES layout (differentiable in both)."""
        import torch

        rho = self.rho * (1.0 + delta)
        u = mach * self.c
        v = self.v_in
        E = self.p / (self.gamma - 1.0) + 0.5 * rho * (u * u + v * v)
        return torch.cat([rho, rho * u, rho * v, E])

    def inlet_params(self, mach):
        """torch [8]: supersonic-inlet parameters (rho, M c, v_in, 0, p)"""
        import torch

        z = torch.zeros((), dtype=torch.float64, device=mach.device)
        return torch.stack([z + self.rho, mach * self.c, z + self.v_in, z, z + self.p, z, z, z])

    def J(self, mach, delta, dts):
        """Differentiable objective (autograd through the adjoint)."""
        x_ic = self.initial_state(mach, delta)
        return rkdg_objective(x_ic, self.inlet_params(mach), self.prob, dts, self.inlet_slot)

    def schedule(self, nsteps, cfl=0.1, device="cuda"):
        import torch

        m = torch.tensor(self.mach, dtype=torch.float64, device=device)
        d = torch.zeros(self.nodes, dtype=torch.float64, device=device)
        self.prob.set_bc(self.inlet_slot, self.inlet_params(m).cpu().numpy())
        x0 = self.prob.limit(self.initial_state(m, d))
        return self.prob.dt_schedule(x0, nsteps, cfl)
