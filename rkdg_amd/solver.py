"""Host-side mirror of the reference's operator interfaces for the RKDG path.

The names, constructor arguments and call semantics follow the reference so a
caller ports line by line; the work runs on the GPU through the C-ABI
(include/rkdg_capi.h):

* ``FE_Evolution``            src/fe_evolution/fe_evolution.hpp:27-82
  (``Mult``, ``UpdateInverseMassMatrix``/``UpdateAfluxPointer`` after re-meshing)
* ``Limiter``                 src/limiter/limiter.hpp:24-116 (``update``)
* ``ExplicitRKLimitedSolver`` src/ode_solver/rk_explicit_limited.hpp:24-60
  (``Init``, ``Step`` -- Step does not advance t, the driver does)
* ``run``                     the time loop of main.cpp:467-553, 701, 733
  (first dt from face wave speeds, dt = cfl*hmin/lambda/(2p+1), the step uses
  min(dt, t_end - t) but t advances by the recomputed dt)
* ``Ensemble``/``run_batch``  that time loop for many independent instances at
  once, on the device (no reference counterpart: the reference runs one case
  per process)

Vectors are torch tensors on the GPU (device memory only); the byNODES layout
x[v*NE*nd + e*nd + i] is the reference's.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import scheme as S
from .capi import BatchContext, Context


class FE_Evolution:
    """TimeDependentOperator: y = M^-1 [face terms + (F(u), grad w)] (fe_evolution.cpp:102-216)."""

    def __init__(self, mesh, order: int, scheme: S.Scheme, dense: bool = False):
        self.ctx = Context(mesh, order, scheme, dense=dense)
        self.dense = dense
        self.order = order
        self.max_char_speed = 0.0
        self.t = 0.0

    @property
    def width(self) -> int:
        return self.ctx.state_size

    def SetTime(self, t: float) -> None:
        self.t = t

    def Mult(self, x, y) -> None:
        """Writes y; max_char_speed is reset and updated as in the reference (fe_evolution.cpp:105)."""
        self.max_char_speed = self.ctx.rhs(x, y)

    def FaceMaxCharSpeed(self, x) -> float:
        """A.Mult(sol, z) used only for its wave speed before the first step (main.cpp:473-484)."""
        return self.ctx.face_lambda(x)

    def UpdateMesh(self, mesh, scheme: S.Scheme = None) -> None:
        """After a re-mesh: UpdateInverseMassMatrix + UpdateAfluxPointer (main.cpp:610-611)."""
        sch = scheme if scheme is not None else self.ctx.scheme
        self.ctx.close()
        self.ctx = Context(mesh, self.order, sch, dense=self.dense)

    UpdateInverseMassMatrix = UpdateMesh

    def hmin(self) -> float:
        return self.ctx.hmin()


class Limiter:
    """Limiter::update(Vector&) (limiter.cpp:22-94) with the scheme's indicator/limiter."""

    def __init__(self, evolution: FE_Evolution):
        self.f = evolution

    def update(self, x, troubled=None) -> None:
        self.f.ctx.limit(x, troubled)


class ExplicitRKLimitedSolver:
    """Butcher RK with limiting after each stage and at step end; the packed
    a[l++] consumption of rk_explicit_limited.cpp:51-58."""

    def __init__(self, s: int, a, b, c, limiter: Limiter):
        self.tableau = (s, list(a), list(b), list(c))
        self.limiter = limiter
        self.f = None

    @classmethod
    def from_order(cls, order: int, limiter: Limiter, literal: bool = False):
        """time.order 2|3 as CaseManager::loadTimeSolver; literal=True takes the reference's
        literal tableau (case_manager.cpp:829-884), otherwise SSP-RK3 for order 3."""
        name = {2: "ref_literal_rk2", 3: "ref_literal_rk3" if literal else "ssprk3"}[order]
        s, a, b, c = S.TABLEAUS[name]
        return cls(s, a, b, c, limiter)

    def Init(self, f: FE_Evolution) -> None:
        self.f = f

    def Step(self, x, t: float, dt: float):
        """Advances x by one step of size dt in place; returns (t, dt) unchanged and
        leaves the last stage's wave speed in f.max_char_speed."""
        self.f.max_char_speed = self.f.ctx.step(x, dt, self.tableau, t=t)
        return t, dt


class SSPRK3LimitedSolver(ExplicitRKLimitedSolver):
    """SSP-RK3 in the Shu-Osher form with the limiter applied to each stage
    value (the classical limited RKDG step; no reference counterpart). Without
    limiting it equals ExplicitRKLimitedSolver with the SSP-RK3 tableau to
    rounding; with limiting the two forms differ (the Butcher form limits
    stage inputs it does not reuse)."""

    def __init__(self, limiter: Limiter):
        super().__init__(*S.TABLEAUS["ssprk3"], limiter=limiter)

    def Step(self, x, t: float, dt: float):
        self.f.max_char_speed = self.f.ctx.step_ssprk3(x, dt, t=t)
        return t, dt


@dataclass
class RunStats:
    steps: int
    t: float
    dt: float
    max_char_speed: float


def run(f: FE_Evolution, solver: ExplicitRKLimitedSolver, limiter: Limiter, sol, t_final: float,
        cfl: float, t0: float = 0.0, dt_const: float = None, max_steps: int = None,
        callback=None, dt0: float = None) -> RunStats:
    """The time loop of main.cpp: limits the projected state (main.cpp:265), takes
    the first dt from face wave speeds (:467-484), then steps with
    dt_real = min(dt, t_end - t) (:512), recomputes dt from the last RHS
    (:538-553) and advances t by that dt (:701) until t >= t_end - 1e-8 dt (:733).
    ``dt_const`` replaces the dynamic dt (time.type: constant); ``dt0`` is the
    first dt of a restarted run (the frame's saved dt, main.cpp:485-488)."""
    limiter.update(sol)
    solver.Init(f)
    p = f.order
    hmin = f.hmin()
    t = t0
    if dt_const is not None:
        dt = dt_const
    elif dt0 is not None:
        dt = dt0
    else:
        lam = f.FaceMaxCharSpeed(sol)
        dt = cfl * hmin / lam / (2 * p + 1)
    steps = 0
    done = False
    while not done:
        dt_real = min(dt, t_final - t)
        solver.Step(sol, t, dt_real)
        if dt_const is None:
            dt = cfl * hmin / f.max_char_speed / (2 * p + 1)
        t += dt
        steps += 1
        if callback is not None:
            callback(steps, t, dt, sol)
        done = t >= t_final - 1e-8 * dt or (max_steps is not None and steps >= max_steps)
    return RunStats(steps, t, dt, f.max_char_speed)


class Ensemble:
    """Independent instances (same element type, order and scheme; boundary
    parameters may differ) advanced together: one launch per kernel per stage
    for all of them, each with its own dt, t and stop time."""

    def __init__(self, meshes, order: int, schemes):
        self.ctx = BatchContext(meshes, order, schemes)
        self.order = order
        self.count = len(meshes)
        self.offsets = self.ctx.offsets
        self.nd = self.ctx.info.nodes_per_element
        self.neq = self.ctx.info.num_equations

    def pack(self, states):
        """Per-instance byNODES states (each neq*NE_i*nd) -> the union state."""
        parts = [np.asarray(s).reshape(self.neq, -1) for s in states]
        return np.ascontiguousarray(np.concatenate(parts, axis=1).ravel())

    def unpack(self, x):
        """Union state (host array) -> per-instance states."""
        u = np.asarray(x).reshape(self.neq, -1)
        o = self.offsets * self.nd
        return [np.ascontiguousarray(u[:, o[i]:o[i + 1]].ravel()) for i in range(self.count)]

    def node_coords(self):
        xyz = self.ctx.node_coords()
        return [xyz[self.offsets[i]:self.offsets[i + 1]] for i in range(self.count)]


def run_batch(ens: Ensemble, x, t_final, cfl: float, tableau: str = "ssprk3", max_steps: int = 0,
              chunk: int = 16, t0=0.0):
    """``run`` for every instance of the ensemble: limits x, takes each first dt
    from the face wave speeds, then replays captured steps (``chunk`` per host
    check) until every instance has reached its t_final (or max_steps).
    Returns (t, dt, steps) per instance."""
    tab = S.TABLEAUS[tableau]
    ens.ctx.init(x, t_final, cfl, t0=t0, max_steps=max_steps)
    while ens.ctx.advance(x, tab, chunk) > 0:
        pass
    t, dt, steps, _ = ens.ctx.status()
    return t, dt, steps
