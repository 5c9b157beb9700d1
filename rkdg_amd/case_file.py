"""YAML case files for the RKDG path: the settings half of CaseManager
(src/case_manager/case_manager.cpp) and the initial conditions of
src/initial_conditions/ic.hpp.

``load_case(path)`` reads a case directory's ``settings.yml`` with the keys the
reference reads and returns a ``cases.Case`` whose mesh, scheme, initial state,
time settings and output settings are the ones the reference would run:

* ``mesh``: ``file`` (Gmsh, relative to the case directory),
  ``serialRefLevels`` + ``parallelRefLevels`` uniform refinements,
  ``adaptive`` with ``maxElemError``, ``totalErrorFraction``, ``hysteresis``,
  ``nonConformingLimit`` (1) and ``preferConformingRefinement`` (false) --
  case_manager.cpp:450-466: non-conforming refinement of quadrilateral and
  hexahedral meshes (``amr.run_adaptive``);
* ``physics``: ``type`` (``covolume`` reads ``covolume``, anything else is a
  perfect gas), ``gamma``, ``R`` (default 1) -- case_manager.cpp:83-101;
* ``internalField``: ``constant``, ``planeBreakup``, ``sphericalBreakup``,
  ``densityPulse`` -- case_manager.cpp:121-215, ic.hpp;
* ``boundaryField``: one entry per patch, matched BY NAME against the mesh's
  $PhysicalNames; a name the mesh lacks is skipped, as the reference does
  (case_manager.cpp:681-695); ``type`` and its parameters as
  case_manager.cpp:699-815; boundary faces without a patch get no flux;
* ``spatial``: ``polyOrder``, ``riemannSolver`` (Rusanov, LLF, HLL, HLLC),
  ``limiter.type`` (Multi, FinDiff), ``limiter.indicator`` (Nowhere,
  Everywhere, BJ, Shu -- with FinDiff), ``limiter.haveLastHope`` (default 1),
  ``limiter.linearize`` (default 0; p = 1) -- case_manager.cpp:523-612;
* ``time``: ``order`` (2 or 3, the reference's literal tableaus,
  case_manager.cpp:822-891), ``tEnd``, ``dt``, ``type`` (``constant`` keeps dt,
  anything else takes dt from ``CoMax``; :943-958), ``restart``,
  ``restartCycle``, ``nSavedFrames``;
* ``postProcess``: ``visSteps``, ``checkTotalEnergy``, ``writeIndicators``.

Settings the reference exits on (an unknown Riemann solver, limiter,
indicator, boundary type or RK order) raise ``CaseError``. Keys beyond the
reference's, for workloads it cannot express:

* ``spatial.limiter.phi`` (michalak | bj | venkat) and ``venkatK``: the
  BJ-family limiter function (the reference compiles Michalak in);
* ``spatial.waveSpeeds`` (Toro | Einfeldt);
* ``time.tableau`` (a name in ``scheme.TABLEAUS``) in place of ``order``;
* ``mesh.maxRefLevel``: a cap on the refinement level of adaptive runs (the
  reference has none; its ZZ estimator does not fall with h at a shock).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np
import yaml

from . import gmsh
from . import mesh as M
from . import scheme as S
from .cases import Case


class CaseError(ValueError):
    pass


# ic.hpp:485 writes pi as this literal
_PI_IC = 3.14169265


@dataclass
class FileCase(Case):
    """A case read from a settings file: ``Case`` plus the file's time and output
    settings and the mesh's physical names."""

    dt: float = 0.0
    time_type: str = "dynamic"
    restart: bool = False
    restart_cycle: int = 0
    n_saved_frames: int = 0
    vis_steps: int = 0
    check_total_energy: bool = False
    write_indicators: bool = False
    case_dir: str = ""
    bdr_names: dict = field(default_factory=dict)
    skipped_patches: list = field(default_factory=list)
    # adaptive refinement (mesh.adaptive, case_manager.cpp:450-466)
    adaptive: bool = False
    max_elem_error: float = 0.0
    total_error_fraction: float = 0.0
    hysteresis: float = 0.0
    max_ref_level: int = None

    @property
    def dt_const(self):
        """dt for ``time.type: constant`` (cfl = -1 in main.cpp:467-484), else None."""
        return self.dt if self.cfl < 0 else None


def _get(d, *keys, default=None, required=False):
    cur = d
    for k in keys:
        if not isinstance(cur, dict) or k not in cur or cur[k] is None:
            if required:
                raise CaseError("missing setting " + ".".join(keys))
            return default
        cur = cur[k]
    return cur


def _vec3(v, name):
    v = list(v)
    if len(v) != 3:
        raise CaseError(f"{name} needs three components")
    return [float(t) for t in v]


def _state(node, name):
    """(rho, u, v, w, p) of an internalField state block."""
    return (float(_get(node, "rho", default=0.0)), *_vec3(_get(node, "U", required=True), name + ".U"),
            float(_get(node, "p", required=True)))


def _energy(rho, u, v, w, p, gamma, b):
    """ComputeEnergy (physics.cpp:128-131)."""
    return p / (gamma - 1.0) * (1.0 - rho * b) + 0.5 * rho * (u * u + v * v + w * w)


class InitialCondition:
    """IC::setIC as a function of the node coordinates (..., dim): returns the
    conservative state flattened byNODES, [rho, rho u, rho v, (rho w), rho E]."""

    def __init__(self, kind: str, gamma: float, covolume: float):
        self.kind = kind
        self.gamma = gamma
        self.covolume = covolume

    def primitives(self, X):
        raise NotImplementedError

    def __call__(self, xyz):
        X = np.asarray(xyz, np.float64)
        dim = X.shape[-1]
        X = X.reshape(-1, dim)
        rho, u, v, w, p = (np.broadcast_to(np.asarray(a, np.float64), (len(X),)) for a in self.primitives(X))
        E = _energy(rho, u, v, w if dim == 3 else 0.0, p, self.gamma, self.covolume)
        comps = [rho, rho * u, rho * v] + ([rho * w] if dim == 3 else []) + [E]
        return np.ascontiguousarray(np.concatenate(comps))


class ICConstant(InitialCondition):
    """ic.hpp:54-128."""

    def __init__(self, state, gamma, covolume):
        super().__init__("constant", gamma, covolume)
        self.state = state

    def primitives(self, X):
        return self.state


class ICPlaneBreakup(InitialCondition):
    """ic.hpp:134-270: the left state where normal . (x - origin) < 0, else the right."""

    def __init__(self, left, right, origin, normal, gamma, covolume):
        super().__init__("planeBreakup", gamma, covolume)
        self.left, self.right = left, right
        self.origin, self.normal = np.asarray(origin), np.asarray(normal)

    def primitives(self, X):
        dim = X.shape[1]
        side = ((X - self.origin[:dim]) * self.normal[:dim]).sum(axis=1) < 0.0
        return [np.where(side, a, b) for a, b in zip(self.left, self.right)]


class ICSphericalBreakup(InitialCondition):
    """ic.hpp:276-422: the inner state inside the sphere of ``radius``; where
    |x - o|^2 - r^2 < 0.04 r^2 the value is the linear ramp from r to 1.2 r;
    outside that, the outer state."""

    def __init__(self, inner, outer, origin, radius, gamma, covolume):
        super().__init__("sphericalBreakup", gamma, covolume)
        self.inner, self.outer = inner, outer
        self.origin = np.asarray(origin)
        self.r2 = radius * radius
        self.r2d = 0.2 * 0.2 * self.r2

    def primitives(self, X):
        dim = X.shape[1]
        d2 = ((X - self.origin[:dim]) ** 2).sum(axis=1)
        s = d2 - self.r2
        dist = np.sqrt(d2)
        r1 = math.sqrt(self.r2)
        r2 = r1 + math.sqrt(self.r2d)
        out = []
        for a, b in zip(self.inner, self.outer):
            ramp = (dist - r1) * (b - a) / (r2 - r1) + a
            out.append(np.where(s < 0.0, a, np.where(s < self.r2d, ramp, b)))
        return out


class ICDensityPulse(InitialCondition):
    """ic.hpp:428-537: rho = (1 - (gamma-1) eps^2 exp(|x - o|^2 - 1) / (8 gamma pi^2))^(1/(gamma-1))
    with the source's literal for pi; velocity and pressure constant."""

    def __init__(self, state, origin, epsilon, gamma, covolume):
        super().__init__("densityPulse", gamma, covolume)
        self.state = state
        self.origin = np.asarray(origin)
        self.epsilon = epsilon

    def primitives(self, X):
        dim = X.shape[1]
        g = self.gamma
        q = ((X - self.origin[:dim]) ** 2).sum(axis=1) - 1.0
        rho = (1.0 - (g - 1.0) * self.epsilon * self.epsilon * np.exp(q) / (8.0 * g * _PI_IC * _PI_IC)) ** (1.0 / (g - 1.0))
        _, u, v, w, p = self.state
        return rho, u, v, w, p


def make_ic(field_node, gamma, covolume) -> InitialCondition:
    kind = _get(field_node, "type", required=True)
    if kind == "constant":
        return ICConstant(_state(field_node, "internalField"), gamma, covolume)
    if kind == "planeBreakup":
        return ICPlaneBreakup(_state(_get(field_node, "left", required=True), "left"),
                              _state(_get(field_node, "right", required=True), "right"),
                              _vec3(_get(field_node, "origin", required=True), "origin"),
                              _vec3(_get(field_node, "normal", required=True), "normal"), gamma, covolume)
    if kind == "sphericalBreakup":
        return ICSphericalBreakup(_state(_get(field_node, "left", required=True), "left"),
                                  _state(_get(field_node, "right", required=True), "right"),
                                  _vec3(_get(field_node, "origin", required=True), "origin"),
                                  float(_get(field_node, "radius", required=True)), gamma, covolume)
    if kind == "densityPulse":
        return ICDensityPulse(_state(field_node, "internalField"),
                              _vec3(_get(field_node, "origin", required=True), "origin"),
                              float(_get(field_node, "epsilon", required=True)), gamma, covolume)
    # the reference prints a message and carries on without an initial
    # condition (case_manager.cpp:205-213); a state cannot be made without one
    raise CaseError(f"unknown initial condition type {kind!r} (constant, planeBreakup, sphericalBreakup, densityPulse)")


def _bc_params(typ, node):
    """The parameter tuple of each boundary type (case_manager.cpp:702-801)."""
    f = lambda k: float(_get(node, k, required=True))  # noqa: E731
    if typ in ("slip", "outlet"):
        return ()
    if typ == "subsonicInletTotalPressure":
        return (f("pTot"), f("TTot"))
    if typ == "subsonicInletFixedPressure":
        return (f("pFix"), f("TTot"))
    if typ == "totalPressureFixedTemperatureOutlet":
        return (f("pTot"), f("TFix"))
    if typ == "fixedPressureOutlet":
        return (f("pFix"),)
    if typ in ("supersonicInlet", "charOutlet"):
        return (f("rhoI"), *_vec3(_get(node, "UI", required=True), "UI"), f("pI"))
    raise CaseError(f"wrong boundary type {typ!r}")


_INDICATORS = {"Nowhere": S.IND_NOWHERE, "Everywhere": S.IND_EVERYWHERE, "BJ": S.IND_BJ, "Shu": S.IND_SHU}
_LIMITERS = {"FinDiff": S.LIM_FINDIFF, "Multi": S.LIM_MULTI}
_PHI = {"michalak": S.PHI_MICHALAK, "bj": S.PHI_BJ, "venkat": S.PHI_VENKAT}
_SPEEDS = {"Toro": S.TORO, "Einfeldt": S.EINFELDT}


def _flag(v, default):
    if v is None:
        return default
    if isinstance(v, str):
        return v.strip().lower() in ("1", "true", "yes", "on")
    return bool(v)


class _UniqueKeyLoader(yaml.SafeLoader):
    """SafeLoader that refuses a key given twice in one mapping: the reference's
    parser keeps both (two boundaryField entries for one patch add two
    integrators and count its flux twice); a dict cannot hold both."""


def _unique_mapping(loader, node, deep=False):
    seen = set()
    for k, _ in node.value:
        key = loader.construct_object(k, deep=deep)
        if key in seen:
            raise CaseError(f"setting {key!r} appears twice in one block (line {k.start_mark.line + 1})")
        seen.add(key)
    return loader.construct_mapping(node, deep)


_UniqueKeyLoader.add_constructor(yaml.resolver.BaseResolver.DEFAULT_MAPPING_TAG, _unique_mapping)


def load_case(path) -> FileCase:
    """Reads a case: ``path`` is a settings file or a directory holding
    ``settings.yml`` (the reference reads ``<cwd>/<file>``, case_manager.cpp:38-41)."""
    if os.path.isdir(path):
        path = os.path.join(path, "settings.yml")
    case_dir = os.path.dirname(os.path.abspath(path))
    with open(path) as fh:
        st = yaml.load(fh, Loader=_UniqueKeyLoader) or {}

    # physics (case_manager.cpp:83-101)
    ph_type = _get(st, "physics", "type", default="perfectGas")
    gamma = float(_get(st, "physics", "gamma", required=True))
    covolume = float(_get(st, "physics", "covolume", required=True)) if ph_type == "covolume" else 0.0
    R = float(_get(st, "physics", "R", default=1.0))

    # mesh (case_manager.cpp:217-297)
    mfile = _get(st, "mesh", "file", required=True)
    adaptive = _flag(_get(st, "mesh", "adaptive"), False)
    amr_kw = {}
    if adaptive:
        ncl = int(_get(st, "mesh", "nonConformingLimit", default=1))
        if ncl != 1:
            raise CaseError("mesh.nonConformingLimit: one level of hanging nodes is built")
        if _flag(_get(st, "mesh", "preferConformingRefinement"), False):
            raise CaseError("mesh.preferConformingRefinement: quadrilaterals and hexahedra refine non-conformingly")
        mre = _get(st, "mesh", "maxRefLevel")
        amr_kw = dict(adaptive=True, max_elem_error=float(_get(st, "mesh", "maxElemError", required=True)),
                      total_error_fraction=float(_get(st, "mesh", "totalErrorFraction", default=0.0)),
                      hysteresis=float(_get(st, "mesh", "hysteresis", default=0.0)),
                      max_ref_level=None if mre is None else int(mre))
    gm = gmsh.read_gmsh(os.path.join(case_dir, mfile))
    levels = int(_get(st, "mesh", "serialRefLevels", default=0)) + int(_get(st, "mesh", "parallelRefLevels", default=0))
    mesh = gmsh.uniform_refinement(gm.mesh, levels) if levels > 0 else gm.mesh

    # initial state (case_manager.cpp:121-215)
    ic = make_ic(_get(st, "internalField", required=True), gamma, covolume)

    # spatial (case_manager.cpp:523-612)
    order = int(_get(st, "spatial", "polyOrder", required=True))
    rs = _get(st, "spatial", "riemannSolver", required=True)
    if rs not in S.RIEMANN_NAMES:
        raise CaseError(f"unknown Riemann solver type {rs!r}")
    lim = _get(st, "spatial", "limiter", default={}) or {}
    ind_name = lim.get("indicator", "notype")
    if ind_name not in _INDICATORS:
        raise CaseError(f"unknown indicator type {ind_name!r} (Nowhere, Everywhere, BJ)")
    lim_name = lim.get("type")
    if lim_name not in _LIMITERS:
        raise CaseError(f"unknown limiter type {lim_name!r} (FinDiff, Multi)")
    if ind_name == "Shu" and lim_name != "FinDiff":
        # IndicatorShu writes no minValues, which the multiplier reads
        # (indicator_shu.cpp:68-73, limiter_multiplier.cpp:35)
        raise CaseError("the Shu indicator sets no multiplier: pair it with the FinDiff limiter")
    phi = _PHI[str(lim.get("phi", "michalak")).lower()]
    speeds = _SPEEDS[_get(st, "spatial", "waveSpeeds", default="Toro")]

    # boundaries: patch names against $PhysicalNames (case_manager.cpp:614-820)
    bcs, skipped = [], []
    for name, node in (_get(st, "boundaryField", default={}) or {}).items():
        name = str(name)
        if name not in gm.bdr_names:
            skipped.append(name)
            continue
        typ = _get(node, "type", required=True)
        if typ not in S.BC_NAMES:
            raise CaseError(f"wrong boundary type {typ!r} for patch {name}")
        attr = gm.bdr_names[name]
        bcs.append(S.BC(attr, S.BC_NAMES[typ], _bc_params(typ, node)))

    sch = S.Scheme(riemann=S.RIEMANN_NAMES[rs], wave_speeds=speeds, gamma=gamma, gas_constant=R,
                   covolume=covolume, bcs=bcs, indicator=_INDICATORS[ind_name], phi=phi,
                   venkat_k=float(lim.get("venkatK", 0.0)), limiter=_LIMITERS[lim_name],
                   last_hope=int(_flag(lim.get("haveLastHope"), True)),
                   linearize=int(_flag(lim.get("linearize"), False)))

    # time (case_manager.cpp:822-891, 943-958)
    tname = _get(st, "time", "tableau")
    if tname is None:
        rk_order = int(_get(st, "time", "order", required=True))
        if rk_order not in (2, 3):
            raise CaseError(f"wrong RK scheme order {rk_order}")
        tname = {2: "ref_literal_rk2", 3: "ref_literal_rk3"}[rk_order]
    elif tname not in S.TABLEAUS:
        raise CaseError(f"unknown tableau {tname!r}")
    t_end = float(_get(st, "time", "tEnd", required=True))
    dt = float(_get(st, "time", "dt", default=0.0))
    ttype = _get(st, "time", "type", default="dynamic")
    cfl = -1.0 if ttype == "constant" else float(_get(st, "time", "CoMax", required=True))

    name = os.path.basename(case_dir)
    if adaptive and mesh.geom == M.GEOM_TRI:
        raise CaseError("mesh.adaptive: non-conforming refinement is built for quadrilaterals and hexahedra")
    if adaptive and _flag(_get(st, "time", "restart"), False):
        raise CaseError("time.restart with mesh.adaptive: restart frames hold no refinement tree")
    if adaptive and ind_name == "Shu":
        raise CaseError("mesh.adaptive: the Shu indicator is not built for non-conforming faces")
    return FileCase(name, mesh, order, sch, t_end, cfl, tableau=tname, params={"ic": ic}, **amr_kw,
                    dt=dt, time_type=ttype,
                    restart=_flag(_get(st, "time", "restart"), False),
                    restart_cycle=int(_get(st, "time", "restartCycle", default=0)),
                    n_saved_frames=int(_get(st, "time", "nSavedFrames", default=0)),
                    vis_steps=int(_get(st, "postProcess", "visSteps", default=0)),
                    check_total_energy=_flag(_get(st, "postProcess", "checkTotalEnergy"), False),
                    write_indicators=_flag(_get(st, "postProcess", "writeIndicators"), False),
                    case_dir=case_dir, bdr_names=dict(gm.bdr_names), skipped_patches=skipped)
