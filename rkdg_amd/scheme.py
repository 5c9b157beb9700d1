"""Scheme description for the RKDG path: the choices the reference makes through
its RiemannSolver, BoundaryIntegrator, Indicator/Limiter and EoS plug-ins
(case_manager.cpp:523-612, 614-820), as plain enums and parameters.

Reference names map onto fields as follows:

* ``spatial.riemannSolver``: Rusanov | LLF | HLL | HLLC -> ``riemann``
* wave-speed estimate (ComputeToroCharSpeeds, physics.cpp:309-350, used by all
  rotated solvers; ComputeEinfeldtCharSpeeds, :269-296, selectable here)
* ``physics.{gamma, R, covolume}`` -> ``gamma, gas_constant, covolume``
* ``boundaryField.<name>.type`` + params -> ``bcs`` (attribute -> BC)
* ``spatial.limiter.{type, indicator, haveLastHope}`` -> ``limiter, indicator, last_hope``
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

RUSANOV, LLF, HLL, HLLC = 0, 1, 2, 3
TORO, EINFELDT = 0, 1

BC_SLIP = 0
BC_OUTLET = 1
BC_SUPERSONIC_INLET = 2
BC_SUBSONIC_INLET_TOTAL_PRESSURE = 3
BC_SUBSONIC_INLET_FIXED_PRESSURE = 4
BC_FIXED_PRESSURE_OUTLET = 5
BC_TOTAL_PRESSURE_OUTLET = 6
BC_CHAR_OUTLET = 7

IND_NOWHERE, IND_EVERYWHERE, IND_BJ, IND_SHU = 0, 1, 2, 3
PHI_MICHALAK, PHI_BJ, PHI_VENKAT = 0, 1, 2
LIM_NONE, LIM_FINDIFF, LIM_MULTI = 0, 1, 2

MAX_BC = 16

RIEMANN_NAMES = {"Rusanov": RUSANOV, "LLF": LLF, "HLL": HLL, "HLLC": HLLC}
BC_NAMES = {
    "slip": BC_SLIP,
    "outlet": BC_OUTLET,
    "supersonicInlet": BC_SUPERSONIC_INLET,
    "subsonicInletTotalPressure": BC_SUBSONIC_INLET_TOTAL_PRESSURE,
    "subsonicInletFixedPressure": BC_SUBSONIC_INLET_FIXED_PRESSURE,
    "fixedPressureOutlet": BC_FIXED_PRESSURE_OUTLET,
    "totalPressureFixedTemperatureOutlet": BC_TOTAL_PRESSURE_OUTLET,
    "charOutlet": BC_CHAR_OUTLET,
}


@dataclass
class BC:
    """One boundary patch: attribute, type and its parameters.

    Parameter order per type: supersonicInlet (rhoI, UI[0], UI[1], UI[2], pI);
    subsonicInletTotalPressure (pTot, TTot); subsonicInletFixedPressure (pFix, TTot);
    fixedPressureOutlet (pFix,); totalPressureFixedTemperatureOutlet (pTot, TFix);
    charOutlet (rhoI, UI[0], UI[1], UI[2], pI).
    """

    attr: int
    type: int
    params: tuple = ()


@dataclass
class Scheme:
    riemann: int = RUSANOV
    wave_speeds: int = TORO
    hllc_abs: int = 1
    gamma: float = 1.4
    gas_constant: float = 1.0
    covolume: float = 0.0
    bcs: List[BC] = field(default_factory=list)
    indicator: int = IND_NOWHERE
    phi: int = PHI_MICHALAK
    venkat_k: float = 0.0
    limiter: int = LIM_NONE
    last_hope: int = 1
    pow_nd: int = 1
    linearize: int = 0


# RK tableaus consumed by the packed a[l++] rule (rk_explicit_limited.cpp:51-58).
# ref_literal_rk2/rk3 are the literals of case_manager.cpp:829-884; ssprk3 is the
# intended Shu-Osher SSP-RK3 written as a Butcher tableau.
TABLEAUS = {
    "ref_literal_rk2": (2, [0.0, 1.0, 0.0], [0.5, 0.5], [0.0, 1.0]),
    "ref_literal_rk3": (3, [0.0, 1.0, 0.0, 0.25, 0.25, 0.0],
                        [0.1666666666666666, 0.1666666666666666, 0.6666666666666666], [0.0, 0.5]),
    "ssprk3": (3, [1.0, 0.25, 0.25], [1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0], [1.0, 0.5]),
    "euler": (1, [], [1.0], []),
}
