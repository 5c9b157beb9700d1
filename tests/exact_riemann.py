"""Exact solution of the 1D Riemann problem for the ideal-gas Euler equations
(Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, ch. 4):
Newton iteration on the pressure function, then sampling. Used to pin the
oracle's wave-speed estimate and its Sod shock-tube solution.
"""
import math


def _fk(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    # rarefaction
    r = p / pk
    return (2.0 * ck / (g - 1.0)) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0), \
        (1.0 / (rho * ck)) * r ** (-(g + 1.0) / (2.0 * g))


def star_state(rL, uL, pL, rR, uR, pR, g=1.4):
    cL = math.sqrt(g * pL / rL)
    cR = math.sqrt(g * pR / rR)
    p = max(1e-12, 0.5 * (pL + pR))
    for _ in range(200):
        fL, dL = _fk(p, rL, pL, cL, g)
        fR, dR = _fk(p, rR, pR, cR, g)
        dp = (fL + fR + (uR - uL)) / (dL + dR)
        pn = max(1e-14, p - dp)
        if abs(pn - p) < 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fL, _ = _fk(p, rL, pL, cL, g)
    fR, _ = _fk(p, rR, pR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(s, rL, uL, pL, rR, uR, pR, g=1.4):
    """Primitive (rho, u, p) at similarity coordinate s = x/t."""
    p, u = star_state(rL, uL, pL, rR, uR, pR, g)
    cL = math.sqrt(g * pL / rL)
    cR = math.sqrt(g * pR / rR)
    if s <= u:
        if p > pL:
            SL = uL - cL * math.sqrt((g + 1) / (2 * g) * p / pL + (g - 1) / (2 * g))
            if s <= SL:
                return rL, uL, pL
            return rL * ((p / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pL + 1)), u, p
        SHL = uL - cL
        if s <= SHL:
            return rL, uL, pL
        cs = cL * (p / pL) ** ((g - 1) / (2 * g))
        STL = u - cs
        if s >= STL:
            return rL * (p / pL) ** (1 / g), u, p
        uu = 2 / (g + 1) * (cL + (g - 1) / 2 * uL + s)
        c = 2 / (g + 1) * (cL + (g - 1) / 2 * (uL - s))
        rr = rL * (c / cL) ** (2 / (g - 1))
        return rr, uu, pL * (c / cL) ** (2 * g / (g - 1))
    if p > pR:
        SR = uR + cR * math.sqrt((g + 1) / (2 * g) * p / pR + (g - 1) / (2 * g))
        if s >= SR:
            return rR, uR, pR
        return rR * ((p / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pR + 1)), u, p
    SHR = uR + cR
    if s >= SHR:
        return rR, uR, pR
    cs = cR * (p / pR) ** ((g - 1) / (2 * g))
    STR = u + cs
    if s <= STR:
        return rR * (p / pR) ** (1 / g), u, p
    uu = 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + s)
    c = 2 / (g + 1) * (cR - (g - 1) / 2 * (uR - s))
    rr = rR * (c / cR) ** (2 / (g - 1))
    return rr, uu, pR * (c / cR) ** (2 * g / (g - 1))
