"""Committed fixtures: Toro's published exact Riemann solutions pin the exact
solver the Sod test measures the oracle against; the oracle vectors
(tests/golden/oracle_vectors.npz, written by tests/golden/make_vectors.py)
pin the oracle against drift (CPU) and are the parity target of the HIP path
(GPU, through the C-ABI) on five cases covering every element family, Riemann
solver family, limiter and boundary-condition group of the path."""
import json
import math
import os

import numpy as np
import pytest

from tests import golden_cases as G
from tests.exact_riemann import star_state
from tests.helpers import rel_linf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL_GPU = 1e-12


def _fixture():
    return G.load(os.path.join(ROOT, G.FIXTURE))


def test_exact_riemann_matches_toro_table():
    d = json.load(open(os.path.join(ROOT, "tests/golden/toro_exact_riemann.json")))
    g = d["gamma"]
    m = (g - 1.0) / (g + 1.0)

    def rho_star(rk, pk, ps):  # shock (Rankine-Hugoniot) or rarefaction (isentrope)
        return rk * (ps / pk + m) / (m * ps / pk + 1.0) if ps > pk else rk * (ps / pk) ** (1.0 / g)

    for t in d["tests"]:
        rL, uL, pL, rR, uR, pR = t["in"]
        ps, us = star_state(rL, uL, pL, rR, uR, pR, g)
        got = [ps, us, rho_star(rL, pL, ps), rho_star(rR, pR, ps)]
        for v, s in zip(got, t["out"]):
            decimals = len(s.split(".")[1]) if "." in s else 0
            # the printed value is rounded to its last digit; the inputs of test 5
            # are themselves 6-digit roundings of test 3 and 4 outputs
            tol = 0.5 * 10.0 ** (-decimals) + 1e-5 * abs(float(s))
            assert abs(v - float(s)) <= tol, (t["in"], s, v)


def test_oracle_reproduces_golden_vectors():
    from oracle.refcpu import RefSolver
    from rkdg_amd import scheme as S

    for name, c in _fixture().items():
        ref = RefSolver(c["mesh"], c["p"], c["scheme"])
        neq = c["mesh"].dim + 2
        y, lam = ref.mult(c["x"])
        assert rel_linf(y, c["y"], neq) <= 1e-14, name
        assert abs(lam - c["lam"]) <= 1e-15 * c["lam"], name
        xl, _, _ = ref.limit(c["x"])
        assert rel_linf(xl, c["xl"], neq) <= 1e-14, name
        xs = c["x"].copy()
        for _ in range(c["steps"]):
            xs, lam_s = ref.step(xs, c["dt"], S.TABLEAUS["ssprk3"])
        assert rel_linf(xs, c["xs"], neq) <= 1e-14, name
        assert abs(lam_s - c["lam_s"]) <= 1e-15 * c["lam_s"], name


def test_golden_cases_match_their_definitions():
    """The fixture was written from the current case definitions (meshes and
    schemes), so a changed definition cannot silently test something else."""
    fx = _fixture()
    defs = G.cases()
    assert sorted(fx) == sorted(defs)
    for name, (mk, p, sch, _, dt, steps) in defs.items():
        m = mk()
        c = fx[name]
        assert c["p"] == p and c["dt"] == dt and c["steps"] == steps
        assert G.scheme_to_json(c["scheme"]) == G.scheme_to_json(sch)
        assert np.array_equal(c["mesh"].elem_verts, m.elem_verts)
        assert np.array_equal(c["mesh"].elem_coords, m.elem_coords)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(G.cases()))
def test_hip_matches_golden_vectors(name, cuda):
    import torch
    from rkdg_amd import scheme as S
    from rkdg_amd.capi import Context

    c = _fixture()[name]
    m, neq = c["mesh"], c["mesh"].dim + 2
    ctx = Context(m, c["p"], c["scheme"])
    xd = torch.from_numpy(c["x"].copy()).to(cuda)
    yd = torch.empty_like(xd)
    lam = ctx.rhs(xd, yd)
    torch.cuda.synchronize()
    err_y = rel_linf(yd.cpu().numpy(), c["y"], neq)
    tr = torch.zeros(m.num_elements, dtype=torch.int8, device=cuda)
    xl = xd.clone()
    ctx.limit(xl, tr)
    torch.cuda.synchronize()
    err_l = rel_linf(xl.cpu().numpy(), c["xl"], neq)
    for _ in range(c["steps"]):
        lam_s = ctx.step(xd, c["dt"], S.TABLEAUS["ssprk3"])
    torch.cuda.synchronize()
    err_s = rel_linf(xd.cpu().numpy(), c["xs"], neq)
    print(f"[golden {name}] rhs {err_y:.2e} limiter {err_l:.2e} steps {err_s:.2e} "
          f"lam {abs(lam - c['lam']) / c['lam']:.2e} {abs(lam_s - c['lam_s']) / c['lam_s']:.2e}", flush=True)
    assert err_y <= TOL_GPU and err_l <= TOL_GPU and err_s <= TOL_GPU
    assert abs(lam - c["lam"]) <= 1e-13 * c["lam"]
    assert abs(lam_s - c["lam_s"]) <= 1e-12 * c["lam_s"]
    ctx.close()
