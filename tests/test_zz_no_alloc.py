"""rkdg_step_noalloc (stage vectors in the static per-device scratch array)
against rkdg_step through the C-ABI: bit-identical states and the same wave
speed on two calls in a row, the second on a smaller context, and a clear
error for a state beyond the scratch limit."""
import ctypes

import numpy as np
import pytest

from rkdg_amd import capi
from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from rkdg_amd.capi import Context
from tests.helpers import smooth_state

pytestmark = pytest.mark.gpu


def _pair(cuda, m, order, sch, seed, steps=2):
    import torch

    ctx = Context(m, order, sch)
    x0 = torch.from_numpy(smooth_state(ctx.node_coords(), seed=seed)).to(cuda)
    xa, xb = x0.clone(), x0.clone()
    lam_dev = torch.zeros(1, dtype=torch.float64, device=cuda)
    tab = S.TABLEAUS["ssprk3"]
    for _ in range(steps):
        lam_a = ctx.step(xa, 1e-3, tab)
        ctx.step_noalloc(xb, 1e-3, tab, lam_dev)
        torch.cuda.synchronize()
        ctx.check_error()
        assert float(lam_dev.item()) == lam_a
    assert torch.equal(xa, xb)
    ctx.close()
    return lam_a


def test_noalloc_step_matches_step_twice_second_smaller(cuda):
    sch = S.Scheme(riemann=S.RUSANOV, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0, limiter=S.LIM_MULTI)
    print("noalloc: Q2 hex 6^3 then 4^3", flush=True)
    _pair(cuda, M.structured_hex(6, 6, 6), 2, sch, seed=3)
    _pair(cuda, M.structured_hex(4, 4, 4), 2, sch, seed=4)


def test_noalloc_step_dense_quads_and_triangles(cuda):
    bcs = [S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(2, S.BC_OUTLET), S.BC(4, S.BC_OUTLET)]
    sch = S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, phi=S.PHI_MICHALAK, limiter=S.LIM_MULTI, bcs=bcs)
    print("noalloc: Q1 quads 20x6 then P2 triangles", flush=True)
    _pair(cuda, M.structured_quad(20, 6, 0.0, 1.0, 0.0, 0.3), 1, sch, seed=5)
    tri = M.jittered_delaunay_tri(8, 4, 0.0, 1.0, 0.0, 0.5, seed=7)
    _pair(cuda, tri, 2, sch, seed=6)


def test_noalloc_rejects_states_beyond_the_scratch(cuda):
    import torch

    sch = S.Scheme(riemann=S.RUSANOV)
    ctx = Context(M.structured_hex(32, 32, 32), 2, sch)
    assert ctx.info.state_size > (1 << 22)
    x = torch.ones(ctx.info.state_size, dtype=torch.float64, device=cuda)
    with pytest.raises(capi.RKDGError) as ei:
        ctx.step_noalloc(x, 1e-3, S.TABLEAUS["ssprk3"])
    assert ei.value.code == capi.RKDG_E_ARG and "allocation-free limit" in str(ei.value)
    ctx.close()
