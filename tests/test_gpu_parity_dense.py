"""GPU parity of the dense small-element kernels (affine triangles P1-P3 on
jittered Delaunay meshes, bilinear quadrilaterals Q1-Q2 on nozzle meshes)
against the CPU oracle, through the C-ABI. Same bars as the tensor tests:
relative L-infinity per conserved variable <= 1e-12, wave speeds to 1e-13.
"""
import numpy as np
import pytest

from oracle.refcpu import RefSolver
from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from tests.helpers import conservative, rel_linf, smooth_state

TOL = 1e-12
pytestmark = pytest.mark.gpu


def _ctx(mesh, p, sch):
    from rkdg_amd.capi import Context

    return Context(mesh, p, sch)


def _dev(x, cuda):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x)).to(cuda)


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def _tri_mesh():
    return M.jittered_delaunay_tri(9, 5, 0.0, 1.0, 0.0, 1.0, jitter=0.3, seed=7)


def _nozzle(ne=120, r=0.45):
    return M.nozzle_quad(ne, r, length=1.0)


def _tri_bcs():
    return [S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(2, S.BC_OUTLET),
            S.BC(4, S.BC_SUPERSONIC_INLET, (1.3, 0.9, 0.05, 0.0, 1.2))]


def _nozzle_bcs():
    return [S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(2, S.BC_FIXED_PRESSURE_OUTLET, (0.8,)),
            S.BC(4, S.BC_SUBSONIC_INLET_TOTAL_PRESSURE, (1.6, 1.2))]


def _state(ref, seed, amp=1e-3):
    return smooth_state(ref.node_coords(), seed=seed, amp=amp)


def _rhs_check(mesh, p, sch, cuda, seed=0):
    import torch

    ref = RefSolver(mesh, p, sch)
    x = _state(ref, seed)
    y_ref, lam_ref = ref.mult(x)
    ctx = _ctx(mesh, p, sch)
    np.testing.assert_allclose(ctx.node_coords(), ref.node_coords(), rtol=0, atol=1e-14)
    xd = _dev(x, cuda)
    yd = torch.empty_like(xd)
    lam = ctx.rhs(xd, yd)
    err = rel_linf(_host(yd), y_ref, 4)
    return err, lam, lam_ref


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("rs", [S.RUSANOV, S.LLF, S.HLL, S.HLLC])
def test_triangle_rhs_parity(p, rs, cuda):
    err, lam, lam_ref = _rhs_check(_tri_mesh(), p, S.Scheme(riemann=rs, bcs=_tri_bcs()), cuda, seed=p)
    print(f"[tri p={p} rs={rs}] rel Linf {err:.3e}", flush=True)
    assert err <= TOL
    assert abs(lam - lam_ref) <= 1e-13 * lam_ref


def test_triangle_rhs_einfeldt_periodic_free(cuda):
    m = _tri_mesh()
    sch = S.Scheme(riemann=S.HLLC, wave_speeds=S.EINFELDT, bcs=[S.BC(a, S.BC_OUTLET) for a in (1, 2, 3, 4)])
    err, _, _ = _rhs_check(m, 2, sch, cuda, seed=3)
    assert err <= TOL


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("rs", [S.RUSANOV, S.HLL, S.HLLC])
def test_bilinear_quad_rhs_parity(p, rs, cuda):
    m = _nozzle()
    ctxinfo = _ctx(m, p, S.Scheme(riemann=rs, bcs=_nozzle_bcs())).info
    assert ctxinfo.affine == 0
    err, lam, lam_ref = _rhs_check(m, p, S.Scheme(riemann=rs, bcs=_nozzle_bcs()), cuda, seed=10 + p)
    print(f"[nozzle p={p} rs={rs}] rel Linf {err:.3e}", flush=True)
    assert err <= TOL
    assert abs(lam - lam_ref) <= 1e-13 * lam_ref


LIMS = [
    S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, phi=S.PHI_MICHALAK, limiter=S.LIM_MULTI),
    S.Scheme(riemann=S.HLL, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0, limiter=S.LIM_MULTI),
    S.Scheme(riemann=S.HLL, indicator=S.IND_BJ, phi=S.PHI_BJ, limiter=S.LIM_FINDIFF),
]


def _jump_state(ref, seed):
    xyz = ref.node_coords()
    x = smooth_state(xyz, seed=seed, amp=1e-2).reshape(4, -1)
    X = xyz.reshape(-1, 2)
    jump = (X[:, 0] > 0.45).astype(float)
    x[0] *= 1 + 2 * jump
    x[3] *= 1 + 3 * jump
    return np.ascontiguousarray(x.reshape(-1))


@pytest.mark.parametrize("li", range(len(LIMS)))
@pytest.mark.parametrize("kind", ["tri", "nozzle"])
def test_dense_limiter_parity(li, kind, cuda):
    import torch

    m, bcs = (_tri_mesh(), _tri_bcs()) if kind == "tri" else (_nozzle(), _nozzle_bcs())
    sch = S.Scheme(**{**LIMS[li].__dict__, "bcs": bcs})
    p = 2 if kind == "tri" else 1
    ref = RefSolver(m, p, sch)
    x = _jump_state(ref, li)
    x_ref, _, mv = ref.limit(x)
    ctx = _ctx(m, p, sch)
    xd = _dev(x, cuda)
    tr = torch.zeros(m.num_elements, dtype=torch.int8, device=cuda)
    ctx.limit(xd, tr)
    err = rel_linf(_host(xd), x_ref, 4)
    print(f"[dense limiter {kind} {li}] rel Linf {err:.3e}", flush=True)
    assert err <= TOL
    if sch.limiter == S.LIM_MULTI:
        flags = _host(tr) & 1
        near = np.abs(mv - 1.0) < 1e-12
        assert np.all((flags == (mv < 1.0)) | near)


@pytest.mark.parametrize("kind", ["tri", "nozzle"])
def test_dense_step_parity(kind, cuda):
    if kind == "tri":
        m, bcs, p = _tri_mesh(), _tri_bcs(), 2
    else:
        m, bcs, p = _nozzle(), _nozzle_bcs(), 1
    sch = S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, limiter=S.LIM_MULTI, bcs=bcs)
    ref = RefSolver(m, p, sch)
    # the projected initial state is limited before time stepping (main.cpp:264-265)
    x = ref.limit(_jump_state(ref, 5))[0]
    ctx = _ctx(m, p, sch)
    xd = _dev(x, cuda)
    dt = 0.1 * ref.hmin() / ref.face_mult(x)[1] / (2 * p + 1)
    for _ in range(3):
        x, lam_ref = ref.step(x, dt, S.TABLEAUS["ssprk3"])
        lam = ctx.step(xd, dt, S.TABLEAUS["ssprk3"])
    err = rel_linf(_host(xd), x, 4)
    print(f"[dense step {kind}] rel Linf {err:.3e}", flush=True)
    assert err <= TOL
    assert abs(lam - lam_ref) <= 1e-12 * lam_ref


def test_dense_conservation_and_hmin(cuda):
    import torch

    m = _tri_mesh()
    sch = S.Scheme(riemann=S.HLLC, bcs=[S.BC(a, S.BC_SLIP) for a in (1, 2, 3, 4)])
    ref = RefSolver(m, 2, sch)
    ctx = _ctx(m, 2, sch)
    assert abs(ctx.hmin() - ref.hmin()) <= 1e-14 * ref.hmin()
    # closed box with slip walls: mass and energy are conserved exactly
    x = _state(ref, 9, amp=1e-2)
    xd = _dev(x, cuda)
    yd = torch.empty_like(xd)
    ctx.rhs(xd, yd)
    means = torch.empty(4 * m.num_elements, dtype=torch.float64, device=cuda)
    ctx.means(yd, means)
    a = m.elem_coords[:, 1] - m.elem_coords[:, 0]
    b = m.elem_coords[:, 2] - m.elem_coords[:, 0]
    area = 0.5 * np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
    tot = (_host(means).reshape(4, -1) * area).sum(axis=1)
    scale = np.abs(_host(yd)).reshape(4, -1).max(axis=1) * area.sum()
    assert abs(tot[0]) <= 1e-13 * scale[0] and abs(tot[3]) <= 1e-13 * scale[3]
