"""GPU parity of curved (trilinear, non-parallelepiped) hexahedra against the
CPU oracle, through the C-ABI.

The element maps are trilinear with every corner moved by a seeded random
offset of up to 15% of the spacing, so adj(J), det J, the face normals and
|n| vary inside each element and the tensor path takes its curved form (dense
inverse mass, adj(J) per node, normal per face point). Bars as for the affine
meshes: relative L-infinity per conserved variable <= 1e-12 for the
right-hand side, the limiter and RK steps; wave speeds to 1e-13. Beyond the
oracle, two properties the exact scheme fixes on any such mesh: a uniform
state has a zero right-hand side (the volume rule at the p+1 Gauss points and
the face rule integrate the metric identities exactly for trilinear maps) and
on a periodic mesh the integral of the right-hand side vanishes.
"""
import numpy as np
import pytest

from oracle.refcpu import RefSolver
from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from tests.helpers import conservative, curved_hex, rel_linf, scramble_orientations, smooth_state

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


def _rhs_check(mesh, p, sch, cuda, seed=0, amp=1e-3):
    """Relative L-infinity error of the GPU right-hand side against the oracle
    evaluated in long double, and the bar for it: 1e-12, or twice the double
    oracle's own error where that is larger (at p = 3 the dense inverse mass of
    a curved element amplifies rounding to a few 1e-12 in either code)."""
    import torch

    ref = RefSolver(mesh, p, sch)
    x = smooth_state(ref.node_coords(), seed=seed, amp=amp)
    y_ref, lam_ref = ref.mult(x)
    y_ld, _ = RefSolver(mesh, p, sch, precision="long double").mult(x)
    ctx = _ctx(mesh, p, sch)
    assert ctx.info.affine == 0
    np.testing.assert_allclose(ctx.node_coords(), ref.node_coords(), rtol=0, atol=1e-14)
    xd = _dev(x, cuda)
    yd = torch.empty_like(xd)
    lam = ctx.rhs(xd, yd)
    y = _host(yd)
    err, err_ref = rel_linf(y, y_ld, 5), rel_linf(y_ref, y_ld, 5)
    print(f"  vs long double: GPU {err:.2e} double oracle {err_ref:.2e}; GPU vs double oracle "
          f"{rel_linf(y, y_ref, 5):.2e}", flush=True)
    return err, max(TOL, 2.0 * err_ref), lam, lam_ref


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("rs", [S.RUSANOV, S.LLF, S.HLL, S.HLLC])
def test_curved_hex_rhs_parity(p, rs, cuda):
    m = curved_hex(4, 3, 3, seed=p)
    err, bar, lam, lam_ref = _rhs_check(m, p, S.Scheme(riemann=rs), cuda, seed=p)
    print(f"[curved p={p} rs={rs}] rel Linf {err:.3e} bar {bar:.1e} lam {lam} ref {lam_ref}", flush=True)
    assert err <= bar
    assert abs(lam - lam_ref) <= 1e-13 * lam_ref


@pytest.mark.parametrize("rs", [S.RUSANOV, S.HLLC])
def test_curved_hex_scrambled_orientation(rs, cuda):
    m = scramble_orientations(curved_hex(4, 3, 3, seed=5), seed=6)
    err, bar, _, _ = _rhs_check(m, 2, S.Scheme(riemann=rs), cuda, seed=6)
    print(f"[curved scrambled rs={rs}] rel Linf {err:.3e}", flush=True)
    assert err <= bar


@pytest.mark.parametrize("rs", [S.RUSANOV, S.LLF, S.HLLC])
@pytest.mark.parametrize("planes", ["x", "xz"])
def test_curved_hex_boundaries(rs, planes, cuda):
    """Inlet, outlet and slip walls on the planes x = 0, 1 (and z = 0, 1).

    On the z planes the reference's 3D rotation (rs_basic.cpp:21-28) is
    ill-conditioned: a unit normal (0, 0, 1 - ulp), which the normalisation
    nor * (1 / |nor|) yields for some |nor|, gives cp = sqrt(1 - n_z^2) = 1.5e-8
    inside the rotation, so the result depends on the last bit of |nor| at the
    1e-8 level, in the double oracle as in the reference. There the bar is twice
    the double oracle's own distance from the long double oracle; on the x
    planes (n_z = 0) it is 1e-12."""
    if planes == "x":
        m = curved_hex(4, 3, 3, seed=8, periodic=(False, True, True))
        bcs = [S.BC(5, S.BC_SUPERSONIC_INLET, (1.1, 0.5, 0.0, 0.1, 1.0)), S.BC(3, S.BC_OUTLET)]
        if rs == S.LLF:
            bcs = [S.BC(5, S.BC_SLIP), S.BC(3, S.BC_SLIP)]
    else:
        m = curved_hex(4, 3, 3, seed=7, periodic=(False, True, False))
        bcs = [S.BC(1, S.BC_SLIP), S.BC(6, S.BC_SLIP), S.BC(5, S.BC_SUPERSONIC_INLET, (1.1, 0.5, 0.0, 0.1, 1.0)),
               S.BC(3, S.BC_OUTLET)]
    err, bar, lam, lam_ref = _rhs_check(m, 2, S.Scheme(riemann=rs, bcs=bcs), cuda, seed=7)
    print(f"[curved bc {planes} rs={rs}] rel Linf {err:.3e} bar {bar:.1e}", flush=True)
    if planes == "x":
        assert bar == TOL
    assert err <= bar
    assert abs(lam - lam_ref) <= 1e-13 * lam_ref


@pytest.mark.parametrize("p", [1, 2, 3])
def test_curved_hex_free_stream(p, cuda):
    """A uniform flow has a zero right-hand side on the curved mesh (round-off)."""
    import torch

    m = curved_hex(4, 4, 3, seed=11)
    sch = S.Scheme(riemann=S.HLLC)
    ctx = _ctx(m, p, sch)
    nd = (p + 1) ** 3
    shape = (m.num_elements, nd)
    u = np.stack(conservative(np.full(shape, 1.2), [np.full(shape, 0.4), np.full(shape, -0.3),
                                                    np.full(shape, 0.25)], np.full(shape, 0.9)))
    x = np.ascontiguousarray(u.reshape(-1))
    xd = _dev(x, cuda)
    yd = torch.empty_like(xd)
    ctx.rhs(xd, yd)
    y = np.abs(_host(yd)).max()
    y_ref, _ = RefSolver(m, p, sch).mult(x)
    # the cancellation runs through the dense inverse mass: round-off of a few
    # 1e-12 on an O(1) state; a wrong normal or metric term leaves O(1)
    print(f"[curved free stream p={p}] max |y| gpu {y:.3e} oracle {np.abs(y_ref).max():.3e}", flush=True)
    assert y <= 2e-11
    assert np.abs(y_ref).max() <= 2e-11


def test_curved_hex_conservation(cuda):
    """On a periodic curved mesh the integral of the right-hand side vanishes."""
    import torch

    m = curved_hex(5, 4, 4, seed=12)
    sch = S.Scheme(riemann=S.HLLC)
    ctx = _ctx(m, 2, sch)
    ref = RefSolver(m, 2, sch)
    x = smooth_state(ref.node_coords(), seed=12, amp=1e-2)
    xd = _dev(x, cuda)
    yd = torch.empty_like(xd)
    ctx.rhs(xd, yd)
    tot = ctx.totals(yd)
    scale = np.abs(_host(yd).reshape(5, -1)).max(axis=1) / m.num_elements
    print("[curved conservation]", tot / scale, flush=True)
    assert np.all(np.abs(tot) <= 1e-12 * scale)
    # the element volumes sum to the box's
    tot1 = ctx.totals(_dev(np.ones_like(x), cuda))
    assert np.all(np.abs(tot1 - 1.0) <= 1e-13)


def _shocky_state(xyz, seed):
    """Smooth state with a density and energy jump between the elements whose
    centres lie left and right of x = 0.5 (so the jump sits on element faces)."""
    x = smooth_state(xyz, seed=seed, amp=1e-2)
    U = x.reshape(5, -1)
    centre = xyz[..., 0].mean(axis=1)
    jump = np.repeat((centre > 0.5).astype(float), xyz.shape[1])
    U[0] *= 1.0 + 2.0 * jump
    U[-1] *= 1.0 + 4.0 * jump
    return np.ascontiguousarray(U.reshape(-1))


CURVED_LIMITERS = [
    S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, phi=S.PHI_MICHALAK, limiter=S.LIM_MULTI),
    S.Scheme(riemann=S.RUSANOV, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0, limiter=S.LIM_MULTI),
    S.Scheme(riemann=S.HLL, indicator=S.IND_BJ, phi=S.PHI_MICHALAK, limiter=S.LIM_FINDIFF),
]


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("li", range(len(CURVED_LIMITERS)))
def test_curved_hex_limiter_parity(li, p, cuda):
    import torch

    sch = CURVED_LIMITERS[li]
    m = curved_hex(6, 4, 3, seed=20 + li)
    ref = RefSolver(m, p, sch)
    x = _shocky_state(ref.node_coords(), seed=li)
    x_ref, _, mv = ref.limit(x)
    ctx = _ctx(m, p, sch)
    xd = _dev(x, cuda)
    tr = torch.zeros(m.num_elements, dtype=torch.int8, device=cuda)
    ctx.limit(xd, tr)
    err = rel_linf(_host(xd), x_ref, 5)
    flags = _host(tr)
    print(f"[curved limiter {li} p={p}] rel Linf {err:.3e} troubled {int((flags & 1).sum())}", flush=True)
    assert err <= TOL
    if sch.limiter == S.LIM_MULTI:
        near = np.abs(mv - 1.0) < 1e-12
        assert np.all(((flags & 1) == (mv < 1.0)) | near)


@pytest.mark.parametrize("tab,p", [("ssprk3", 1), ("ssprk3", 2), ("ssprk3", 3), ("ref_literal_rk2", 2)])
def test_curved_hex_step_parity(tab, p, cuda):
    """Limited RK steps (the SSP-RK3 one applies the limiter on load)."""
    m = curved_hex(5, 4, 3, seed=30)
    sch = S.Scheme(riemann=S.HLLC if p < 3 else S.RUSANOV, indicator=S.IND_BJ, phi=S.PHI_MICHALAK,
                   limiter=S.LIM_MULTI)
    ref = RefSolver(m, p, sch)
    x = _shocky_state(ref.node_coords(), seed=3)
    ctx = _ctx(m, p, sch)
    xd = _dev(x, cuda)
    for _ in range(2):
        x, lam_ref = ref.step(x, 1e-3, S.TABLEAUS[tab])
        lam = ctx.step(xd, 1e-3, S.TABLEAUS[tab])
    err = rel_linf(_host(xd), x, 5)
    print(f"[curved step {tab} p={p}] rel Linf {err:.3e} lam {lam} {lam_ref}", flush=True)
    assert err <= TOL
    assert abs(lam - lam_ref) <= 1e-12 * lam_ref


def test_curved_hex_hmin_and_volume(cuda):
    m = curved_hex(4, 3, 3, seed=40)
    sch = S.Scheme(riemann=S.HLLC)
    ctx = _ctx(m, 2, sch)
    ref = RefSolver(m, 2, sch)
    assert abs(ctx.hmin() - ref.hmin()) <= 1e-14 * ref.hmin()


def test_curved_nc_free_stream_and_conservation(cuda):
    """A curved mesh refined with hanging nodes (non-conforming faces whose
    normals vary per point): a uniform flow stays uniform and the integral of
    the right-hand side of a smooth state vanishes."""
    import torch

    from rkdg_amd import amr

    t = amr.NCMesh(curved_hex(4, 4, 3, seed=70))
    t.refine([5, 6, 30])
    m = t.mesh()
    assert len(m.nc) > 0
    sch = S.Scheme(riemann=S.RUSANOV)
    ctx = _ctx(m, 2, sch)
    n = m.num_elements * 27
    x = np.concatenate([np.full(n, 1.3), np.full(n, 0.26), np.full(n, -0.2), np.full(n, 0.13), np.full(n, 2.1)])
    xd = _dev(x, cuda)
    yd = torch.empty_like(xd)
    ctx.rhs(xd, yd)
    fs = float(yd.abs().max())
    xs = smooth_state(ctx.node_coords(), seed=71, amp=1e-2)
    xd = _dev(xs, cuda)
    ctx.rhs(xd, yd)
    tot = ctx.totals(yd)
    y_ref, _ = RefSolver(m, 2, sch).mult(xs)
    err = rel_linf(_host(yd), y_ref, 5)
    scale = np.abs(_host(yd).reshape(5, -1)).max(axis=1) / m.num_elements
    print(f"[curved nc] NE {m.num_elements} slave faces {len(m.nc)}: free stream {fs:.2e}, "
          f"conservation {tot / scale}, RHS vs oracle {err:.2e}", flush=True)
    assert fs <= 2e-11
    assert np.all(np.abs(tot) <= 1e-12 * scale)
    assert err <= TOL


@pytest.mark.parametrize("p", [1, 3])
def test_curved_hex_walls_other_orders(p, cuda):
    """Inlet, outlet and slip walls on the x planes at p = 1 and 3 (the
    boundary instantiations of the curved kernel at those orders)."""
    m = curved_hex(4, 3, 3, seed=100 + p, periodic=(False, True, True))
    bcs = [S.BC(5, S.BC_SUPERSONIC_INLET, (1.1, 0.5, 0.0, 0.1, 1.0)), S.BC(3, S.BC_SLIP)]
    err, bar, lam, lam_ref = _rhs_check(m, p, S.Scheme(riemann=S.HLL, bcs=bcs), cuda, seed=p)
    print(f"[curved walls p={p}] rel Linf {err:.3e} bar {bar:.1e}", flush=True)
    assert err <= bar
    assert abs(lam - lam_ref) <= 1e-13 * lam_ref


@pytest.mark.parametrize("which", ["einfeldt", "covolume"])
def test_curved_hex_scheme_variants(which, cuda):
    """Einfeldt's wave speeds with HLLC, and the covolume equation of state
    with HLL and Rusanov, on the curved mesh."""
    m = curved_hex(4, 3, 3, seed=110)
    schemes = ([S.Scheme(riemann=S.HLLC, wave_speeds=S.EINFELDT)] if which == "einfeldt" else
               [S.Scheme(riemann=S.HLL, covolume=0.05), S.Scheme(riemann=S.RUSANOV, covolume=0.05)])
    for sch in schemes:
        err, bar, lam, lam_ref = _rhs_check(m, 2, sch, cuda, seed=9)
        print(f"[curved {which} rs={sch.riemann}] rel Linf {err:.3e} bar {bar:.1e}", flush=True)
        assert err <= bar
        assert abs(lam - lam_ref) <= 1e-13 * lam_ref
