"""The remaining limiter options on the GPU against the oracle: the
multiplier's linear reconstruction (linearize, limiter_multiplier.cpp:101-178)
and Shu's indicator on neighbour-extrapolated means with the FinDiff limiter
(indicator_shu.cpp:13-74, averager.cpp:152-273), on every element family the
path builds; the limited state to 1e-12 per conserved variable, identical
troubled cells, and RK steps."""
import numpy as np
import pytest

from oracle.refcpu import RefSolver
from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from tests.helpers import all_outlet, conservative, curved_hex, rel_linf, smooth_state

pytestmark = pytest.mark.gpu
TOL = 1e-12
NOP = (False, False, False)


def _state(xyz, seed, jump=(1.0, 2.0)):
    """Smooth flow scaled to the unit box, plus a density and energy jump."""
    lo = xyz.reshape(-1, xyz.shape[-1]).min(axis=0)
    hi = xyz.reshape(-1, xyz.shape[-1]).max(axis=0)
    u = (xyz - lo) / (hi - lo)
    x = smooth_state(u, seed=seed, amp=0.0)
    U = x.reshape(xyz.shape[-1] + 2, -1)
    j = (u[..., 0].ravel() > 0.45).astype(float)
    U[0] *= 1.0 + jump[0] * j
    U[-1] *= 1.0 + jump[1] * j
    return np.ascontiguousarray(U.reshape(-1))


def _mild_state(xyz):
    """A gently varying flow (2 % waves) with a density and pressure jump at
    x = 0.45 of the box: Shu's indicator (Ck = 0.03) flags the jump only."""
    dim = xyz.shape[-1]
    lo = xyz.reshape(-1, dim).min(axis=0)
    hi = xyz.reshape(-1, dim).max(axis=0)
    u = (xyz - lo) / (hi - lo)
    X = [u[..., d].ravel() for d in range(dim)]
    tp = 2 * np.pi
    j = (X[0] > 0.45).astype(float)
    rho = (1.0 + 0.02 * np.sin(tp * X[0]) * np.cos(tp * X[1])) * (1.0 + j)
    p = (1.0 + 0.02 * np.cos(tp * (X[0] + X[1]))) * (1.0 + 3.0 * j)
    vel = [0.1 + 0 * rho, 0.05 + 0 * rho] + ([0.02 + 0 * rho] if dim == 3 else [])
    return np.ascontiguousarray(np.concatenate([np.ravel(c) for c in conservative(rho, vel, p)]))


MESHES = {
    "quad_q1": (lambda: M.structured_quad(10, 5, 0.0, 1.0, 0.0, 0.5), 1, False),
    "quad_q1_bilinear": (lambda: M.structured_quad(10, 5, warp=lambda X, Y: (X + 0.04 * np.sin(3 * Y), Y + 0.03 * X * Y)),
                         1, False),
    "quad_q1_dense": (lambda: M.structured_quad(10, 5), 1, True),
    "hex_q1": (lambda: M.structured_hex(5, 4, 3, periodic=NOP), 1, False),
    "quad_q2": (lambda: M.structured_quad(8, 4), 2, False),
    "tri_p2": (lambda: M.jittered_delaunay_tri(6, 4, 0.0, 1.0, 0.0, 1.0, seed=7), 2, False),
    "hex_q2": (lambda: M.structured_hex(4, 3, 3, periodic=NOP), 2, False),
    "quad_q2_bilinear": (lambda: M.structured_quad(8, 4, warp=lambda X, Y: (X + 0.04 * np.sin(3 * Y), Y)), 2, False),
    # curved (trilinear) hexahedra, corners moved by up to 2 % of the spacing:
    # Shu's extrapolated means take k's Jacobian determinant at the last rule
    # point for its mean (averager.cpp:262-266), so stronger curvature flags
    # every cell of these small meshes
    "hex_q1_curved": (lambda: curved_hex(5, 4, 3, amp=0.02, seed=50, periodic=NOP), 1, False),
    "hex_q2_curved": (lambda: curved_hex(4, 3, 3, amp=0.02, seed=52, periodic=NOP), 2, False),
}


def _riemann(name):
    # curved meshes with walls on the z planes: Rusanov, whose flux needs no
    # rotation (the reference's 3D rotation turns an n_z of 1 - ulp into a
    # 1e-8 term, on a last bit that differs between codes)
    return S.RUSANOV if "curved" in name else S.HLLC


LINEARIZE = ["quad_q1", "quad_q1_bilinear", "quad_q1_dense", "hex_q1", "tri_p2", "hex_q1_curved"]


def _limit_pair(mesh, p, sch, dense, seed, cuda):
    import torch

    from rkdg_amd.capi import Context

    ref = RefSolver(mesh, p, sch)
    x = _state(ref.node_coords(), seed) if seed is not None else _mild_state(ref.node_coords())
    x_ref, ind, mv = ref.limit(x)
    ctx = Context(mesh, p, sch, dense=dense)
    xd = torch.from_numpy(x).to(cuda)
    tr = torch.zeros(mesh.num_elements, dtype=torch.int8, device=cuda)
    ctx.limit(xd, tr)
    return ref, ctx, x, x_ref, ind, mv, xd, tr.cpu().numpy()


@pytest.mark.parametrize("name", LINEARIZE)
def test_linearize_matches_oracle(name, cuda):
    import torch

    mk, p, dense = MESHES[name]
    m = mk()
    sch = S.Scheme(riemann=_riemann(name), indicator=S.IND_BJ, phi=S.PHI_MICHALAK, limiter=S.LIM_MULTI, linearize=1,
                   bcs=all_outlet(m))
    ref, ctx, x, x_ref, ind, mv, xd, tr = _limit_pair(m, p, sch, dense, 1, cuda)
    neq = m.dim + 2
    err = rel_linf(xd.cpu().numpy(), x_ref, neq)
    changed = int((np.abs(x_ref - x).reshape(neq, m.num_elements, -1).max(axis=(0, 2)) > 0).sum())
    print(f"[linearize {name}] limiter rel Linf {err:.2e}, elements changed {changed}, "
          f"troubled {int((tr & 1).sum())}", flush=True)
    assert err <= TOL
    near = np.abs(mv - 1.0) < 1e-12
    assert np.all(((tr & 1) == (mv < 1.0)) | near)
    if m.geom != M.GEOM_TRI:
        assert changed == m.num_elements  # every element is linearized
    # two steps of SSP-RK3 with the limiter in every stage
    xs = x_ref.copy()
    xg = torch.from_numpy(x_ref.copy()).to(cuda)
    dt = 0.05 * ref.hmin() / ref.mult(x_ref)[1] / (2 * p + 1)
    for _ in range(2):
        xs, _ = ref.step(xs, dt, S.TABLEAUS["ssprk3"])
        ctx.step(xg, dt, S.TABLEAUS["ssprk3"])
    assert rel_linf(xg.cpu().numpy(), xs, neq) <= TOL


@pytest.mark.parametrize("name", list(MESHES))
def test_shu_findiff_matches_oracle(name, cuda):
    import torch

    mk, p, dense = MESHES[name]
    m = mk()
    sch = S.Scheme(riemann=_riemann(name), indicator=S.IND_SHU, limiter=S.LIM_FINDIFF, bcs=all_outlet(m))
    ref, ctx, x, x_ref, ind, mv, xd, tr = _limit_pair(m, p, sch, dense, None, cuda)
    neq = m.dim + 2
    err = rel_linf(xd.cpu().numpy(), x_ref, neq)
    last = ind.reshape(neq, -1)[-1]
    want = (last < 0.999).astype(np.int8)
    near = np.abs(last - 0.999) < 1e-12
    print(f"[shu {name}] limiter rel Linf {err:.2e}, troubled {int(tr.sum())} (oracle {int(want.sum())} "
          f"of {m.num_elements})", flush=True)
    assert err <= TOL
    assert np.all((tr == want) | near)
    assert 0 < want.sum() < m.num_elements
    xs = x_ref.copy()
    xg = torch.from_numpy(x_ref.copy()).to(cuda)
    dt = 0.05 * ref.hmin() / ref.mult(x_ref)[1] / (2 * p + 1)
    for _ in range(2):
        xs, _ = ref.step(xs, dt, S.TABLEAUS["ssprk3"])
        ctx.step(xg, dt, S.TABLEAUS["ssprk3"])
    assert rel_linf(xg.cpu().numpy(), xs, neq) <= TOL
