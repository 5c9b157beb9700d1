"""Post-step diagnostics on the GPU through the C-ABI: conservation totals,
primitive fields and their time integrals, evaluation at reference points,
and the ParaView/restart output of a short run."""
import numpy as np
import pytest

from oracle import refcpu as R
from oracle.refcpu import RefSolver
from rkdg_amd import io
from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from tests.helpers import smooth_state

pytestmark = pytest.mark.gpu

KINDS = {
    "hex_q2": (lambda: M.structured_hex(4, 3, 3), 2, False, []),
    "quad_q2_bc": (lambda: M.structured_quad(6, 4), 2, False,
                   [S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(2, S.BC_OUTLET), S.BC(4, S.BC_OUTLET)]),
    "quad_q1_bilinear": (lambda: M.structured_quad(6, 4, warp=lambda X, Y: (X + 0.05 * np.sin(3 * Y) * X, Y)), 1,
                         False, [S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(2, S.BC_OUTLET), S.BC(4, S.BC_OUTLET)]),
    "tri_p2": (lambda: M.jittered_delaunay_tri(5, 4, 0.0, 1.0, 0.0, 1.0, seed=4), 2, False,
               [S.BC(a, S.BC_SLIP) for a in (1, 2, 3, 4)]),
}


def _volumes(mesh):
    X = mesh.elem_coords
    if mesh.geom == M.GEOM_TRI:
        return 0.5 * np.abs((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1])
                            - (X[:, 1, 1] - X[:, 0, 1]) * (X[:, 2, 0] - X[:, 0, 0]))
    if mesh.geom == M.GEOM_QUAD:
        x, y = X[..., 0], X[..., 1]
        return 0.5 * np.abs((x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y).sum(axis=1))
    J = np.stack([X[:, 1] - X[:, 0], X[:, 3] - X[:, 0], X[:, 4] - X[:, 0]], axis=2)
    return np.abs(np.linalg.det(J))


@pytest.mark.parametrize("kind", list(KINDS))
def test_totals_primitives_and_interpolation(kind, cuda):
    import torch

    from rkdg_amd.capi import Context

    mk, p, dense, bcs = KINDS[kind]
    m = mk()
    sch = S.Scheme(riemann=S.HLLC, bcs=bcs, gas_constant=0.8, indicator=S.IND_BJ, limiter=S.LIM_MULTI)
    ref = RefSolver(m, p, sch)
    ctx = Context(m, p, sch, dense=dense)
    neq, dim = m.dim + 2, m.dim
    x = smooth_state(ref.node_coords(), seed=3)
    xd = torch.from_numpy(x).to(cuda)
    # totals: sum_e |e| mean_e with the oracle's means and the elements' measures
    want = (ref.means(x).reshape(neq, -1) * _volumes(m)[None]).sum(axis=1)
    got = ctx.totals(xd)
    print(f"[io {kind}] totals {got} (oracle means x volumes {want})", flush=True)
    # a component can sum to nearly zero (the z-momentum here), so the bound is
    # relative to the sum of magnitudes
    scale = (np.abs(ref.means(x).reshape(neq, -1)) * _volumes(m)[None]).sum(axis=1)
    assert np.all(np.abs(got - want) <= 1e-13 * scale), (got - want) / scale
    # primitives at the nodes, and their dt-weighted accumulation
    nd = ref.nd
    prim = torch.empty((dim + 3) * m.num_elements * nd, dtype=torch.float64, device=cuda)
    acc = torch.zeros_like(prim)
    ctx.primitives(xd, prim=prim, acc=acc, weight=0.5)
    ctx.primitives(xd, acc=acc, weight=0.25)
    U = x.reshape(neq, -1)
    f = io.primitive_fields(U, dim, sch.gamma, sch.gas_constant, sch.covolume)
    want = np.concatenate([f["rho"][None], f["U"].T, f["p"][None], f["T"][None]])
    np.testing.assert_allclose(prim.cpu().numpy().reshape(dim + 3, -1), want, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(acc.cpu().numpy().reshape(dim + 3, -1), 0.75 * want, rtol=1e-14, atol=1e-15)
    # evaluation at points: the nodes give the state back (Lagrange basis); random
    # reference points match the oracle's shape functions
    rng = np.random.default_rng(1)
    pts = rng.uniform(0, 1, (7, dim))
    if m.geom == M.GEOM_TRI:
        pts[:, :2] *= 0.5
    out = torch.empty(neq * m.num_elements * len(pts), dtype=torch.float64, device=cuda)
    ctx.interpolate(xd, pts, out)
    phi = np.stack([R.shape(m.geom, p, np.pad(q, (0, 3 - dim)))[0] for q in pts])
    want = np.einsum("qk,vek->veq", phi, x.reshape(neq, m.num_elements, nd))
    np.testing.assert_allclose(out.cpu().numpy().reshape(neq, m.num_elements, -1), want, rtol=0,
                               atol=1e-13 * np.abs(x).max())


def test_totals_conserved_over_periodic_steps(cuda):
    import torch

    from rkdg_amd.capi import Context

    m = M.structured_hex(4, 4, 4)
    sch = S.Scheme(riemann=S.RUSANOV, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0, limiter=S.LIM_MULTI)
    ctx = Context(m, 2, sch)
    xd = torch.from_numpy(smooth_state(ctx.node_coords(), seed=9)).to(cuda)
    t0 = ctx.totals(xd)
    for _ in range(5):
        ctx.step(xd, 2e-3, S.TABLEAUS["ssprk3"])
    t1 = ctx.totals(xd)
    drift = np.abs(t1 - t0) / np.abs(t0).max()
    print(f"[io] totals drift over 5 periodic steps {drift}", flush=True)
    assert drift[0] <= 1e-14 and drift[-1] <= 1e-14 and drift.max() <= 1e-13


def test_postprocessor_run_writes_paraview_and_restarts(tmp_path, cuda):
    import torch

    from rkdg_amd import cases
    from rkdg_amd.solver import ExplicitRKLimitedSolver, FE_Evolution, Limiter, run

    cs = cases.sod2d(nx=20, ny=2)
    f = FE_Evolution(cs.mesh, cs.order, cs.scheme)
    lim = Limiter(f)
    sol = ExplicitRKLimitedSolver(*S.TABLEAUS["ssprk3"], limiter=lim)
    x = torch.from_numpy(cs.initial_state(f.ctx.node_coords())).to(cuda)
    pp = io.PostProcessor(f.ctx, str(tmp_path), vis_steps=5, n_saved=2, lod=2, check_totals=True)
    stats = run(f, sol, lim, x, 0.02, cs.cfl, callback=pp)
    pp.write(stats.steps, stats.t, stats.dt, x)
    assert pp.restart.frames()[-1] == stats.steps and len(pp.restart.frames()) <= 2
    cyc, t, dt, xs = pp.restart.load()
    np.testing.assert_array_equal(xs, x.cpu().numpy())
    r = io.read_vtu(pp.series[-1][1])
    assert len(r["types"]) == cs.mesh.num_elements * 4 and len(r["Points"]) == cs.mesh.num_elements * 9
    # lod 2 points include the element vertices: their values are the state there
    assert np.all(r["rho"] > 0) and np.all(r["p"] > 0)
    mass = [rec["totals"][0] for rec in pp.log]
    print(f"[io] Sod 20x2: {stats.steps} steps, mass {mass[0]:.15g} .. {mass[-1]:.15g}", flush=True)
    means = pp.means(stats.t)
    assert means.shape == (5, f.ctx.info.num_dofs) and np.all(means[0] > 0)
