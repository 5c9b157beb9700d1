"""GPU against the oracle on the meshes the reference ships with its cases
(tests/golden/meshes, read by rkdg_amd.gmsh) and on cases loaded from YAML
settings: right-hand side, limiter and RK steps to 1e-12 per conserved
variable, then short runs of the time loop."""
import os
import shutil

import numpy as np
import pytest

from oracle import refcpu as R
from oracle.refcpu import RefSolver
from rkdg_amd import case_file as CF
from rkdg_amd import gmsh
from rkdg_amd import scheme as S
from tests.helpers import rel_linf, smooth_state

pytestmark = pytest.mark.gpu

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")
TOL = 1e-12

CASES = [("sod2D_sod2d.msh", 1), ("sod2D_sod2d.msh", 2), ("sod1d_covolume_sod2d.msh", 1),
         ("sod1d_covolume_sod2d.msh", 2), ("sod1d_trian.msh", 2), ("sod1d_quad6.msh", 3),
         ("jet_overexpanded0p25.msh.gz", 1), ("jet_overexpanded0p25.msh.gz", 2)]


def _scheme(gm, rs=S.HLLC):
    """The first patch a supersonic inlet, the second an outlet, the rest slip walls."""
    names = sorted(gm.bdr_names, key=gm.bdr_names.get)
    types = [(S.BC_SUPERSONIC_INLET, (1.2, 0.4, 0.05, 0.0, 1.1)), (S.BC_OUTLET, ())]
    bcs = [S.BC(gm.bdr_names[n], *(types[i] if i < 2 else (S.BC_SLIP, ()))) for i, n in enumerate(names)]
    return S.Scheme(riemann=rs, indicator=S.IND_BJ, phi=S.PHI_MICHALAK, limiter=S.LIM_MULTI, bcs=bcs)


def _state(xyz, seed, jump_rho=2.0, jump_e=4.0):
    """Smooth flow scaled to the mesh's box plus a jump across its middle (the
    nodes are Gauss-Legendre points, so traces extrapolate: a large jump inside
    an element gives unphysical face states)."""
    lo = xyz.reshape(-1, xyz.shape[-1]).min(axis=0)
    hi = xyz.reshape(-1, xyz.shape[-1]).max(axis=0)
    u = (xyz - lo) / (hi - lo)
    x = smooth_state(u, seed=seed, amp=1e-2)
    U = x.reshape(xyz.shape[-1] + 2, -1)
    jump = (u[..., 0].ravel() > 0.5).astype(float)
    U[0] *= 1.0 + jump_rho * jump
    U[-1] *= 1.0 + jump_e * jump
    return np.ascontiguousarray(U.reshape(-1))


def _per_var(a, b, neq):
    a, b = a.reshape(neq, -1), b.reshape(neq, -1)
    return [float(np.abs(a[v] - b[v]).max() / np.abs(b[v]).max()) for v in range(neq)]


@pytest.mark.parametrize("fname,p", CASES, ids=[f"{c[0].split('.')[0]}-p{c[1]}" for c in CASES])
def test_shipped_mesh_parity(fname, p, cuda):
    import torch

    from rkdg_amd.capi import Context

    gm = gmsh.read_gmsh(os.path.join(MESHES, fname))
    m = gm.mesh
    sch = _scheme(gm)
    neq = m.dim + 2
    ref = RefSolver(m, p, sch)
    ctx = Context(m, p, sch)
    x = _state(ref.node_coords(), seed=p, **({"jump_rho": 0.5, "jump_e": 1.0} if fname.startswith("jet") else {}))
    xd = torch.from_numpy(x).to(cuda)
    # right-hand side
    # the bar: the oracle's own rounding, measured against its long-double
    # build, or 1e-12, whichever is larger (per conserved variable)
    y_ref, lam_ref = ref.mult(x)
    y_ld, _ = RefSolver(m, p, sch, precision="long double").mult(x)
    yd = torch.empty_like(xd)
    lam = ctx.rhs(xd, yd)
    y = yd.cpu().numpy()
    e_gpu = _per_var(y, y_ld, neq)
    e_ref = _per_var(y_ref, y_ld, neq)
    print(f"[case-mesh {fname} p{p}] NE {m.num_elements} rhs vs long double: GPU {['%.1e' % e for e in e_gpu]} "
          f"oracle {['%.1e' % e for e in e_ref]}; GPU vs oracle {rel_linf(y, y_ref, neq):.1e}", flush=True)
    assert all(eg <= max(2.0 * er, TOL) for eg, er in zip(e_gpu, e_ref))
    assert abs(lam - lam_ref) <= 1e-13 * lam_ref
    # limiter with the troubled flags
    x_lim, _, mv = ref.limit(x)
    xl = xd.clone()
    tr = torch.zeros(m.num_elements, dtype=torch.int8, device=cuda)
    ctx.limit(xl, tr)
    err = rel_linf(xl.cpu().numpy(), x_lim, neq)
    flags = tr.cpu().numpy()
    near = np.abs(mv - 1.0) < 1e-12
    print(f"[case-mesh {fname} p{p}] limiter {err:.2e} troubled {int((flags & 1).sum())} "
          f"(oracle {int((mv < 1).sum())})", flush=True)
    assert err <= TOL and np.all(((flags & 1) == (mv < 1.0)) | near)
    # three steps of each tableau from the limited state
    dt = 0.1 * ref.hmin() / lam_ref / (2 * p + 1)
    for tab in ("ssprk3", "ref_literal_rk2"):
        xr = x_lim.copy()
        xg = torch.from_numpy(x_lim.copy()).to(cuda)
        for _ in range(3):
            xr, lr = ref.step(xr, dt, S.TABLEAUS[tab])
            lg = ctx.step(xg, dt, S.TABLEAUS[tab])
        err = rel_linf(xg.cpu().numpy(), xr, neq)
        print(f"[case-mesh {fname} p{p}] 3 steps {tab} {err:.2e}", flush=True)
        assert err <= TOL and abs(lg - lr) <= 1e-12 * lr


def _copy_case(tmp_path, mesh_name, text):
    d = tmp_path / "case"
    d.mkdir()
    shutil.copyfile(os.path.join(MESHES, mesh_name), d / mesh_name)
    (d / "settings.yml").write_text(text)
    return str(d)


JET = """
mesh:
    file: jet_overexpanded0p25.msh.gz
    serialRefLevels: 0
    parallelRefLevels: 0
    adaptive: false
physics:
    type:   perfectGas
    gamma:  1.4
    R: 287.0
internalField:
    type:       constant
    rho:        1.205
    U:          [0,0,0]
    p:          101325
boundaryField:
    inlet:
        type:   supersonicInlet
        rhoI:   2.188
        UI:     [649.4,0,0]
        pI:     52120
    outlet:
        type:   outlet
    wall:
        type:   slip
spatial:
    polyOrder:      1
    riemannSolver:  HLL
    limiter:
        type:       Multi
        indicator:  BJ
time:
    order:  2
    tEnd:   1e-5
    dt:     1e-10
    type:   dynamic
    CoMax:  0.1
postProcess:
    visSteps:   10
"""

COVOLUME = """
mesh:
    file: sod1d_covolume_sod2d.msh
    serialRefLevels: 1
    parallelRefLevels: 0
physics:
    type:   covolume
    gamma:  1.3
    covolume: 0.001
internalField:
    type:       planeBreakup
    origin:     [0.4,0,0]
    normal:     [1,0,0]
    left:
        rho:    100.0
        U:      [0,0,0]
        p:      1e+8
    right:
        rho:    1.0
        U:      [0,0,0]
        p:      0.1
boundaryField:
    Patch1:
        type:   slip
    Patch2:
        type:   slip
    Patch3:
        type:   slip
    Patch4:
        type:   slip
spatial:
    polyOrder:      1
    riemannSolver:  HLL
    limiter:
        type:       Multi
        indicator:  BJ
time:
    order:  3
    tEnd:   0.0002
    dt:     1e-6
    type:   dynamic
    CoMax:  0.1
postProcess:
    visSteps:   500
"""


def _errs(U, Ur, neq):
    """Per-variable max error; the momentum components on the momentum
    magnitude (the transverse one is round-off in these nearly
    one-dimensional flows)."""
    U, Ur = U.reshape(neq, -1), Ur.reshape(neq, -1)
    mscale = np.abs(Ur[1:neq - 1]).max()
    return [np.abs(U[v] - Ur[v]).max() / (mscale if 0 < v < neq - 1 else np.abs(Ur[v]).max()) for v in range(neq)]


@pytest.mark.parametrize("name,mesh_name,text,steps", [
    ("jet", "jet_overexpanded0p25.msh.gz", JET, 20),
    ("covolume_sod", "sod1d_covolume_sod2d.msh", COVOLUME, 20),
])
def test_case_file_run_matches_oracle_loop(name, mesh_name, text, steps, tmp_path, cuda):
    """The time loop (first dt from the faces, dt from the last stage, t advanced
    by the recomputed dt) on the GPU through solver.run against the oracle's
    loop, from the case's own initial state: the same step count and time,
    one GPU step from the oracle's state at 1e-12 at checkpoints, and the
    final states apart by no more than the rounding differences the limiter's
    switches carry forward over the run."""
    import torch

    from rkdg_amd.solver import ExplicitRKLimitedSolver, FE_Evolution, Limiter, run

    cs = CF.load_case(_copy_case(tmp_path, mesh_name, text))
    tab = S.TABLEAUS[cs.tableau]
    neq = cs.mesh.dim + 2
    ref = RefSolver(cs.mesh, cs.order, cs.scheme)
    x0 = cs.initial_state(ref.node_coords())
    keep = {}

    def cb(k, t, dt, x):
        if k in (1, steps // 2, steps - 1):
            keep[k] = (x.copy(), min(dt, cs.t_final - t), t)

    xr, st = R.run(ref, x0, cs.t_final, cs.cfl, tab, max_steps=steps, callback=cb)
    f = FE_Evolution(cs.mesh, cs.order, cs.scheme)
    lim = Limiter(f)
    sol = ExplicitRKLimitedSolver(*tab, limiter=lim)
    for k, (xk, dtk, tk) in sorted(keep.items()):
        xg = torch.from_numpy(xk.copy()).to(cuda)
        sol.Init(f)
        sol.Step(xg, tk, dtk)
        xo, _ = ref.step(xk, dtk, tab, tk)
        e = _errs(xg.cpu().numpy(), xo, neq)
        print(f"[case-run {name}] one step from the oracle's state after step {k}: {['%.1e' % v for v in e]}",
              flush=True)
        assert max(e) <= TOL
    xg = torch.from_numpy(x0.copy()).to(cuda)
    stats = run(f, sol, lim, xg, cs.t_final, cs.cfl, max_steps=steps)
    errs = _errs(xg.cpu().numpy(), xr, neq)
    print(f"[case-run {name}] NE {cs.mesh.num_elements} steps {stats.steps}/{st['steps']} "
          f"t {stats.t:.6e}/{st['t']:.6e} final {['%.1e' % e for e in errs]}", flush=True)
    assert stats.steps == st["steps"]
    assert abs(stats.t - st["t"]) <= 1e-12 * st["t"]
    assert max(errs) <= 1e-9


def test_run_case_driver_with_restart(tmp_path, cuda):
    """python -m rkdg_amd.run_case on the jet case: 6 steps straight, and 4 steps
    then a restart from frame 4 for 2 more (the frame's state, time and dt;
    the limiter runs on the restored state as after any projection)."""
    from rkdg_amd import run_case as RC

    text = JET.replace("visSteps:   10", "visSteps:   2\n    checkTotalEnergy: 1")
    text = text.replace("order:  2\n", "order:  2\n    nSavedFrames: 2\n")
    (tmp_path / "a").mkdir()
    d1 = _copy_case(tmp_path / "a", "jet_overexpanded0p25.msh.gz", text)
    _, s6, pp6, x6 = RC.run_case(d1, max_steps=6, quiet=True)
    assert s6.steps == 6 and pp6.restart.frames() == [4, 6]
    assert len(pp6.series) == 3 and all(os.path.exists(p) for _, p in pp6.series)
    (tmp_path / "b").mkdir()
    d2 = _copy_case(tmp_path / "b", "jet_overexpanded0p25.msh.gz", text)
    _, s4, pp4, _ = RC.run_case(d2, max_steps=4, quiet=True)
    assert pp4.restart.frames() == [2, 4]
    st = open(os.path.join(d2, "settings.yml")).read().replace("time:\n", "time:\n    restart: true\n    restartCycle: 4\n")
    open(os.path.join(d2, "settings.yml"), "w").write(st)
    _, s2, pp2, x2 = RC.run_case(d2, max_steps=2, quiet=True)
    assert pp2.restart.frames() == [4, 6]
    assert pp2.log[-1]["step"] == 6 and abs(pp2.log[-1]["t"] - s6.t) <= 1e-15 * s6.t
    a, b = x2.cpu().numpy(), x6.cpu().numpy()
    err = rel_linf(a, b, 4)
    print(f"[run_case] restart vs straight rel Linf {err:.2e}, t {s6.t:.9e}", flush=True)
    assert err <= 1e-10


def test_run_case_adaptive_matches_oracle_loop(tmp_path, cuda):
    """mesh.adaptive through python -m rkdg_amd.run_case (the reference's keys,
    cases/sod2d/settings.yml:1-14, plus a level cap) on the shipped Sod quad
    mesh refined twice: the refine / repeat / coarsen loop on the GPU against
    the same loop on the oracle."""
    from rkdg_amd import amr
    from rkdg_amd import case_file as CF
    from rkdg_amd import run_case as RC

    text = COVOLUME.replace("file: sod1d_covolume_sod2d.msh", "file: sod1d_quad.msh")
    text = text.replace("serialRefLevels: 1", "serialRefLevels: 2\n    adaptive: 1\n    totalErrorFraction: 0.0\n"
                        "    maxElemError: 0.1\n    hysteresis: 0.5\n    preferConformingRefinement: false\n"
                        "    nonConformingLimit: 1\n    maxRefLevel: 2")
    d = _copy_case(tmp_path, "sod1d_quad.msh", text)
    cs, stats, pp, x = RC.run_case(d, max_steps=8, quiet=True)
    ro = amr.run_adaptive(cs.mesh, cs.order, cs.scheme, cs.initial_state, cs.t_final, cs.cfl,
                          S.TABLEAUS[cs.tableau], cs.max_elem_error, hysteresis=cs.hysteresis, max_steps=8,
                          solver="oracle", max_level=cs.max_ref_level)
    errs = _errs(x.cpu().numpy(), ro.x, 4)
    print(f"[case-run adaptive] elements per step {stats.elements} (oracle {ro.elements}), t {stats.t:.9e} / "
          f"{ro.t:.9e}, final {['%.1e' % e for e in errs]}", flush=True)
    assert stats.elements == ro.elements and max(stats.elements) > cs.mesh.num_elements
    assert len(stats.mesh.nc) > 0
    assert abs(stats.t - ro.t) <= 1e-12 * ro.t
    assert max(errs) <= 1e-9
    assert len(pp.series) == 1 and os.path.exists(pp.series[0][1])


CURVED_CASE = """
mesh:
    file: curved.msh
    serialRefLevels: 0
    parallelRefLevels: 0
physics:
    type:   perfectGas
    gamma:  1.4
internalField:
    type:       planeBreakup
    origin:     [0.5,0,0]
    normal:     [1,0,0]
    left:
        rho:    1.0
        U:      [0,0,0]
        p:      1.0
    right:
        rho:    0.5
        U:      [0,0,0]
        p:      0.4
boundaryField:
    Patch1:
        type:   outlet
    Patch6:
        type:   outlet
    Patch2:
        type:   slip
    Patch3:
        type:   slip
    Patch4:
        type:   slip
    Patch5:
        type:   slip
spatial:
    polyOrder:      2
    riemannSolver:  Rusanov
    limiter:
        type:       Multi
        indicator:  BJ
time:
    order:  3
    tEnd:   0.01
    dt:     1e-4
    type:   dynamic
    CoMax:  0.1
postProcess:
    visSteps:   500
"""


def test_curved_hex_case_directory(tmp_path, cuda):
    """A case directory whose Gmsh mesh holds curved (trilinear) hexahedra,
    read through the case layer (MSH 2.2 written here, the physical names
    mapped to boundary conditions) and stepped on the GPU against the oracle:
    right-hand side, limiter with its troubled cells, three limited SSP-RK3
    steps. Outlets on the z planes and Rusanov keep the reference's 3D rotation
    out of the faces whose normal is the z axis."""
    import torch

    from rkdg_amd.capi import Context
    from tests.helpers import curved_hex

    d = tmp_path / "case"
    d.mkdir()
    gmsh.write_gmsh22(str(d / "curved.msh"), curved_hex(6, 4, 3, seed=95, periodic=(False, False, False)))
    (d / "settings.yml").write_text(CURVED_CASE)
    cs = CF.load_case(str(d))
    m = cs.mesh
    ctx = Context(m, cs.order, cs.scheme)
    assert ctx.info.affine == 0
    ref = RefSolver(m, cs.order, cs.scheme)
    x0 = cs.initial_state(ref.node_coords())  # the breakup plane cuts through curved cells
    x, _, mv = ref.limit(x0)
    xl = torch.from_numpy(x0.copy()).to(cuda)
    tr = torch.zeros(m.num_elements, dtype=torch.int8, device=cuda)
    ctx.limit(xl, tr)
    el = rel_linf(xl.cpu().numpy(), x, 5)
    xd = torch.from_numpy(x.copy()).to(cuda)
    y_ref, lam_ref = ref.mult(x)
    yd = torch.empty_like(xd)
    lam = ctx.rhs(xd, yd)
    err = rel_linf(yd.cpu().numpy(), y_ref, 5)
    dt = 0.1 * ref.hmin() / lam_ref / 5
    xr, xg = x.copy(), xd.clone()
    for _ in range(3):
        xr, _ = ref.step(xr, dt, S.TABLEAUS["ssprk3"])
        ctx.step(xg, dt, S.TABLEAUS["ssprk3"])
    es = rel_linf(xg.cpu().numpy(), xr, 5)
    print(f"[case curved] NE {m.num_elements}: rhs {err:.2e}, limiter {el:.2e} troubled "
          f"{int((tr.cpu().numpy() & 1).sum())} (oracle {int((mv < 1).sum())}), 3 steps {es:.2e}", flush=True)
    assert err <= TOL and abs(lam - lam_ref) <= 1e-13 * lam_ref
    near = np.abs(mv - 1.0) < 1e-12
    assert el <= TOL and np.all(((tr.cpu().numpy() & 1) == (mv < 1.0)) | near)
    assert int((mv < 1.0).sum()) > 0
    assert es <= TOL
