"""GPU checks of the discrete adjoint (hand-written backward kernels of the
dense 2D path) against central finite differences of the forward kernels:
dot-product tests of the right-hand side, limiter and RK-step transposes
(<dF/dx . d, w> by differences vs <d, (dF/dx)^T w>), boundary-parameter
adjoints, the sqrt(N)-checkpointed gradient of the time-integrated step-face
pressure, and the torch.autograd.Function harness.

Tolerances: the forward maps are smooth at the sampled states (no branch of a
Riemann solver, limiter or stencil bound switches within +-h), so central
differences with h ~ 1e-6 agree to ~1e-8 relative; the tests ask for 1e-6.
The sampled flow has a transverse velocity bounded away from zero: at v = 0
the LLF speed max(|SL|, |SR|) and HLLC's star-side choice (S* = 0) have
kinks, where one-sided derivatives differ and differences cannot agree; the
forward-facing-step runs use a 0.1 transverse inflow velocity for that reason.
"""
import numpy as np
import pytest

from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from tests.helpers import smooth_state

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _ffs_scheme(limiter=S.LIM_MULTI, phi=S.PHI_VENKAT, rs=S.HLLC, ws=S.TORO):
    from rkdg_amd.adjoint import ffs_scheme

    sch = ffs_scheme()
    sch.limiter, sch.phi, sch.riemann, sch.wave_speeds = limiter, phi, rs, ws
    return sch


def _ctx(mesh, sch):
    from rkdg_amd.capi import Context

    return Context(mesh, 1, sch, dense=True)


def _flow_state(ctx, seed, amp=0.05):
    """supersonic inflow-like state with smooth perturbations"""
    xyz = ctx.node_coords()
    X, Y = xyz[..., 0].ravel(), xyz[..., 1].ravel()
    rng = np.random.default_rng(seed)
    k = rng.uniform(1, 3, 4)
    rho = 1.4 * (1 + amp * np.sin(k[0] * X + 0.3) * np.cos(k[1] * Y))
    u = 3.0 * (1 + amp * np.cos(k[2] * X * Y))
    v = 0.5 * (1 + amp * np.sin(k[3] * Y + X))
    p = 1.0 * (1 + amp * np.cos(k[1] * X - k[0] * Y))
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    return np.concatenate([rho, rho * u, rho * v, E])


def _dev(a, cuda):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


@pytest.mark.parametrize("rs", [S.RUSANOV, S.LLF, S.HLL, S.HLLC])
def test_rhs_adjoint_dot_product(cuda, rs):
    import torch

    m = M.forward_step_quad(15, 5)
    ctx = _ctx(m, _ffs_scheme(rs=rs))
    x = _dev(_flow_state(ctx, 1), cuda)
    rng = np.random.default_rng(2)
    d = _dev(rng.normal(size=x.numel()), cuda)
    w = _dev(rng.normal(size=x.numel()), cuda)
    h = 1e-6
    yp, ym = torch.empty_like(x), torch.empty_like(x)
    ctx.rhs(x + h * d, yp)
    ctx.rhs(x - h * d, ym)
    fd = float(((yp - ym) / (2 * h) * w).sum())
    xb = torch.empty_like(x)
    ctx.rhs_adjoint(x, w, xb)
    ad = float((d * xb).sum())
    print(f"rhs rs={rs}: fd {fd:.12e} adjoint {ad:.12e}", flush=True)
    assert abs(fd - ad) <= TOL * max(1.0, abs(fd))


def test_rhs_adjoint_triangles_and_addend(cuda):
    import torch

    m = M.jittered_delaunay_tri(9, 5, 0.0, 1.0, 0.0, 1.0, jitter=0.3, seed=7)
    bcs = [S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(2, S.BC_OUTLET),
           S.BC(4, S.BC_SUPERSONIC_INLET, (1.3, 2.9, 0.05, 0.0, 1.2))]
    for p in (1, 2):
        from rkdg_amd.capi import Context

        ctx = Context(m, p, S.Scheme(riemann=S.HLLC, wave_speeds=S.EINFELDT, bcs=bcs))
        x = _dev(_flow_state(ctx, 3 + p), cuda)
        rng = np.random.default_rng(p)
        d = _dev(rng.normal(size=x.numel()), cuda)
        w = _dev(rng.normal(size=x.numel()), cuda)
        a = _dev(rng.normal(size=x.numel()), cuda)
        h = 1e-6
        yp, ym = torch.empty_like(x), torch.empty_like(x)
        ctx.rhs(x + h * d, yp)
        ctx.rhs(x - h * d, ym)
        fd = float(((yp - ym) / (2 * h) * w).sum())
        xb = torch.empty_like(x)
        ctx.rhs_adjoint(x, w, xb, addend=a)
        ad = float((d * (xb - a)).sum())
        assert abs(fd - ad) <= TOL * max(1.0, abs(fd))


def test_rhs_adjoint_bc_parameters(cuda):
    import torch

    m = M.forward_step_quad(15, 5)
    sch = _ffs_scheme()
    ctx = _ctx(m, sch)
    x = _dev(_flow_state(ctx, 4), cuda)
    w = _dev(np.random.default_rng(9).normal(size=x.numel()), cuda)
    pbar = torch.zeros(S.MAX_BC * 8, dtype=torch.float64, device=cuda)
    xb = torch.empty_like(x)
    ctx.rhs_adjoint(x, w, xb, pbar=pbar)
    pbar = pbar.cpu().numpy().reshape(S.MAX_BC, 8)
    base = np.array(sch.bcs[0].params + (0.0,) * 3, dtype=np.float64)
    h = 1e-6
    for k in (0, 1, 4):
        vals = []
        for sgn in (1, -1):
            pp = base.copy()
            pp[k] += sgn * h
            ctx.set_bc_params(0, pp)
            y = torch.empty_like(x)
            ctx.rhs(x, y)
            vals.append(float((y * w).sum()))
        ctx.set_bc_params(0, base)
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - pbar[0, k]) <= TOL * max(1.0, abs(fd)), (k, fd, pbar[0, k])


@pytest.mark.parametrize("phi", [S.PHI_VENKAT, S.PHI_MICHALAK, S.PHI_BJ])
def test_limit_adjoint_dot_product(cuda, phi):
    import torch

    m = M.forward_step_quad(15, 5)
    ctx = _ctx(m, _ffs_scheme(phi=phi))
    # a state with strong sub-cell gradients so that many elements are limited
    base = _flow_state(ctx, 5, amp=0.05)
    rng = np.random.default_rng(11)
    x = _dev(base * (1 + 0.08 * rng.uniform(-1, 1, base.size)), cuda)
    troubled = torch.zeros(m.num_elements, dtype=torch.int8, device=cuda)
    y0 = x.clone()
    ctx.limit(y0, troubled)
    nlim = int((troubled.cpu().numpy() & 1).sum())
    assert nlim > 10
    d = _dev(rng.normal(size=x.numel()), cuda)
    w = _dev(rng.normal(size=x.numel()), cuda)
    h = 1e-7
    yp, ym = x + h * d, x - h * d
    ctx.limit(yp)
    ctx.limit(ym)
    fd = float(((yp - ym) / (2 * h) * w).sum())
    xb = torch.empty_like(x)
    ctx.limit_adjoint(x, w, xb)
    ad = float((d * xb).sum())
    print(f"limit phi={phi}: {nlim} limited, fd {fd:.12e} adjoint {ad:.12e}", flush=True)
    assert abs(fd - ad) <= 1e-5 * max(1.0, abs(fd))


def test_step_adjoint_dot_product(cuda):
    import torch

    m = M.forward_step_quad(15, 5)
    ctx = _ctx(m, _ffs_scheme())
    x = _dev(_flow_state(ctx, 6), cuda)
    ctx.limit(x)
    tab = S.TABLEAUS["ssprk3"]
    dt = 0.1 * ctx.hmin() / ctx.face_lambda(x) / 3
    rng = np.random.default_rng(12)
    d = _dev(rng.normal(size=x.numel()), cuda)
    w = _dev(rng.normal(size=x.numel()), cuda)
    h = 1e-7
    yp, ym = x + h * d, x - h * d
    ctx.step(yp, dt, tab)
    ctx.step(ym, dt, tab)
    fd = float(((yp - ym) / (2 * h) * w).sum())
    xb = torch.empty_like(x)
    ctx.step_adjoint(x, dt, tab, w, xb)
    ad = float((d * xb).sum())
    print(f"step: fd {fd:.12e} adjoint {ad:.12e}", flush=True)
    assert abs(fd - ad) <= 1e-5 * max(1.0, abs(fd))


def test_objective_gradient(cuda):
    import torch

    m = M.forward_step_quad(15, 5)
    ctx = _ctx(m, _ffs_scheme())
    x = _dev(_flow_state(ctx, 7), cuda)
    slot = 4  # step front
    J = torch.zeros(1, dtype=torch.float64, device=cuda)
    g = torch.zeros_like(x)
    ctx.objective(x, slot, 0.7, J, g)
    # pressure 1 +- 5 % integrated over the step front (height 0.2), times 0.7
    assert 0.7 * 0.2 * 0.9 < float(J) < 0.7 * 0.2 * 1.1
    d = _dev(np.random.default_rng(3).normal(size=x.numel()), cuda)
    h = 1e-6
    Jp = torch.zeros(1, dtype=torch.float64, device=cuda)
    Jm = torch.zeros(1, dtype=torch.float64, device=cuda)
    ctx.objective(x + h * d, slot, 0.7, Jp)
    ctx.objective(x - h * d, slot, 0.7, Jm)
    fd = float((Jp - Jm) / (2 * h))
    assert abs(fd - float((g * d).sum())) <= TOL * max(1.0, abs(fd))


@pytest.mark.parametrize("every", [0, 7])
def test_checkpointed_gradient_vs_finite_differences(cuda, every):
    import torch

    from rkdg_amd.adjoint import FFSProblem

    # 120 steps on the coarse mesh: the inflow reaches the step, so the inlet
    # parameters have an O(1) effect on J
    P = FFSProblem(15, 5, v_in=0.1)
    dts = P.schedule(120)
    mach = torch.tensor(3.0, dtype=torch.float64, device=cuda)
    delta = torch.from_numpy(P.perturbation()).to(cuda)
    prob = P.prob
    x_ic = P.initial_state(mach, delta)
    prob.set_bc(P.inlet_slot, P.inlet_params(mach).cpu().numpy())
    x0 = prob.limit(x_ic)
    J, x0bar, pbar = prob.gradient(x0, dts, every=every)
    assert J > 0.0
    assert abs(J - prob.objective(x0, dts)) <= 1e-13 * abs(J)
    # direction in the initial state
    rng = np.random.default_rng(21)
    dvec = _dev(rng.normal(size=x0.numel()) * 1e-2 * np.abs(x0.cpu().numpy()), cuda)
    h = 1e-6
    fd = (prob.objective(x0 + h * dvec, dts) - prob.objective(x0 - h * dvec, dts)) / (2 * h)
    ad = float((x0bar * dvec).sum())
    print(f"dJ/dx0.d: fd {fd:.12e} adjoint {ad:.12e}", flush=True)
    assert ad != 0.0
    assert abs(fd - ad) <= TOL * max(abs(fd), 1e-12)
    # inflow velocity parameter
    base = P.inlet_params(mach).cpu().numpy()
    vals = []
    for sgn in (1, -1):
        pp = base.copy()
        pp[1] += sgn * h
        prob.set_bc(P.inlet_slot, pp)
        vals.append(prob.objective(x0, dts))
    prob.set_bc(P.inlet_slot, base)
    fd = (vals[0] - vals[1]) / (2 * h)
    print(f"dJ/du_in: fd {fd:.12e} adjoint {pbar[P.inlet_slot, 1]:.12e}", flush=True)
    assert pbar[P.inlet_slot, 1] != 0.0
    assert abs(fd - pbar[P.inlet_slot, 1]) <= TOL * max(abs(fd), 1e-12)


def test_autograd_harness(cuda):
    import torch

    from rkdg_amd.adjoint import FFSProblem

    P = FFSProblem(15, 5, v_in=0.1)
    dts = P.schedule(12)
    mach = torch.tensor(3.0, dtype=torch.float64, device=cuda, requires_grad=True)
    delta = torch.from_numpy(P.perturbation()).to(cuda).requires_grad_(True)
    J = P.J(mach, delta, dts)
    J.backward()
    gm, gd = float(mach.grad), delta.grad.clone()
    h = 1e-6
    with torch.no_grad():
        fdm = (float(P.J(mach + h, delta, dts)) - float(P.J(mach - h, delta, dts))) / (2 * h)
        dvec = torch.from_numpy(np.random.default_rng(4).normal(size=delta.numel())).to(cuda)
        hd = 1e-5
        fdd = (float(P.J(mach, delta + hd * dvec, dts)) - float(P.J(mach, delta - hd * dvec, dts))) / (2 * hd)
    print(f"dJ/dM fd {fdm:.12e} adjoint {gm:.12e}; dJ/ddelta.d fd {fdd:.12e} adjoint {float((gd*dvec).sum()):.12e}",
          flush=True)
    assert gm != 0.0
    assert abs(fdm - gm) <= TOL * abs(fdm)
    assert abs(fdd - float((gd * dvec).sum())) <= TOL * max(abs(fdd), 1e-12)
