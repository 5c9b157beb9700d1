"""Benchmark: DOF-stage updates/s of the RKDG step on one MI355X.

Workload (one GPU): 3D periodic acoustic pulse on [0,1]^3, 96^3 hex mesh
(884,736 elements), Q2 DG (27 nodes x 5 variables = 119,439,360 DOF), Rusanov
flux, Venkatakrishnan limiter (K = 5) on the multiplier limiter, SSP-RK3, CFL
0.2, FP64. A "step" is one RK step = 3 stages of limiter + right-hand side +
stage update; the metric counts DOF x stages per second, inputs resident in
HBM. The headline steps are the reference's limited Butcher-form step with the
SSP-RK3 tableau (rkdg_step: each stage input limited, the stage sums formed
unlimited); the Shu-Osher form with each stage value limited
(rkdg_step_ssprk3, a different limited method) is timed beside it.

The dominant kernel (the fused right-hand-side/stage kernel) is timed with HIP
events on the context's stream inside the timed region; its roofline uses the
algorithmic bytes of each launch (arrays it must read and write once, plus
geometry and connectivity). The CPU baseline is the oracle (a restatement of
the reference path with its data structures) timed on this host on a smaller
mesh of the same workload.

The same line carries the batched-ensemble figure (256 nozzle instances,
batched vs sequential solves, bit-identical check) under "ensemble", and the
adjoint/forward time ratio of the discrete adjoint (Mach 3 forward-facing
step, 2000 steps) with its finite-difference checks under "adjoint", the
double Mach reflection mesh's right-hand side and steps under "dmr", and the
curved (trilinear) Q2 hexahedra of a 48^3 box under "curved".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 96] [--no-cpu]
                       [--no-ensemble | --ensemble-only] [--no-dmr] [--no-curved]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "DOF-stage updates/s (FP64, Q2 hex) & % HBM roofline; adjoint/forward time ratio"


def pulse_state(xyz, gamma=1.4, amp=1e-3, width=0.05):
    """rho = 1 + A g, p = 1/gamma + A g, u = 0, g = exp(-|x - 1/2|^2 / w^2)."""
    r2 = ((xyz - 0.5) ** 2).sum(axis=-1).ravel()
    g = np.exp(-r2 / (width * width))
    rho = 1.0 + amp * g
    p = 1.0 / gamma + amp * g
    E = p / (gamma - 1.0)
    z = np.zeros_like(rho)
    return np.ascontiguousarray(np.concatenate([rho, z, z, z, E]))


def scheme():
    from rkdg_amd import scheme as S

    return S.Scheme(riemann=S.RUSANOV, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0,
                    limiter=S.LIM_MULTI, last_hope=1, pow_nd=1)


def stage_bytes_shu_osher(neq, ndofs, ne, dim):
    """Algorithmic bytes of the three stage-kernel launches of one SSP-RK3 step
    in the Shu-Osher form: stage 0 reads x and writes u1; stage 1 reads u1 and
    x and writes u2; stage 2 reads u2 and x and writes x. Per element as
    stage_bytes: geometry, neighbour and face-code tables, the means written."""
    vec = 8 * neq * ndofs
    per_el = 8 * (dim * dim + 1) + 2 * 4 * 2 * dim + 8 * neq
    return [a * vec + per_el * ne for a in (2, 3, 3)]


def stage_bytes(neq, ndofs, ne, dim, s=3):
    """Algorithmic bytes of the s stage-kernel launches of one limited SSP-RK3
    step as rkdg_step runs it (the tableau's stage sums arranged as
    W1 = x + dt k0, W2 = 3/4 x + 1/4 (W1 + dt k1), x = 1/3 x + 2/3 (W2 + dt k2),
    each stage input limited as the kernel loads it).

    Stage 0 reads x and writes W1; stage 1 reads W1 and x and writes W2; stage
    2 reads W2 and x and writes x. Every launch also reads the per-element
    geometry (adj(J), det J), the neighbour and face-code tables and writes
    the element means of its output; stages 1 and 2 read the limiter factor
    and the means of their input.
    """
    vec = 8 * neq * ndofs
    per_el = 8 * (dim * dim + 1) + 2 * 4 * 2 * dim + 8 * neq
    lim = 8 + 8 * neq
    arrays = [(2, 0), (3, lim), (3, lim)]
    return [a * vec + (per_el + extra) * ne for a, extra in arrays[:s]]


def limiter_bytes(neq, ndofs, ne, dim):
    """Algorithmic bytes of the limiter launches of one step: two factor passes
    (read the stage value, write one factor per element) and the final pass
    (read and write x); each reads the element's means and the deciding
    variable's means of its 2*dim face neighbours."""
    vec = 8 * neq * ndofs
    means = 8 * neq * ne + 8 * 2 * dim * ne
    return [vec + means + 8 * ne, vec + means + 8 * ne, 2 * vec + means]


def measured_traffic(n):
    """HBM bytes per stage-kernel launch from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    when it was taken on this mesh size; None otherwise."""
    path = os.path.join(ROOT, "profiles", "traffic_rhs_stage_q2hex96.md")
    if n != 96 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch ="):
            return float(line.split("=", 1)[1])
    return None


# FP64 work of the Q2 hex right-hand side per element: 64 lanes x (2 FMA + MUL +
# ADD) per wave from the SQ_INSTS_VALU_{FMA,MUL,ADD}_F64 counters of the plain
# RHS at 96^3, two elements per wave (idle lanes included: an upper bound on
# useful FLOPs; the stage epilogue's few axpy FLOPs per DOF are left out)
RHS_FLOP_PER_ELEMENT = 45952
FP64_PEAK_TFLOPS = 78.6


def copy_bandwidth(dev, n=1 << 27, reps=5):
    """Attainable HBM bandwidth of a plain streaming kernel on this GPU: a
    1 GiB read + 1 GiB write elementwise op (torch), timed with events on
    torch's stream, best of `reps`; GB/s counting both directions."""
    import torch

    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    best = None
    for _ in range(reps):
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        torch.mul(src, 1.0, out=dst)
        t1.record()
        t1.synchronize()
        ms = t0.elapsed_time(t1)
        best = ms if best is None else min(best, ms)
    del src, dst
    torch.cuda.empty_cache()
    return 2 * 8 * n / (best * 1e-3) / 1e9


def host_cpu_info():
    """CPU model, logical CPUs, sockets, physical cores and memory of this host."""
    model, sockets, cores = "", set(), set()
    phys = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and not model:
                    model = v
                elif k == "physical id":
                    phys = v
                    sockets.add(v)
                elif k == "core id":
                    cores.add((phys, v))
    except OSError:
        pass
    mem_gb = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    mem_gb = round(int(line.split()[1]) / 2**20, 1)
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "sockets": len(sockets) or None,
            "physical_cores": len(cores) or None, "mem_gib": mem_gb}


def cpu_baseline(n, steps, threads, warmup=1, mults=5):
    """The oracle on the same workload at n^3: ``steps`` SSP-RK3 steps after
    ``warmup`` steps, then the median of ``mults`` Mult calls. Returns
    (DOF-stage/s, seconds of the timed steps, elements); the Mult median and
    the setup time are left on the function object."""
    from oracle.refcpu import RefSolver, lib
    from rkdg_amd import mesh as M
    from rkdg_amd import scheme as S

    lib().ref_set_threads(threads)
    t0 = time.perf_counter()
    m = M.structured_hex(n, n, n)
    ref = RefSolver(m, 2, scheme(), threads=threads)
    x = pulse_state(ref.node_coords())
    cpu_baseline.setup_s = time.perf_counter() - t0
    _, lam = ref.face_mult(x)
    dt = 0.2 * ref.hmin() / lam / 5.0
    for _ in range(warmup):
        x, lam = ref.step(x, dt, S.TABLEAUS["ssprk3"])
    t0 = time.perf_counter()
    for _ in range(steps):
        x, lam = ref.step(x, dt, S.TABLEAUS["ssprk3"])
    el = time.perf_counter() - t0
    dofs = 5 * m.num_elements * 27
    # the right-hand side alone: median of the Mult calls after the steps above
    mt = []
    for _ in range(mults):
        t1 = time.perf_counter()
        ref.mult(x)
        mt.append(time.perf_counter() - t1)
    cpu_baseline.mult_median_s = sorted(mt)[len(mt) // 2]
    del ref
    return dofs * 3 * steps / el, el, m.num_elements


def ensemble_bench(dev, count=256, steps=20, warmup=3, seed=12345, sequential=True):
    """The batched ensemble: ``count`` converging-diverging nozzles (P1 quads,
    HLL, Michalak multiplier limiter, SSP-RK3, per-instance dynamic dt), element
    counts log-uniform in [2000, 40000]. Every instance takes warmup + steps
    steps; the batch (one captured HIP graph per step for all instances) and the
    sequential solves (the scalar time loop, one instance after another, on the
    same kernels) are timed over the last ``steps`` steps and their final states
    compared bit for bit."""
    import torch

    from rkdg_amd import cases
    from rkdg_amd import scheme as S
    from rkdg_amd.solver import Ensemble, ExplicitRKLimitedSolver, FE_Evolution, Limiter

    t0 = time.perf_counter()
    css = [cases.nozzle(ne, ar, pr) for ne, ar, pr in cases.nozzle_ensemble(count, seed)]
    tab = S.TABLEAUS["ssprk3"]
    ens = Ensemble([c.mesh for c in css], 1, [c.scheme for c in css])
    xyz = ens.node_coords()
    x = torch.from_numpy(ens.pack([c.initial_state(p) for c, p in zip(css, xyz)])).to(dev)
    ne_total = int(ens.offsets[-1])
    dof = ens.neq * ne_total * ens.nd
    print(f"[ensemble] {count} instances, {ne_total} elements, {dof} DOF, setup "
          f"{time.perf_counter() - t0:.1f}s", flush=True)
    best, ms = ens.ctx.tune_grid(x, reps=3)
    ens.ctx.init(x, np.full(count, 1e30), css[0].cfl, max_steps=warmup + steps)
    ens.ctx.advance(x, tab, warmup)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    left = ens.ctx.advance(x, tab, steps)
    torch.cuda.synchronize()
    batch_s = time.perf_counter() - t1
    _, _, nsteps, _ = ens.ctx.status()
    assert left == 0 and (nsteps == warmup + steps).all()
    xb = x.cpu().numpy()
    states = ens.unpack(xb)
    del x
    ens.ctx.close()

    if not sequential:
        # the batched leg alone (a kernel profile of the batched launches)
        out = {"instances": count, "elements": ne_total, "dof": dof, "steps_timed": steps,
               "batched_dof_stage_per_s": dof * 3 * steps / batch_s, "batched_ms_per_step": batch_s * 1e3 / steps,
               "grid_cap": best}
        print(f"[ensemble] batched {out['batched_ms_per_step']:.2f} ms/step (sequential leg skipped), grid {best}",
              flush=True)
        return out
    seq_s = 0.0
    identical = True
    for i, cs in enumerate(css):
        f = FE_Evolution(cs.mesh, 1, cs.scheme, dense=True)
        lim = Limiter(f)
        sol = ExplicitRKLimitedSolver.from_order(3, lim)
        sol.Init(f)
        xi = torch.from_numpy(cs.initial_state(f.ctx.node_coords())).to(dev)
        lim.update(xi)
        hmin = f.hmin()
        dt = cs.cfl * hmin / f.FaceMaxCharSpeed(xi) / 3
        t = 0.0
        for k in range(warmup + steps):
            if k == warmup:
                torch.cuda.synchronize()
                ts = time.perf_counter()
            sol.Step(xi, t, dt)
            dt = cs.cfl * hmin / f.max_char_speed / 3
            t += dt
        torch.cuda.synchronize()
        seq_s += time.perf_counter() - ts
        identical = identical and np.array_equal(xi.cpu().numpy(), states[i])
        f.ctx.close()
    work = dof * 3 * steps
    out = {
        "instances": count,
        "elements": ne_total,
        "dof": dof,
        "steps_timed": steps,
        "batched_dof_stage_per_s": work / batch_s,
        "sequential_dof_stage_per_s": work / seq_s,
        "ratio": seq_s / batch_s,
        "batched_ms_per_step": batch_s * 1e3 / steps,
        "sequential_ms_per_step": seq_s * 1e3 / steps,
        "bit_identical": bool(identical),
        "grid_cap": best,
        "tuner_ms": {str(k): v for k, v in ms.items()},
    }
    print(f"[ensemble] batched {out['batched_ms_per_step']:.2f} ms/step, sequential "
          f"{out['sequential_ms_per_step']:.2f} ms/step, ratio {out['ratio']:.2f}, "
          f"identical {identical}, grid {best}", flush=True)
    return out


def dmr_bench(dev, reps=20, steps=5):
    """The double Mach reflection mesh as configured (about 120k P2 triangles
    from a jittered Delaunay grid of 490x123 cells; HLLC with Einfeldt speeds,
    Michalak, SSP-RK3, CFL 0.1): the plain right-hand side timed with HIP
    events on the context's stream (reps launches) against its roofline, and
    steps of the reference dt loop. The shock is blended over one cell width
    (with the discontinuous nodal data the scheme stops on an unphysical face
    state in its first step, as the reference does)."""
    import torch

    from rkdg_amd import cases
    from rkdg_amd import scheme as S
    from rkdg_amd.capi import Context

    t0 = time.perf_counter()
    cs = cases.double_mach(shock_width=4.0 / 490)
    ctx = Context(cs.mesh, cs.order, cs.scheme)
    ne = cs.mesh.num_elements
    x = torch.from_numpy(cs.initial_state(ctx.node_coords())).to(dev)
    ctx.limit(x)
    dof = ctx.info.state_size
    y = torch.empty_like(x)
    ctx.rhs(x, y, want_lambda=False)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    ctx.profile(True)
    for _ in range(reps):
        ctx.rhs(x, y, want_lambda=False)
    ms, n, _, _ = ctx.profile_read()
    ctx.profile(False)
    avg = ms / max(1, n)
    # 16 B per DOF + the affine-triangle record (adj J, det J, 3 area normals:
    # 11 doubles) and the neighbour/face-code tables (3 faces x 2 int32)
    nbytes = 16 * dof + (11 * 8 + 24) * ne
    tab = S.TABLEAUS["ssprk3"]
    hmin = ctx.hmin()
    lam = ctx.face_lambda(x)
    dt = cs.cfl * hmin / lam / 5.0
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        lam = ctx.step(x, dt, tab)
        dt = cs.cfl * hmin / lam / 5.0
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    ctx.check_error()
    ctx.close()
    out = {"workload": f"double Mach reflection mesh, {ne} P2 triangles ({dof} DOF), HLLC + Einfeldt, "
                       f"Michalak, SSP-RK3, CFL 0.1",
           "elements": ne, "dof": dof, "rhs_avg_launch_ms": avg, "rhs_launches": n,
           "rhs_algorithmic_bytes": nbytes, "rhs_achieved_gbs": nbytes / (avg * 1e-3) / 1e9,
           "rhs_frac_of_hbm": nbytes / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "rhs_dof_per_s": dof / (avg * 1e-3),
           "steps": steps, "ms_per_step": el * 1e3 / steps, "dof_stage_per_s": dof * 3 * steps / el,
           "setup_s": setup}
    print(f"[dmr] {ne} P2 triangles: RHS {avg:.3f} ms ({out['rhs_achieved_gbs']:.0f} GB/s, "
          f"{out['rhs_frac_of_hbm'] * 100:.1f} % of HBM), step {out['ms_per_step']:.3f} ms", flush=True)
    return out


def curved_bench(dev, n=48, reps=20, steps=3):
    """Curved (trilinear) hexahedra: the n^3 periodic box with every vertex
    moved by a seeded U(-0.15, 0.15) of the spacing per coordinate, Q2, the
    headline's scheme (Rusanov, Venkatakrishnan K = 5, SSP-RK3, CFL 0.2). The
    plain right-hand side timed with HIP events on the context's stream against
    its roofline (16 B per DOF, the element's CurvedLayout record and the face
    tables), and steps of the reference dt loop."""
    import torch

    from rkdg_amd import mesh as M
    from rkdg_amd import scheme as S
    from rkdg_amd.capi import Context

    t0 = time.perf_counter()
    m = M.structured_hex(n, n, n)
    rng = np.random.default_rng(0)
    nv = int(m.elem_verts.max()) + 1
    m.elem_coords = np.ascontiguousarray(m.elem_coords + rng.uniform(-0.15, 0.15, (nv, 3))[m.elem_verts] / n)
    sch = S.Scheme(riemann=S.RUSANOV, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0, limiter=S.LIM_MULTI)
    ctx = Context(m, 2, sch)
    if ctx.info.affine:
        raise RuntimeError("the curved box came out affine")
    ne = m.num_elements
    dof = ctx.info.state_size
    xyz = ctx.node_coords()
    r2 = ((xyz - 0.5) ** 2).sum(axis=-1)
    rho = 1.0 + 0.1 * np.exp(-r2 / 0.01)
    p = 1.0 + 0.1 * np.exp(-r2 / 0.01)
    z = np.zeros_like(rho)
    x = torch.from_numpy(np.concatenate([rho.ravel(), z.ravel(), z.ravel(), z.ravel(),
                                         (p / 0.4).ravel()])).to(dev)
    y = torch.empty_like(x)
    ctx.rhs(x, y, want_lambda=False)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    ctx.profile(True)
    for _ in range(reps):
        ctx.rhs(x, y, want_lambda=False)
    ms, cnt, _, _ = ctx.profile_read()
    ctx.profile(False)
    avg = ms / max(1, cnt)
    # CurvedLayout<2>: adj(J) at 27 nodes, the 27x27 inverse mass, 27 mean
    # weights, 6 x 16 face points x 5, the volume: 1480 doubles; with the
    # affine block (10 doubles) and nbr/fcode (6 x 2 int32)
    nbytes = 16 * dof + (1480 * 8 + 80 + 48) * ne
    tab = S.TABLEAUS["ssprk3"]
    hmin = ctx.hmin()
    lam = ctx.face_lambda(x)
    dt = 0.2 * hmin / lam / 5.0
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        lam = ctx.step(x, dt, tab)
        dt = 0.2 * hmin / lam / 5.0
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    ctx.check_error()
    ctx.close()
    out = {"workload": f"{n}^3 periodic box of curved (trilinear) Q2 hexahedra, corners moved by up to 15 % "
                       f"of the spacing; Rusanov, Venkatakrishnan K=5, SSP-RK3",
           "elements": ne, "dof": dof, "rhs_avg_launch_ms": avg, "rhs_launches": cnt,
           "rhs_algorithmic_bytes": nbytes, "rhs_achieved_gbs": nbytes / (avg * 1e-3) / 1e9,
           "rhs_frac_of_hbm": nbytes / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "steps": steps, "ms_per_step": el * 1e3 / steps, "dof_stage_per_s": dof * 3 * steps / el,
           "setup_s": setup}
    print(f"[curved] {ne} curved Q2 hexes: RHS {avg:.3f} ms ({out['rhs_achieved_gbs']:.0f} GB/s, "
          f"{out['rhs_frac_of_hbm'] * 100:.1f} % of HBM), step {out['ms_per_step']:.3f} ms, "
          f"setup {setup:.1f} s", flush=True)
    return out


def adjoint_bench(dev, nsteps=2000, nx=240, ny=80, oracle_steps=200, fd_steps=1000, threads=None):
    """The discrete adjoint on the Mach 3 forward-facing step (240x80 minus the
    step, P1 quads, HLLC, Venkatakrishnan K=5, SSP-RK3): objective = pressure
    integrated over the step face and over time; parameters = inflow Mach and
    a seeded 1e-3 density perturbation field. The dt schedule of one dynamic run
    is frozen. Times one forward solve and one gradient (sqrt(N) checkpointed
    adjoint, its own forward pass included) over nsteps steps, and checks the
    gradient by central differences (h = 1e-6) of the GPU forward solve over
    fd_steps steps and of the oracle's forward solve over oracle_steps steps.
    Beyond about 1000 steps J(M) stops being differentiable at the scale of h
    (limiter thresholds and resets switch inside +-h: the differences change by
    orders of magnitude with h), so differences cannot check a gradient there;
    the full horizon is reported for the time ratio only."""
    import torch

    from rkdg_amd.adjoint import FFSProblem

    P = FFSProblem(nx, ny)
    prob = P.prob
    dts = P.schedule(nsteps, device=dev)
    mach = torch.tensor(3.0, dtype=torch.float64, device=dev)
    delta = torch.from_numpy(P.perturbation()).to(dev)

    def x0_of(m, d):
        prob.set_bc(P.inlet_slot, P.inlet_params(m).cpu().numpy())
        return prob.limit(P.initial_state(m, d))

    x0 = x0_of(mach, delta)
    prob.gradient(x0, dts[:50])  # allocations and warm-up
    prob.objective(x0, dts[:50])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = prob.objective(x0, dts)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    Ja, x0bar, pbar = prob.gradient(x0, dts)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    # full parameter gradient through the harness (initial limiter + IC map)
    dfd = dts[:fd_steps]
    mr = mach.clone().requires_grad_(True)
    dr = delta.clone().requires_grad_(True)
    Jt = P.J(mr, dr, dfd)
    Jt.backward()
    gm = float(mr.grad)
    # a smooth direction in the density field: a per-node random one moves
    # limiter diffs of the 1e-3 noise across the |diff| = 1e-3 threshold within
    # +-h (J jumps there), so its differences only agree for much smaller h
    xyz = prob.ctx.node_coords().reshape(-1, 2)
    dvec = torch.from_numpy(np.sin(2.0 * xyz[:, 0]) * np.cos(3.0 * xyz[:, 1])).to(dev)
    gd = float((dr.grad * dvec).sum())
    h = 1e-6
    with torch.no_grad():
        fdm = (float(P.J(mach + h, delta, dfd)) - float(P.J(mach - h, delta, dfd))) / (2 * h)
        fdd = (float(P.J(mach, delta + h * dvec, dfd)) - float(P.J(mach, delta - h * dvec, dfd))) / (2 * h)
    out = {
        "workload": f"Mach 3 forward-facing step, {P.mesh.num_elements} P1 quads, HLLC, Venkatakrishnan K=5, "
                    f"SSP-RK3, {nsteps} steps (frozen dt schedule)",
        "steps": nsteps,
        "J": J,
        "forward_s": t_fwd,
        "gradient_s": t_adj,
        "adjoint_forward_ratio": t_adj / t_fwd,
        "fd_steps": fd_steps,
        "dJ_dMach": gm,
        "fd_dJ_dMach": fdm,
        "rel_err_mach": abs(gm - fdm) / abs(fdm),
        "dJ_ddelta_dot_d": gd,
        "fd_dJ_ddelta_dot_d": fdd,
        "rel_err_delta": abs(gd - fdd) / abs(fdd),
        "J_adjoint_pass_matches": abs(Ja - J) <= 1e-13 * abs(J),
    }
    print(f"[adjoint] forward {t_fwd:.3f}s gradient {t_adj:.3f}s ratio {t_adj / t_fwd:.2f}; "
          f"over {fd_steps} steps dJ/dM {gm:.9e} "
          f"(fd {fdm:.9e}, rel {out['rel_err_mach']:.2e}); dJ/ddelta.d {gd:.9e} (fd {fdd:.9e}, "
          f"rel {out['rel_err_delta']:.2e})", flush=True)
    if oracle_steps:
        from oracle.refcpu import RefSolver
        from rkdg_amd.adjoint import ffs_scheme

        th = threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ds = dts[:oracle_steps]
        tab = None
        from rkdg_amd import scheme as S

        tab = S.TABLEAUS["ssprk3"]

        def oracle_J(m):
            ref = RefSolver(P.mesh, 1, ffs_scheme(m), threads=th)
            xi = P.initial_state(torch.tensor(m, dtype=torch.float64), delta.cpu()).numpy()
            x = ref.limit(xi)[0]
            total = 0.0
            for n in range(len(ds)):
                x, _ = ref.step(x, ds[n], tab)
                total += ds[n] * ref.patch_pressure(x, 5)
            return total

        t0 = time.perf_counter()
        fdo = (oracle_J(3.0 + h) - oracle_J(3.0 - h)) / (2 * h)
        t_or = time.perf_counter() - t0
        mr = mach.clone().requires_grad_(True)
        Jt = P.J(mr, delta, ds)
        Jt.backward()
        gmo = float(mr.grad)
        out.update({"oracle_steps": oracle_steps, "oracle_fd_dJ_dMach": fdo, "dJ_dMach_at_oracle_steps": gmo,
                    "rel_err_mach_vs_oracle_fd": abs(gmo - fdo) / abs(fdo), "oracle_fd_s": t_or,
                    "oracle_threads": th})
        print(f"[adjoint] oracle FD over {oracle_steps} steps: {fdo:.9e} vs adjoint {gmo:.9e} "
              f"(rel {out['rel_err_mach_vs_oracle_fd']:.2e}, {t_or:.1f}s)", flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--cfl", type=float, default=0.2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=96,
                    help="mesh of the oracle's timed sample (96: the benchmark mesh itself, about 40 GB of host "
                         "memory for the oracle's CSR Aflux and dense inverse masses; 64 when the host has less)")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-n2", type=int, default=48, help="second oracle sample (0: none)")
    ap.add_argument("--cpu-steps2", type=int, default=8)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB read + 1 GiB write elementwise kernel (3 launches), "
                         "the reference for scaling FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--ensemble-count", type=int, default=256)
    ap.add_argument("--ensemble-batched-only", action="store_true",
                    help="with --ensemble-only: skip the sequential solves (profiles of the batched launches)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-dmr", action="store_true")
    ap.add_argument("--no-curved", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--oracle-steps", type=int, default=200)
    ap.add_argument("--profile-rhs", action="store_true",
                    help="only run the plain right-hand side 3 times and the limiter once on the 96^3 "
                         "workload (the driver of the counter profiles), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the RKDG path is built and measured on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import torch

    from rkdg_amd import mesh as M
    from rkdg_amd import scheme as S
    from rkdg_amd.capi import Context

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_bench(dev, args.ensemble_count,
                                                     sequential=not args.ensemble_batched_only)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench(dev, args.adjoint_steps, oracle_steps=args.oracle_steps)}),
              flush=True)
        return
    n = args.n
    t0 = time.perf_counter()
    m = M.structured_hex(n, n, n)
    ctx = Context(m, 2, scheme())
    if args.profile_rhs:
        x = torch.from_numpy(pulse_state(ctx.node_coords())).to(dev)
        y = torch.empty_like(x)
        for _ in range(3):
            ctx.rhs(x, y, want_lambda=False)
        ctx.limit(x)
        torch.cuda.synchronize()
        print(f"[bench] profile driver: 3 right-hand sides and 1 limiter pass at {n}^3", flush=True)
        return
    info = ctx.info
    x0 = pulse_state(ctx.node_coords())
    print(f"[bench] mesh {n}^3 Q2 hex: NE={info.num_elements} DOF={info.state_size} "
          f"setup {time.perf_counter() - t0:.1f}s", flush=True)
    x = torch.from_numpy(x0).to(dev)
    tab = S.TABLEAUS["ssprk3"]
    hmin = ctx.hmin()
    lam = ctx.face_lambda(x)
    ctx.limit(x)
    dt = args.cfl * hmin / lam / (2 * 2 + 1)
    for _ in range(args.warmup):
        lam = ctx.step(x, dt, tab)
        dt = args.cfl * hmin / lam / 5.0
    torch.cuda.synchronize()
    # the Shu-Osher form with each stage value limited (rkdg_step_ssprk3),
    # timed the same way from a copy of the state
    xb = x.clone()
    dtb = dt
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lamb = ctx.step_ssprk3(xb, dtb)
        dtb = args.cfl * hmin / lamb / 5.0
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t0
    rhs_ms_b, rhs_n_b, lim_ms_b, lim_n_b = ctx.profile_read()
    ctx.profile(False)
    del xb
    # the headline: the reference's limited Butcher-form step (rkdg_step with
    # the SSP-RK3 tableau) in the reference time loop (each step's last-stage
    # wave speed sets the next dt)
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lam = ctx.step(x, dt, tab)
        dt = args.cfl * hmin / lam / 5.0
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rhs_ms, rhs_n, lim_ms, lim_n = ctx.profile_read()
    ctx.profile(False)
    ctx.check_error()
    lam = ctx.face_lambda(x)
    # the plain right-hand side (rkdg_rhs, FE_Evolution::Mult): read x, write y,
    # HIP events on the context's stream around each of 10 launches
    y = torch.empty_like(x)
    ctx.rhs(x, y, want_lambda=False)
    torch.cuda.synchronize()
    ctx.profile(True)
    for _ in range(10):
        ctx.rhs(x, y, want_lambda=False)
    plain_ms, plain_n, _, _ = ctx.profile_read()
    ctx.profile(False)
    del y
    plain_avg_ms = plain_ms / max(1, plain_n)

    copy_gbs = copy_bandwidth(dev)
    dofs = info.state_size
    ms_step = el * 1e3 / args.steps
    value = dofs * 3 * args.steps / el
    sb = stage_bytes(5, info.num_dofs, info.num_elements, 3)
    bytes_per_launch = sum(sb) / len(sb)
    sbb = stage_bytes_shu_osher(5, info.num_dofs, info.num_elements, 3)
    bytes_per_launch_b = sum(sbb) / len(sbb)
    rhs_avg_ms_b = rhs_ms_b / max(1, rhs_n_b)
    shu_osher = {"ms_per_step": el_b * 1e3 / args.steps, "value": info.state_size * 3 * args.steps / el_b,
               "stage_kernel_avg_ms": rhs_avg_ms_b, "algorithmic_bytes_per_launch": bytes_per_launch_b,
               "achieved_gbs": bytes_per_launch_b / (rhs_avg_ms_b * 1e-3) / 1e9,
               "frac": bytes_per_launch_b / (rhs_avg_ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS,
               "limiter_avg_ms": lim_ms_b / max(1, lim_n_b)}
    rhs_avg_ms = rhs_ms / max(1, rhs_n)
    step_bytes = sum(sb) + sum(limiter_bytes(5, info.num_dofs, info.num_elements, 3))
    lim_avg_ms = lim_ms / max(1, lim_n)
    achieved = bytes_per_launch / (rhs_avg_ms * 1e-3) / 1e9
    lb = limiter_bytes(5, info.num_dofs, info.num_elements, 3)
    lim_bytes = sum(lb) / len(lb)
    rhs_bytes = 16 * dofs + (80 + 48) * info.num_elements
    traffic = measured_traffic(n)
    print(f"[bench] stage kernel avg {rhs_avg_ms:.3f} ms over {rhs_n} launches "
          f"({achieved:.0f} GB/s algorithmic; plain copy {copy_gbs:.0f} GB/s), limiter avg {lim_avg_ms:.3f} ms "
          f"({lim_bytes / (lim_avg_ms * 1e-3) / 1e9:.0f} GB/s), step {ms_step:.3f} ms, "
          f"lambda {lam:.6f}; plain RHS {plain_avg_ms:.3f} ms ({rhs_bytes / (plain_avg_ms * 1e-3) / 1e9:.0f} GB/s, "
          f"{rhs_bytes / (plain_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS * 100:.1f} % of HBM)", flush=True)

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        host = host_cpu_info()
        cn = args.cpu_n
        if cn == 96 and (host["mem_gib"] or 0) < 64:
            cn = 64  # the oracle's 96^3 structures need about 40 GB
        cv, cel, cne = cpu_baseline(cn, args.cpu_steps, threads, warmup=0)
        cpu_mult_s, cpu_setup_s = cpu_baseline.mult_median_s, cpu_baseline.setup_s
        print(f"[bench] cpu baseline at {cn}^3: {cv:.3e} DOF-stage/s ({cel:.1f}s, {threads} threads), Mult "
              f"median {cpu_mult_s:.2f}s, setup {cpu_setup_s:.0f}s", flush=True)
        second = None
        if args.cpu_n2:
            cv2, cel2, cne2 = cpu_baseline(args.cpu_n2, args.cpu_steps2, threads)
            second = {"value": cv2, "elements": cne2, "mult_median_s": cpu_baseline.mult_median_s,
                      "sample": f"{args.cpu_n2}^3 elements, {args.cpu_steps2} SSP-RK3 steps after 1 warm-up "
                                f"step, {cel2:.1f} s, OpenMP {threads} threads"}
        cv1, cel1, cne1 = cpu_baseline(24, 1, 1)
        from oracle.refcpu import lib as _oracle_lib

        _oracle_lib().ref_set_threads(threads)
        cpu = {"value": cv, "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
               "sample": f"oracle (CSR Aflux, dense inverse mass, per-point shape evaluation) on the "
                         f"benchmark workload at {cn}^3 elements ({cne} elements, {5 * cne * 27} DOF): "
                         f"{args.cpu_steps} SSP-RK3 step(s) of 3 stages (limiter + Mult + update), "
                         f"{cel:.1f} s, OpenMP {threads} threads; then the median of 5 Mult calls",
               "host": host,
               "mult_median_s": cpu_mult_s,
               "mult_dof_per_s": 5 * cne * 27 / cpu_mult_s,
               "setup_s": cpu_setup_s,
               "second_sample": second,
               "value_1_thread": cv1,
               "sample_1_thread": f"24^3 elements ({cne1}), 1 step after 1 warm-up step, {cel1:.1f} s"}
        print(f"[bench] cpu baseline {cv:.3e} DOF-stage/s ({cel:.1f}s, {threads} threads); "
              f"1 thread {cv1:.3e} ({cel1:.1f}s)", flush=True)

    ens = None
    if not (args.no_ensemble and args.no_adjoint and args.no_dmr and args.no_curved):
        del x
        ctx.close()
        torch.cuda.empty_cache()
    dmr = None
    if not args.no_dmr:
        dmr = dmr_bench(dev)
    curved = None
    if not args.no_curved:
        curved = curved_bench(dev)
    if not args.no_ensemble:
        ens = ensemble_bench(dev, args.ensemble_count)
    adj = None
    if not args.no_adjoint:
        adj = adjoint_bench(dev, args.adjoint_steps, oracle_steps=0 if args.no_cpu else args.oracle_steps)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D periodic acoustic pulse, {n}^3 Q2 hex ({info.num_elements} elements, "
                        f"{dofs} DOF), Rusanov, Venkatakrishnan K=5 multiplier limiter, SSP-RK3, "
                        f"CFL {args.cfl}",
            "elements": info.num_elements,
            "dof": dofs,
            "stages_per_step": 3,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "tensor_rhs_kernel<3,2> (fused RHS + SSP-RK3 stage update, limiter on load)",
            "avg_launch_ms": rhs_avg_ms,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "copy_gbs": copy_gbs,
            "frac_of_copy": achieved / copy_gbs,
            "fp64": {
                "flop_per_element": RHS_FLOP_PER_ELEMENT,
                "achieved_tflops": RHS_FLOP_PER_ELEMENT * info.num_elements / (rhs_avg_ms * 1e-3) / 1e12,
                "peak_tflops": FP64_PEAK_TFLOPS,
                "frac": RHS_FLOP_PER_ELEMENT * info.num_elements / (rhs_avg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            },
            # the plain rkdg_rhs (FE_Evolution::Mult) against its HBM bound:
            # 16 B per DOF (read x, write y) + geometry 80 B and
            # connectivity 48 B per element, timed with HIP events
            "rhs": {
                "kernel": "tensor_rhs_kernel<3,2> as rkdg_rhs (y = RHS(x))",
                "avg_launch_ms": plain_avg_ms,
                "algorithmic_bytes_per_launch": rhs_bytes,
                "achieved": rhs_bytes / (plain_avg_ms * 1e-3) / 1e9,
                "frac": rhs_bytes / (plain_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            },
            # the whole step against the stage ceilings: 24 B per DOF-stage (the
            # survey's low-storage stage) and the bytes of the step as run here
            # (stage and limiter launches, per DOF-stage)
            "step": {
                "dof_stage_per_s": value,
                "ceiling_24B": HBM_PEAK_GBS * 1e9 / 24.0,
                "frac_of_24B_ceiling": value / (HBM_PEAK_GBS * 1e9 / 24.0),
                "bytes_per_dof_stage_as_run": step_bytes / (3 * dofs),
                "ceiling_as_run": HBM_PEAK_GBS * 1e9 / (step_bytes / (3 * dofs)),
                "frac_of_as_run_ceiling": value / (HBM_PEAK_GBS * 1e9 / (step_bytes / (3 * dofs))),
                "limiter_avg_ms": lim_avg_ms,
            },
            "shu_osher_form": shu_osher,
        },
        "cpu_baseline": cpu,
        "dmr": dmr,
        "curved": curved,
        "ensemble": ens,
        "adjoint": adj,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
