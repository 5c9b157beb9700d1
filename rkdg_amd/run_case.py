"""Runs a case directory the way the reference's executable does (main.cpp):
read ``settings.yml`` and its Gmsh mesh, project the initial state (or load
the restart frame ``time.restartCycle`` when ``time.restart`` is set), then
the time loop with the limiter after projection, dt from CoMax (or constant),
ParaView output and restart frames every ``postProcess.visSteps`` steps, and
one log line per step (main.cpp:776-785).

    python -m rkdg_amd.run_case CASE_DIR [--max-steps N] [--out DIR] [--lod L]

The GPU does the work; without the HIP library this stops with its error.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import case_file as CF
from . import io
from . import scheme as S


def run_case(case_dir: str, out_dir: str = None, max_steps: int = None, lod: int = 1, quiet: bool = False,
             device: int = 0):
    import torch

    from .solver import ExplicitRKLimitedSolver, FE_Evolution, Limiter, run

    cs = CF.load_case(case_dir)
    out_dir = out_dir or cs.case_dir
    cuda = torch.device("cuda", device)
    torch.cuda.set_device(cuda)
    if cs.adaptive:
        return _run_adaptive_case(cs, out_dir, max_steps, lod, quiet, cuda)
    f = FE_Evolution(cs.mesh, cs.order, cs.scheme)
    lim = Limiter(f)
    sol = ExplicitRKLimitedSolver(*S.TABLEAUS[cs.tableau], limiter=lim)
    pp = io.PostProcessor(f.ctx, out_dir, vis_steps=max(cs.vis_steps, 1), n_saved=max(cs.n_saved_frames, 1),
                          lod=lod, check_totals=cs.check_total_energy)
    if not quiet:
        print(f"case {cs.name}: {cs.mesh.num_elements} cells, order {cs.order}, "
              f"{f.ctx.info.num_dofs * (cs.mesh.dim + 2)} unknowns, tableau {cs.tableau}"
              + (f", patches without a physical group: {cs.skipped_patches}" if cs.skipped_patches else ""),
              flush=True)
    t0, step0, dt_const, dt0 = 0.0, 0, cs.dt_const, None
    if cs.restart:
        # case_manager.cpp:229-253, 460-475, 484-487: state, time, cycle and dt of the frame
        step0, t0, dt0, x = io.RestartFrames(out_dir, restart_cycle=cs.restart_cycle).load(cs.restart_cycle)
        x = torch.from_numpy(np.ascontiguousarray(x)).to(cuda)
    else:
        x = torch.from_numpy(cs.initial_state(f.ctx.node_coords())).to(cuda)
    clock = time.perf_counter()
    last = [clock]

    def callback(steps, t, dt, xs):
        done = t >= cs.t_final - 1e-8 * dt
        pp(step0 + steps, t, dt, xs, done=done)
        if not quiet:
            now = time.perf_counter()
            e_tot = pp.log[-1]["totals"][-1] if cs.check_total_energy else 0.0
            print(f"Time step: {step0 + steps}\tdt: {dt:.6g}\tPhys time: {t:.6g} \tExecution time: "
                  f"{now - clock:.6g} \tReal dt time: {now - last[0]:.6g} \t E_total: {e_tot:.18g}", flush=True)
            last[0] = now

    stats = run(f, sol, lim, x, cs.t_final, cs.cfl, t0=t0, dt_const=dt_const, max_steps=max_steps,
                callback=callback, dt0=dt0)
    # a run cut short by max_steps still leaves its last state on disk
    last_step = step0 + stats.steps
    if not stats.t >= cs.t_final - 1e-8 * stats.dt and last_step % pp.vis_steps != 0:
        pp.write(last_step, stats.t, stats.dt, x)
    return cs, stats, pp, x


def _run_adaptive_case(cs, out_dir, max_steps, lod, quiet, cuda):
    """mesh.adaptive: the reference's refinement loop around every step
    (main.cpp:253-386, 505-700) through amr.run_adaptive, the GPU context
    re-created after every change of the mesh; one log line per step with the
    element count, and the final state written on the final mesh."""
    import types

    import torch

    from . import amr
    from .capi import Context

    clock = time.perf_counter()

    def callback(steps, t, dt, mesh, x):
        if not quiet:
            print(f"Time step: {steps}\tdt: {dt:.6g}\tPhys time: {t:.6g} \tExecution time: "
                  f"{time.perf_counter() - clock:.6g} \tElements: {mesh.num_elements} "
                  f"(non-conforming faces {len(mesh.nc)})", flush=True)

    r = amr.run_adaptive(cs.mesh, cs.order, cs.scheme, cs.initial_state, cs.t_final, cs.cfl,
                         S.TABLEAUS[cs.tableau], cs.max_elem_error, hysteresis=cs.hysteresis,
                         total_error_fraction=cs.total_error_fraction, max_steps=max_steps, solver="gpu",
                         max_level=cs.max_ref_level, dt_const=cs.dt_const, callback=callback)
    ctx = Context(r.mesh, cs.order, cs.scheme)
    pp = io.PostProcessor(ctx, out_dir, vis_steps=1, n_saved=max(cs.n_saved_frames, 1), lod=lod,
                          check_totals=cs.check_total_energy)
    x = torch.from_numpy(np.ascontiguousarray(r.x)).to(cuda)
    pp.write(r.steps, r.t, r.dt, x)
    stats = types.SimpleNamespace(steps=r.steps, t=r.t, dt=r.dt, elements=r.elements, mesh=r.mesh)
    return cs, stats, pp, x


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("case_dir")
    ap.add_argument("--out", default=None, help="output directory (default: the case directory)")
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--lod", type=int, default=1, help="ParaView sub-cells per element edge")
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    if not os.path.isdir(a.case_dir) and not os.path.isfile(a.case_dir):
        print(f"Case file {a.case_dir} is not found", file=sys.stderr)
        return 1
    _, stats, _, _ = run_case(a.case_dir, a.out, a.max_steps, a.lod, a.quiet)
    print(f"done: {stats.steps} steps, t = {stats.t:.9g}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
