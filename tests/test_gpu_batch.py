"""Batched ensembles on the GPU: instances of uneven size advanced together
(union mesh, per-instance dt/t/stop time/boundary parameters, captured HIP
graph per step, persistent right-hand-side grid) must give bit-identical states,
times and step counts to the same instances solved one after another with the
scalar time loop on the same kernels.
"""
import numpy as np
import pytest

from rkdg_amd import cases
from rkdg_amd import scheme as S

pytestmark = pytest.mark.gpu


def _instances(count, seed, ne_lo=150, ne_hi=900):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        ne = int(round(np.exp(rng.uniform(np.log(ne_lo), np.log(ne_hi)))))
        out.append(cases.nozzle(ne, rng.uniform(0.3, 0.6), rng.uniform(1.5, 8.0)))
    return out


def _sequential(cs, t_final, max_steps, cuda):
    import torch

    from rkdg_amd.solver import ExplicitRKLimitedSolver, FE_Evolution, Limiter, run

    f = FE_Evolution(cs.mesh, cs.order, cs.scheme, dense=True)
    lim = Limiter(f)
    sol = ExplicitRKLimitedSolver.from_order(3, lim)
    x = torch.from_numpy(cs.initial_state(f.ctx.node_coords())).to(cuda)
    st = run(f, sol, lim, x, t_final, cs.cfl, max_steps=max_steps)
    torch.cuda.synchronize()
    return x.cpu().numpy(), st


def _batch(css, t_finals, max_steps, cuda, use_graph=True, grid_cap=0, chunk=5):
    import torch

    from rkdg_amd.solver import Ensemble, run_batch

    ens = Ensemble([c.mesh for c in css], 1, [c.scheme for c in css])
    ens.ctx.set_options(use_graph=use_graph, grid_cap=grid_cap)
    xyz = ens.node_coords()
    x = torch.from_numpy(ens.pack([c.initial_state(p) for c, p in zip(css, xyz)])).to(cuda)
    t, dt, steps = run_batch(ens, x, t_finals, css[0].cfl, max_steps=max_steps, chunk=chunk)
    torch.cuda.synchronize()
    return ens.unpack(x.cpu().numpy()), t, dt, steps, ens


@pytest.mark.parametrize("use_graph,grid_cap", [(True, 0), (True, 7), (False, 0)])
def test_batch_bit_identical_to_sequential(cuda, use_graph, grid_cap):
    css = _instances(5, seed=11)
    # uneven stop times: instances finish after different step counts
    t_finals = np.array([0.004, 0.009, 0.006, 0.012, 0.0075])
    states, t, dt, steps, ens = _batch(css, t_finals, 0, cuda, use_graph, grid_cap)
    assert len(set(steps.tolist())) > 1
    for i, cs in enumerate(css):
        x_ref, st = _sequential(cs, t_finals[i], None, cuda)
        print(f"instance {i}: NE={cs.mesh.num_elements} steps={st.steps} t={st.t}", flush=True)
        assert steps[i] == st.steps
        assert t[i] == st.t and dt[i] == st.dt
        assert np.array_equal(states[i], x_ref), f"instance {i} differs: max {np.abs(states[i]-x_ref).max()}"


def test_batch_max_steps_and_boundary_parameters(cuda):
    # same mesh twice with different pressure ratios: the boundary parameters are per instance
    a = cases.nozzle(300, 0.4, 2.0)
    b = cases.nozzle(300, 0.4, 6.0)
    states, t, dt, steps, _ = _batch([a, b], [1.0, 1.0], 12, cuda)
    assert steps.tolist() == [12, 12]
    assert not np.array_equal(states[0], states[1])
    for i, cs in enumerate([a, b]):
        x_ref, st = _sequential(cs, 1.0, 12, cuda)
        assert np.array_equal(states[i], x_ref)
        assert t[i] == st.t


def test_tune_grid_keeps_results(cuda):
    import torch

    from rkdg_amd.solver import Ensemble

    css = _instances(4, seed=3)
    ens = Ensemble([c.mesh for c in css], 1, [c.scheme for c in css])
    xyz = ens.node_coords()
    x = torch.from_numpy(ens.pack([c.initial_state(p) for c, p in zip(css, xyz)])).to(cuda)
    y0 = torch.empty_like(x)
    y1 = torch.empty_like(x)
    lam0 = ens.ctx.rhs(x, y0)
    best, ms = ens.ctx.tune_grid(x, candidates=(3, 64, 256), reps=2)
    assert best in (0, 3, 64, 256) and all(v > 0 for v in ms.values())
    lam1 = ens.ctx.rhs(x, y1)
    assert lam0 == lam1
    assert torch.equal(y0, y1)
