"""Write tests/golden/oracle_vectors.npz: for every case of tests/golden_cases.py,
the mesh arrays, the seeded input state and the oracle's outputs on it (the
right-hand side and its wave speed, the limited state, and the state and wave
speed after `steps` limited SSP-RK3 steps of size dt).

Run from the repository root after building the oracle:
    python tests/golden/make_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.refcpu import RefSolver  # noqa: E402
from rkdg_amd import scheme as S  # noqa: E402
from tests import golden_cases as G  # noqa: E402


def main():
    arrays = {}
    names = []
    for name, (mk, p, sch, state, dt, steps) in G.cases().items():
        m = mk()
        ref = RefSolver(m, p, sch)
        x = state(ref.node_coords())
        y, lam = ref.mult(x)
        xl, _, _ = ref.limit(x)
        xs = x.copy()
        lam_s = 0.0
        for _ in range(steps):
            xs, lam_s = ref.step(xs, dt, S.TABLEAUS["ssprk3"])
        for k, v in dict(geom=m.geom, elem_coords=m.elem_coords, elem_verts=m.elem_verts,
                         bdr_verts=m.bdr_verts, bdr_attr=m.bdr_attr, p=p,
                         scheme=G.scheme_to_json(sch), dt=dt, steps=steps,
                         x=x, y=y, lam=lam, xl=xl, xs=xs, lam_s=lam_s).items():
            arrays[f"{name}/{k}"] = np.asarray(v)
        names.append(name)
        print(f"{name}: NE={m.num_elements} p={p} |y|={np.abs(y).max():.3e} lam={lam:.6f} "
              f"limited {np.abs(xl - x).max():.3e} steps lam={lam_s:.6f}")
    arrays["cases"] = np.array(names)
    np.savez_compressed(os.path.join(ROOT, G.FIXTURE), **arrays)


if __name__ == "__main__":
    main()
