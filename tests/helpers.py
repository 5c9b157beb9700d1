"""Shared helpers for the RKDG tests: states, meshes, comparisons."""
from __future__ import annotations

import itertools

import numpy as np

from rkdg_amd import mesh as M
from rkdg_amd import scheme as S


def conservative(rho, vel, p, gamma=1.4, b=0.0):
    """(rho, u[dim], p) -> (rho, rho*u, rho*E) with the covolume EoS."""
    E = p / (gamma - 1.0) * (1.0 - rho * b) + 0.5 * rho * sum(v * v for v in vel)
    return [rho] + [rho * v for v in vel] + [E]


def smooth_state(xyz, seed=0, amp=1e-3, gamma=1.4):
    """Smooth periodic flow on [0,1]^d plus seeded nodal noise; byNODES layout."""
    rng = np.random.default_rng(seed)
    ne, nd, dim = xyz.shape
    X = [xyz[..., d] for d in range(dim)]
    tp = 2 * np.pi
    rho = 1.0 + 0.2 * np.sin(tp * X[0]) * np.cos(tp * X[1])
    vel = [0.3 * np.cos(tp * X[1]) + 0.1, -0.2 * np.sin(tp * X[0]) + 0.05]
    if dim == 3:
        rho = rho + 0.1 * np.sin(tp * X[2])
        vel.append(0.15 * np.cos(tp * (X[0] + X[2])))
    p = 1.0 + 0.1 * np.cos(tp * (X[0] + X[1]))
    U = conservative(rho, vel, p, gamma)
    U = np.stack([u.reshape(-1) for u in U])
    U = U * (1.0 + amp * rng.uniform(-1, 1, U.shape))
    return np.ascontiguousarray(U.reshape(-1))


def rel_linf(a, b, neq):
    a = a.reshape(neq, -1)
    b = b.reshape(neq, -1)
    out = []
    for v in range(neq):
        den = np.abs(b[v]).max()
        out.append(np.abs(a[v] - b[v]).max() / (den if den > 0 else 1.0))
    return max(out)


def cube_rotations():
    """The 24 proper rotations of the reference cube as vertex permutations
    (MFEM hex vertex order)."""
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                      [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    mats = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product([1, -1], repeat=3):
            R = np.zeros((3, 3))
            for i in range(3):
                R[i, perm[i]] = signs[i]
            if np.linalg.det(R) > 0:
                mats.append(R)
    perms = []
    for R in mats:
        p = []
        for v in range(8):
            w = R @ (verts[v] - 0.5) + 0.5
            p.append(int(np.argmin(np.abs(verts - w).sum(axis=1))))
        perms.append(p)
    return perms


def scramble_orientations(mesh, seed=1):
    """Re-number each element's vertices by a random proper symmetry of its
    reference element: same geometry, non-trivial face orientations."""
    rng = np.random.default_rng(seed)
    if mesh.geom == M.GEOM_HEX:
        perms = cube_rotations()
    elif mesh.geom == M.GEOM_QUAD:
        perms = [[(v + k) % 4 for v in range(4)] for k in range(4)]
    else:
        perms = [[(v + k) % 3 for v in range(3)] for k in range(3)]
    coords = mesh.elem_coords.copy()
    verts = mesh.elem_verts.copy()
    for e in range(mesh.num_elements):
        p = perms[rng.integers(len(perms))]
        coords[e] = mesh.elem_coords[e][p]
        verts[e] = mesh.elem_verts[e][p]
    return M.Mesh(mesh.geom, coords, verts, mesh.bdr_verts, mesh.bdr_attr, name=mesh.name + "-rot")


def all_outlet(mesh):
    return [S.BC(int(a), S.BC_OUTLET) for a in sorted(set(mesh.bdr_attr.tolist()))]


def curved_hex(nx, ny, nz, amp=0.15, seed=0, periodic=(True, True, True)):
    """structured_hex with each topological vertex moved by U(-amp, amp) times
    the spacing per coordinate; vertices on a periodic seam move together."""
    m = M.structured_hex(nx, ny, nz, periodic=periodic)
    rng = np.random.default_rng(seed)
    nv = int(m.elem_verts.max()) + 1
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    d = rng.uniform(-amp, amp, size=(nv, 3)) * h
    if not all(periodic):
        # boundary vertices stay on their planes (the boundary attributes keep their meaning)
        X = m.elem_coords.reshape(-1, 3)
        V = m.elem_verts.reshape(-1)
        for ax in range(3):
            if not periodic[ax]:
                on = (np.abs(X[:, ax]) < 1e-12) | (np.abs(X[:, ax] - 1.0) < 1e-12)
                d[V[on], ax] = 0.0
    m.elem_coords = np.ascontiguousarray(m.elem_coords + d[m.elem_verts])
    m.name += "_curved"
    return m
