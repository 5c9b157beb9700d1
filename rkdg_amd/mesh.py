"""Host-side mesh descriptions and generators for the RKDG path.

A mesh is described the way the C-ABI consumes it (the analogue of an MFEM
mesh with discontinuous nodes):

* ``elem_coords`` (NE, nv, dim) float64 -- vertex coordinates per element, in
  MFEM's reference vertex order (quad: (0,0),(1,0),(1,1),(0,1); triangle:
  (0,0),(1,0),(0,1); hex: bottom face counter-clockwise then top face);
* ``elem_verts`` (NE, nv) int64 -- topological vertex ids; periodic meshes use
  identified ids so that faces across the period match;
* ``bdr_verts`` (NB, nvf) int64 and ``bdr_attr`` (NB,) int32 -- boundary
  elements with 1-based attributes (MFEM ``bdr_attributes``).

Generators cover the workloads the RKDG path is exercised on: structured quads
(Sod tube, forward-facing step), periodic hexes (acoustic pulse), jittered
Delaunay triangles (double Mach reflection) and converging-diverging nozzles.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GEOM_QUAD = 0
GEOM_TRI = 1
GEOM_HEX = 2

_NV = {GEOM_QUAD: 4, GEOM_TRI: 3, GEOM_HEX: 8}
_DIM = {GEOM_QUAD: 2, GEOM_TRI: 2, GEOM_HEX: 3}
_NVF = {GEOM_QUAD: 2, GEOM_TRI: 2, GEOM_HEX: 4}


@dataclass
class Mesh:
    geom: int
    elem_coords: np.ndarray
    elem_verts: np.ndarray
    bdr_verts: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int64))
    bdr_attr: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    name: str = ""
    # non-conforming faces (rows: slave element, its local face, master element,
    # its local face): the slave's face covers part of the coarser master's face
    nc: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int64))
    # per record, the slave face's corners (in the slave's local face order) in
    # the master face's parametrisation (face vertex 0 at 0; 2D one, 3D two
    # coordinates per corner): MFEM's slave point matrix
    nc_param: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float64))

    def __post_init__(self):
        self.elem_coords = np.ascontiguousarray(self.elem_coords, dtype=np.float64)
        self.elem_verts = np.ascontiguousarray(self.elem_verts, dtype=np.int64)
        nvf = _NVF[self.geom]
        self.bdr_verts = np.ascontiguousarray(np.asarray(self.bdr_verts, dtype=np.int64).reshape(-1, nvf))
        self.bdr_attr = np.ascontiguousarray(self.bdr_attr, dtype=np.int32)
        self.nc = np.ascontiguousarray(np.asarray(self.nc, dtype=np.int64).reshape(-1, 4))
        npar = nvf * (_DIM[self.geom] - 1)
        self.nc_param = np.ascontiguousarray(np.asarray(self.nc_param, dtype=np.float64).reshape(-1, npar))
        assert len(self.nc_param) == len(self.nc)
        assert self.elem_coords.shape == (self.num_elements, _NV[self.geom], _DIM[self.geom])
        assert self.elem_verts.shape == (self.num_elements, _NV[self.geom])

    @property
    def dim(self) -> int:
        return _DIM[self.geom]

    @property
    def num_elements(self) -> int:
        return int(self.elem_verts.shape[0])

    def num_equations(self) -> int:
        return self.dim + 2


def structured_quad(nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=1.0, periodic=(False, False),
                    warp=None, remove=None) -> Mesh:
    """nx*ny quads on a box, x fastest. Attributes: 1 bottom, 2 right, 3 top, 4 left.

    ``warp(X, Y) -> (X, Y)`` maps the vertex grid (e.g. nozzle walls);
    ``remove(xc, yc) -> bool mask`` deletes elements by centroid (step geometry).
    """
    for n, per in ((nx, periodic[0]), (ny, periodic[1])):
        if per and n < 3:
            raise ValueError("a periodic direction needs at least 3 elements")
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")  # (ny+1, nx+1)
    if warp is not None:
        X, Y = warp(X, Y)
    px, py = periodic

    def vid(i, j):
        if px:
            i = i % nx
        if py:
            j = j % ny
        return i + (nx + 1) * j

    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    i = i.ravel()
    j = j.ravel()
    corners = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
    coords = np.stack([np.stack([X[cj, ci], Y[cj, ci]], axis=-1) for ci, cj in corners], axis=1)
    verts = np.stack([vid(ci, cj) for ci, cj in corners], axis=1)
    keep = np.ones(len(i), bool)
    if remove is not None:
        xc = coords[:, :, 0].mean(axis=1)
        yc = coords[:, :, 1].mean(axis=1)
        keep = ~remove(xc, yc)
    coords = coords[keep]
    verts = verts[keep]
    bv, ba = _boundary_from_topology(GEOM_QUAD, verts, coords, _box_attr_2d(x0, x1, y0, y1))
    return Mesh(GEOM_QUAD, coords, verts, bv, ba, name=f"quad{nx}x{ny}")


def _box_attr_2d(x0, x1, y0, y1):
    def attr(mid, nrm):
        # outward normal of a straight boundary edge decides the attribute
        if nrm[1] < -0.5:
            return 1
        if nrm[0] > 0.5:
            return 2
        if nrm[1] > 0.5:
            return 3
        return 4
    return attr


_QUAD_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0)]
_TRI_EDGES = [(0, 1), (1, 2), (2, 0)]
_HEX_FACES = [(3, 2, 1, 0), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)]


def _local_faces(geom):
    return {GEOM_QUAD: _QUAD_EDGES, GEOM_TRI: _TRI_EDGES, GEOM_HEX: _HEX_FACES}[geom]


def _boundary_from_topology(geom, verts, coords, attr_fn):
    """Faces seen once are boundary faces (in element, then local-face order);
    attribute from attr_fn(midpoint, outward normal)."""
    lf = _local_faces(geom)
    nvf = len(lf[0])
    loc = np.array(lf, np.int64)                       # (nf, nvf)
    fv = verts[:, loc].reshape(-1, nvf)                # (NE*nf, nvf), element-major
    keys = np.sort(fv, axis=1)
    order = np.lexsort(keys.T[::-1])
    ks = keys[order]
    new_group = np.ones(len(ks), bool)
    new_group[1:] = np.any(ks[1:] != ks[:-1], axis=1)
    gid = np.cumsum(new_group) - 1
    cnt = np.bincount(gid)
    single = np.empty(len(ks), bool)
    single[order] = cnt[gid] == 1
    once = np.flatnonzero(single)
    bv, ba = [], []
    nf = len(lf)
    for idx in once:
        e, f = divmod(int(idx), nf)
        pts = coords[e, list(lf[f])]
        mid = pts.mean(axis=0)
        if pts.shape[1] == 2:
            t = pts[1] - pts[0]
            n = np.array([t[1], -t[0]])
        else:
            n = np.cross(pts[1] - pts[0], pts[3] - pts[0])
        n = n / np.linalg.norm(n)
        bv.append(fv[idx])
        ba.append(attr_fn(mid, n))
    return np.array(bv, np.int64).reshape(-1, nvf), np.array(ba, np.int32)


def structured_hex(nx, ny, nz, box=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)),
                   periodic=(True, True, True)) -> Mesh:
    """nx*ny*nz hexes, x fastest then y then z. Attributes (MFEM Cartesian):
    1 z-min, 2 y-min, 3 x-max, 4 y-max, 5 x-min, 6 z-max."""
    (x0, x1), (y0, y1), (z0, z1) = box
    for n, per in zip((nx, ny, nz), periodic):
        if per and n < 3:
            raise ValueError("a periodic direction needs at least 3 elements")
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    zs = np.linspace(z0, z1, nz + 1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    k = k.ravel()
    px, py, pz = periodic
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    coords = np.empty((len(i), 8, 3))
    verts = np.empty((len(i), 8), np.int64)
    for v, (a, b, c) in enumerate(offs):
        ii, jj, kk = i + a, j + b, k + c
        coords[:, v, 0] = xs[ii]
        coords[:, v, 1] = ys[jj]
        coords[:, v, 2] = zs[kk]
        ti = ii % nx if px else ii
        tj = jj % ny if py else jj
        tk = kk % nz if pz else kk
        verts[:, v] = ti + (nx + 1) * (tj + (ny + 1) * tk)
    if px and py and pz:
        bv = np.zeros((0, 4), np.int64)
        ba = np.zeros((0,), np.int32)
    else:
        def attr(mid, n):
            if n[2] < -0.5:
                return 1
            if n[1] < -0.5:
                return 2
            if n[0] > 0.5:
                return 3
            if n[1] > 0.5:
                return 4
            if n[0] < -0.5:
                return 5
            return 6
        bv, ba = _boundary_from_topology(GEOM_HEX, verts, coords, attr)
    return Mesh(GEOM_HEX, coords, verts, bv, ba, name=f"hex{nx}x{ny}x{nz}")


def jittered_delaunay_tri(nx, ny, x0=0.0, x1=4.0, y0=0.0, y1=1.0, jitter=0.3, seed=20240601) -> Mesh:
    """Delaunay triangulation of a uniform (nx+1)*(ny+1) vertex grid whose interior
    vertices are jittered by U(-jitter*h, jitter*h). Attributes: 1 bottom, 2 right,
    3 top, 4 left."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    hx = (x1 - x0) / nx
    hy = (y1 - y0) / ny
    interior = np.zeros_like(X, bool)
    interior[1:-1, 1:-1] = True
    X = X + np.where(interior, rng.uniform(-jitter * hx, jitter * hx, X.shape), 0.0)
    Y = Y + np.where(interior, rng.uniform(-jitter * hy, jitter * hy, Y.shape), 0.0)
    pts = np.stack([X.ravel(), Y.ravel()], axis=1)
    tri = Delaunay(pts)
    simp = tri.simplices.astype(np.int64)
    # counter-clockwise orientation
    a, b, c = pts[simp[:, 0]], pts[simp[:, 1]], pts[simp[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    flip = cross < 0
    simp[flip] = simp[flip][:, [0, 2, 1]]
    area = np.abs(cross) / 2
    simp = simp[area > 1e-14 * hx * hy]
    coords = pts[simp]
    bv, ba = _boundary_from_topology(GEOM_TRI, simp, coords, _box_attr_2d(x0, x1, y0, y1))
    return Mesh(GEOM_TRI, coords, simp, bv, ba, name=f"tri{len(simp)}")


def relabel_boundary(m: Mesh, rule) -> Mesh:
    """A copy of m whose boundary element k gets attribute rule(attr_k,
    midpoint_k) (midpoint of its vertices), e.g. to split one side of a box
    into two boundary conditions."""
    vx = {}
    for e in range(m.num_elements):
        for v in range(m.elem_verts.shape[1]):
            vx[int(m.elem_verts[e, v])] = m.elem_coords[e, v]
    ba = m.bdr_attr.copy()
    for k in range(len(ba)):
        mid = np.mean([vx[int(v)] for v in m.bdr_verts[k]], axis=0)
        ba[k] = rule(int(ba[k]), mid)
    out = Mesh(m.geom, m.elem_coords, m.elem_verts, m.bdr_verts, ba, name=m.name, nc=m.nc, nc_param=m.nc_param)
    return out


def nozzle_quad(num_elements, area_ratio, length=1.0, inlet_half_height=0.5) -> Mesh:
    """Converging-diverging planar nozzle on a structured quad grid with about
    ``num_elements`` cells (aspect kept near length:height). The half-height is
    h(x) = H (1 - (1-r) sin^2(pi x / L)) ... symmetric about the throat at L/2 with
    throat/inlet ratio r. Attributes: 1 bottom wall, 2 outlet, 3 top wall, 4 inlet."""
    H = inlet_half_height
    ny = max(2, int(round(np.sqrt(num_elements * 2 * H / length))))
    nx = max(2, int(round(num_elements / ny)))

    def warp(X, Y):
        s = np.sin(np.pi * X / length) ** 2
        h = H * (1.0 - (1.0 - area_ratio) * s)
        return X, (Y - 0.5) * 2.0 * h

    m = structured_quad(nx, ny, 0.0, length, 0.0, 1.0, warp=warp)
    m.name = f"nozzle{nx}x{ny}"
    return m


def forward_step_quad(nx=240, ny=80, length=3.0, height=1.0, step_x=0.6, step_h=0.2) -> Mesh:
    """Mach-3 forward-facing step: nx*ny quads on [0,L]x[0,H] minus [step_x,L]x[0,step_h].
    Attributes: 1 floor (outward normal -y), 2 outlet (+x), 3 top wall (+y),
    4 inlet (-x), 5 step front (x = step_x, normal +x), 6 step top (y = step_h)."""
    def remove(xc, yc):
        return (xc > step_x) & (yc < step_h)

    m = structured_quad(nx, ny, 0.0, length, 0.0, height, remove=remove)
    # re-label the step's front face (x == step_x; the domain's outward normal
    # there points +x, into the step) as attribute 5 and its top as 6
    bv, ba = m.bdr_verts, m.bdr_attr.copy()
    vx = {}
    for e in range(m.num_elements):
        for v in range(4):
            vx[int(m.elem_verts[e, v])] = m.elem_coords[e, v]
    for b in range(len(ba)):
        p0 = vx[int(bv[b, 0])]
        p1 = vx[int(bv[b, 1])]
        if ba[b] == 2 and abs(p0[0] - step_x) < 1e-12 and abs(p1[0] - step_x) < 1e-12:
            ba[b] = 5
        if ba[b] == 1 and abs(p0[1] - step_h) < 1e-12 and abs(p1[1] - step_h) < 1e-12:
            ba[b] = 6
    m.bdr_attr = ba
    m.name = f"ffs{m.num_elements}"
    return m
