"""Non-conforming (hanging-node) refinement of quadrilateral and hexahedral
meshes on the host: the mesh side of the reference's adaptive runs
(main.cpp:253-386 and :564-696, MFEM's NCMesh behind ThresholdRefiner /
ThresholdDerefiner, amr.cpp:3-133), producing what the C-ABI takes after each
refinement: the leaf elements (coordinates, topological vertex ids, boundary
faces) and the list of non-conforming faces (slave element and face, master
element and face), as MFEM's NCMesh master/slave lists give them.

Refinement is isotropic (4 children in 2D, 8 in 3D); the children of an
element are ordered x fastest. New vertices get topological ids from the
vertices they are built from (an edge midpoint from its two end vertices, a
face centre from its four corners), so that neighbours, periodic
identifications included, agree on them. Coordinates of new vertices are the
averages of those corners, i.e. the element's multilinear map at the child
corners; each element keeps its own copy of its vertex coordinates, as the
C-ABI mesh does.

The level difference of face neighbours is kept at most one (MFEM's
nonConformingLimit 1): refinement propagates to coarser neighbours.

Solution transfer on refinement: the children take the parent's polynomial at
their nodes, which is exact (MFEM's refinement operator for nodal DG spaces).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import mesh as M

_REF_CORNERS = {
    M.GEOM_QUAD: np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.int64),
    M.GEOM_HEX: np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                          [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.int64),
}


@dataclass
class _Leaf:
    root: int
    path: tuple
    X: np.ndarray      # (nv, dim) vertex coordinates
    V: tuple           # topological vertex ids

    @property
    def level(self):
        return len(self.path)


def _gl_nodes(n):
    x, _ = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0)


def _lagrange_matrix(nodes, pts):
    """L[i, k] = l_k(pts[i]) for the Lagrange basis on `nodes`."""
    L = np.ones((len(pts), len(nodes)))
    for k, xk in enumerate(nodes):
        for m, xm in enumerate(nodes):
            if m != k:
                L[:, k] *= (pts - xm) / (xk - xm)
    return L


class NCMesh:
    """A refinement forest over a conforming quadrilateral or hexahedral mesh."""

    def __init__(self, base: M.Mesh):
        if base.geom not in (M.GEOM_QUAD, M.GEOM_HEX):
            raise ValueError("non-conforming refinement is built for quadrilaterals and hexahedra")
        if len(base.nc):
            raise ValueError("the base mesh must be conforming")
        self.geom = base.geom
        self.dim = base.dim
        self.nv = 4 if self.dim == 2 else 8
        self.faces = M._local_faces(self.geom)
        self.leaves = {}
        for e in range(base.num_elements):
            self.leaves[(e, ())] = _Leaf(e, (), base.elem_coords[e].copy(), tuple(int(v) for v in base.elem_verts[e]))
        self._new = {}  # sorted tuple of corner ids -> id of the vertex built from them
        self._built_from = {}
        self._next = int(base.elem_verts.max()) + 1 if base.num_elements else 0
        self.bdr = {}
        for bv, ba in zip(base.bdr_verts, base.bdr_attr):
            self.bdr[tuple(sorted(int(v) for v in bv))] = int(ba)
        self.name = base.name

    # -- topology helpers ------------------------------------------------
    def _vid(self, corners):
        key = tuple(sorted(corners))
        if len(key) == 1:
            return key[0]
        v = self._new.get(key)
        if v is None:
            v = self._next
            self._next += 1
            self._new[key] = v
            self._built_from[v] = key
        return v

    def _face_key(self, leaf, lf):
        return tuple(sorted(leaf.V[c] for c in self.faces[lf]))

    def _subfaces(self, key_corners):
        """The sub-face keys of a face (corners in face order) once split, or None
        if the face is not split."""
        if self.dim == 2:
            a, b = key_corners
            m = self._new.get(tuple(sorted((a, b))))
            if m is None:
                return None
            return [tuple(sorted((a, m))), tuple(sorted((m, b)))]
        a, b, c, d = key_corners
        f = self._new.get(tuple(sorted((a, b, c, d))))
        if f is None:
            return None
        mab = self._new[tuple(sorted((a, b)))]
        mbc = self._new[tuple(sorted((b, c)))]
        mcd = self._new[tuple(sorted((c, d)))]
        mda = self._new[tuple(sorted((d, a)))]
        return [tuple(sorted(q)) for q in ((a, mab, f, mda), (mab, b, mbc, f), (f, mbc, c, mcd), (mda, f, mcd, d))]

    def _face_corners(self, leaf, lf):
        return tuple(leaf.V[c] for c in self.faces[lf])

    # -- refinement ------------------------------------------------------
    def _refine_leaf(self, key):
        leaf = self.leaves.pop(key)
        R = _REF_CORNERS[self.geom]
        nch = 1 << self.dim
        for ch in range(nch):
            cb = np.array([(ch >> d) & 1 for d in range(self.dim)])
            X = np.zeros((self.nv, self.dim))
            V = []
            for k in range(self.nv):
                p2 = cb + R[k]  # the child corner in parent coordinates, times 2: 0, 1 or 2 per axis
                half = [d for d in range(self.dim) if p2[d] == 1]
                base = np.where(p2 == 2, 1, 0)
                corners = []
                for s in range(1 << len(half)):
                    q = base.copy()
                    for i, d in enumerate(half):
                        q[d] = (s >> i) & 1
                    corners.append(int(np.flatnonzero((R == q).all(axis=1))[0]))
                ids = [leaf.V[c] for c in corners]
                order = np.argsort(ids, kind="stable")
                if len(corners) == 1:
                    X[k] = leaf.X[corners[0]]
                elif len(corners) == 2:
                    X[k] = 0.5 * (leaf.X[corners[order[0]]] + leaf.X[corners[order[1]]])
                elif len(corners) == 4:
                    c4 = [leaf.X[corners[o]] for o in order]
                    X[k] = 0.25 * ((c4[0] + c4[1]) + (c4[2] + c4[3]))
                else:
                    c8 = [leaf.X[corners[o]] for o in order]
                    X[k] = 0.125 * (((c8[0] + c8[1]) + (c8[2] + c8[3])) + ((c8[4] + c8[5]) + (c8[6] + c8[7])))
                if len(corners) == 8:
                    V.append(None)  # the element centre: one id for the children of this element
                else:
                    V.append(self._vid(ids))
            self.leaves[(leaf.root, leaf.path + (ch,))] = _Leaf(leaf.root, leaf.path + (ch,), X, V)
        if self.dim == 3:
            cid = self._next
            self._next += 1
            for ch in range(nch):
                lk = self.leaves[(leaf.root, leaf.path + (ch,))]
                lk.V = tuple(cid if v is None else v for v in lk.V)
        else:
            for ch in range(nch):
                lk = self.leaves[(leaf.root, leaf.path + (ch,))]
                lk.V = tuple(lk.V)
        # boundary faces of the parent split into the children's faces
        for lf in range(len(self.faces)):
            fk = self._face_key(leaf, lf)
            attr = self.bdr.pop(fk, None)
            if attr is not None:
                for sk in self._subfaces(self._face_corners(leaf, lf)):
                    self.bdr[sk] = attr

    def _needs_refinement(self, leaf):
        """A face neighbour two levels finer (a sub-face of this leaf's face split again)."""
        for lf in range(len(self.faces)):
            subs = self._subfaces(self._face_corners(leaf, lf))
            if subs is None:
                continue
            for sk in subs:
                if self._subfaces_key(sk) is not None:
                    return True
        return False

    def _subfaces_key(self, key):
        if self.dim == 2:
            return self._new.get(key)
        return self._new.get(key)  # a quarter face is split when its centre exists

    def refine(self, elements):
        """Refine the leaves with these indices (in the current leaf order),
        then their coarser neighbours as the level limit requires. Returns the
        transfer map: for each new leaf, (old leaf index, scale, offset) such
        that the new leaf's reference point xi is old-leaf point offset + scale * xi."""
        old = self.ordered_keys()
        old_index = {k: i for i, k in enumerate(old)}
        for i in sorted(set(int(e) for e in elements)):
            k = old[i]
            if k in self.leaves:
                self._refine_leaf(k)
        changed = True
        while changed:
            changed = False
            for k in list(self.leaves):
                if k in self.leaves and self._needs_refinement(self.leaves[k]):
                    self._refine_leaf(k)
                    changed = True
        return self._transfer(old_index)

    def _transfer(self, old_index):
        """For each current leaf: the old leaf it came from and the affine map
        of reference coordinates (scale, offset per axis) from the new leaf to
        the old one."""
        out = []
        for k in self.ordered_keys():
            root, path = k
            if k in old_index:
                out.append((old_index[k], 1.0, np.zeros(self.dim)))
                continue
            # an old ancestor: the longest prefix that was an old leaf
            for n in range(len(path) - 1, -1, -1):
                anc = (root, path[:n])
                if anc in old_index:
                    sub = path[n:]
                    scale = 0.5 ** len(sub)
                    off = np.zeros(self.dim)
                    for lvl, c in enumerate(sub):
                        for d in range(self.dim):
                            off[d] += ((c >> d) & 1) * 0.5 ** (lvl + 1)
                    out.append((old_index[anc], scale, off))
                    break
            else:
                out.append((-1, 1.0, np.zeros(self.dim)))
        return out

    # -- output ----------------------------------------------------------
    def ordered_keys(self):
        return sorted(self.leaves, key=lambda k: (k[0], k[1]))

    def mesh(self) -> M.Mesh:
        keys = self.ordered_keys()
        leaves = [self.leaves[k] for k in keys]
        X = np.array([lf.X for lf in leaves])
        V = np.array([lf.V for lf in leaves], np.int64)
        face_map = {}
        for e, lv in enumerate(leaves):
            for lf in range(len(self.faces)):
                face_map.setdefault(self._face_key(lv, lf), []).append((e, lf))
        nc, par = [], []
        bv, ba = [], []
        covered = set()
        for e, lv in enumerate(leaves):
            for lf in range(len(self.faces)):
                fk = self._face_key(lv, lf)
                if len(face_map[fk]) == 2:
                    continue
                if len(face_map[fk]) > 2:
                    raise ValueError("face shared by more than two leaves")
                if fk in self.bdr:
                    bv.append([lv.V[c] for c in self.faces[lf]])
                    ba.append(self.bdr[fk])
                    continue
                subs = self._subfaces(self._face_corners(lv, lf))
                if subs is None:
                    continue  # a slave face, recorded from its master
                mcorn = self._face_corners(lv, lf)
                for sk in subs:
                    owners = face_map.get(sk, [])
                    if len(owners) != 1:
                        raise ValueError("a master face whose sub-face is not a single leaf face "
                                         "(level difference above one)")
                    es, lfs = owners[0]
                    nc.append((es, lfs, e, lf))
                    par.append([self._face_param(mcorn, v) for v in self._face_corners(leaves[es], lfs)])
                    covered.add((es, lfs))
        for e, lv in enumerate(leaves):
            for lf in range(len(self.faces)):
                fk = self._face_key(lv, lf)
                if len(face_map[fk]) == 1 and fk not in self.bdr and self._subfaces(self._face_corners(lv, lf)) is None \
                        and (e, lf) not in covered:
                    raise ValueError(f"leaf {e} face {lf} has no neighbour, master or boundary element")
        nvf = 2 if self.dim == 2 else 4
        return M.Mesh(self.geom, X, V, np.array(bv, np.int64).reshape(-1, nvf), np.array(ba, np.int32),
                      name=f"{self.name}-nc{len(leaves)}", nc=np.array(nc, np.int64).reshape(-1, 4),
                      nc_param=np.array(par, np.float64).reshape(len(nc), nvf * (self.dim - 1)))

    def _face_param(self, mcorn, v):
        """Coordinates of vertex v in the parametrisation of the face with
        corners mcorn (face order; corner 0 at the origin): a corner of the
        face, or a vertex built from corners of the face (their average)."""
        unit = [(0.0,), (1.0,)] if self.dim == 2 else [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
        if v in mcorn:
            return unit[mcorn.index(v)]
        key = self._built_from.get(v)
        if key is None or any(c not in mcorn for c in key):
            raise ValueError("a slave corner that is not built from the master face's corners")
        pts = np.array([unit[mcorn.index(c)] for c in key])
        return tuple(pts.mean(axis=0))

    def levels(self):
        return np.array([len(k[1]) for k in self.ordered_keys()], np.int64)


def prolong_state(x_old, transfer, order, dim, neq, ne_old):
    """States of the new leaves from the old ones (refinement: the old
    polynomial at the new nodes). x is byNODES: [v][e][node], nodes x fastest."""
    N = order + 1
    nd = N ** dim
    nodes = _gl_nodes(N)
    xo = np.asarray(x_old).reshape(neq, ne_old, nd)
    out = np.zeros((neq, len(transfer), nd))
    for i, (src, scale, off) in enumerate(transfer):
        if src < 0:
            raise ValueError("a leaf without an old ancestor")
        if scale == 1.0:
            out[:, i] = xo[:, src]
            continue
        mats = [_lagrange_matrix(nodes, off[d] + scale * nodes) for d in range(dim)]
        u = xo[:, src].reshape((neq,) + (N,) * dim)  # [v][z][y][x]
        # apply per axis: x is the last array axis
        u = np.einsum("ia,v...a->v...i", mats[0], u)
        if dim == 2:
            u = np.einsum("jb,vba->vja", mats[1], u)
        else:
            u = np.einsum("jb,vcba->vcja", mats[1], u)
            u = np.einsum("kc,vcba->vkba", mats[2], u)
        out[:, i] = u.reshape(neq, nd)
    return out.reshape(-1)
