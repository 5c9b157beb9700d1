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


def _is_affine(X):
    """A parallelogram / parallelepiped: X(v) = X0 + sum_r R(v)_r (X(e_r) - X0)
    to 1e-12 of the coordinates."""
    dim = X.shape[1]
    R = _REF_CORNERS[M.GEOM_QUAD if dim == 2 else M.GEOM_HEX].astype(float)
    E = np.stack([X[c] - X[0] for c in [1, 3, 4][:dim]])  # [r][i]
    aff = X[0][None, :] + R @ E
    return bool(np.all(np.abs(aff - X) <= 1e-12 * max(float(np.abs(X).max()), 1.0)))


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
                # a parent of old leaves (derefinement): its children's old indices
                kids = [old_index.get((root, path + (c,)), -1) for c in range(1 << self.dim)]
                if min(kids) < 0:
                    raise ValueError("a leaf without old ancestor or old children")
                out.append((-1, kids, None))
        return out

    def derefine(self, parents):
        """Coarsen the sibling groups whose parents are given as (root, path)
        keys, when all their children are leaves and the coarsened parent keeps
        the level limit with its face neighbours (MFEM's DerefineByError with
        nc_limit 1). Returns (transfer, the parents coarsened)."""
        old = self.ordered_keys()
        old_index = {k: i for i, k in enumerate(old)}
        nch = 1 << self.dim
        R = _REF_CORNERS[self.geom]
        done = []
        for root, path in parents:
            path = tuple(path)
            kids = [(root, path + (c,)) for c in range(nch)]
            if not all(k in self.leaves for k in kids):
                continue
            ch = [self.leaves[k] for k in kids]
            X = np.zeros((self.nv, self.dim))
            V = []
            for k in range(self.nv):
                child = sum(int(R[k][d]) << d for d in range(self.dim))  # the child at corner k
                X[k] = ch[child].X[k]
                V.append(ch[child].V[k])
            parent = _Leaf(root, path, X, tuple(V))
            for k in kids:
                self.leaves.pop(k)
            self.leaves[(root, path)] = parent
            saved = self._prune_vertices()
            if self._needs_refinement(parent):
                self.leaves.pop((root, path))
                for k, c in zip(kids, ch):
                    self.leaves[k] = c
                self._new.update(saved[0])
                self._built_from.update(saved[1])
                continue
            for lf in range(len(self.faces)):
                fk = self._face_key(parent, lf)
                subs = self._subfaces_of_leaf_face(ch, lf)
                if subs and all(sk in self.bdr for sk in subs):
                    attr = self.bdr[subs[0]]
                    for sk in subs:
                        self.bdr.pop(sk)
                    self.bdr[fk] = attr
            done.append((root, path))
        return self._transfer(old_index), done

    def _subfaces_of_leaf_face(self, children, lf):
        """Keys of the children's faces that lie on local face lf of their parent."""
        r, c = self._face_axis(lf)
        out = []
        for i, lv in enumerate(children):
            if ((i >> r) & 1) == c:
                out.append(self._face_key(lv, lf))
        return out

    def _face_axis(self, lf):
        """(normal axis, side) of local face lf in the reference element."""
        R = _REF_CORNERS[self.geom]
        pts = R[list(self.faces[lf])]
        for d in range(self.dim):
            if np.all(pts[:, d] == pts[0, d]):
                return d, int(pts[0, d])
        raise ValueError("face not on a coordinate plane")

    def _prune_vertices(self):
        """Forget built vertices that no leaf uses any more (so a face whose
        neighbour was coarsened reads as conforming again); returns what was
        removed, to restore it."""
        used = set()
        for lv in self.leaves.values():
            used.update(lv.V)
        gone_new = {k: v for k, v in self._new.items() if v not in used}
        gone_bf = {v: self._built_from[v] for v in gone_new.values() if v in self._built_from}
        for k in gone_new:
            self._new.pop(k)
        for v in gone_bf:
            self._built_from.pop(v)
        return gone_new, gone_bf

    def derefinable(self):
        """(root, path) of every sibling group whose children are all leaves."""
        nch = 1 << self.dim
        out, seen = [], set()
        for root, path in self.ordered_keys():
            if not path:
                continue
            par = (root, path[:-1])
            if par in seen:
                continue
            seen.add(par)
            if all((root, path[:-1] + (c,)) in self.leaves for c in range(nch)):
                out.append(par)
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


def _apply_axes(mats, u, dim):
    """Apply the 1D matrix mats[d] along reference axis d of nodal values
    u[v][z][y][x] (x fastest)."""
    u = np.einsum("ia,v...a->v...i", mats[0], u)
    if dim == 2:
        return np.einsum("jb,vba->vja", mats[1], u)
    u = np.einsum("jb,vcba->vcja", mats[1], u)
    return np.einsum("kc,vcba->vkba", mats[2], u)


def _tensor_eval(mats):
    """Rows: points of the tensor grid (first axis fastest); columns: the tensor
    basis (first axis fastest), from the 1D matrices mats[d][point, node]."""
    out = mats[0]
    for m in mats[1:]:
        out = np.einsum("qa,pb->pqba", out, m).reshape(out.shape[0] * m.shape[0], out.shape[1] * m.shape[1])
    return out


def _multilinear_det(X, pts):
    """det J of the element's multilinear map (vertices X, MFEM order) at the
    reference points pts [n][dim]."""
    dim = pts.shape[1]
    R = _REF_CORNERS[M.GEOM_QUAD if dim == 2 else M.GEOM_HEX]
    J = np.zeros((len(pts), dim, dim))
    for k in range(len(R)):
        for r in range(dim):
            f = np.ones(len(pts))
            for d in range(dim):
                f = f * ((1.0 if R[k][d] else -1.0) if d == r else (pts[:, d] if R[k][d] else 1.0 - pts[:, d]))
            J[:, :, r] += f[:, None] * X[k][None, :]
    return np.linalg.det(J)


def _curved_restriction(kid_states, Xp, order, dim):
    """The L2 projection onto the parent of its children's polynomials, weighted
    by the parent's Jacobian determinant (a child's map is the restriction of
    the parent's): M_P u_P = sum_c int_c phi_P u_c det J, with the Gauss rule of
    p + 2 points per axis on every child, exact for the trilinear map."""
    N = order + 1
    nodes = _gl_nodes(N)
    nq = order + 2
    t = _gl_nodes(nq)
    wq = np.polynomial.legendre.leggauss(nq)[1] / 2.0
    Wt = wq
    for _ in range(dim - 1):
        Wt = np.multiply.outer(wq, Wt)
    Wt = Wt.ravel() / (1 << dim)
    grids = np.meshgrid(*([t] * dim), indexing="ij")
    grid = np.stack([g.transpose(*range(dim)[::-1]).ravel() for g in grids], axis=-1)  # first axis fastest
    Psi = _tensor_eval([_lagrange_matrix(nodes, t)] * dim)  # child basis at the child's points
    nd = N ** dim
    Mp = np.zeros((nd, nd))
    rhs = np.zeros((kid_states[0].shape[0], nd))
    for c, uc in enumerate(kid_states):
        o = np.array([0.5 * ((c >> d) & 1) for d in range(dim)])
        pts = o + 0.5 * grid
        Phi = _tensor_eval([_lagrange_matrix(nodes, o[d] + 0.5 * t) for d in range(dim)])
        w = Wt * _multilinear_det(Xp, pts)
        Mp += Phi.T @ (w[:, None] * Phi)
        rhs += (uc @ Psi.T * w[None, :]) @ Phi
    return np.linalg.solve(Mp, rhs.T).T


def transfer_state(x_old, transfer, order, dim, neq, ne_old, coords_new=None):
    """States of the new leaves from the old ones. x is byNODES: [v][e][node],
    nodes x fastest. A refined element's children take its polynomial at their
    nodes (exact). A coarsened parent takes the L2 projection of its children,
    which keeps every element's integral and so the conserved totals: on a
    parallelogram / parallelepiped the parent's Gauss-Legendre mass matrix is
    diagonal and exact and the children's integrals are exact with their own
    rule; on a curved parent (its vertices in coords_new, the new mesh's
    element coordinates) the projection is weighted by det J."""
    N = order + 1
    nd = N ** dim
    nodes = _gl_nodes(N)
    w = np.polynomial.legendre.leggauss(N)[1] / 2.0
    xo = np.asarray(x_old).reshape(neq, ne_old, nd)
    out = np.zeros((neq, len(transfer), nd))
    for i, (src, scale, off) in enumerate(transfer):
        if src < 0:
            kids = scale
            if coords_new is not None and not _is_affine(coords_new[i]):
                out[:, i] = _curved_restriction([xo[:, k] for k in kids], coords_new[i], order, dim)
                continue
            acc = np.zeros((neq,) + (N,) * dim)
            for c, k in enumerate(kids):
                mats = []
                for d in range(dim):
                    o = 0.5 * ((c >> d) & 1)
                    # R[i][k] = (1/2) w_k psi_i(o + x_k / 2) / w_i
                    mats.append(0.5 * _lagrange_matrix(nodes, o + 0.5 * nodes).T * w[None, :] / w[:, None])
                acc += _apply_axes(mats, xo[:, k].reshape((neq,) + (N,) * dim), dim)
            out[:, i] = acc.reshape(neq, nd)
            continue
        if scale == 1.0:
            out[:, i] = xo[:, src]
            continue
        mats = [_lagrange_matrix(nodes, off[d] + scale * nodes) for d in range(dim)]
        out[:, i] = _apply_axes(mats, xo[:, src].reshape((neq,) + (N,) * dim), dim).reshape(neq, nd)
    return out.reshape(-1)


prolong_state = transfer_state


# ---------------------------------------------------------------------------
# Error indicator and the threshold refinement loop


def zz_indicator(mesh, order, x, var=0):
    """Zienkiewicz-Zhu-style error indicator of one conserved variable
    (density by default) per element: the L2 norm over the element of the
    difference between its DG gradient and a recovered continuous gradient
    (the element-mean gradients averaged at each vertex, interpolated
    multilinearly). This stands in for MFEM's L2ZienkiewiczZhuEstimator on the
    density (main.cpp:237-251), whose recovery is a global L2 projection onto
    a Raviart-Thomas space that is not rebuilt here: the two estimate the same
    quantity, not with the same numbers. Quadrilaterals and hexahedra, the
    Jacobian of the element's multilinear map taken at every node."""
    dim = mesh.dim
    N = order + 1
    nd = N ** dim
    NE = mesh.num_elements
    nodes = _gl_nodes(N)
    w1 = np.polynomial.legendre.leggauss(N)[1] / 2.0
    D = np.zeros((N, N))  # D[a, b] = psi_b'(x_a)
    for b in range(N):
        for a in range(N):
            s = 0.0
            for m in range(N):
                if m == b:
                    continue
                t = 1.0 / (nodes[b] - nodes[m])
                for j in range(N):
                    if j not in (b, m):
                        t *= (nodes[a] - nodes[j]) / (nodes[b] - nodes[j])
                s += t
            D[a, b] = s
    u = np.asarray(x).reshape(dim + 2, NE, nd)[var].reshape((NE,) + (N,) * dim)  # [e][z][y][x]
    gref = []
    for d in range(dim):
        ax = u.ndim - 1 - d  # x is the last axis
        gref.append(np.moveaxis(np.tensordot(u, D, axes=([ax], [1])), -1, ax).reshape(NE, nd))
    gref = np.stack(gref, axis=-1)  # [e][node][d] reference gradient
    X = mesh.elem_coords
    R = _REF_CORNERS[mesh.geom]
    grids = np.meshgrid(*([nodes] * dim), indexing="ij")  # [x][y][z] ordering
    xi = np.stack([g.transpose(*range(dim)[::-1]).ravel() for g in grids], axis=-1)  # nodes x fastest
    nv = X.shape[1]
    # the multilinear map's Jacobian at every node (constant on affine elements)
    dN = np.ones((nd, nv, dim))  # dN[n, k, r] = d N_k / d xi_r at node n
    for k in range(nv):
        for r in range(dim):
            for d in range(dim):
                f = (1.0 if R[k][d] else -1.0) if d == r else (xi[:, d] if R[k][d] else 1.0 - xi[:, d])
                dN[:, k, r] *= f
    J = np.einsum("nkr,eki->enir", dN, X)  # [e][node][i][r]
    Jinv = np.linalg.inv(J)
    grad = np.einsum("enr,enri->eni", gref, Jinv)  # J^-T applied: g_x = sum_r dxi_r/dx_i g_r
    det = np.abs(np.linalg.det(J))  # [e][node]
    W = w1
    for _ in range(dim - 1):
        W = np.multiply.outer(w1, W)
    W = W.ravel()  # node weights, x fastest
    gmean = np.einsum("n,en,eni->ei", W, det, grad) / np.einsum("n,en->e", W, det)[:, None]
    V = mesh.elem_verts
    ids, inv = np.unique(V.ravel(), return_inverse=True)
    acc = np.zeros((len(ids), dim))
    cnt = np.zeros(len(ids))
    np.add.at(acc, inv, np.repeat(gmean, nv, axis=0))
    np.add.at(cnt, inv, 1.0)
    Gv = (acc / cnt[:, None])[inv].reshape(NE, nv, dim)
    # multilinear interpolation of the corner values at the nodes
    Nk = np.ones((nd, nv))
    for k in range(nv):
        for d in range(dim):
            Nk[:, k] *= xi[:, d] if R[k][d] else 1.0 - xi[:, d]
    Gn = np.einsum("nk,eki->eni", Nk, Gv)
    err2 = np.einsum("n,en,eni->e", W, det, (Gn - grad) ** 2)
    return np.sqrt(err2)


def refine_marks(eta, max_elem_error, total_error_fraction=0.0, total_norm_p=np.inf):
    """ThresholdRefiner's marking (the reference sets the local goal
    maxElemError and the total fraction totalErrorFraction,
    case_manager.cpp:450-466): elements whose indicator exceeds
    max(total_error_fraction * total * NE^(-1/p), max_elem_error)."""
    ne = len(eta)
    if np.isinf(total_norm_p):
        total = float(np.max(eta)) if ne else 0.0
        scale = 1.0
    else:
        total = float(np.sum(eta ** total_norm_p) ** (1.0 / total_norm_p))
        scale = ne ** (-1.0 / total_norm_p)
    thr = max(total_error_fraction * total * scale, max_elem_error)
    return np.flatnonzero(eta > thr)


def derefine_groups(tree, eta, threshold):
    """ThresholdDerefiner's choice: sibling groups whose children's indicators
    sum below the threshold (hysteresis * maxElemError)."""
    keys = tree.ordered_keys()
    index = {k: i for i, k in enumerate(keys)}
    nch = 1 << tree.dim
    out = []
    for root, path in tree.derefinable():
        s = sum(eta[index[(root, tuple(path) + (c,))]] for c in range(nch))
        if s < threshold:
            out.append((root, path))
    return out


class _GpuSolver:
    """The GPU path on one mesh of the adaptive run (rkdg_create_nc after each
    change of the mesh, the state handed over through host memory)."""

    def __init__(self, mesh, order, scheme, device=0):
        import torch

        from .capi import Context

        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.ctx = Context(mesh, order, scheme)

    def node_coords(self):
        return self.ctx.node_coords()

    def hmin(self):
        return self.ctx.hmin()

    def _d(self, x):
        return self.torch.from_numpy(np.ascontiguousarray(x)).to(self.dev)

    def limit(self, x):
        xd = self._d(x)
        self.ctx.limit(xd)
        return xd.cpu().numpy()

    def face_lambda(self, x):
        return self.ctx.face_lambda(self._d(x))

    def step(self, x, dt, tableau):
        xd = self._d(x)
        lam = self.ctx.step(xd, dt, tableau)
        return xd.cpu().numpy(), lam

    def close(self):
        self.ctx.close()


class _OracleSolver:
    """The oracle on one mesh of the adaptive run (the checker of the GPU run)."""

    def __init__(self, mesh, order, scheme):
        from oracle.refcpu import RefSolver

        self.ref = RefSolver(mesh, order, scheme)

    def node_coords(self):
        return self.ref.node_coords()

    def hmin(self):
        return self.ref.hmin()

    def limit(self, x):
        return self.ref.limit(x)[0]

    def face_lambda(self, x):
        return self.ref.face_mult(x)[1]

    def step(self, x, dt, tableau):
        return self.ref.step(x, dt, tableau)

    def close(self):
        pass


@dataclass
class AdaptiveResult:
    x: np.ndarray
    mesh: M.Mesh
    tree: NCMesh
    t: float
    dt: float
    steps: int
    elements: list      # element count after each step
    refinements: int    # refinement iterations over the run


def run_adaptive(base, order, scheme, initial_state, t_final, cfl, tableau, max_elem_error, hysteresis=0.5,
                 total_error_fraction=0.0, max_steps=None, solver="gpu", advance_time_per_retry=True,
                 hmin_from_base=True, max_level=None, dt_const=None, callback=None):
    """The reference's adaptive time loop (main.cpp:253-386 and :505-700) on
    the host, with the right-hand side, limiter and RK step on the chosen
    solver ("gpu": rkdg_create_nc and the HIP kernels; "oracle": the CPU
    restatement, as the checker). Per time step: step, estimate on the new
    state, refine the marked elements and repeat the step from the old state
    on the refined mesh until nothing is marked, then coarsen the groups below
    hysteresis * max_elem_error and limit. The reference advances t on every
    repeat of the step (Step takes t by reference inside the refinement loop,
    main.cpp:514); advance_time_per_retry=False advances it once per step. It
    also takes hmin of the mesh it read, before any refinement
    (main.cpp:191-201), for every dt; hmin_from_base=False takes the current
    mesh's. max_level caps the refinement level (an indicator that does not
    fall with h, as at a shock, would refine without end; MFEM caps it through
    its element budget). dt_const keeps dt fixed (cfl <= 0 in the
    reference). callback(steps, t, dt, mesh, x) after every step.
    initial_state(xyz) -> state on the nodes xyz [NE][nd][dim]."""
    make = (lambda m: _GpuSolver(m, order, scheme)) if solver == "gpu" else (lambda m: _OracleSolver(m, order, scheme))

    def marks_of(mesh, x):
        mk = refine_marks(zz_indicator(mesh, order, x), max_elem_error, total_error_fraction)
        if max_level is not None:
            mk = mk[tree.levels()[mk] < max_level]
        return mk

    dim = base.dim
    neq = dim + 2
    tree = NCMesh(base)
    refinements = 0
    hmin0 = None
    # initial refinement on the projected initial condition (main.cpp:253-320)
    while True:
        mesh = tree.mesh()
        S_ = make(mesh)
        if hmin0 is None:
            hmin0 = S_.hmin()
        x = S_.limit(initial_state(S_.node_coords()))
        marks = marks_of(mesh, x)
        if len(marks) == 0:
            break
        S_.close()
        tree.refine(marks)
        refinements += 1

    def coarsen(x, S_, mesh):
        groups = derefine_groups(tree, zz_indicator(mesh, order, x), hysteresis * max_elem_error)
        if not groups:
            return x, S_, mesh
        ne_old = mesh.num_elements
        tr, done = tree.derefine(groups)
        if not done:
            return x, S_, mesh
        mesh = tree.mesh()
        x = transfer_state(x, tr, order, dim, neq, ne_old, mesh.elem_coords)
        S_.close()
        return x, make(mesh), mesh

    x, S_, mesh = coarsen(x, S_, mesh)
    x = S_.limit(x)
    hmin = (lambda: hmin0) if hmin_from_base else (lambda: S_.hmin())
    lam = S_.face_lambda(x)
    dt = dt_const if dt_const is not None else cfl * hmin() / lam / (2 * order + 1)
    t, steps, counts = 0.0, 0, []
    while t < t_final - 1e-8 * dt and (max_steps is None or steps < max_steps):
        dt_real = min(dt, t_final - t)
        x_old = x
        t_start = t
        while True:
            x_new, lam = S_.step(x_old, dt_real, tableau)
            t = t + dt_real if advance_time_per_retry else t_start + dt_real
            if dt_const is None:
                dt = cfl * hmin() / lam / (2 * order + 1)
            marks = marks_of(mesh, x_new)
            if len(marks) == 0:
                x = x_new
                break
            ne_old = mesh.num_elements
            tr = tree.refine(marks)
            refinements += 1
            mesh = tree.mesh()
            x_old = transfer_state(x_old, tr, order, dim, neq, ne_old, mesh.elem_coords)
            S_.close()
            S_ = make(mesh)
        x, S_, mesh = coarsen(x, S_, mesh)
        x = S_.limit(x)
        steps += 1
        counts.append(mesh.num_elements)
        if callback is not None:
            callback(steps, t, dt, mesh, x)
    S_.close()
    return AdaptiveResult(x, mesh, tree, t, dt, steps, counts, refinements)
