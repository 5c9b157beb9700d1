"""Gmsh mesh files for the RKDG path: the mesh half of CaseManager::loadMesh
(src/case_manager/case_manager.cpp:217-297) with readPhysicalNames (:299-352)
and minimizeAttributes (:354-448).

Reads MSH 2.2 and MSH 4.1, ASCII or binary (MFEM's reader, which the
reference calls, takes both encodings), into the C-ABI's mesh description
(``mesh.Mesh``): elements of the highest dimension present become cells,
elements one dimension lower become boundary elements, points and anything
else are ignored. A cell's attribute is its physical tag (MSH 2.2: the first
element tag; MSH 4.1: the first physical tag of its entity; 1 without one).

What MFEM's mesh constructor does to such a file, as the reference calls it
(``Mesh(stream, 1, 1)``, fix_orientation on):

* inverted 2D cells are turned over (triangle: vertices 0 and 1 swapped;
  quadrilateral: vertices 1 and 3 swapped); an inverted hexahedron is an error;
* triangle meshes are marked for refinement: each triangle's vertices are
  rotated so that the edge 0-1 is its longest (ties keep the earlier edge).
  This fixes the local node order of a P2 triangle, nothing else;
* ``$Periodic`` node links identify the slave nodes with their masters, so
  faces across the period are interior faces.

Then minimizeAttributes renumbers the attributes: the sorted distinct cell
attributes become 1..n, the sorted distinct boundary attributes n+1..n+m, and
the $PhysicalNames maps follow.

Curved (higher-order) geometry and meshes mixing cell types are not part of
the path's mesh description and raise ``GmshError``.
"""
from __future__ import annotations

import gzip
import shlex
from dataclasses import dataclass, field

import numpy as np

from . import mesh as M


class GmshError(ValueError):
    pass


# MSH element type -> (dimension, number of nodes); only straight-sided
# first-order cells enter the path
_ELEM = {1: (1, 2), 2: (2, 3), 3: (2, 4), 4: (3, 4), 5: (3, 8), 6: (3, 6), 7: (3, 5), 15: (0, 1),
         8: (1, 3), 9: (2, 6), 10: (2, 9), 11: (3, 10), 12: (3, 27), 16: (2, 8), 17: (3, 20)}
_CELL_GEOM = {2: M.GEOM_TRI, 3: M.GEOM_QUAD, 5: M.GEOM_HEX}
_FACE_TYPE = {M.GEOM_TRI: 1, M.GEOM_QUAD: 1, M.GEOM_HEX: 3}


@dataclass
class GmshMesh:
    """A read mesh: the C-ABI description plus the names of its physical groups.

    ``cell_names``/``bdr_names`` map a physical name to its attribute after
    minimizeAttributes (the reference's map_phys_names_tag_{2D,3D} and
    map_phys_names_tag_{1D,2D}); ``elem_attr`` holds the renumbered cell
    attributes; ``version`` is the file's MSH version."""

    mesh: M.Mesh
    elem_attr: np.ndarray
    cell_names: dict = field(default_factory=dict)
    bdr_names: dict = field(default_factory=dict)
    version: str = "2.2"
    periodic: bool = False


def _sections(text):
    """Yields (name, lines) for every $Name ... $EndName block."""
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        s = lines[i].strip()
        if s.startswith("$") and not s.startswith("$End"):
            name = s[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != "$End" + name:
                j += 1
            if j == len(lines):
                raise GmshError(f"section ${name} has no $End{name}")
            yield name, lines[i + 1:j]
            i = j + 1
        else:
            i += 1


def _physical_names(lines):
    """$PhysicalNames: (dim, tag, name); quotes are removed (case_manager.cpp:322-325)."""
    n = int(lines[0].split()[0])
    out = []
    for ln in lines[1:1 + n]:
        parts = shlex.split(ln)
        out.append((int(parts[0]), int(parts[1]), parts[2]))
    return out


def _read_v2(secs):
    if "Nodes" not in secs or "Elements" not in secs:
        raise GmshError("MSH 2.2 file without $Nodes or $Elements")
    nl = secs["Nodes"]
    n = int(nl[0])
    node = np.array([ln.split() for ln in nl[1:1 + n]], dtype=np.float64)
    ids = node[:, 0].astype(np.int64)
    xyz = node[:, 1:4]
    elems = []
    el = secs["Elements"]
    ne = int(el[0])
    for ln in el[1:1 + ne]:
        v = [int(t) for t in ln.split()]
        et, nt = v[1], v[2]
        tags = v[3:3 + nt]
        phys = tags[0] if nt > 0 and tags[0] > 0 else 1
        elems.append((et, phys, v[3 + nt:]))
    links = _periodic_v2(secs["Periodic"]) if "Periodic" in secs else []
    return ids, xyz, elems, links


def _periodic_v2(pl):
    """$Periodic of MSH 2.2 (ASCII, also in binary files): (slave, master) node pairs."""
    links = []
    k = 1
    for _ in range(int(pl[0])):
        k += 1  # dim slave master
        if pl[k].strip().startswith("Affine"):
            k += 1
        cnt = int(pl[k])
        k += 1
        for ln in pl[k:k + cnt]:
            a, b = ln.split()[:2]
            links.append((int(a), int(b)))
        k += cnt
    return links


def _read_v41(secs):
    if "Nodes" not in secs or "Elements" not in secs:
        raise GmshError("MSH 4.1 file without $Nodes or $Elements")
    ent_phys = {}
    if "Entities" in secs:
        el = secs["Entities"]
        counts = [int(t) for t in el[0].split()]
        k = 1
        for dim, cnt in enumerate(counts):
            for ln in el[k:k + cnt]:
                v = ln.split()
                tag = int(v[0])
                off = 4 if dim == 0 else 7
                npt = int(v[off])
                ent_phys[(dim, tag)] = [int(t) for t in v[off + 1:off + 1 + npt]]
            k += cnt
    nl = secs["Nodes"]
    nblocks = int(nl[0].split()[0])
    ids, xyz = [], []
    k = 1
    for _ in range(nblocks):
        _, _, param, cnt = (int(t) for t in nl[k].split())
        if param:
            raise GmshError("parametric node coordinates are not supported")
        ids.extend(int(t) for t in nl[k + 1:k + 1 + cnt])
        xyz.extend([float(t) for t in ln.split()[:3]] for ln in nl[k + 1 + cnt:k + 1 + 2 * cnt])
        k += 1 + 2 * cnt
    el = secs["Elements"]
    nblocks = int(el[0].split()[0])
    elems = []
    k = 1
    for _ in range(nblocks):
        edim, etag, et, cnt = (int(t) for t in el[k].split())
        ph = ent_phys.get((edim, etag), [])
        phys = ph[0] if ph and ph[0] > 0 else 1
        for ln in el[k + 1:k + 1 + cnt]:
            v = [int(t) for t in ln.split()]
            elems.append((et, phys, v[1:]))
        k += 1 + cnt
    links = []
    if "Periodic" in secs:
        pl = secs["Periodic"]
        k = 1
        for _ in range(int(pl[0])):
            k += 1  # dim tag master
            naff = int(pl[k].split()[0])
            k += 1 if naff == 0 or len(pl[k].split()) > 1 else 2
            cnt = int(pl[k])
            k += 1
            for ln in pl[k:k + cnt]:
                a, b = ln.split()[:2]
                links.append((int(a), int(b)))
            k += cnt
    return np.asarray(ids, np.int64), np.asarray(xyz, np.float64).reshape(-1, 3), elems, links


class _Cursor:
    """Position in the bytes of a binary MSH file."""

    def __init__(self, data: bytes, pos: int = 0):
        self.data = data
        self.pos = pos

    def line(self) -> str:
        end = self.data.index(b"\n", self.pos)
        out = self.data[self.pos:end].decode("ascii", "replace").strip()
        self.pos = end + 1
        return out

    def take(self, dtype, count):
        dt = np.dtype(dtype)
        n = dt.itemsize * count
        if self.pos + n > len(self.data):
            raise GmshError("binary section ends early")
        out = np.frombuffer(self.data, dtype=dt, count=count, offset=self.pos)
        self.pos += n
        return out

    def skip_to_end(self, name: str):
        """Moves past the line $End<name> that follows the section's data."""
        tag = ("$End" + name).encode()
        at = self.data.find(tag, self.pos)
        if at < 0:
            raise GmshError(f"section ${name} has no $End{name}")
        self.pos = at + len(tag)
        nl = self.data.find(b"\n", self.pos)
        self.pos = len(self.data) if nl < 0 else nl + 1


def _read_binary(data: bytes, path):
    """Binary MSH 2.2 or 4.1 (the Gmsh reference manual's layouts; numbers in
    the file's byte order, which the integer 1 after the format line gives).
    Returns (version, ids, xyz, elems, links, names)."""
    cur = _Cursor(data)
    version = None
    bo = "<"
    ids = xyz = None
    elems, links, names = [], [], []
    ent_phys = {}

    def i4(n):
        return cur.take(bo + "i4", n)

    def sz(n):
        return cur.take(bo + "u8", n).astype(np.int64)

    while True:
        try:
            ln = cur.line()
        except ValueError:
            break
        if not ln.startswith("$") or ln.startswith("$End"):
            continue
        name = ln[1:]
        if name == "MeshFormat":
            fmt = cur.line().split()
            version, dsize = fmt[0], int(fmt[2])
            if dsize != 8:
                raise GmshError(f"{path}: binary MSH with data size {dsize} (8 is read)")
            one = cur.data[cur.pos:cur.pos + 4]
            if np.frombuffer(one, "<i4")[0] == 1:
                bo = "<"
            elif np.frombuffer(one, ">i4")[0] == 1:
                bo = ">"
            else:
                raise GmshError(f"{path}: binary MSH without the integer 1 after the format line")
            cur.pos += 4
            cur.skip_to_end(name)
        elif name == "PhysicalNames":
            n = int(cur.line().split()[0])
            names = _physical_names([str(n)] + [cur.line() for _ in range(n)])
            cur.skip_to_end(name)
        elif version is None:
            raise GmshError(f"{path}: ${name} before $MeshFormat")
        elif version.startswith("2"):
            if name == "Nodes":
                n = int(cur.line().split()[0])
                rec = cur.take(np.dtype([("id", bo + "i4"), ("x", bo + "f8", (3,))]), n)
                ids = rec["id"].astype(np.int64)
                xyz = np.array(rec["x"], np.float64)
                cur.skip_to_end(name)
            elif name == "Elements":
                ne = int(cur.line().split()[0])
                got = 0
                while got < ne:
                    et, nfollow, ntags = (int(t) for t in i4(3))
                    if et not in _ELEM:
                        raise GmshError(f"{path}: unknown element type {et}")
                    nn = _ELEM[et][1]
                    blk = i4(nfollow * (1 + ntags + nn)).reshape(nfollow, 1 + ntags + nn)
                    for row in blk:
                        phys = int(row[1]) if ntags > 0 and row[1] > 0 else 1
                        elems.append((et, phys, [int(t) for t in row[1 + ntags:]]))
                    got += nfollow
                cur.skip_to_end(name)
            elif name == "Periodic":
                # the MSH 2 writer keeps this section in ASCII in binary files
                start = cur.pos
                cur.skip_to_end(name)
                body = data[start:cur.pos].decode("ascii").splitlines()
                links = _periodic_v2([b for b in body if b.strip() and not b.startswith("$End")])
            else:
                cur.skip_to_end(name)
        elif version.startswith("4.1"):
            if name == "Entities":
                counts = sz(4)
                for dim, cnt in enumerate(counts):
                    for _ in range(int(cnt)):
                        tag = int(i4(1)[0])
                        cur.take(bo + "f8", 3 if dim == 0 else 6)
                        nph = int(sz(1)[0])
                        ent_phys[(dim, tag)] = [int(t) for t in i4(nph)]
                        if dim > 0:
                            nb = int(sz(1)[0])
                            i4(nb)
                cur.skip_to_end(name)
            elif name == "Nodes":
                nblocks, nnodes = (int(t) for t in sz(4)[:2])
                id_parts, x_parts = [], []
                for _ in range(nblocks):
                    _, _, param = (int(t) for t in i4(3))
                    cnt = int(sz(1)[0])
                    if param:
                        raise GmshError("parametric node coordinates are not supported")
                    id_parts.append(sz(cnt))
                    x_parts.append(cur.take(bo + "f8", 3 * cnt).reshape(cnt, 3))
                ids = np.concatenate(id_parts) if id_parts else np.zeros(0, np.int64)
                xyz = np.concatenate(x_parts).astype(np.float64) if x_parts else np.zeros((0, 3))
                if len(ids) != nnodes:
                    raise GmshError(f"{path}: $Nodes holds {len(ids)} nodes, its header {nnodes}")
                cur.skip_to_end(name)
            elif name == "Elements":
                nblocks = int(sz(4)[0])
                for _ in range(nblocks):
                    edim, etag, et = (int(t) for t in i4(3))
                    cnt = int(sz(1)[0])
                    if et not in _ELEM:
                        raise GmshError(f"{path}: unknown element type {et}")
                    nn = _ELEM[et][1]
                    blk = sz(cnt * (1 + nn)).reshape(cnt, 1 + nn)
                    ph = ent_phys.get((edim, etag), [])
                    phys = ph[0] if ph and ph[0] > 0 else 1
                    for row in blk:
                        elems.append((et, phys, [int(t) for t in row[1:]]))
                cur.skip_to_end(name)
            elif name == "Periodic":
                nlinks = int(sz(1)[0])
                for _ in range(nlinks):
                    i4(3)
                    naff = int(sz(1)[0])
                    cur.take(bo + "f8", naff)
                    cnt = int(sz(1)[0])
                    pairs = sz(2 * cnt).reshape(cnt, 2)
                    links.extend((int(a), int(b)) for a, b in pairs)
                cur.skip_to_end(name)
            else:
                cur.skip_to_end(name)
        else:
            raise GmshError(f"{path}: MSH version {version} (2.2 and 4.1 are read)")
    if version is None:
        raise GmshError(f"{path}: no $MeshFormat")
    if ids is None or not elems:
        raise GmshError(f"{path}: MSH file without $Nodes or $Elements")
    return version, ids, xyz, elems, links, names


def _orient_2d(geom, verts, X):
    """CheckElementOrientation(fix_it = true) for 2D cells: the Jacobian at the
    cell centre decides; inverted cells get two vertices exchanged."""
    if geom == M.GEOM_TRI:
        J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]], axis=2)
        swap = (0, 1)
    else:
        # d/dxi and d/deta of the bilinear map at (1/2, 1/2)
        J = np.stack([0.5 * (X[:, 1] + X[:, 2] - X[:, 0] - X[:, 3]),
                      0.5 * (X[:, 3] + X[:, 2] - X[:, 0] - X[:, 1])], axis=2)
        swap = (1, 3)
    det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
    bad = det < 0.0
    a, b = swap
    for arr in (verts, X):
        tmp = arr[bad, a].copy()
        arr[bad, a] = arr[bad, b]
        arr[bad, b] = tmp
    return int(bad.sum())


def _mark_triangles(verts, X):
    """Mesh::MarkTriMeshForRefinement: rotate (v0, v1, v2) so that 0-1 is the
    longest edge; the comparisons keep the earlier edge on ties."""
    d0 = ((X[:, 1] - X[:, 0]) ** 2).sum(axis=1)
    d1 = ((X[:, 2] - X[:, 1]) ** 2).sum(axis=1)
    d2 = ((X[:, 2] - X[:, 0]) ** 2).sum(axis=1)
    shift = np.where(d0 >= d1, np.where(d0 >= d2, 0, 2), np.where(d1 >= d2, 1, 2))
    for s, perm in ((1, [1, 2, 0]), (2, [2, 0, 1])):
        sel = shift == s
        verts[sel] = verts[sel][:, perm]
        X[sel] = X[sel][:, perm]


def _hex_jacobians_ok(X):
    """Every corner Jacobian of the trilinear map positive (C-ABI requirement)."""
    corners = [(0, 1, 3, 4), (1, 2, 0, 5), (2, 3, 1, 6), (3, 0, 2, 7),
               (4, 7, 5, 0), (5, 4, 6, 1), (6, 5, 7, 2), (7, 6, 4, 3)]
    for c, a, b, d in corners:
        J = np.stack([X[:, a] - X[:, c], X[:, b] - X[:, c], X[:, d] - X[:, c]], axis=2)
        if np.any(np.linalg.det(J) <= 0.0):
            return False
    return True


def read_gmsh(path, fix_orientation: bool = True, mark_triangles: bool = True) -> GmshMesh:
    """Reads an MSH 2.2 or 4.1 file, ASCII or binary, gzip-compressed when its
    name ends in .gz (case_manager.cpp:255-263, 298-448)."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rb") as fh:
        data = fh.read()
    at = data.find(b"$MeshFormat")
    if at < 0:
        raise GmshError(f"{path}: no $MeshFormat")
    fmt_line = data[at:].split(b"\n", 2)[1].split()
    if len(fmt_line) < 2:
        raise GmshError(f"{path}: malformed $MeshFormat")
    if int(fmt_line[1]) == 1:
        version, ids, xyz, elems, links, names = _read_binary(data, path)
    else:
        secs = {}
        for name, lines in _sections(data.decode("utf-8", "replace")):
            secs.setdefault(name, lines)
        version = secs["MeshFormat"][0].split()[0]
        if version.startswith("2"):
            ids, xyz, elems, links = _read_v2(secs)
        elif version.startswith("4.1"):
            ids, xyz, elems, links = _read_v41(secs)
        else:
            raise GmshError(f"{path}: MSH version {version} (2.2 and 4.1 are read)")
        names = _physical_names(secs["PhysicalNames"]) if "PhysicalNames" in secs else []

    for et, _, nodes in elems:
        if et not in _ELEM:
            raise GmshError(f"{path}: unknown element type {et}")
        if _ELEM[et][1] != len(nodes):
            raise GmshError(f"{path}: element of type {et} with {len(nodes)} nodes")
    dims = {_ELEM[et][0] for et, _, _ in elems}
    if not dims:
        raise GmshError(f"{path}: no elements")
    dim = max(dims)
    cell_types = {et for et, _, _ in elems if _ELEM[et][0] == dim}
    if len(cell_types) != 1:
        raise GmshError(f"{path}: cells of types {sorted(cell_types)}; one cell type per mesh is supported")
    ct = cell_types.pop()
    if ct not in _CELL_GEOM:
        raise GmshError(f"{path}: cell type {ct} is not a first-order triangle, quadrilateral or hexahedron")
    geom = _CELL_GEOM[ct]
    ft = _FACE_TYPE[geom]

    pos = np.full(int(ids.max()) + 1, -1, np.int64)
    pos[ids] = np.arange(len(ids))
    # periodic identification: every slave node points at its final master
    topo = np.arange(len(ids), dtype=np.int64)
    if links:
        for a, b in links:
            topo[pos[a]] = pos[b]
        for _ in range(8):  # resolve chains (corner nodes may be linked twice)
            nxt = topo[topo]
            if np.array_equal(nxt, topo):
                break
            topo = nxt

    cells = [(phys, nodes) for et, phys, nodes in elems if et == ct]
    cverts = np.array([[pos[n] for n in nodes] for _, nodes in cells], np.int64)
    cattr = np.array([phys for phys, _ in cells], np.int64)
    X = xyz[cverts][:, :, :M._DIM[geom]].copy()
    if geom != M.GEOM_HEX and np.any(np.abs(xyz[:, 2]) > 0) and np.ptp(xyz[:, 2]) > 0:
        raise GmshError(f"{path}: a 2D mesh that is not in the z = const plane")
    if geom != M.GEOM_HEX and fix_orientation:
        _orient_2d(geom, cverts, X)
    if geom == M.GEOM_HEX and not _hex_jacobians_ok(X):
        raise GmshError(f"{path}: inverted or degenerate hexahedra")
    if geom == M.GEOM_TRI and mark_triangles:
        _mark_triangles(cverts, X)

    bdr = [(phys, nodes) for et, phys, nodes in elems if _ELEM[et][0] == dim - 1]
    for et, _, _ in elems:
        if _ELEM[et][0] == dim - 1 and et != ft:
            raise GmshError(f"{path}: boundary element type {et} does not match the cells' faces")
    bverts = np.array([[topo[pos[n]] for n in nodes] for _, nodes in bdr], np.int64).reshape(-1, M._NVF[geom])
    battr = np.array([phys for phys, _ in bdr], np.int64)

    # minimizeAttributes (case_manager.cpp:354-448)
    cell_set = np.unique(cattr)
    bdr_set = np.unique(battr)
    ncell = len(cell_set)
    cnew = np.searchsorted(cell_set, cattr) + 1
    bnew = np.searchsorted(bdr_set, battr) + 1 + ncell if len(battr) else battr
    cell_names, bdr_names = {}, {}
    for d, tag, name in names:
        if d == dim - 1:
            hit = np.nonzero(bdr_set == tag)[0]
            bdr_names[name] = int(hit[0]) + 1 + ncell if len(hit) else tag
        elif d == dim:
            hit = np.nonzero(cell_set == tag)[0]
            cell_names[name] = int(hit[0]) + 1 if len(hit) else tag
        else:
            # case_manager.cpp:329-342
            raise GmshError(f"physical group {name} of dimension {d} in a {dim}D mesh")

    # boundary elements on interior faces (periodic links, internal physical
    # lines) carry no boundary integrator: MFEM's GetBdrFaceTransformations
    # gives none for an interior face. They count in the numbering above and
    # are dropped here.
    cv = topo[cverts]
    if len(bverts):
        lf = M._local_faces(geom)
        keys, counts = np.unique(np.sort(cv[:, np.asarray(lf)].reshape(-1, M._NVF[geom]), axis=1),
                                 axis=0, return_counts=True)
        open_keys = {tuple(k) for k in keys[counts == 1].tolist()}
        keep = np.array([tuple(k) in open_keys for k in np.sort(bverts, axis=1).tolist()], bool)
        bverts, bnew = bverts[keep], np.asarray(bnew)[keep]

    m = M.Mesh(geom, X, cv, bverts, np.asarray(bnew, np.int32), name=str(path))
    return GmshMesh(m, cnew.astype(np.int32), cell_names, bdr_names, version, bool(links))


def write_gmsh22(path, m: M.Mesh, bdr_names: dict = None, cell_name: str = "Volume") -> None:
    """Writes a mesh description as ASCII MSH 2.2 (nodes per element corner;
    periodic meshes are written unlinked). For tests and for moving the
    generated workloads' meshes into the reference's case format."""
    dim = m.dim
    X = m.elem_coords
    verts = m.elem_verts
    nv = verts.max() + 1
    coords = np.zeros((nv, 3))
    coords[verts.ravel(), :dim] = X.reshape(-1, dim)
    et = {M.GEOM_TRI: 2, M.GEOM_QUAD: 3, M.GEOM_HEX: 5}[m.geom]
    ft = _FACE_TYPE[m.geom]
    battrs = sorted(set(int(a) for a in m.bdr_attr))
    cell_tag = max(battrs, default=0) + 1
    names = dict(bdr_names or {a: f"Patch{a}" for a in battrs})
    with open(path, "w") as fh:
        fh.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$PhysicalNames\n")
        fh.write(f"{len(battrs) + 1}\n")
        for a in battrs:
            fh.write(f'{dim - 1} {a} "{names[a]}"\n')
        fh.write(f'{dim} {cell_tag} "{cell_name}"\n$EndPhysicalNames\n$Nodes\n{nv}\n')
        for i in range(nv):
            fh.write(f"{i + 1} " + " ".join(repr(float(c)) for c in coords[i]) + "\n")
        fh.write("$EndNodes\n$Elements\n")
        fh.write(f"{len(m.bdr_attr) + m.num_elements}\n")
        k = 1
        for b, a in zip(m.bdr_verts, m.bdr_attr):
            fh.write(f"{k} {ft} 2 {a} {a} " + " ".join(str(v + 1) for v in b) + "\n")
            k += 1
        for e in range(m.num_elements):
            fh.write(f"{k} {et} 2 {cell_tag} {cell_tag} " + " ".join(str(v + 1) for v in verts[e]) + "\n")
            k += 1
        fh.write("$EndElements\n")


def uniform_refinement(m: M.Mesh, levels: int = 1) -> M.Mesh:
    """Mesh::UniformRefinement for straight-sided cells (serialRefLevels,
    parallelRefLevels; case_manager.cpp:266-279): quadrilaterals and hexahedra
    split at their mid-points into 4 and 8, triangles into 4 at the edge
    mid-points. Boundary elements split with their faces and keep their
    attributes. Children are numbered cell by cell (MFEM numbers them in its own
    order; the state layout differs by that permutation only)."""
    for _ in range(levels):
        m = _refine_once(m)
    return m


def _refine_once(m: M.Mesh) -> M.Mesh:
    X, V = m.elem_coords, m.elem_verts
    NE = m.num_elements
    # topological ids of new points: keyed by the sorted vertex set they split
    keymap = {}
    nxt = [int(V.max()) + 1]

    def pid(*vs):
        k = tuple(sorted(int(v) for v in vs))
        if len(k) == 1:
            return k[0]
        if k not in keymap:
            keymap[k] = nxt[0]
            nxt[0] += 1
        return keymap[k]

    cells_v, cells_x = [], []
    bv, ba = [], []
    if m.geom == M.GEOM_QUAD:
        # points on the reference square: corners, edge mid-points, centre
        ref = {(0, 0): (0,), (2, 0): (1,), (2, 2): (2,), (0, 2): (3,), (1, 0): (0, 1), (2, 1): (1, 2),
               (1, 2): (2, 3), (0, 1): (3, 0), (1, 1): (0, 1, 2, 3)}
        kids = [[(0, 0), (1, 0), (1, 1), (0, 1)], [(1, 0), (2, 0), (2, 1), (1, 1)],
                [(1, 1), (2, 1), (2, 2), (1, 2)], [(0, 1), (1, 1), (1, 2), (0, 2)]]
        for e in range(NE):
            for kid in kids:
                cells_v.append([pid(*[V[e, c] for c in ref[p]]) for p in kid])
                cells_x.append([X[e, list(ref[p])].mean(axis=0) for p in kid])
        for b, a in zip(m.bdr_verts, m.bdr_attr):
            mid = pid(b[0], b[1])
            bv += [[b[0], mid], [mid, b[1]]]
            ba += [a, a]
    elif m.geom == M.GEOM_TRI:
        ref = {0: (0,), 1: (1,), 2: (2,), 3: (0, 1), 4: (1, 2), 5: (2, 0)}
        kids = [[0, 3, 5], [3, 1, 4], [5, 4, 2], [3, 4, 5]]
        for e in range(NE):
            for kid in kids:
                cells_v.append([pid(*[V[e, c] for c in ref[p]]) for p in kid])
                cells_x.append([X[e, list(ref[p])].mean(axis=0) for p in kid])
        for b, a in zip(m.bdr_verts, m.bdr_attr):
            mid = pid(b[0], b[1])
            bv += [[b[0], mid], [mid, b[1]]]
            ba += [a, a]
    else:
        # reference cube corners in MFEM's order; a point (i, j, k) in {0,1,2}^3
        # is the mean of the corners its half-integer coordinates span
        corner = {(0, 0, 0): 0, (2, 0, 0): 1, (2, 2, 0): 2, (0, 2, 0): 3,
                  (0, 0, 2): 4, (2, 0, 2): 5, (2, 2, 2): 6, (0, 2, 2): 7}

        def span(p):
            opts = [(0, 2) if c == 1 else (c,) for c in p]
            return tuple(corner[(a, b, c)] for a in opts[0] for b in opts[1] for c in opts[2])

        order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        for e in range(NE):
            for ox in (0, 1):
                for oy in (0, 1):
                    for oz in (0, 1):
                        pts = [(ox + a, oy + b, oz + c) for a, b, c in order]
                        cells_v.append([pid(*[V[e, c] for c in span(p)]) for p in pts])
                        cells_x.append([X[e, list(span(p))].mean(axis=0) for p in pts])
        for b, a in zip(m.bdr_verts, m.bdr_attr):
            q = list(b)
            m01, m12, m23, m30 = pid(q[0], q[1]), pid(q[1], q[2]), pid(q[2], q[3]), pid(q[3], q[0])
            c = pid(*q)
            bv += [[q[0], m01, c, m30], [m01, q[1], m12, c], [c, m12, q[2], m23], [m30, c, m23, q[3]]]
            ba += [a] * 4
    return M.Mesh(m.geom, np.array(cells_x), np.array(cells_v, np.int64),
                  np.array(bv, np.int64).reshape(-1, M._NVF[m.geom]), np.array(ba, np.int32), name=m.name)
