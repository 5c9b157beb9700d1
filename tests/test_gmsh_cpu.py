"""Gmsh meshes and YAML case files (CPU): the reader against the meshes the
reference ships (tests/golden/meshes), hand-written MSH 4.1 and $Periodic
files, MFEM's orientation fix and triangle marking, uniform refinement against
the structured generators, and the case settings with the reference's
initial conditions in closed form."""
import math
import os
import shutil

import numpy as np
import pytest

from oracle.refcpu import RefSolver
from rkdg_amd import case_file as CF
from rkdg_amd import gmsh
from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from tests.helpers import smooth_state

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")

# (file, geometry, cells, boundary elements, bounding box); counts read off the
# files' $Elements sections
SHIPPED = [
    ("sod2D_sod2d.msh", M.GEOM_QUAD, 100, 202, ((0.0, -0.1), (1.0, 0.1))),
    ("sod1d_covolume_sod2d.msh", M.GEOM_TRI, 200, 202, ((0.0, 0.0), (1.0, 0.1))),
    ("sod1d_quad.msh", M.GEOM_QUAD, 5, 12, ((0.0, 0.0), (1.0, 0.1))),
    ("sod1d_quad6.msh", M.GEOM_QUAD, 6, 14, ((0.0, 0.0), (1.0, 0.1))),
    ("sod1d_trian.msh", M.GEOM_TRI, 12, 14, ((0.0, 0.0), (1.0, 0.1))),
    ("jet_overexpanded0p25.msh.gz", M.GEOM_QUAD, 27263, 672, ((0.0, -0.072225), (0.107, 0.072225))),
]


def _areas(m):
    X = m.elem_coords
    if m.geom == M.GEOM_TRI:
        return 0.5 * ((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1])
                      - (X[:, 1, 1] - X[:, 0, 1]) * (X[:, 2, 0] - X[:, 0, 0]))
    x, y = X[..., 0], X[..., 1]
    return 0.5 * (x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y).sum(axis=1)


def _open_faces(m):
    """Faces with one adjacent cell, as sorted vertex tuples."""
    lf = M._local_faces(m.geom)
    cnt = {}
    for e in range(m.num_elements):
        for f in lf:
            k = tuple(sorted(int(m.elem_verts[e, q]) for q in f))
            cnt[k] = cnt.get(k, 0) + 1
    return {k for k, c in cnt.items() if c == 1}, cnt


@pytest.mark.parametrize("fname,geom,ne,nb,box", SHIPPED, ids=[s[0] for s in SHIPPED])
def test_shipped_meshes(fname, geom, ne, nb, box):
    g = gmsh.read_gmsh(os.path.join(MESHES, fname))
    m = g.mesh
    assert m.geom == geom and m.num_elements == ne and len(m.bdr_attr) == nb
    lo, hi = np.asarray(box[0]), np.asarray(box[1])
    np.testing.assert_allclose(m.elem_coords.min(axis=(0, 1)), lo, atol=1e-12)
    np.testing.assert_allclose(m.elem_coords.max(axis=(0, 1)), hi, atol=1e-12)
    # minimizeAttributes: one volume group -> 1; boundary groups after it
    assert set(g.elem_attr.tolist()) == {1} and g.cell_names == {"Volume": 1}
    nbg = len(g.bdr_names)
    assert sorted(g.bdr_names.values()) == list(range(2, 2 + nbg))
    assert sorted(set(m.bdr_attr.tolist())) == list(range(2, 2 + nbg))
    # orientation fixed: every cell counter-clockwise; cells tile the box
    # (the jet domain is not a box: its area is checked positive only)
    a = _areas(m)
    assert np.all(a > 0)
    if not fname.startswith("jet"):
        np.testing.assert_allclose(a.sum(), np.prod(hi - lo), rtol=1e-12)
    # every boundary element is an open face and every open face has one
    open_f, cnt = _open_faces(m)
    bset = {tuple(sorted(int(v) for v in b)) for b in m.bdr_verts}
    assert bset == open_f and max(cnt.values()) == 2
    if m.geom == M.GEOM_TRI:
        X = m.elem_coords
        d = [((X[:, (i + 1) % 3] - X[:, i]) ** 2).sum(axis=1) for i in range(3)]
        assert np.all(d[0] >= d[1]) and np.all(d[0] >= d[2])


def test_physical_name_maps_follow_renumbering():
    # sod1d_covolume: "Patch4" is physical line 1 -> the first boundary attribute
    g = gmsh.read_gmsh(os.path.join(MESHES, "sod1d_covolume_sod2d.msh"))
    assert g.bdr_names == {"Patch4": 2, "Patch3": 3, "Patch2": 4, "Patch1": 5}
    m = g.mesh
    # Patch4 (physical line 1) is one of the long edges: 100 segments on y = const
    coords = {}
    for e in range(m.num_elements):
        for q in range(3):
            coords[int(m.elem_verts[e, q])] = m.elem_coords[e, q]
    ys = np.array([[coords[int(v)][1] for v in b] for b in m.bdr_verts[m.bdr_attr == 2]])
    assert len(ys) == 100 and np.ptp(ys) == 0.0


MSH41 = """$MeshFormat
4.1 0 8
$EndMeshFormat
$PhysicalNames
3
1 7 "wall"
1 9 "side"
2 4 "fluid"
$EndPhysicalNames
$Entities
0 2 1 0
1 0 0 0 2 0 0 1 7 0
2 0 0 0 0 1 0 1 9 0
1 0 0 0 2 1 0 1 4 2 1 2
$EndEntities
$Nodes
1 6 1 6
2 1 0 6
1
2
3
4
5
6
0 0 0
1 0 0
2 0 0
0 1 0
1 1 0
2 1 0
$EndNodes
$Elements
3 6 1 6
1 1 1 2
1 1 2
2 2 3
1 2 1 2
3 3 6
4 4 1
2 1 3 2
5 1 2 5 4
6 2 3 6 5
$EndElements
"""


def test_msh41_hand_written(tmp_path):
    p = tmp_path / "two.msh"
    p.write_text(MSH41)
    g = gmsh.read_gmsh(str(p))
    m = g.mesh
    assert g.version == "4.1" and m.geom == M.GEOM_QUAD and m.num_elements == 2
    np.testing.assert_array_equal(m.elem_coords[1], [[1, 0], [2, 0], [2, 1], [1, 1]])
    # physical tags 7 and 9 -> 2 and 3 after the one cell group
    assert g.bdr_names == {"wall": 2, "side": 3} and g.cell_names == {"fluid": 1}
    np.testing.assert_array_equal(m.bdr_attr, [2, 2, 3, 3])
    assert {tuple(sorted(b)) for b in m.bdr_verts[:2].tolist()} == {(0, 1), (1, 2)}


def test_msh22_round_trip(tmp_path):
    for mk in (lambda: M.structured_quad(5, 3, 0.0, 2.0, 0.0, 1.0),
               lambda: M.jittered_delaunay_tri(5, 4, 0.0, 1.0, 0.0, 1.0, seed=3),
               lambda: M.structured_hex(3, 2, 2, periodic=(False, False, False))):
        m = mk()
        p = tmp_path / "m.msh"
        gmsh.write_gmsh22(str(p), m)
        g = gmsh.read_gmsh(str(p), mark_triangles=False)
        r = g.mesh
        assert r.geom == m.geom and r.num_elements == m.num_elements
        np.testing.assert_array_equal(r.elem_coords, m.elem_coords)
        np.testing.assert_array_equal(r.elem_verts, m.elem_verts)
        # attributes: the sorted distinct boundary tags follow the one cell group
        tags = sorted(set(m.bdr_attr.tolist()))
        np.testing.assert_array_equal(r.bdr_attr, [tags.index(a) + 2 for a in m.bdr_attr])


# ---------------------------------------------------------------- binary MSH
# Encoded here by hand from the Gmsh reference manual's layouts, independently
# of the reader: MSH 2.2 (int32 records, $PhysicalNames and $Periodic in ASCII)
# and MSH 4.1 (int32 tags and flags, size_t counts and node/element tags).

def _msh22_binary(path, m, links=(), bo="<"):
    import struct

    dim = m.dim
    nv = int(m.elem_verts.max()) + 1
    coords = np.zeros((nv, 3))
    coords[m.elem_verts.ravel(), :dim] = m.elem_coords.reshape(-1, dim)
    et = {M.GEOM_TRI: 2, M.GEOM_QUAD: 3, M.GEOM_HEX: 5}[m.geom]
    ft = {M.GEOM_TRI: 1, M.GEOM_QUAD: 1, M.GEOM_HEX: 3}[m.geom]
    battrs = sorted(set(int(a) for a in m.bdr_attr))
    cell_tag = max(battrs, default=0) + 1
    out = bytearray(b"$MeshFormat\n2.2 1 8\n" + struct.pack(bo + "i", 1) + b"\n$EndMeshFormat\n")
    out += b"$PhysicalNames\n%d\n" % (len(battrs) + 1)
    for a in battrs:
        out += b'%d %d "Patch%d"\n' % (dim - 1, a, a)
    out += b'%d %d "Volume"\n$EndPhysicalNames\n' % (dim, cell_tag)
    out += b"$Nodes\n%d\n" % nv
    for i in range(nv):
        out += struct.pack(bo + "i3d", i + 1, *coords[i])
    out += b"\n$EndNodes\n$Elements\n%d\n" % (len(m.bdr_attr) + m.num_elements)
    nb = len(m.bdr_attr)
    if nb:
        out += struct.pack(bo + "3i", ft, nb, 2)
        for k, (b, a) in enumerate(zip(m.bdr_verts, m.bdr_attr)):
            out += struct.pack(bo + "%di" % (3 + len(b)), k + 1, int(a), int(a), *(int(v) + 1 for v in b))
    out += struct.pack(bo + "3i", et, m.num_elements, 2)
    for e in range(m.num_elements):
        vs = m.elem_verts[e]
        out += struct.pack(bo + "%di" % (3 + len(vs)), nb + e + 1, cell_tag, cell_tag, *(int(v) + 1 for v in vs))
    out += b"\n$EndElements\n"
    if links:
        out += b"$Periodic\n1\n1 1 2\n%d\n" % len(links)
        for a, b in links:
            out += b"%d %d\n" % (a + 1, b + 1)
        out += b"$EndPeriodic\n"
    with open(path, "wb") as fh:
        fh.write(bytes(out))


@pytest.mark.parametrize("bo", ["<", ">"])
def test_msh22_binary_matches_ascii(tmp_path, bo):
    for mk in (lambda: M.structured_quad(5, 3, 0.0, 2.0, 0.0, 1.0),
               lambda: M.jittered_delaunay_tri(5, 4, 0.0, 1.0, 0.0, 1.0, seed=3),
               lambda: M.structured_hex(3, 2, 2, periodic=(False, False, False))):
        m = mk()
        pa, pb = tmp_path / "a.msh", tmp_path / "b.msh"
        gmsh.write_gmsh22(str(pa), m)
        _msh22_binary(str(pb), m, bo=bo)
        ga, gb = gmsh.read_gmsh(str(pa)), gmsh.read_gmsh(str(pb))
        for f in ("elem_coords", "elem_verts", "bdr_verts", "bdr_attr"):
            np.testing.assert_array_equal(getattr(gb.mesh, f), getattr(ga.mesh, f))
        np.testing.assert_array_equal(gb.elem_attr, ga.elem_attr)
        assert gb.bdr_names == ga.bdr_names and gb.cell_names == ga.cell_names


def test_msh22_binary_periodic(tmp_path):
    n = 4
    plain = M.structured_quad(n, n)
    per = M.structured_quad(n, n, periodic=(True, True))
    vid = lambda i, j: j * (n + 1) + i  # noqa: E731
    links = [(vid(n, j), vid(0, j)) for j in range(n + 1)] + [(vid(i, n), vid(i, 0)) for i in range(n + 1)]
    p = tmp_path / "per.msh"
    _msh22_binary(str(p), plain, links)
    g = gmsh.read_gmsh(str(p))
    assert g.periodic
    np.testing.assert_array_equal(g.mesh.elem_coords, per.elem_coords)
    assert not _open_faces(g.mesh)[0]


def _msh41_binary(path, periodic=False):
    """The two-quad mesh of MSH41 above in binary MSH 4.1, optionally with the
    right edge's nodes linked to the left edge's."""
    import struct

    out = bytearray(b"$MeshFormat\n4.1 1 8\n" + struct.pack("<i", 1) + b"\n$EndMeshFormat\n")
    out += b'$PhysicalNames\n3\n1 7 "wall"\n1 9 "side"\n2 4 "fluid"\n$EndPhysicalNames\n'
    out += b"$Entities\n" + struct.pack("<4Q", 0, 2, 1, 0)
    # curves 1 (phys 7) and 2 (phys 9): tag, bbox, phys tags, bounding points
    out += struct.pack("<i6dQiQ", 1, 0, 0, 0, 2, 0, 0, 1, 7, 0)
    out += struct.pack("<i6dQiQ", 2, 0, 0, 0, 0, 1, 0, 1, 9, 0)
    # surface 1 (phys 4) bounded by curves 1, 2
    out += struct.pack("<i6dQiQ2i", 1, 0, 0, 0, 2, 1, 0, 1, 4, 2, 1, 2)
    out += b"\n$EndEntities\n$Nodes\n" + struct.pack("<4Q", 1, 6, 1, 6)
    out += struct.pack("<3iQ", 2, 1, 0, 6) + struct.pack("<6Q", 1, 2, 3, 4, 5, 6)
    out += struct.pack("<18d", 0, 0, 0, 1, 0, 0, 2, 0, 0, 0, 1, 0, 1, 1, 0, 2, 1, 0)
    out += b"\n$EndNodes\n$Elements\n" + struct.pack("<4Q", 3, 6, 1, 6)
    out += struct.pack("<3iQ", 1, 1, 1, 2) + struct.pack("<6Q", 1, 1, 2, 2, 2, 3)
    out += struct.pack("<3iQ", 1, 2, 1, 2) + struct.pack("<6Q", 3, 3, 6, 4, 4, 1)
    out += struct.pack("<3iQ", 2, 1, 3, 2) + struct.pack("<10Q", 5, 1, 2, 5, 4, 6, 2, 3, 6, 5)
    out += b"\n$EndElements\n"
    if periodic:
        out += b"$Periodic\n" + struct.pack("<Q", 1)
        out += struct.pack("<3iQ", 1, 2, 2, 0)  # curve 2 -> itself as master, no affine map
        out += struct.pack("<Q4Q", 2, 3, 1, 6, 4)
        out += b"\n$EndPeriodic\n"
    with open(path, "wb") as fh:
        fh.write(bytes(out))


def test_msh41_binary_matches_ascii(tmp_path):
    pa, pb = tmp_path / "a.msh", tmp_path / "b.msh"
    pa.write_text(MSH41)
    _msh41_binary(str(pb))
    ga, gb = gmsh.read_gmsh(str(pa)), gmsh.read_gmsh(str(pb))
    assert gb.version == "4.1"
    for f in ("elem_coords", "elem_verts", "bdr_verts", "bdr_attr"):
        np.testing.assert_array_equal(getattr(gb.mesh, f), getattr(ga.mesh, f))
    assert gb.bdr_names == {"wall": 2, "side": 3} and gb.cell_names == {"fluid": 1}
    # node 3 -> 1 and 6 -> 4: the right edge is the left edge, the cells join there
    gp = gmsh.read_gmsh(str(_write_tmp(tmp_path, periodic=True)))
    assert gp.periodic
    np.testing.assert_array_equal(gp.mesh.elem_verts[1], [1, 0, 3, 4])


def _write_tmp(tmp_path, periodic):
    p = tmp_path / "p.msh"
    _msh41_binary(str(p), periodic=periodic)
    return p


def test_binary_msh_errors(tmp_path):
    import struct

    p = tmp_path / "bad.msh"
    p.write_bytes(b"$MeshFormat\n2.2 1 4\n" + struct.pack("<i", 1) + b"\n$EndMeshFormat\n")
    with pytest.raises(gmsh.GmshError, match="data size"):
        gmsh.read_gmsh(str(p))
    p.write_bytes(b"$MeshFormat\n2.2 1 8\n" + struct.pack("<i", 7) + b"\n$EndMeshFormat\n")
    with pytest.raises(gmsh.GmshError, match="integer 1"):
        gmsh.read_gmsh(str(p))
    p.write_bytes(b"$MeshFormat\n2.2 1 8\n" + struct.pack("<i", 1) + b"\n$EndMeshFormat\n$Nodes\n5\n"
                  + struct.pack("<i3d", 1, 0, 0, 0) + b"\n$EndNodes\n")
    with pytest.raises(gmsh.GmshError, match="ends early"):
        gmsh.read_gmsh(str(p))


def test_orientation_fix(tmp_path):
    m = M.structured_quad(3, 2)
    t = M.jittered_delaunay_tri(3, 3, 0.0, 1.0, 0.0, 1.0, seed=1)
    for mm, swap in ((m, (1, 3)), (t, (0, 1))):
        bad = mm.elem_verts.copy()
        X = mm.elem_coords.copy()
        # turn cells 0 and 2 over (reverse their vertex order)
        for e in (0, 2):
            bad[e] = bad[e][::-1]
            X[e] = X[e][::-1]
        flipped = M.Mesh(mm.geom, X, bad, mm.bdr_verts, mm.bdr_attr)
        p = tmp_path / "f.msh"
        gmsh.write_gmsh22(str(p), flipped)
        r = gmsh.read_gmsh(str(p), mark_triangles=False).mesh
        assert np.all(_areas(r) > 0)
        a, b = swap
        for e in (0, 2):
            want = X[e].copy()
            want[[a, b]] = want[[b, a]]
            np.testing.assert_array_equal(r.elem_coords[e], want)
        for e in (1, 3):
            np.testing.assert_array_equal(r.elem_coords[e], mm.elem_coords[e])


def _append_periodic_22(path, links):
    with open(path, "a") as fh:
        fh.write(f"$Periodic\n1\n1 1 2\n{len(links)}\n")
        for a, b in links:
            fh.write(f"{a + 1} {b + 1}\n")
        fh.write("$EndPeriodic\n")


def test_periodic_links_match_generator(tmp_path):
    n = 4
    plain = M.structured_quad(n, n)
    per = M.structured_quad(n, n, periodic=(True, True))
    p = tmp_path / "per.msh"
    gmsh.write_gmsh22(str(p), plain)
    # node ids of structured_quad: j * (n + 1) + i; right -> left, top -> bottom
    vid = lambda i, j: j * (n + 1) + i  # noqa: E731
    links = [(vid(n, j), vid(0, j)) for j in range(n + 1)] + [(vid(i, n), vid(i, 0)) for i in range(n + 1)]
    _append_periodic_22(str(p), links)
    g = gmsh.read_gmsh(str(p))
    assert g.periodic
    r = g.mesh
    np.testing.assert_array_equal(r.elem_coords, per.elem_coords)
    open_f, _ = _open_faces(r)
    assert not open_f
    # the oracle sees the same operator on both descriptions
    sch = S.Scheme(riemann=S.HLLC)
    a, b = RefSolver(r, 2, sch), RefSolver(per, 2, sch)
    x = smooth_state(a.node_coords(), seed=2)
    np.testing.assert_array_equal(a.mult(x)[0], b.mult(x)[0])


def _match_by_centroid(ma, mb):
    ca = ma.elem_coords.mean(axis=1)
    cb = mb.elem_coords.mean(axis=1)
    ka = np.lexsort(np.round(ca, 12).T[::-1])
    kb = np.lexsort(np.round(cb, 12).T[::-1])
    perm = np.empty(len(ka), np.int64)
    perm[kb] = ka
    return perm  # mb element i == ma element perm[i]


@pytest.mark.parametrize("kind", ["quad", "hex", "tri"])
def test_uniform_refinement(kind):
    if kind == "quad":
        coarse, fine, p = M.structured_quad(3, 2, 0.0, 1.5, 0.0, 1.0), M.structured_quad(6, 4, 0.0, 1.5, 0.0, 1.0), 2
    elif kind == "hex":
        nop = (False, False, False)
        coarse, fine, p = M.structured_hex(2, 2, 1, periodic=nop), M.structured_hex(4, 4, 2, periodic=nop), 1
    else:
        coarse, fine, p = M.jittered_delaunay_tri(3, 3, 0.0, 1.0, 0.0, 1.0, seed=5), None, 2
    r = gmsh.uniform_refinement(coarse, 1)
    k = 4 if coarse.dim == 2 else 8
    assert r.num_elements == k * coarse.num_elements
    assert len(r.bdr_attr) == (2 if coarse.dim == 2 else 4) * len(coarse.bdr_attr)
    for a in set(coarse.bdr_attr.tolist()):
        assert (r.bdr_attr == a).sum() == (k // 2) * (coarse.bdr_attr == a).sum()
    open_f, cnt = _open_faces(r)
    bset = {tuple(sorted(int(v) for v in b)) for b in r.bdr_verts}
    assert bset == open_f and max(cnt.values()) == 2
    bcs = [S.BC(a, S.BC_SLIP) for a in sorted(set(r.bdr_attr.tolist()))]
    sch = S.Scheme(riemann=S.HLLC, bcs=bcs)
    ref_r = RefSolver(r, p, sch)
    x = smooth_state(ref_r.node_coords(), seed=4)
    y = ref_r.mult(x)[0]
    if fine is None:
        a = _areas(coarse)
        np.testing.assert_allclose(_areas(r).reshape(-1, 4), np.repeat(a[:, None] / 4, 4, axis=1), rtol=1e-13)
        return
    # the children are the fine generator's cells, each with its local node order
    perm = _match_by_centroid(r, fine)
    np.testing.assert_array_equal(r.elem_coords[perm], fine.elem_coords)
    ref_f = RefSolver(fine, p, S.Scheme(riemann=S.HLLC, bcs=bcs))
    neq, nd = r.dim + 2, ref_r.nd
    xf = x.reshape(neq, -1, nd)[:, perm].reshape(-1)
    yf = ref_f.mult(xf)[0]
    np.testing.assert_allclose(y.reshape(neq, -1, nd)[:, perm].reshape(-1), yf, rtol=0,
                               atol=1e-12 * np.abs(yf).max())


def _case_dir(tmp_path, mesh_name, settings):
    d = tmp_path / "case"
    d.mkdir(exist_ok=True)
    shutil.copyfile(os.path.join(MESHES, mesh_name), d / mesh_name)
    (d / "settings.yml").write_text(settings.replace("@MESH@", mesh_name))
    return str(d)


SETTINGS = """
mesh:
    file: @MESH@
    serialRefLevels: 0
    parallelRefLevels: 0
    adaptive: false
physics:
    type:   covolume
    gamma:  1.3
    covolume: 0.001
    R: 287.0
internalField:
    type:       planeBreakup
    origin:     [0.4,0,0]
    normal:     [1,0,0]
    left:
        rho:    100.0
        U:      [1,0,0]
        p:      1e+8
    right:
        rho:    1.0
        U:      [0,0,0]
        p:      0.1
boundaryField:
    Patch1:
        tag: 1
        type:   slip
    Patch2:
        type:   supersonicInlet
        rhoI:   2.0
        UI:     [3.0, 0.5, 7.0]
        pI:     4.0
    Patch3:
        type:   fixedPressureOutlet
        pFix:   0.5
    notInMesh:
        type:   slip
spatial:
    polyOrder:      2
    riemannSolver:  HLLC
    limiter:
        type:       Multi
        indicator:  BJ
time:
    restart: false
    restartCycle:  100
    nSavedFrames: 10
    order:  3
    tEnd:   0.0002
    dt:     1e-6
    type:   dynamic
    CoMax:  0.1
postProcess:
    format:     vtk
    visSteps:   500
    checkTotalEnergy: 1
"""


def test_case_file_settings(tmp_path):
    cs = CF.load_case(_case_dir(tmp_path, "sod1d_covolume_sod2d.msh", SETTINGS))
    sch = cs.scheme
    assert (sch.gamma, sch.covolume, sch.gas_constant) == (1.3, 0.001, 287.0)
    assert sch.riemann == S.HLLC and sch.limiter == S.LIM_MULTI and sch.indicator == S.IND_BJ
    assert sch.last_hope == 1 and cs.order == 2
    assert cs.tableau == "ref_literal_rk3" and cs.cfl == 0.1 and cs.dt_const is None
    assert (cs.t_final, cs.dt, cs.n_saved_frames, cs.vis_steps) == (0.0002, 1e-6, 10, 500)
    assert cs.check_total_energy and not cs.write_indicators and not cs.restart
    # patches by name: Patch1 -> 5, Patch2 -> 4, Patch3 -> 3 in this mesh
    got = {(b.attr, b.type, tuple(b.params)) for b in sch.bcs}
    assert got == {(5, S.BC_SLIP, ()), (4, S.BC_SUPERSONIC_INLET, (2.0, 3.0, 0.5, 7.0, 4.0)),
                   (3, S.BC_FIXED_PRESSURE_OUTLET, (0.5,))}
    assert cs.skipped_patches == ["notInMesh"]
    # initial state: left of x = 0.4 (strictly), covolume energy
    ref = RefSolver(cs.mesh, cs.order, sch)
    xyz = ref.node_coords()
    U = cs.initial_state(xyz).reshape(4, -1)
    X = xyz[..., 0].ravel()
    left = X < 0.4
    rho = np.where(left, 100.0, 1.0)
    p = np.where(left, 1e8, 0.1)
    u = np.where(left, 1.0, 0.0)
    np.testing.assert_array_equal(U[0], rho)
    np.testing.assert_array_equal(U[1], rho * u)
    np.testing.assert_array_equal(U[2], 0 * rho)
    # the pressure the reference's EoS reads back (physics.cpp:101-112)
    pr = (1.3 - 1.0) * (U[3] - 0.5 * U[1] ** 2 / U[0]) / (1.0 - U[0] * 0.001)
    np.testing.assert_allclose(pr, p, rtol=1e-12)


@pytest.mark.parametrize("edit,msg", [
    (("type:       Multi", "type:       BJ"), "unknown limiter type"),
    (("indicator:  BJ", "indicator:  Shu2"), "unknown indicator"),
    (("riemannSolver:  HLLC", "riemannSolver:  Roe"), "unknown Riemann solver"),
    (("order:  3", "order:  4"), "wrong RK scheme order"),
    (("type:   fixedPressureOutlet", "type:   inlet"), "wrong boundary type"),
    (("notInMesh:", "Patch1:"), "appears twice"),
    (("adaptive: false", "adaptive: true"), "maxElemError"),
    (("adaptive: false", "adaptive: true\n    maxElemError: 1e-3\n    nonConformingLimit: 2"), "nonConformingLimit"),
    (("adaptive: false", "adaptive: true\n    maxElemError: 1e-3\n    preferConformingRefinement: true"),
     "preferConformingRefinement"),
])
def test_case_file_errors(tmp_path, edit, msg):
    with pytest.raises(CF.CaseError, match=msg):
        CF.load_case(_case_dir(tmp_path, "sod1d_covolume_sod2d.msh", SETTINGS.replace(*edit)))


def test_case_file_adaptive_settings(tmp_path):
    # the keys of the reference's adaptive cases (cases/sod2d/settings.yml:1-14)
    s = SETTINGS.replace("adaptive: false", "adaptive: 1\n    totalErrorFraction: 0.0\n    maxElemError: 10e-4\n"
                         "    hysteresis: 0.5\n    preferConformingRefinement: false\n    nonConformingLimit: 1\n"
                         "    maxRefLevel: 3")
    cs = CF.load_case(_case_dir(tmp_path, "sod1d_quad.msh", s))
    assert cs.adaptive and cs.max_elem_error == 1e-3 and cs.hysteresis == 0.5 and cs.total_error_fraction == 0.0
    assert cs.max_ref_level == 3


def test_case_file_constant_time_and_refinement(tmp_path):
    s = SETTINGS.replace("type:   dynamic", "type:   constant").replace("serialRefLevels: 0", "serialRefLevels: 1")
    cs = CF.load_case(_case_dir(tmp_path, "sod1d_quad.msh", s))
    assert cs.cfl < 0 and cs.dt_const == 1e-6
    assert cs.mesh.num_elements == 20 and len(cs.mesh.bdr_attr) == 24


def test_initial_conditions_closed_forms():
    g, b = 1.4, 0.0
    X = np.array([[0.1, 0.2], [0.5, 0.0], [0.9, 0.3]])
    ic = CF.ICPlaneBreakup((1.0, 0.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.0, 0.1),
                           (0.5, 0.0, 0.0), (1.0, 0.0, 0.0), g, b)
    U = ic(X).reshape(4, -1)
    np.testing.assert_array_equal(U[0], [1.0, 0.125, 0.125])  # on the plane: the right state
    np.testing.assert_array_equal(U[3], [1.0 / (g - 1.0), 0.1 / (g - 1.0), 0.1 / (g - 1.0)])
    # spherical: inside, on the ramp (|x|^2 - r^2 < 0.04 r^2, ramp to 1.2 r), outside
    r = 0.5
    sp = CF.ICSphericalBreakup((2.0, 0.0, 0.0, 0.0, 3.0), (1.0, 0.0, 0.0, 0.0, 1.0), (0.0, 0.0, 0.0), r, g, b)
    d_ramp = r * math.sqrt(1.03)
    P = np.array([[0.3, 0.0, 0.0], [d_ramp, 0.0, 0.0], [0.6, 0.0, 0.0]])
    U = sp(P).reshape(5, -1)
    ramp = (d_ramp - r) * (1.0 - 2.0) / (0.2 * r) + 2.0
    np.testing.assert_allclose(U[0], [2.0, ramp, 1.0], rtol=1e-15)
    # density pulse: rho at |x - o| = 1 from the formula with the source's pi literal
    eps = 5.0
    dp = CF.ICDensityPulse((0.0, 0.3, -0.2, 0.0, 1.0), (0.0, 0.0, 0.0), eps, g, b)
    U = dp(np.array([[1.0, 0.0], [0.0, 0.0]])).reshape(4, -1)
    pi_ = 3.14169265
    want = (1.0 - (g - 1.0) * eps * eps / (8.0 * g * pi_ * pi_)) ** (1.0 / (g - 1.0))
    assert abs(U[0, 0] - want) <= 1e-15 * want
    want0 = (1.0 - (g - 1.0) * eps * eps * math.exp(-1.0) / (8.0 * g * pi_ * pi_)) ** (1.0 / (g - 1.0))
    assert abs(U[0, 1] - want0) <= 1e-15 * want0
    np.testing.assert_allclose(U[1], 0.3 * U[0], rtol=1e-15)
    # constant state in 3D keeps w; in 2D w is dropped from the energy
    c = CF.ICConstant((1.2, 1.0, 2.0, 3.0, 1.0), g, b)
    U3 = c(np.zeros((1, 3))).reshape(5)
    U2 = c(np.zeros((1, 2))).reshape(4)
    assert U3[4] == 1.0 / (g - 1.0) + 0.5 * 1.2 * 14.0 and U2[3] == 1.0 / (g - 1.0) + 0.5 * 1.2 * 5.0
