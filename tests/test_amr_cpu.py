"""Non-conforming refinement on the host (rkdg_amd/amr.py) and the oracle's
non-conforming faces, on the CPU.

The refiner: leaf counts, the one-level limit propagating to coarser
neighbours, boundary faces split with their attribute, every open face either
a boundary face, a master or a recorded slave, point matrices at the dyadic
positions of the refinement, exact solution transfer.

The oracle's slave-face integration (face_integrator.cpp on an NCMesh slave
face) and non-conforming stencil (limiter.cpp:166-217) are pinned by:
- free-stream preservation and discrete conservation on meshes with one and
  two levels of hanging nodes;
- exactness: for a state whose flux is a polynomial the face rules integrate
  exactly (density constant, velocity and pressure linear, p = 2), every
  element's right-hand side depends on the element alone, so an element of
  the non-conforming mesh has the right-hand side it has in the uniformly
  refined (conforming) mesh or in the base mesh. A wrong point matrix, slave
  weight or normal breaks this by O(1), which the last test shows on a
  corrupted matrix;
- the limiter's stencil: a cell whose only out-of-range neighbour lies across
  a master face is limited only when the non-conforming neighbours count.
"""
import numpy as np
import pytest

from oracle.refcpu import RefSolver
from rkdg_amd import amr
from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from tests.helpers import smooth_state


def _outlet(m):
    return S.Scheme(riemann=S.HLLC, bcs=[S.BC(int(a), S.BC_OUTLET) for a in sorted(set(m.bdr_attr.tolist()))])


def test_refine_counts_and_limit_propagation():
    base = M.structured_quad(4, 4)
    t = amr.NCMesh(base)
    t.refine([5])
    m1 = t.mesh()
    assert m1.num_elements == 16 - 1 + 4
    # the four children of element 5 touch its four neighbours through 8 slave faces
    assert len(m1.nc) == 8
    # refining a child at the corner next to element 0 again: its neighbours
    # 1 and 4 (level 0) would see level 2, so they are refined too
    keys = t.ordered_keys()
    child = keys.index((5, (0,)))
    t.refine([child])
    m2 = t.mesh()
    lv = t.levels()
    assert lv.max() == 2
    assert (1, ()) not in t.leaves and (4, ()) not in t.leaves
    assert m2.num_elements == 19 - 1 + 4 + 2 * 3
    # every slave face is one level finer than its master
    for s, _, mm, _ in m2.nc:
        assert lv[s] == lv[mm] + 1


def test_boundary_faces_split_with_attributes():
    base = M.structured_quad(3, 2)
    t = amr.NCMesh(base)
    t.refine([0])
    m = t.mesh()
    assert len(m.bdr_attr) == len(base.bdr_attr) + 2
    assert sorted(set(m.bdr_attr.tolist())) == sorted(set(base.bdr_attr.tolist()))
    # total boundary length unchanged
    def blen(mm):
        vx = {}
        for e in range(mm.num_elements):
            for v in range(4):
                vx[int(mm.elem_verts[e, v])] = mm.elem_coords[e, v]
        return sum(np.linalg.norm(vx[int(b[1])] - vx[int(b[0])]) for b in mm.bdr_verts)
    assert abs(blen(m) - blen(base)) < 1e-14


@pytest.mark.parametrize("geom", ["quad", "hex"])
def test_point_matrices_are_dyadic_and_cover_the_master(geom):
    if geom == "quad":
        t = amr.NCMesh(M.structured_quad(5, 5, periodic=(True, True)))
        t.refine([6, 12])
        t.refine([0])
    else:
        t = amr.NCMesh(M.structured_hex(3, 3, 3, periodic=(False, False, False)))
        t.refine([13])
    m = t.mesh()
    assert set(np.unique(m.nc_param).tolist()) <= {0.0, 0.5, 1.0}
    fd = m.dim - 1
    nvf = 2 if m.dim == 2 else 4
    # the slaves of each master face tile it: their parameter boxes have area
    # 2^-(dim-1) each and sum to one
    area = {}
    for (s, ls, mm, lm), par in zip(m.nc, m.nc_param):
        P = par.reshape(nvf, fd)
        ext = np.prod(P.max(axis=0) - P.min(axis=0))
        assert abs(ext - 0.5 ** fd) < 1e-15
        area[(mm, lm)] = area.get((mm, lm), 0.0) + ext
    assert all(abs(a - 1.0) < 1e-15 for a in area.values())


def test_prolongation_is_exact_for_polynomials():
    base = M.structured_quad(3, 3)
    p = 2
    ref0 = RefSolver(base, p, _outlet(base))
    X = ref0.node_coords()

    def field(xyz):
        x, y = xyz[..., 0].ravel(), xyz[..., 1].ravel()
        return np.concatenate([1 + 0.3 * x * x - 0.2 * x * y, 0.1 + x * y, -0.2 + y * y, 2.0 + 0.5 * x])

    t = amr.NCMesh(base)
    tr = t.refine([4, 8])
    m = t.mesh()
    x1 = amr.prolong_state(field(X), tr, p, 2, 4, base.num_elements)
    X1 = RefSolver(m, p, _outlet(m)).node_coords()
    assert np.abs(x1 - field(X1)).max() < 1e-14


def _poly_state(xyz, dim):
    X = xyz.reshape(-1, dim)
    rho = np.ones(len(X))
    vel = [0.3 + 0.1 * X[:, 0], -0.2 + 0.05 * X[:, 1]] + ([0.1 + 0.07 * X[:, 2]] if dim == 3 else [])
    p = 1.0 + 0.1 * X[:, 0] + 0.2 * X[:, 1] + (0.05 * X[:, 2] if dim == 3 else 0)
    E = p / 0.4 + 0.5 * rho * sum(v * v for v in vel)
    return np.concatenate([rho] + [rho * v for v in vel] + [E])


def _rhs_by_centroid(m, sch):
    ref = RefSolver(m, 2, sch)
    y, _ = ref.mult(_poly_state(ref.node_coords(), m.dim))
    Y = y.reshape(m.dim + 2, m.num_elements, -1)
    cen = m.elem_coords.mean(axis=1)
    return {tuple(np.round(c, 9)): Y[:, e] for e, c in enumerate(cen)}


def _exactness(base, marks, sch, nc_param=None):
    t = amr.NCMesh(base)
    t.refine(marks)
    mB = t.mesh()
    if nc_param is not None:
        mB = M.Mesh(mB.geom, mB.elem_coords, mB.elem_verts, mB.bdr_verts, mB.bdr_attr, nc=mB.nc,
                    nc_param=nc_param(mB.nc_param.copy()))
    full = amr.NCMesh(base)
    full.refine(range(base.num_elements))
    mA = full.mesh()
    assert len(mA.nc) == 0 and len(mB.nc) > 0
    rB, rA, r0 = _rhs_by_centroid(mB, sch), _rhs_by_centroid(mA, sch), _rhs_by_centroid(base, sch)
    dA = max(np.abs(rB[k] - rA[k]).max() / np.abs(rA[k]).max() for k in rB if k in rA)
    d0 = max(np.abs(rB[k] - r0[k]).max() / np.abs(r0[k]).max() for k in rB if k in r0)
    return dA, d0


@pytest.mark.parametrize("geom", ["quad", "hex"])
def test_oracle_nc_faces_exact_for_polynomial_flux(geom):
    if geom == "quad":
        base = M.structured_quad(5, 4)
        marks = [6, 7, 12]
        sch = S.Scheme(riemann=S.HLLC, bcs=[S.BC(a, S.BC_OUTLET) for a in (1, 2, 3, 4)])
    else:
        base = M.structured_hex(3, 3, 3, periodic=(False, False, False))
        marks = [13, 4]
        sch = S.Scheme(riemann=S.RUSANOV, bcs=[S.BC(a, S.BC_OUTLET) for a in range(1, 7)])
    dA, d0 = _exactness(base, marks, sch)
    # fine elements equal their copies in the uniformly refined mesh, coarse ones their base copies
    assert dA < 1e-12 and d0 < 1e-12


def test_oracle_nc_exactness_fails_on_a_corrupted_point_matrix():
    base = M.structured_quad(5, 4)
    sch = S.Scheme(riemann=S.HLLC, bcs=[S.BC(a, S.BC_OUTLET) for a in (1, 2, 3, 4)])
    dA, d0 = _exactness(base, [6, 7, 12], sch, nc_param=lambda q: q[:, ::-1])
    assert dA > 1e-2 or d0 > 1e-2


@pytest.mark.parametrize("geom,p", [("quad", 1), ("quad", 3), ("hex", 1), ("hex", 2)])
def test_oracle_nc_free_stream_and_conservation(geom, p):
    if geom == "quad":
        t = amr.NCMesh(M.structured_quad(6, 6, periodic=(True, True)))
        t.refine([14, 15])
        t.refine([0])
    else:
        t = amr.NCMesh(M.structured_hex(4, 4, 4))
        t.refine([21, 22])
    m = t.mesh()
    dim = m.dim
    ref = RefSolver(m, p, S.Scheme(riemann=S.HLLC))
    n = m.num_elements * (p + 1) ** dim
    rho, pr = 1.2, 0.9
    vel = [0.3, -0.2, 0.1][:dim]
    E = pr / 0.4 + 0.5 * rho * sum(v * v for v in vel)
    x = np.concatenate([np.full(n, rho)] + [np.full(n, rho * v) for v in vel] + [np.full(n, E)])
    y, _ = ref.mult(x)
    assert np.abs(y).max() < 1e-12
    y, _ = ref.mult(smooth_state(ref.node_coords(), seed=1))
    w = np.polynomial.legendre.leggauss(p + 1)[1] / 2
    W = w
    for _ in range(dim - 1):
        W = np.multiply.outer(w, W)
    X = m.elem_coords
    J = np.stack([X[:, 1] - X[:, 0], X[:, 3] - X[:, 0]] + ([X[:, 4] - X[:, 0]] if dim == 3 else []), axis=-1)
    det = np.abs(np.linalg.det(J))
    Y = y.reshape(dim + 2, m.num_elements, -1)
    tot = (Y * W.ravel()[None, None] * det[None, :, None]).sum(axis=(1, 2))
    mag = (np.abs(Y) * W.ravel()[None, None] * det[None, :, None]).sum(axis=(1, 2))
    assert np.abs(tot / mag).max() < 1e-14


def test_oracle_nc_stencil_includes_the_fine_neighbours():
    # a 3x1 strip refined at element 1: the coarse element 0 carries a ramp
    # rho = 1 + 0.3 x (mean 1.05, 1.1 at its right face, a master face), its
    # fine neighbours across that face the constant 1.06. The only indicator
    # points of element 0 are on the master face (its other faces are on the
    # boundary) and the only stencil neighbours are the fine ones, so
    # y = (1.06 - 1.05) / (1.1 - 1.05) = 0.2 exactly when both count
    base = M.structured_quad(3, 1)
    t = amr.NCMesh(base)
    t.refine([1])
    m = t.mesh()
    sch = S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, phi=S.PHI_BJ, limiter=S.LIM_MULTI,
                   bcs=[S.BC(a, S.BC_OUTLET) for a in (1, 2, 3, 4)])
    ref = RefSolver(m, 1, sch)
    X = ref.node_coords()
    x0 = X[..., 0].ravel()
    rho = np.where(x0 < 1.0 / 3.0, 1.0 + 0.3 * x0, 1.06)
    z = np.zeros_like(rho)
    x = np.concatenate([rho, z, z, 2.5 + 0.0 * rho])
    _, _, mv = ref.limit(x)
    assert abs(mv[0] - 0.2) < 1e-12


def test_indicator_vanishes_on_linear_fields_and_sees_a_kink():
    base = M.structured_quad(6, 6)
    t = amr.NCMesh(base)
    t.refine([14])
    m = t.mesh()
    X = RefSolver(m, 2, _outlet(m)).node_coords()
    x0, y0 = X[..., 0].ravel(), X[..., 1].ravel()
    one = np.ones_like(x0)
    lin = np.concatenate([1 + 0.3 * x0 - 0.2 * y0, one, one, one])
    assert amr.zz_indicator(m, 2, lin).max() < 1e-13
    kink = np.concatenate([1 + 0.5 * np.abs(x0 - 0.5), one, one, one])
    eta = amr.zz_indicator(m, 2, kink)
    # elements touching the line x = 0.5 (through the recovered vertex
    # gradients) see the kink; the others have a constant gradient and zero
    xs = m.elem_coords[:, :, 0]
    near = (xs.max(axis=1) >= 0.5 - 1e-12) & (xs.min(axis=1) <= 0.5 + 1e-12)
    assert eta[near].min() > 1e-3 and eta[~near].max() < 1e-13


def test_derefine_round_trip_is_exact_and_restores_the_mesh():
    base = M.structured_quad(4, 3)
    p = 2
    ref0 = RefSolver(base, p, _outlet(base))
    X = ref0.node_coords()
    x0 = X[..., 0].ravel()
    y0 = X[..., 1].ravel()
    field = np.concatenate([1 + 0.3 * x0 * x0 - 0.2 * x0 * y0, 0.1 + x0 * y0, -0.2 + y0 * y0, 2.0 + 0.5 * x0])
    t = amr.NCMesh(base)
    tr = t.refine([5, 6])
    m1 = t.mesh()
    x1 = amr.transfer_state(field, tr, p, 2, 4, base.num_elements)
    # derefinement: the L2 projection of the prolonged polynomial is the polynomial
    tr2, done = t.derefine(t.derefinable())
    assert sorted(done) == [(5, ()), (6, ())]
    m2 = t.mesh()
    assert m2.num_elements == base.num_elements and len(m2.nc) == 0
    assert len(m2.bdr_attr) == len(base.bdr_attr)
    x2 = amr.transfer_state(x1, tr2, p, 2, 4, m1.num_elements)
    assert np.abs(x2 - field).max() < 1e-14


def test_restriction_keeps_the_means():
    base = M.structured_quad(3, 3)
    p = 2
    t = amr.NCMesh(base)
    t.refine([4])
    m1 = t.mesh()
    rng = np.random.default_rng(3)
    x1 = 1.0 + rng.uniform(-0.3, 0.3, 4 * m1.num_elements * 9)
    tr, done = t.derefine(t.derefinable())
    m2 = t.mesh()
    x2 = amr.transfer_state(x1, tr, p, 2, 4, m1.num_elements)
    r1, r2 = RefSolver(m1, p, _outlet(m1)), RefSolver(m2, p, _outlet(m2))
    mean1 = r1.means(x1).reshape(4, -1)
    mean2 = r2.means(x2).reshape(4, -1)
    # the coarsened element (index 4 in the base order) averages its four children (4..7 before)
    assert np.abs(mean2[:, 4] - mean1[:, 4:8].mean(axis=1)).max() < 1e-14


def test_derefine_keeps_the_level_limit():
    base = M.structured_quad(4, 4)
    t = amr.NCMesh(base)
    t.refine([5])
    child = t.ordered_keys().index((5, (0,)))
    t.refine([child])
    # element 1 was refined to keep the limit next to the level-2 children of
    # element 5's child 0; coarsening it again would break the limit
    tr, done = t.derefine([(1, ())])
    assert done == [] and (1, (0,)) in t.leaves
    t.mesh()
    # element 5's child 0 itself can be coarsened
    tr, done = t.derefine([(5, (0,))])
    assert done == [(5, (0,))]
    t.mesh()


def test_adaptive_loop_with_the_oracle_refines_at_the_jump():
    from tests.helpers import conservative

    base = M.structured_quad(16, 4, 0.0, 1.0, 0.0, 0.25)
    sch = S.Scheme(riemann=S.HLLC, indicator=S.IND_BJ, phi=S.PHI_BJ, limiter=S.LIM_MULTI,
                   bcs=[S.BC(1, S.BC_SLIP), S.BC(3, S.BC_SLIP), S.BC(2, S.BC_OUTLET), S.BC(4, S.BC_OUTLET)])

    def ic(xyz):
        x = xyz[..., 0].ravel()
        rho = np.where(x < 0.5, 1.0, 0.125)
        p = np.where(x < 0.5, 1.0, 0.1)
        return np.concatenate(conservative(rho, [0 * x, 0 * x], p))

    r = amr.run_adaptive(base, 1, sch, ic, 0.05, 0.2, S.TABLEAUS["ssprk3"], max_elem_error=0.004,
                         max_steps=6, solver="oracle", max_level=2)
    lv = r.tree.levels()
    cen = r.mesh.elem_coords.mean(axis=1)
    assert lv.max() == 2 and len(r.mesh.nc) > 0
    # refined elements sit around the jump, none at the ends of the tube
    assert np.all(lv[(cen[:, 0] < 0.2) | (cen[:, 0] > 0.9)] == 0)
    assert r.steps == 6 and r.refinements >= 1


@pytest.mark.parametrize("p", [1, 2])
def test_curved_derefinement_is_the_weighted_l2_projection(p):
    """On curved (trilinear) hexahedra, coarsening keeps each conserved total
    (the integrals with the oracle's element mass matrices) and undoes a
    refinement exactly: a parent's polynomial prolonged to its children and
    restricted again comes back. At p = 1, det J (degree 2 per axis) is not in
    the space, and the unweighted restriction loses the totals."""
    from tests.helpers import curved_hex

    def totals(ref, m, x):
        Y = x.reshape(5, m.num_elements, -1)
        out, mag = np.zeros(5), np.zeros(5)
        for e in range(m.num_elements):
            ones = np.linalg.inv(ref.mass_inverse(e)).sum(axis=0)
            out += Y[:, e, :] @ ones
            mag += np.abs(Y[:, e, :]) @ np.abs(ones)
        return out, mag

    base = curved_hex(3, 3, 3, seed=80)
    t = amr.NCMesh(base)
    t.refine([4, 13])
    m1 = t.mesh()
    r1 = RefSolver(m1, p, S.Scheme())
    x1 = smooth_state(r1.node_coords(), seed=3, amp=1e-2)
    T1, mag = totals(r1, m1, x1)
    tr, done = t.derefine([(4, ()), (13, ())])
    assert len(done) == 2
    m2 = t.mesh()
    r2 = RefSolver(m2, p, S.Scheme())
    x2 = amr.transfer_state(x1, tr, p, 3, 5, m1.num_elements, m2.elem_coords)
    T2, _ = totals(r2, m2, x2)
    assert np.all(np.abs(T2 - T1) <= 1e-14 * mag)
    x2a = amr.transfer_state(x1, tr, p, 3, 5, m1.num_elements)  # unweighted
    assert np.abs(x2a - x2).max() > 1e-8
    if p == 1:
        T2a, _ = totals(r2, m2, x2a)
        assert np.abs(T2a - T1).max() > 1e-10 * mag.max()
    t2 = amr.NCMesh(base)
    r0 = RefSolver(base, p, S.Scheme())
    x0 = smooth_state(r0.node_coords(), seed=5, amp=1e-2)
    tr = t2.refine([4])
    ma = t2.mesh()
    xa = amr.transfer_state(x0, tr, p, 3, 5, base.num_elements, ma.elem_coords)
    tr, _ = t2.derefine([(4, ())])
    mb = t2.mesh()
    xb = amr.transfer_state(xa, tr, p, 3, 5, ma.num_elements, mb.elem_coords)
    np.testing.assert_allclose(xb, x0, rtol=0, atol=1e-13)


def test_zz_indicator_vanishes_for_a_linear_field_on_curved_hexahedra():
    """A density linear in x is trilinear in the reference coordinates of a
    trilinear element, so its nodal interpolant is exact: every element's DG
    gradient, taken with the Jacobian at each node, is the constant gradient,
    the recovered gradient equals it and the indicator vanishes."""
    from tests.helpers import curved_hex

    m = curved_hex(4, 3, 3, seed=85, periodic=(False, False, False))
    p = 2
    ref = RefSolver(m, p, S.Scheme())
    X = ref.node_coords()
    rho = 1.0 + 0.3 * X[..., 0] - 0.2 * X[..., 1] + 0.1 * X[..., 2]
    x = np.concatenate([rho.ravel()] + [np.zeros(rho.size)] * 3 + [np.full(rho.size, 2.5)])
    eta = amr.zz_indicator(m, p, x)
    assert eta.max() < 1e-13
    # the elements are not parallelepipeds: the Jacobian differs between corners
    X = m.elem_coords
    J0 = np.stack([X[:, 1] - X[:, 0], X[:, 3] - X[:, 0], X[:, 4] - X[:, 0]], axis=-1)
    J7 = np.stack([X[:, 6] - X[:, 7], X[:, 6] - X[:, 5], X[:, 6] - X[:, 2]], axis=-1)
    assert np.abs(J0 - J7).max() > 1e-3  # the elements are not parallelepipeds
