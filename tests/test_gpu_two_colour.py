"""The two-colour (solver/receiver) form of the tensor right-hand side, which
solves every face once (RKDG_CREATE_TWO_COLOUR, on periodic meshes of even
extents). It is checked against the oracle (RHS, wave speed, face-only wave
speed, limited steps) and against the one-kernel form on the same mesh: the
two forms evaluate the same fluxes (the
canonical Elem1/Elem2 argument order) and the receiver's share is the
solver's lift negated, mapped into the receiver's face frame; they agree to
rounding (the separately compiled kernels may contract the accumulation's
multiply-adds differently), and the wave speeds are equal.
Meshes with an odd periodic extent, or with boundary faces, keep the
one-kernel form."""
import numpy as np
import pytest

from oracle.refcpu import RefSolver
from rkdg_amd import mesh as M
from rkdg_amd import scheme as S
from tests.helpers import rel_linf, scramble_orientations, smooth_state

TOL = 1e-12

pytestmark = pytest.mark.gpu

CASES = {
    "hex_q2_4x4x4": (lambda: M.structured_hex(4, 4, 4), 2, True),
    "hex_q2_6x4x4_skew": (lambda: _skew(M.structured_hex(6, 4, 4)), 2, True),
    "hex_q1_4x4x6": (lambda: M.structured_hex(4, 4, 6), 1, True),
    "hex_q3_4x4x4": (lambda: M.structured_hex(4, 4, 4), 3, True),
    "quad_q1_8x6": (lambda: M.structured_quad(8, 6, periodic=(True, True)), 1, True),
    "quad_q2_6x4": (lambda: M.structured_quad(6, 4, periodic=(True, True)), 2, True),
    "quad_q3_4x4": (lambda: M.structured_quad(4, 4, periodic=(True, True)), 3, True),
    "hex_q2_scrambled": (lambda: scramble_orientations(M.structured_hex(4, 4, 4), seed=5), 2, False),
    "quad_q2_scrambled": (lambda: scramble_orientations(M.structured_quad(6, 6, periodic=(True, True)), seed=6),
                          2, False),
}


def _skew(m):
    A = np.array([[1.0, 0.2, 0.1], [0.0, 0.9, -0.15], [0.05, 0.0, 1.1]])
    m.elem_coords = np.ascontiguousarray(m.elem_coords @ A.T)
    return m


def _dev(x, cuda):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x)).to(cuda)


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("rs", [S.RUSANOV, S.HLLC])
def test_two_colour_rhs_and_steps(name, rs, cuda):
    import torch

    from rkdg_amd.capi import Context

    mk, p, _ = CASES[name]
    m = mk()
    sch = S.Scheme(riemann=rs, indicator=S.IND_BJ, phi=S.PHI_VENKAT, venkat_k=5.0, limiter=S.LIM_MULTI)
    ref = RefSolver(m, p, sch)
    two = Context(m, p, sch, two_colour=True)
    one = Context(m, p, sch)
    assert two.info.two_colour == 1 and one.info.two_colour == 0
    neq = m.dim + 2
    x = smooth_state(ref.node_coords(), seed=11)
    xd = _dev(x, cuda)
    y2, y1 = torch.empty_like(xd), torch.empty_like(xd)
    lam2 = two.rhs(xd, y2)
    lam1 = one.rhs(xd, y1)
    y_ref, lam_ref = ref.mult(x)
    e2 = rel_linf(_host(y2), y_ref, neq)
    d21 = rel_linf(_host(y2), _host(y1), neq)
    print(f"[two-colour {name} rs={rs}] RHS vs oracle {e2:.2e}, vs one kernel {d21:.2e}", flush=True)
    assert e2 <= TOL and abs(lam2 - lam_ref) <= 1e-13 * lam_ref and lam2 == lam1
    assert d21 <= 1e-13
    _, lf_ref = ref.face_mult(x)
    assert abs(two.face_lambda(xd) - lf_ref) <= 1e-13 * lf_ref
    tab = S.TABLEAUS["ssprk3"]
    xa, xb = xd.clone(), xd.clone()
    for _ in range(3):
        x, l_ref = ref.step(x, 2e-3, tab)
        la = two.step(xa, 2e-3, tab)
        lb = one.step(xb, 2e-3, tab)
    e = rel_linf(_host(xa), x, neq)
    print(f"[two-colour {name} rs={rs}] 3 steps vs oracle {e:.2e}", flush=True)
    assert e <= TOL and abs(la - l_ref) <= 1e-12 * l_ref
    assert rel_linf(_host(xa), _host(xb), neq) <= 1e-13 and abs(la - lb) <= 1e-14 * lb
    two.close()
    one.close()


def test_odd_extent_and_boundaries_keep_one_kernel(cuda):
    from rkdg_amd.capi import Context

    assert Context(M.structured_hex(5, 4, 4), 2, S.Scheme(), two_colour=True).info.two_colour == 0
    assert Context(M.structured_hex(4, 4, 4, periodic=(True, True, False)), 2,
                   S.Scheme(bcs=[S.BC(1, S.BC_SLIP), S.BC(6, S.BC_SLIP)]), two_colour=True).info.two_colour == 0
    assert Context(M.structured_hex(4, 4, 4), 2, S.Scheme(), two_colour=True).info.two_colour == 1
    assert Context(M.structured_hex(4, 4, 4), 2, S.Scheme()).info.two_colour == 0
