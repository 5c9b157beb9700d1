"""CPU checks of the C-ABI library: it loads, and exports every symbol that
include/rkdg_capi.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from rkdg_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "rkdg_capi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rkdg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(capi.LIB_PATH)
    names = _declared()
    assert len(names) >= 13
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.EXPORTS)


def test_struct_layouts_match_header(tmp_path):
    """ctypes descriptors have the sizes and offsets the C compiler gives the header."""
    import subprocess

    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "rkdg_capi.h"\n'
        'int main(){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(rkdg_mesh_desc), sizeof(rkdg_scheme_desc),'
        ' sizeof(rkdg_info), offsetof(rkdg_scheme_desc, bc_param), offsetof(rkdg_scheme_desc, pow_nd),'
        ' offsetof(rkdg_info, two_colour)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(capi.MeshDesc), ctypes.sizeof(capi.SchemeDesc), ctypes.sizeof(capi.Info),
            capi.SchemeDesc.bc_param.offset, capi.SchemeDesc.pow_nd.offset, capi.Info.two_colour.offset]
    assert got == want


def test_library_exports_selftest_symbols():
    txt = open(os.path.join(ROOT, "include", "rkdg_selftest.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(rkdg_selftest_[a-z_0-9]+)\s*\(", txt)))
    L = ctypes.CDLL(capi.LIB_PATH)
    assert names and set(names) == set(capi.SELFTEST_EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def _create_rc(mesh, order, scheme_desc):
    L = capi.lib()
    md, keep = capi._mesh_desc(mesh, order)
    out = ctypes.c_void_p()
    rc = L.rkdg_create(ctypes.byref(md), ctypes.byref(scheme_desc), ctypes.byref(out))
    return rc, L.rkdg_last_error().decode(), out


def test_create_rejects_bad_arguments_before_touching_the_device():
    """Every argument check of rkdg_create runs on the host before any device
    work, so it is exercised here without a GPU: unknown plug-in enums, bad gas
    constants, unknown boundary types, empty meshes, unknown geometry and
    orders without kernels all return RKDG_E_ARG / RKDG_E_UNSUPPORTED with a
    message and no context."""
    from rkdg_amd import mesh as M
    from rkdg_amd import scheme as S

    m = M.structured_hex(3, 3, 3)
    cases = [
        (dict(riemann=7), "Riemann"),
        (dict(wave_speeds=2), "wave-speed"),
        (dict(indicator=5), "indicator"),
        (dict(phi=-1), "limiter function"),
        (dict(limiter=3), "limiter"),
        (dict(gamma=1.0), "gamma"),
        (dict(covolume=-0.1), "covolume"),
        (dict(bcs=[S.BC(1, 9)]), "boundary condition type"),
        # Shu's indicator sets no multiplier; linearize is defined for p = 1
        (dict(indicator=S.IND_SHU, limiter=S.LIM_MULTI), "FinDiff"),
        (dict(indicator=S.IND_BJ, limiter=S.LIM_MULTI, linearize=1), "linearize"),
    ]
    for kw, word in cases:
        rc, msg, out = _create_rc(m, 2, capi.scheme_desc(S.Scheme(**kw)))
        assert rc == capi.RKDG_E_ARG and word in msg and not out.value, (kw, rc, msg)
    good = capi.scheme_desc(S.Scheme())
    empty = M.Mesh(M.GEOM_HEX, m.elem_coords[:0], m.elem_verts[:0])
    rc, msg, _ = _create_rc(empty, 2, good)
    assert rc == capi.RKDG_E_ARG and "empty" in msg
    rc, msg, _ = _create_rc(m, 9, good)
    assert rc == capi.RKDG_E_UNSUPPORTED and "order" in msg
    md, keep = capi._mesh_desc(m, 2)
    md.geom = 7
    out = ctypes.c_void_p()
    assert capi.lib().rkdg_create(ctypes.byref(md), ctypes.byref(good), ctypes.byref(out)) == capi.RKDG_E_ARG
    assert "geometry" in capi.lib().rkdg_last_error().decode()


def test_entry_points_reject_null_context():
    L = capi.lib()
    null = None
    assert L.rkdg_rhs(null, null, null, null) == capi.RKDG_E_ARG
    assert L.rkdg_limit(null, null, null) == capi.RKDG_E_ARG
    assert L.rkdg_means(null, null, null) == capi.RKDG_E_ARG
    assert L.rkdg_hmin(null, null) == capi.RKDG_E_ARG
    assert L.rkdg_get_info(null, null) == capi.RKDG_E_ARG
    assert L.rkdg_node_coords(null, null) == capi.RKDG_E_ARG
    assert L.rkdg_destroy(null) == 0


def test_error_codes_match_header():
    txt = open(os.path.join(ROOT, "include", "rkdg_capi.h")).read()
    codes = dict((k, int(v)) for k, v in re.findall(r"#define (RKDG_E_[A-Z]+) \((-?\d+)\)", txt))
    for name in ("RKDG_E_ARG", "RKDG_E_UNPHYSICAL", "RKDG_E_HIP", "RKDG_E_UNSUPPORTED", "RKDG_E_NAN"):
        assert getattr(capi, name) == codes[name], name
    assert len(set(codes.values())) == len(codes)


def test_buffer_validation_rejects_bad_buffers_on_the_host():
    """Every device buffer is checked before its pointer reaches the kernels:
    host tensors, wrong dtypes, strided views, wrong sizes and non-tensors are
    refused with ValueError (no GPU needed: the checks run first)."""
    import pytest
    import torch

    x = torch.zeros(10, dtype=torch.float64)
    with pytest.raises(ValueError, match="on cpu"):
        capi.check_buffer(x, 10, "x")
    with pytest.raises(ValueError, match="torch tensor"):
        capi.check_buffer(12345, 10, "x")
    with pytest.raises(ValueError, match="required"):
        capi.check_buffer(None, 10, "x")
    assert capi.check_buffer(None, 10, "x", nullable=True) is None

    class FakeCuda:
        """A stand-in with the tensor attributes the check reads."""

        def __init__(self, t):
            self.t = t
            self.is_cuda = True
            self.dtype = t.dtype

        def is_contiguous(self):
            return self.t.is_contiguous()

        def numel(self):
            return self.t.numel()

    import unittest.mock as um

    with um.patch.object(torch, "Tensor", FakeCuda):
        with pytest.raises(ValueError, match="dtype"):
            capi.check_buffer(FakeCuda(torch.zeros(10, dtype=torch.float32)), 10, "x")
        with pytest.raises(ValueError, match="contiguous"):
            capi.check_buffer(FakeCuda(torch.zeros(20, dtype=torch.float64)[::2]), 10, "x")
        with pytest.raises(ValueError, match="expected 10"):
            capi.check_buffer(FakeCuda(torch.zeros(9, dtype=torch.float64)), 10, "x")
        with pytest.raises(ValueError, match="int8"):
            capi.check_buffer(FakeCuda(torch.zeros(10, dtype=torch.float64)), 10, "troubled", dtype="int8")
        ok = FakeCuda(torch.zeros(10, dtype=torch.int8))
        assert capi.check_buffer(ok, 10, "troubled", dtype="int8") is ok


def test_tableau_validation():
    import pytest

    from rkdg_amd import scheme as S

    for name, tab in S.TABLEAUS.items():
        assert capi.check_tableau(tab) is tab, name
    with pytest.raises(ValueError, match="stages"):
        capi.check_tableau((5, [0.0] * 10, [0.2] * 5, []))
    with pytest.raises(ValueError, match="a entries"):
        capi.check_tableau((3, [1.0, 0.25], [1 / 6, 1 / 6, 2 / 3], [1.0, 0.5]))
    with pytest.raises(ValueError, match="b entries"):
        capi.check_tableau((3, [1.0, 0.25, 0.25], [0.5, 0.5], [1.0, 0.5]))
