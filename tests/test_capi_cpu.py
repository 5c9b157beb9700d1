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
    return sorted(set(re.findall(r"\b(rkdg_[a-z_]+)\s*\(", txt)))


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
        ' offsetof(rkdg_info, affine)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(capi.MeshDesc), ctypes.sizeof(capi.SchemeDesc), ctypes.sizeof(capi.Info),
            capi.SchemeDesc.bc_param.offset, capi.SchemeDesc.pow_nd.offset, capi.Info.affine.offset]
    assert got == want


def test_library_exports_selftest_symbols():
    txt = open(os.path.join(ROOT, "include", "rkdg_selftest.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(rkdg_selftest_[a-z_0-9]+)\s*\(", txt)))
    L = ctypes.CDLL(capi.LIB_PATH)
    assert names and set(names) == set(capi.SELFTEST_EXPORTS)
    for n in names:
        assert hasattr(L, n), n
