"""Copies the Gmsh meshes the reference ships with its cases into
tests/golden/meshes/ (the large jet mesh gzip-compressed). These are data
files, the reference's only fixtures; the tests read the copies.

    python tests/golden/make_mesh_fixtures.py [/path/to/reference]
"""
import gzip
import os
import shutil
import sys

SRC = {
    "sod2D_sod2d.msh": "cases/sod2D/sod2d.msh",
    "sod1d_covolume_sod2d.msh": "cases/sod1d_covolume/sod2d.msh",
    "sod1d_quad.msh": "cases/sod1d_simple/sod1d_quad.msh",
    "sod1d_quad6.msh": "cases/sod1d_simple/sod1d_quad6.msh",
    "sod1d_trian.msh": "cases/sod1d_simple/sod1d_trian.msh",
    "jet_overexpanded0p25.msh.gz": "cases/overexpandedJet0p25/jet.msh",
}


def main(ref="/root/reference"):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "meshes")
    os.makedirs(out, exist_ok=True)
    for dst, src in SRC.items():
        s = os.path.join(ref, src)
        d = os.path.join(out, dst)
        if dst.endswith(".gz"):
            with open(s, "rb") as fi, gzip.GzipFile(d, "wb", compresslevel=9, mtime=0) as fo:
                shutil.copyfileobj(fi, fo)
        else:
            shutil.copyfile(s, d)
        print(dst, os.path.getsize(d))


if __name__ == "__main__":
    main(*sys.argv[1:])
