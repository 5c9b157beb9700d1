"""Host side of the post-step output (rkdg_amd/io.py): sub-cell grids, the
element maps, VTK XML round trips, the ParaView collection and the rolling
window of restart frames."""
import json
import os

import numpy as np
import pytest

from rkdg_amd import io
from rkdg_amd import mesh as M


@pytest.mark.parametrize("geom", [M.GEOM_TRI, M.GEOM_QUAD, M.GEOM_HEX])
@pytest.mark.parametrize("lod", [1, 2, 3])
def test_subcells_tile_the_reference_element(geom, lod):
    pts, cells, vtype = io.subcell_points(geom, lod)
    dim = pts.shape[1]
    if geom == M.GEOM_TRI:
        assert len(pts) == (lod + 1) * (lod + 2) // 2 and len(cells) == lod * lod and vtype == 5
        a = pts[cells]
        area = 0.5 * np.abs((a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1])
                            - (a[:, 1, 1] - a[:, 0, 1]) * (a[:, 2, 0] - a[:, 0, 0]))
        assert abs(area.sum() - 0.5) < 1e-14 and np.all(area > 0)
    else:
        assert len(pts) == (lod + 1) ** dim and len(cells) == lod ** dim
        # every sub-cell spans a box of side 1/lod from its first vertex
        c = pts[cells]
        side = c.max(axis=1) - c.min(axis=1)
        assert np.allclose(side, 1.0 / lod)
    if lod == 1:  # the element's own vertices in MFEM (= VTK) order
        ref = {M.GEOM_TRI: [[0, 0], [1, 0], [0, 1]], M.GEOM_QUAD: [[0, 0], [1, 0], [1, 1], [0, 1]],
               M.GEOM_HEX: [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]}
        np.testing.assert_array_equal(pts[cells[0]], np.array(ref[geom], float))


def test_element_maps_reproduce_vertices_and_centres():
    for m in (M.structured_quad(3, 2, warp=lambda X, Y: (X + 0.1 * Y * Y, Y)),
              M.structured_hex(2, 2, 2, periodic=(False, False, False)),
              M.jittered_delaunay_tri(3, 3, seed=1)):
        ref, cells, _ = io.subcell_points(m.geom, 1)
        X = io.map_to_physical(m, ref)
        np.testing.assert_allclose(X[:, cells[0]], m.elem_coords, atol=1e-15)
        cen = np.full((1, m.dim), 1.0 / 3 if m.geom == M.GEOM_TRI else 0.5)
        np.testing.assert_allclose(io.map_to_physical(m, cen)[:, 0], m.elem_coords.mean(axis=1), atol=1e-15)


@pytest.mark.parametrize("binary", [True, False])
def test_vtu_round_trip(tmp_path, binary):
    m = M.structured_quad(3, 2)
    ref, sub, vtype = io.subcell_points(m.geom, 2)
    X = io.map_to_physical(m, ref).reshape(-1, 2)
    cells = (sub[None] + len(ref) * np.arange(m.num_elements)[:, None, None]).reshape(-1, 4)
    rho = np.sin(X[:, 0]) + 2.0
    vel = np.stack([X[:, 1], -X[:, 0]], axis=1)
    f = io.write_vtu(str(tmp_path / "a.vtu"), X, cells, vtype, {"rho": rho, "U": vel}, binary=binary,
                     time=0.25, cycle=7)
    r = io.read_vtu(f)
    np.testing.assert_array_equal(r["Points"][:, :2], X)
    assert np.all(r["Points"][:, 2] == 0.0)
    np.testing.assert_array_equal(r["connectivity"], cells.ravel())
    np.testing.assert_array_equal(r["offsets"], 4 * np.arange(1, len(cells) + 1))
    assert np.all(r["types"] == vtype)
    np.testing.assert_array_equal(r["rho"], rho)
    np.testing.assert_array_equal(r["U"][:, :2], vel)
    assert r["TIME"][0] == 0.25 and r["CYCLE"][0] == 7
    pvd = io.write_pvd(str(tmp_path / "s.pvd"), [(0.25, f)])
    txt = open(pvd).read()
    assert 'timestep="0.25"' in txt and 'file="a.vtu"' in txt


def test_primitive_fields_closed_form():
    g, R, b = 1.4, 0.7, 0.05
    rho, u, v, p = 1.3, 0.2, -0.4, 2.1
    E = p / (g - 1) * (1 - rho * b) + 0.5 * rho * (u * u + v * v)
    f = io.primitive_fields(np.array([[rho], [rho * u], [rho * v], [E]]), 2, g, R, b)
    assert abs(f["p"][0] - p) < 1e-14 and np.allclose(f["U"][0], [u, v])
    # the reference's temperature leaves the covolume out (physics.cpp:115-126)
    assert abs(f["T"][0] - p * (1 - rho * b) / rho / R) < 1e-14


def test_restart_frames_rolling_window_and_restart(tmp_path):
    d = str(tmp_path / "case")
    rf = io.RestartFrames(d, n_saved=2)
    for cyc in (10, 20, 30):
        rf.save(cyc, 0.1 * cyc, 1e-3 * cyc, np.arange(5.0) * cyc)
    assert rf.frames() == [20, 30]
    cyc, t, dt, x = rf.load()
    assert cyc == 30 and t == pytest.approx(3.0) and dt == pytest.approx(0.03)
    np.testing.assert_array_equal(x, np.arange(5.0) * 30)
    meta = json.load(open(os.path.join(d, "restart_000020", "meta.json")))
    assert meta["cycle"] == 20 and meta["size"] == 5
    # restarting from cycle 20 drops the newer frame (initializeRestartQueue)
    rf2 = io.RestartFrames(d, n_saved=2, restart_cycle=20)
    assert rf2.frames() == [20]
    with pytest.raises(FileNotFoundError):
        io.RestartFrames(str(tmp_path / "empty")).load()
