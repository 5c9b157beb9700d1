"""Post-step output of the RKDG path: ParaView files, restart frames, the
conservation report and the time means of the primitive fields.

The reference does this in its time loop (main.cpp:388-451, 705-777) with
MFEM's ParaViewDataCollection (levels of detail 1: every element written as a
linear cell with the solution at its vertices), a restart data collection
saved every visSteps steps with a rolling window of nSavedFrames frames
(case_manager.cpp:897-985) and ComputeTotalEnergy
(src/postprocessing/check_total_energy.cpp:3-63). Here the device work goes
through the C-ABI (rkdg_interpolate, rkdg_primitives, rkdg_totals); files are
written on the host:

* :func:`write_vtu` -- an unstructured-grid VTK XML file: every element split
  into ``lod``^dim linear sub-cells (``lod`` = 1: the element itself, as the
  reference writes it) on its own points (the fields are discontinuous), with
  the conserved variables and rho, U, p, T at the points; :func:`write_pvd`
  collects a time series;
* :class:`RestartFrames` -- restart_<cycle> directories with the state, t, dt
  and the cycle, keeping the newest ``n_saved``;
* :class:`PostProcessor` -- the per-step bookkeeping of the reference loop
  (time integrals of U, p, T, rho; totals; output every vis_steps).
"""
from __future__ import annotations

import base64
import json
import os
import shutil
import struct
from typing import List, Optional, Tuple

import numpy as np

from . import mesh as M

VTK_TRIANGLE, VTK_QUAD, VTK_HEXAHEDRON = 5, 9, 12


# ---------------------------------------------------------------- reference sub-grids
def subcell_points(geom: int, lod: int) -> Tuple[np.ndarray, np.ndarray, int]:
    """Reference points of a lod-subdivision of the reference element and the
    linear sub-cells on them: (points [npts][dim], cells [ncells][nv], vtk type).
    Points are equispaced; for lod = 1 they are the element's vertices in MFEM
    order (quad (0,0),(1,0),(1,1),(0,1); hex bottom then top; triangle
    (0,0),(1,0),(0,1)), which is also VTK's vertex order."""
    n = lod
    if geom == M.GEOM_TRI:
        idx = {}
        pts = []
        for j in range(n + 1):
            for i in range(n + 1 - j):
                idx[(i, j)] = len(pts)
                pts.append((i / n, j / n))
        cells = []
        for j in range(n):
            for i in range(n - j):
                cells.append((idx[(i, j)], idx[(i + 1, j)], idx[(i, j + 1)]))
                if i + j + 1 < n:
                    cells.append((idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]))
        return np.array(pts, float), np.array(cells, np.int64), VTK_TRIANGLE
    if geom == M.GEOM_QUAD:
        def pid(i, j):
            return i + (n + 1) * j
        pts = [(i / n, j / n) for j in range(n + 1) for i in range(n + 1)]
        cells = [(pid(i, j), pid(i + 1, j), pid(i + 1, j + 1), pid(i, j + 1))
                 for j in range(n) for i in range(n)]
        return np.array(pts, float), np.array(cells, np.int64), VTK_QUAD

    def pid3(i, j, k):
        return i + (n + 1) * (j + (n + 1) * k)
    pts = [(i / n, j / n, k / n) for k in range(n + 1) for j in range(n + 1) for i in range(n + 1)]
    cells = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                cells.append((pid3(i, j, k), pid3(i + 1, j, k), pid3(i + 1, j + 1, k), pid3(i, j + 1, k),
                              pid3(i, j, k + 1), pid3(i + 1, j, k + 1), pid3(i + 1, j + 1, k + 1),
                              pid3(i, j + 1, k + 1)))
    return np.array(pts, float), np.array(cells, np.int64), VTK_HEXAHEDRON


def map_to_physical(mesh: M.Mesh, ref_pts: np.ndarray) -> np.ndarray:
    """Physical coordinates [NE][npts][dim] of reference points under each
    element's vertex map (affine triangles, bilinear quadrilaterals,
    trilinear hexahedra; MFEM vertex order)."""
    X = mesh.elem_coords
    r = np.asarray(ref_pts, float)
    if mesh.geom == M.GEOM_TRI:
        N = np.stack([1 - r[:, 0] - r[:, 1], r[:, 0], r[:, 1]], axis=1)
    elif mesh.geom == M.GEOM_QUAD:
        x, y = r[:, 0], r[:, 1]
        N = np.stack([(1 - x) * (1 - y), x * (1 - y), x * y, (1 - x) * y], axis=1)
    else:
        x, y, z = r[:, 0], r[:, 1], r[:, 2]
        b = [(1 - x) * (1 - y), x * (1 - y), x * y, (1 - x) * y]
        N = np.stack([bb * (1 - z) for bb in b] + [bb * z for bb in b], axis=1)
    return np.einsum("qv,evd->eqd", N, X)


# ---------------------------------------------------------------- VTK XML
def _data_array(name, arr, ncomp, binary):
    arr = np.ascontiguousarray(arr)
    vtype = {np.dtype(np.float64): "Float64", np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}[arr.dtype]
    head = f'<DataArray type="{vtype}" Name="{name}" NumberOfComponents="{ncomp}" format="{"binary" if binary else "ascii"}">'
    if binary:
        raw = arr.tobytes()
        body = base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")
    else:
        body = " ".join(repr(float(v)) if arr.dtype == np.float64 else str(int(v)) for v in arr.ravel())
    return head + body + "</DataArray>\n"


def write_vtu(path: str, points: np.ndarray, cells: np.ndarray, cell_type: int, point_data: dict,
              binary: bool = True, time: Optional[float] = None, cycle: Optional[int] = None) -> str:
    """Writes an unstructured-grid VTK XML file (points [n][dim], cells
    [ncells][nv] point ids, one VTK cell type, point fields name -> [n] or
    [n][ncomp]); 2D points are padded with z = 0. Returns the path."""
    pts = np.asarray(points, float)
    if pts.shape[1] == 2:
        pts = np.concatenate([pts, np.zeros((len(pts), 1))], axis=1)
    cells = np.asarray(cells, np.int64)
    nv = cells.shape[1]
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt32">\n',
           "<UnstructuredGrid>\n"]
    if time is not None or cycle is not None:
        out.append("<FieldData>\n")
        if time is not None:
            out.append(_data_array("TIME", np.array([time], float), 1, binary).replace(
                'NumberOfComponents="1"', 'NumberOfTuples="1"'))
        if cycle is not None:
            out.append(_data_array("CYCLE", np.array([cycle], np.int64), 1, binary).replace(
                'NumberOfComponents="1"', 'NumberOfTuples="1"'))
        out.append("</FieldData>\n")
    out.append(f'<Piece NumberOfPoints="{len(pts)}" NumberOfCells="{len(cells)}">\n')
    out.append("<PointData>\n")
    for name, val in point_data.items():
        val = np.asarray(val, float)
        ncomp = 1 if val.ndim == 1 else val.shape[1]
        if ncomp == 2:  # vectors as 3 components for ParaView
            val = np.concatenate([val, np.zeros((len(val), 1))], axis=1)
            ncomp = 3
        out.append(_data_array(name, val, ncomp, binary))
    out.append("</PointData>\n<Points>\n")
    out.append(_data_array("Points", pts, 3, binary))
    out.append("</Points>\n<Cells>\n")
    out.append(_data_array("connectivity", cells.ravel(), 1, binary))
    out.append(_data_array("offsets", np.arange(1, len(cells) + 1, dtype=np.int64) * nv, 1, binary))
    out.append(_data_array("types", np.full(len(cells), cell_type, np.uint8), 1, binary))
    out.append("</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as f:
        f.write("".join(out))
    return path


def read_vtu(path: str) -> dict:
    """Reads back what write_vtu writes (points, connectivity, offsets, types
    and every point field) -- for checks of the output."""
    import xml.etree.ElementTree as ET

    root = ET.parse(path).getroot()
    res = {}
    for da in root.iter("DataArray"):
        dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[da.get("type")]
        txt = (da.text or "").strip()
        if da.get("format") == "binary":
            raw = base64.b64decode(txt)
            n = struct.unpack("<I", raw[:4])[0]
            arr = np.frombuffer(raw[4:4 + n], dtype=dt)
        else:
            arr = np.array([dt(float(v)) if dt is np.float64 else dt(int(v)) for v in txt.split()], dtype=dt)
        nc = int(da.get("NumberOfComponents") or 1)
        res[da.get("Name")] = arr.reshape(-1, nc) if nc > 1 else arr
    return res


def write_pvd(path: str, entries: List[Tuple[float, str]]) -> str:
    """A ParaView collection of (time, vtu file) entries."""
    lines = ['<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n',
             "<Collection>\n"]
    for t, f in entries:
        lines.append(f'<DataSet timestep="{t!r}" group="" part="0" file="{os.path.relpath(f, os.path.dirname(path) or ".")}"/>\n')
    lines.append("</Collection>\n</VTKFile>\n")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


def primitive_fields(U: np.ndarray, dim: int, gamma: float, R: float = 1.0, b: float = 0.0) -> dict:
    """rho, U, p, T from conserved values U [neq][n] (physics.cpp:101-126)."""
    rho = U[0]
    mom = U[1:1 + dim]
    E = U[1 + dim]
    ei = E - 0.5 * (mom * mom).sum(axis=0) / rho
    return {"rho": rho, "U": (mom / rho).T, "p": (gamma - 1.0) * ei / (1.0 - rho * b),
            "T": (gamma - 1.0) * ei / rho / R}


def write_solution(path: str, ctx, x, lod: int = 1, time=None, cycle=None, binary: bool = True,
                   extra: Optional[dict] = None) -> str:
    """The state x (device) of context ctx as a VTU file: every element split
    into lod^dim linear cells on its own points, with the conserved variables
    (rho, mom, energy, as the reference registers them) and rho, U, p, T
    evaluated there (rkdg_interpolate). ``extra``: more point fields."""
    import torch

    mesh = ctx.mesh
    info = ctx.info
    ref, sub, vtype = subcell_points(mesh.geom, lod)
    npts = len(ref)
    out = torch.empty(info.num_equations * info.num_elements * npts, dtype=torch.float64, device=x.device)
    ctx.interpolate(x, ref, out)
    U = out.cpu().numpy().reshape(info.num_equations, info.num_elements * npts)
    X = map_to_physical(mesh, ref).reshape(-1, mesh.dim)
    cells = (sub[None, :, :] + npts * np.arange(info.num_elements)[:, None, None]).reshape(-1, sub.shape[1])
    sch = ctx.scheme
    fields = {"rho": U[0], "mom": U[1:1 + mesh.dim].T, "energy": U[1 + mesh.dim]}
    prim = primitive_fields(U, mesh.dim, sch.gamma, sch.gas_constant, sch.covolume)
    fields.update({"U": prim["U"], "p": prim["p"], "T": prim["T"]})
    if extra:
        fields.update(extra)
    return write_vtu(path, X, cells, vtype, fields, binary=binary, time=time, cycle=cycle)


# ---------------------------------------------------------------- restart frames
class RestartFrames:
    """Restart frames in case_dir/restart_<cycle, 6 digits>/ holding the state
    (state.npy) and meta.json (cycle, t, dt); at most n_saved frames are kept,
    the oldest removed first (case_manager.cpp:897-936, 975-985). Opening with
    ``restart_cycle`` removes frames newer than it, as a restarted run does."""

    def __init__(self, case_dir: str, n_saved: int = 2, restart_cycle: Optional[int] = None):
        if n_saved < 1:
            raise ValueError("n_saved must be at least 1")
        self.case_dir = case_dir
        self.n_saved = n_saved
        os.makedirs(case_dir, exist_ok=True)
        cur = f"restart_{restart_cycle:06d}" if restart_cycle is not None else None
        for name in self._frames():
            if cur is not None and name > cur:
                shutil.rmtree(os.path.join(case_dir, name))

    def _frames(self):
        return sorted(d for d in os.listdir(self.case_dir)
                      if d.startswith("restart_") and os.path.isdir(os.path.join(self.case_dir, d)))

    def save(self, cycle: int, t: float, dt: float, x) -> str:
        name = f"restart_{cycle:06d}"
        d = os.path.join(self.case_dir, name)
        os.makedirs(d, exist_ok=True)
        arr = x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)
        tmp = os.path.join(d, "state.npy.tmp")
        with open(tmp, "wb") as f:
            np.save(f, np.ascontiguousarray(arr, np.float64))
        os.replace(tmp, os.path.join(d, "state.npy"))
        with open(os.path.join(d, "meta.json"), "w") as f:
            json.dump({"cycle": int(cycle), "t": float(t), "dt": float(dt), "size": int(arr.size)}, f)
        frames = self._frames()
        while len(frames) > self.n_saved:
            shutil.rmtree(os.path.join(self.case_dir, frames.pop(0)))
        return d

    def frames(self):
        return [int(n.split("_")[1]) for n in self._frames()]

    def load(self, cycle: Optional[int] = None):
        """(cycle, t, dt, state) of the given frame, or of the newest."""
        frames = self._frames()
        if not frames:
            raise FileNotFoundError(f"no restart frame in {self.case_dir}")
        name = f"restart_{cycle:06d}" if cycle is not None else frames[-1]
        d = os.path.join(self.case_dir, name)
        with open(os.path.join(d, "meta.json")) as f:
            meta = json.load(f)
        x = np.load(os.path.join(d, "state.npy"), allow_pickle=False)
        return meta["cycle"], meta["t"], meta["dt"], x


# ---------------------------------------------------------------- the loop's bookkeeping
class PostProcessor:
    """What the reference loop does after each step (main.cpp:700-777): the
    time integrals of U, p, T and rho (weight dt; written divided by t), the
    total of every conserved variable (the E_total of the log line, when
    check_totals), and every vis_steps steps (and at the end) a ParaView file
    and a restart frame. Use as the ``callback`` of solver.run (it is called
    with (steps, t, dt, x))."""

    def __init__(self, ctx, out_dir: str, vis_steps: int = 100, n_saved: int = 2, lod: int = 1,
                 check_totals: bool = False, paraview: bool = True):
        import torch

        self.ctx = ctx
        self.out_dir = out_dir
        self.vis_steps = vis_steps
        self.lod = lod
        self.check_totals = check_totals
        self.paraview = paraview
        os.makedirs(out_dir, exist_ok=True)
        self.restart = RestartFrames(out_dir, n_saved)
        n = (ctx.info.dim + 3) * ctx.info.num_dofs
        self.acc = torch.zeros(n, dtype=torch.float64, device=torch.device("cuda", torch.cuda.current_device()))
        self.series: List[Tuple[float, str]] = []
        self.log: List[dict] = []

    def __call__(self, steps: int, t: float, dt: float, x, done: bool = False):
        self.ctx.primitives(x, acc=self.acc, weight=dt)
        rec = {"step": steps, "t": t, "dt": dt}
        if self.check_totals:
            rec["totals"] = self.ctx.totals(x).tolist()
        self.log.append(rec)
        if done or steps % self.vis_steps == 0:
            self.write(steps, t, dt, x)

    def means(self, t: float):
        """The time means (rho, U, p, T) so far, [(dim+3)][num_dofs] on the host."""
        return (self.acc / t).cpu().numpy().reshape(self.ctx.info.dim + 3, -1)

    def write(self, steps: int, t: float, dt: float, x):
        if self.paraview:
            f = os.path.join(self.out_dir, f"cycle_{steps:06d}.vtu")
            write_solution(f, self.ctx, x, lod=self.lod, time=t, cycle=steps)
            self.series.append((t, f))
            write_pvd(os.path.join(self.out_dir, "solution.pvd"), self.series)
        self.restart.save(steps, t, dt, x)
