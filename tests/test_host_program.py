"""The C host program (zf_host) and its reference-format I/O.

CPU: the mesh, cell and history files round-trip through the host library
in the reference's layouts (column counts and headers of ZebraflowWT.c:346,
:362, :435-446); the lattice-unit membrane constants equal the oracle
helper's.  GPU: zf_host run on a case written in those formats produces, in
FLUIDdomain-%09d.csv and AllCellMemb-%09d.csv, the state the engine reaches
through the C ABI for the same case (to the files' printed precision).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from zebraflow_amd import geometry
from zebraflow_amd import PKG, load_host_library
from zebraflow_amd.engine import (B_PRESSURE_OUTLET, B_VELOCITY_INLET, STORE_F64, Engine,
                                  lattice_config, lattice_ibm_config, lattice_membrane_config,
                                  lattice_repulsion_config)


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def write_case(d, b, cells, xyz, tris, steps, every, extra=""):
    h = load_host_library()
    im, jm, km = b.shape
    bpath = os.path.join(d, "FLUIDdomainBindex.csv")
    assert h.zfh_write_bindex_csv(bpath.encode(), p(np.ascontiguousarray(b)), im, jm, km, 0.0, 0.0,
                                  0.0, 1.0) == 0
    assert h.zfh_write_points(os.path.join(d, "ZFISHRBCNODES.dat").encode(), p(xyz), len(xyz)) == 0
    assert h.zfh_write_rbc_triangles(os.path.join(d, "ZFISHRBCTRIANGLES.dat").encode(),
                                     p(np.ascontiguousarray(tris, np.int32)), len(tris)) == 0
    c = np.ascontiguousarray(cells, np.float64)
    st = np.ones(len(cells), np.int32)
    assert h.zfh_write_cells_csv(os.path.join(d, "AllCellMembStart.csv").encode(), len(cells),
                                 len(xyz), p(c), None, p(st), None, None, None) == 0
    cfg = f"""# tube with a pulsatile inlet and cells
box = {im} {jm} {km}
grid = 0 0 0 1
bindex = {bpath}
units = lattice
tau = 0.8
storage = f64
inlet = 0.02 0.3 40 1 0 0
rbc_nodes = {os.path.join(d, "ZFISHRBCNODES.dat")}
rbc_triangles = {os.path.join(d, "ZFISHRBCTRIANGLES.dat")}
cells = {os.path.join(d, "AllCellMembStart.csv")}
max_cells = {len(cells)}
rbc_interval = 2
wall = surface
steps = {steps}
output_every = {every}
output_dir = {os.path.join(d, "out")}
{extra}
"""
    path = os.path.join(d, "case.cfg")
    open(path, "w").write(cfg)
    return path


def test_mesh_and_cell_files_round_trip(tmp_path):
    h = load_host_library()
    xyz, tris = geometry.rbc_mesh(radius=3.0, level=2)
    f = str(tmp_path / "n.dat").encode()
    assert h.zfh_write_points(f, p(xyz), len(xyz)) == 0
    assert h.zfh_read_points(f, None, 0) == len(xyz)
    back = np.zeros_like(xyz)
    assert h.zfh_read_points(f, p(back), len(xyz)) == len(xyz)
    np.testing.assert_allclose(back, xyz, rtol=1e-14)
    assert h.zfh_read_points(f, p(back), 3) == -2
    t = np.ascontiguousarray(tris, np.int32)
    f = str(tmp_path / "t.dat").encode()
    assert h.zfh_write_rbc_triangles(f, p(t), len(t)) == 0
    tb = np.zeros_like(t)
    assert h.zfh_read_rbc_triangles(f, p(tb), len(t)) == len(t)
    np.testing.assert_array_equal(tb, t)
    # wall mesh files: 0-based polygons on disk, 1-based in memory
    b = geometry.tube((16, 12, 12), axis=0, r=4.0, periodic=True)
    wx, wt, wn = geometry.wall_mesh(b)
    names = [str(tmp_path / n).encode() for n in ("wn.dat", "wc.dat", "wnorm.dat")]
    assert h.zfh_write_wall_mesh(*names, p(wx), len(wx), p(np.ascontiguousarray(wt)), p(wn), len(wt)) == 0
    assert open(names[1], "rb").read().startswith(b"POLYGONS")
    wt2 = np.zeros_like(wt)
    wn2 = np.zeros_like(wn)
    assert h.zfh_read_wall_cells(names[1], p(wt2), len(wt)) == len(wt)
    assert h.zfh_read_wall_normals(names[2], p(wn2), len(wt)) == len(wt)
    np.testing.assert_array_equal(wt2, wt)
    np.testing.assert_allclose(wn2, wn, atol=1e-8)
    # cells: the 21-column AllCellMemb layout
    cells = np.stack([xyz + 5.0, xyz + 9.0])
    st = np.array([1, 1], np.int32)
    vel = np.full_like(cells, 0.25)
    f = str(tmp_path / "cells.csv").encode()
    assert h.zfh_write_cells_csv(f, 2, len(xyz), p(cells), p(vel), p(st), None, None, None) == 0
    lines = open(f).read().splitlines()
    assert lines[0].split(",")[:5] == ["X", "Y", "Z", "Cell", "CellStatus"]
    assert len(lines[0].split(",")) == 21 and all(len(l.split(",")) == 21 for l in lines[1:])
    assert len(lines) == 1 + 2 * len(xyz)
    back = np.zeros_like(cells)
    st2 = np.zeros(2, np.int32)
    assert h.zfh_read_cells_csv(f, len(xyz), p(back), p(st2), 2) == 2
    np.testing.assert_allclose(back, cells, atol=1e-6)  # printed with %f
    assert h.zfh_read_cells_csv(f, len(xyz), p(back), p(st2), 1) == -2
    # the volume/area history appends, header once
    f = str(tmp_path / "VolumeAreaHistory.dat").encode()
    vol, area, cen = np.array([1.0, 2.0]), np.array([3.0, 4.0]), np.zeros((2, 3))
    for it in (100, 200):
        assert h.zfh_append_volume_area(f, it, 2, p(vol), p(area), p(cen), p(st)) == 0
    lines = open(f).read().splitlines()
    assert lines[0] == "Iteration#,Cell,Vol,Area,CenX,CenY,CenZ,CellState,RegressFlag"
    assert len(lines) == 5 and lines[3].startswith("200,1,1.0000000e+00,3.0000000e+00")


def test_lattice_membrane_constants_match_oracle_helper():
    from tests.test_oracle_mesh import lattice_params
    xyz, tris = geometry.rbc_mesh(radius=3.0)
    mp, _, _, _ = lattice_params(xyz, tris)
    mc = lattice_membrane_config(xyz, tris, 7)
    for k in ("kT", "kbend", "kd", "ka", "kv", "area0", "vol0", "force_limit", "sliver_f1",
              "sliver_f2", "x0", "x_max"):
        np.testing.assert_allclose(getattr(mc, k), mp[k], rtol=1e-12, err_msg=k)
    assert mc.max_cells == 7 and mc.len_scale == 1.0


@pytest.mark.gpu
def test_host_program_matches_engine(tmp_path):
    print("case zf_host tube with 2 cells, 12 iterations", flush=True)
    b = geometry.tube((30, 20, 20), axis=0, r=8.0, inlet=B_VELOCITY_INLET, outlet=B_PRESSURE_OUTLET)
    xyz, tris = geometry.rbc_mesh(radius=3.0, level=2)
    cells = geometry.place_cells(b, xyz, 3.0, 1.0, 2, 5)
    assert len(cells) == 2
    steps, every = 12, 6
    cfg = write_case(str(tmp_path), b, cells, xyz, tris, steps, every)
    exe = os.path.join(PKG, "zf_host")
    r = subprocess.run([exe, cfg], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "out"
    assert (out / f"FLUIDdomain-{6:09d}.csv").exists()
    fl = np.genfromtxt(out / f"FLUIDdomain-{steps:09d}.csv", delimiter=",", skip_header=1)
    cm = np.genfromtxt(out / f"AllCellMemb-{steps:09d}.csv", delimiter=",", skip_header=1)
    hist = open(out / "VolumeAreaHistory.dat").read().splitlines()
    assert len(hist) == 1 + 2 * (steps // every)
    # the same case through the C ABI
    cfgz = lattice_config(*b.shape, 0.8, storage=STORE_F64, inlet_u0=0.02, inlet_amp=0.3,
                          inlet_period=40.0, inlet_dir=(1, 0, 0))
    eng = Engine(cfgz)
    eng.upload_lattice(b)
    eng.membrane_setup(lattice_membrane_config(xyz, tris, 2), tris, xyz)
    eng.upload_cells(np.round(cells, 6))  # the start file holds %f coordinates
    eng.ibm_setup(lattice_ibm_config(rbc_interval=2))
    eng.repulsion_setup(lattice_repulsion_config())
    eng.wall_setup(*geometry.wall_mesh(b))
    for it in range(1, steps + 1):
        eng.full_step(it)
    d, u, v, w = eng.macro()
    np.testing.assert_allclose(fl[:, 4], u, rtol=2e-7, atol=1e-12)
    np.testing.assert_allclose(fl[:, 5], v, rtol=2e-7, atol=1e-12)
    np.testing.assert_allclose(fl[:, 7], d, rtol=2e-7)
    f = eng.force()
    np.testing.assert_allclose(fl[:, 8], f[0], rtol=2e-7, atol=1e-14)
    x, _, st = eng.cells()
    np.testing.assert_allclose(cm[:, :3], x.reshape(-1, 3), atol=2e-6)
    assert (cm[:, 4] == 1).all()
    assert np.abs(u).max() > 1e-3
    eng.close()
