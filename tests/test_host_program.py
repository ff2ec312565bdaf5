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


def write_nucleus_files(d, rest, edge, nuclei):
    """NucleusNodes.dat, NucleusEdges.dat and AllCellNucStart.csv as the
    reference reads them (ZebraflowWT.c:1903-1922, :5541-5593)."""
    h = load_host_library()
    assert h.zfh_write_points(os.path.join(d, "NucleusNodes.dat").encode(), p(rest), len(rest)) == 0
    with open(os.path.join(d, "NucleusEdges.dat"), "w") as f:
        f.write(f"LINES {len(edge)} {3 * len(edge)}\n")
        f.writelines(f"{a} {b}\n" for a, b in edge)
    nuc = np.ascontiguousarray(nuclei, np.float64)
    assert h.zfh_write_nuclei_csv(os.path.join(d, "AllCellNucStart.csv").encode(), nuc.shape[0],
                                  nuc.shape[1], p(nuc), None, None, None) == 0
    return "\n".join(f"{k} = {os.path.join(d, f)}" for k, f in (
        ("nucleus_nodes", "NucleusNodes.dat"), ("nucleus_edges", "NucleusEdges.dat"),
        ("nuclei", "AllCellNucStart.csv")))


def write_case(d, b, cells, xyz, tris, steps, every, extra="", out="out"):
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
output_dir = {os.path.join(d, out)}
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
    # nucleus files: NucleusEdges "LINES n 3n" pairs, AllCellNucleus layout
    f = str(tmp_path / "NucleusEdges.dat").encode()
    ned = np.ascontiguousarray([[1, 2], [2, 3], [3, 1], [1, 4]], np.int32)
    open(f, "w").write(f"LINES {len(ned)} {3 * len(ned)}\n" + "".join(f"{a} {b}\n" for a, b in ned))
    assert h.zfh_read_lines(f, None, 0) == 4
    e2 = np.zeros_like(ned)
    assert h.zfh_read_lines(f, p(e2), 4) == 4 and h.zfh_read_lines(f, p(e2), 3) == -2
    np.testing.assert_array_equal(e2, ned)
    nuc = np.stack([xyz[:5] * 0.5 + 5.0, xyz[:5] * 0.5 + 9.0])
    fn = np.full_like(nuc, -1.5e-6)
    st3 = np.array([1, -5000, 2], np.int32)  # the dead middle cell is not written
    nuc3 = np.stack([nuc[0], nuc[0] * 0, nuc[1]])
    f = str(tmp_path / "AllCellNucleus.csv").encode()
    assert h.zfh_write_nuclei_csv(f, 3, 5, p(nuc3), None, p(np.stack([fn[0], fn[0], fn[1]])),
                                  p(st3)) == 0
    lines = open(f).read().splitlines()
    assert lines[0] == "X,Y,Z,Cell,CellStatus,Vx,Vy,Vz,Fx,Fy,Fz,Neigh,Neighflag"
    assert len(lines) == 11 and all(len(l.split(",")) == 13 for l in lines[1:])
    assert lines[6].split(",")[3:5] == ["3", "2"] and lines[6].split(",")[8] == "-1.5000000e-06"
    back = np.zeros_like(nuc)
    assert h.zfh_read_nuclei_csv(f, 5, p(back), 2) == 2
    np.testing.assert_allclose(back, nuc, atol=1e-6)
    assert h.zfh_read_nuclei_csv(f, 5, p(back), 1) == -2
    assert h.zfh_read_nuclei_csv(f, 4, p(back), 3) == -1  # 10 rows are not whole cells of 4
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


@pytest.mark.gpu
def test_host_program_with_nuclei_matches_engine(tmp_path):
    """zf_host reading the nucleus network and the start nuclei in the
    reference's files: its AllCellNucleus output equals the engine's
    nuclei after the same iterations, and a restart file written midway
    continues to the same state."""
    print("case zf_host with nucleated cells, 2 cells, 8 iterations", flush=True)
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mesh_zfish.npz"))
    b = geometry.tube((30, 20, 20), axis=0, r=8.0, inlet=B_VELOCITY_INLET, outlet=B_PRESSURE_OUTLET)
    xyz, tris = geometry.rbc_mesh(radius=3.0, level=2)
    cells = geometry.place_cells(b, xyz, 3.0, 1.0, 2, 5)
    rest = 1.4 * g["nuc_nodes"]
    rng = np.random.default_rng(3)
    nuclei = np.stack([1.05 * rest + 0.02 * rng.standard_normal(rest.shape) + c.mean(0)
                       + [0.0, 0.0, 0.5] for c in cells])
    nuclei = np.round(nuclei, 6)  # as printed with %f
    steps, every = 8, 4
    extra = write_nucleus_files(str(tmp_path), rest, g["nuc_edge"], nuclei)
    cfg = write_case(str(tmp_path), b, cells, xyz, tris, steps, every,
                     extra + "\ncheckpoint_every = 4")
    exe = os.path.join(PKG, "zf_host")
    r = subprocess.run([exe, cfg], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0 and "nucleated cells, 121 nucleus nodes, 544 filaments" in r.stdout
    out = tmp_path / "out"
    nu = np.genfromtxt(out / f"AllCellNucleus-{steps:09d}.csv", delimiter=",", skip_header=1)
    assert nu.shape == (2 * 121, 13)
    cfgz = lattice_config(*b.shape, 0.8, storage=STORE_F64, inlet_u0=0.02, inlet_amp=0.3,
                          inlet_period=40.0, inlet_dir=(1, 0, 0))
    eng = Engine(cfgz)
    eng.upload_lattice(b)
    eng.membrane_setup(lattice_membrane_config(xyz, tris, 2), tris, xyz)
    eng.nucleus_setup(g["nuc_edge"], np.loadtxt(open(tmp_path / "NucleusNodes.dat").readlines()[1:]).reshape(-1, 3))
    eng.upload_cells(np.round(cells, 6))
    eng.upload_nuclei(nuclei)
    eng.ibm_setup(lattice_ibm_config(rbc_interval=2))
    eng.repulsion_setup(lattice_repulsion_config())
    eng.wall_setup(*geometry.wall_mesh(b))
    for it in range(1, steps + 1):
        eng.full_step(it)
    xn, vn, fn = eng.nuclei()
    assert np.abs(xn - nuclei).max() > 1e-6  # the nuclei moved
    np.testing.assert_allclose(nu[:, :3], xn.reshape(-1, 3), atol=2e-6)
    np.testing.assert_allclose(nu[:, 8:11], fn.reshape(-1, 3), rtol=2e-7, atol=1e-14)
    assert np.abs(fn).max() > 0
    # zf_host again from its restart file at iteration 4: the same end state
    st4 = out / f"ZFstate-{4:09d}.bin"
    assert st4.exists()
    cfg2 = write_case(str(tmp_path), b, cells, xyz, tris, steps - 4, every,
                      extra + f"\nrestart = {st4}", out="out2")
    r = subprocess.run([exe, cfg2], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout[-1000:], flush=True)
    assert r.returncode == 0
    nu2 = np.genfromtxt(tmp_path / "out2" / f"AllCellNucleus-{steps:09d}.csv", delimiter=",",
                        skip_header=1)
    np.testing.assert_array_equal(nu2, nu)
    eng.close()
