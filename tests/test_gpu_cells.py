"""GPU cell list, cell-cell and cell-wall repulsion (a9, a14, a15) and the
whole RBC iteration of zf_step against the oracle, through the C ABI.

Cell lists and the W_index map are integers and must match exactly;
repulsion velocities are FP64 on both sides and agree to 1e-12 of their
largest value.  The composed step reads FP32 macro velocities and FP32 body
forces on the GPU, so positions agree to 1e-8 lu and the spread force to
2e-6 of its largest value.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from tests.test_gpu_ibm import ibm_config, macro_of
from tests.test_oracle_cells import membrane_geometry, two_cells, wall_case
from tests.test_oracle_mesh import lattice_params
from zebraflow_amd.engine import STORE_F64, Engine, lattice_config, load_library, membrane_config
from zebraflow_amd._abi import ZfRepulsionConfig

pytestmark = pytest.mark.gpu


def repulsion_config():
    rc = ZfRepulsionConfig()
    load_library().zf_repulsion_config_reference(C.byref(rc))
    r = po.rep_params()
    for name, _ in ZfRepulsionConfig._fields_:
        setattr(rc, name, getattr(r, name))
    return rc


def engine(b, cells, tris, xyz_rest, periodic=(1, 0, 0), wall=None, store=STORE_F64):
    cfg = lattice_config(*b.shape, 0.8, storage=store, periodic=periodic)
    mp, edge, eq, tae = lattice_params(xyz_rest, tris)
    eng = Engine(cfg)
    eng.upload_lattice(b)
    eng.membrane_setup(membrane_config(dict(mp, es=5e-3), len(cells)), tris, xyz_rest)
    eng.upload_cells(cells)
    eng.ibm_setup(ibm_config())
    eng.repulsion_setup(repulsion_config())
    if wall is not None:
        eng.wall_setup(*wall)
    return eng


@pytest.mark.parametrize("gap", [0.3, 0.1])
def test_cell_list_and_cell_repulsion_match_oracle(gap):
    """Two cells pressed together: the nearer, the more facing triangles a
    node lists (past its 16 own slots into pool chunks at 0.1)."""
    print(f"case cell-cell, two cells {gap} apart", flush=True)
    b, cells, tris = two_cells(gap, seed=2)
    xyz_rest = geometry.rbc_mesh(radius=3.0, level=2)[0]
    eng = engine(b, cells, tris, xyz_rest)
    eng.membrane_forces()
    bins = eng.cell_list()
    eng.repulse()
    v = eng.repulsion()
    g = membrane_geometry(cells, tris, xyz_rest)
    p = po.ibm_params(b.shape)
    r = po.rep_params()
    domid, _ = po.dom_setup(b)
    status = np.ones(len(cells), np.int32)
    ref_bins = po.cell_list(p, b, domid, g["tri_centroid"], status, r.bin_cutoff)
    np.testing.assert_array_equal(bins, ref_bins)
    ref = po.cell_repulsion(p, r, b, domid, tris, status, cells, g["node_normal"],
                            g["tri_centroid"], g["tri_normal"], ref_bins)
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(v, ref, rtol=0, atol=1e-12 * scale)
    eng.close()


@pytest.mark.parametrize("offset", [0.1, -0.3])
def test_windex_and_wall_repulsion_match_oracle(offset):
    print(f"case cell-wall, offset {offset}", flush=True)
    b, wall, cells, tris = wall_case(offset)
    xyz_rest = geometry.rbc_mesh(radius=3.0, level=2)[0]
    eng = engine(b, cells, tris, xyz_rest, wall=wall)
    p = po.ibm_params(b.shape)
    r = po.rep_params()
    nwt, windex = po.wall_tables(p, wall[0], wall[1])
    np.testing.assert_array_equal(eng.windex(), windex)
    eng.membrane_forces()
    eng.cell_list()
    eng.repulse()
    v = eng.repulsion()
    near = eng.wall_neighbour()
    status = np.ones(1, np.int32)
    vw, near_ref = po.wall_repulsion(p, r, windex, *wall, nwt, status, cells)
    g = membrane_geometry(cells, tris, xyz_rest)
    domid, _ = po.dom_setup(b)
    bins = po.cell_list(p, b, domid, g["tri_centroid"], status, r.bin_cutoff)
    vc = po.cell_repulsion(p, r, b, domid, tris, status, cells, g["node_normal"],
                           g["tri_centroid"], g["tri_normal"], bins)
    np.testing.assert_array_equal(near, near_ref)
    ref = vw + vc
    scale = np.abs(ref).max()
    assert np.abs(vw).max() > 0
    np.testing.assert_allclose(v, ref, rtol=0, atol=1e-12 * scale)
    eng.close()


def test_full_step_matches_composed_oracle():
    """Three RBC iterations of zf_step (LBM step, membrane forces, cell list,
    interpolation, repulsion, spreading, position update) against the same
    sequence of oracle calls."""
    print("case zf_step, 3 cells, 3 iterations", flush=True)
    b = geometry.tube((48, 30, 30), axis=0, r=12.0, periodic=True)
    xyz, tris = geometry.rbc_mesh(radius=3.0)
    cells = geometry.place_cells(b, xyz, 3.0, 0.6, 3, 11)
    assert len(cells) == 3
    wall = geometry.wall_mesh(b)
    eng = engine(b, cells, tris, xyz, wall=wall)
    rng = np.random.default_rng(4)
    domid, dl = po.dom_setup(b)
    n = len(dl)
    f0 = po.feq(1 + 0.001 * rng.standard_normal(n), 0.01 * rng.standard_normal((3, n)))
    eng.upload_populations(f0)
    o = po.LbmOracle(b, dict(wall_mode=1, visc_scale=0.3, force_dt=1.0), periodic=(1, 0, 0))
    o.set_populations(f0)
    mp, edge, eq, tae = lattice_params(xyz, tris)
    p = po.ibm_params(b.shape)
    r = po.rep_params()
    nwt, windex = po.wall_tables(p, wall[0], wall[1])
    x = cells.copy()
    status = np.ones(len(cells), np.int32)
    for it in range(1, 4):
        eng.full_step(it)
        o.step(it, rbc_step=1)
        g = po.membrane_forces(x, tris, edge, eq, tae, mp)
        bins = po.cell_list(p, b, domid, g["tri_centroid"], status, r.bin_cutoff)
        v = po.ibm_interpolate(p, b, domid, macro_of(o.populations), x)
        vrep = po.cell_repulsion(p, r, b, domid, tris, status, x, g["node_normal"],
                                 g["tri_centroid"], g["tri_normal"], bins)
        vrep += po.wall_repulsion(p, r, windex, *wall, nwt, status, x)[0]
        po.ibm_spread(p, b, domid, n, x, 0.5 * vrep + g["force"], force=o.force)
        x, status, removed = po.advance(p, x, v + vrep, status)
        assert removed == 0
    xg, _, st = eng.cells()
    assert (st > 0).all()
    assert np.abs(x - cells).max() > 1e-4  # the cells moved
    np.testing.assert_allclose(xg, x, rtol=0, atol=1e-8)
    fg = eng.force()
    scale = np.abs(o.force).max()
    assert scale > 0
    np.testing.assert_allclose(fg, o.force, rtol=0, atol=2e-6 * scale)
    np.testing.assert_allclose(eng.populations(), o.populations, rtol=0, atol=1e-9)
    eng.close()


@pytest.mark.parametrize("select", [0, 1])
def test_uvoid_and_wall_layer_interpolation_match_oracle(select):
    """The wall layer's Uvoid (ZebraflowWT.c:4788-4851) built on the device
    equals the oracle's (pinned by brute force), and the interpolation of a
    cell 0.3 lu from the wall, which reaches that layer, matches the
    oracle's interpolation with the Uvoid terms (:8980-8996)."""
    print(f"case uvoid, select {select}", flush=True)
    b, wall, cells, tris = wall_case(0.3)
    xyz_rest = geometry.rbc_mesh(radius=3.0, level=2)[0]
    eng = engine(b, cells, tris, xyz_rest, wall=wall)
    point = (30.0, -5.0, 7.5)
    eng.uvoid_setup(speed=2e-3, select=select, select_point=point)
    p = po.ibm_params(b.shape)
    nwt, _ = po.wall_tables(p, wall[0], wall[1])
    windex, wdist = po.windex_dist(p, wall[0])
    uv = po.uvoid(p, b, windex, wdist, *wall, nwt, speed=2e-3, select=select, point=point)
    np.testing.assert_array_equal(eng.uvoid(), uv)
    assert np.abs(uv).max() > 1e-3
    rng = np.random.default_rng(6)
    domid, dl = po.dom_setup(b)
    n = len(dl)
    f0 = po.feq(1 + 0.001 * rng.standard_normal(n), 0.01 * rng.standard_normal((3, n)))
    eng.upload_populations(f0)
    eng.ibm_interpolate()
    _, v, _ = eng.cells()
    ref = po.ibm_interpolate(p, b, domid, macro_of(f0), cells, windex=windex, uvoid=uv)
    plain = po.ibm_interpolate(p, b, domid, macro_of(f0), cells)
    assert np.abs(ref - plain).max() > 1e-5  # the layer is reached
    np.testing.assert_allclose(v, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # the position transpose includes the layer's terms (central differences
    # of the oracle's interpolation with them)
    lamV = rng.standard_normal(cells.shape)
    _, lx, _ = eng.ibm_vjp(lam_vel=lamV)
    u = macro_of(f0)
    J = lambda x: (lamV * po.ibm_interpolate(p, b, domid, u, x, windex=windex, uvoid=uv)).sum()
    for trial in range(2):
        d = rng.standard_normal(cells.shape)
        h = 1e-6
        fd = (J(cells + h * d) - J(cells - h * d)) / (2 * h)
        np.testing.assert_allclose((lx * d).sum(), fd, rtol=1e-6)
    # off again: the plain interpolation
    eng.uvoid_setup(False)
    eng.ibm_interpolate()
    _, v0, _ = eng.cells()
    np.testing.assert_allclose(v0, plain, rtol=0, atol=1e-12 * np.abs(plain).max())
    eng.close()
