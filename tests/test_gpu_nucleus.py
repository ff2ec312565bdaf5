"""The nucleus network on the GPU through the C ABI: its equilibrium
parameters against the reference's own (goldens), its spring forces and
nucleus-membrane contact against the oracle (pinned in
tests/test_oracle_nucleus.py), and whole RBC iterations of zf_step with
nucleated cells against the composed FP64 oracle.  The cells are the
reference's zebrafish RBC (504 nodes) with its 121-node nucleus, the
nucleus pushed towards one face so that the contact acts."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.test_gpu_cells import repulsion_config
from tests.test_gpu_ibm import ibm_config, macro_of
from tests.test_oracle_mesh import lattice_params
from zebraflow_amd import geometry
from zebraflow_amd.engine import (STORE_F64, Engine, lattice_config, membrane_config,
                                  reference_config)

pytestmark = pytest.mark.gpu

def zfish():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "mesh_zfish.npz"))


def rigid_fit(a, b):
    """R, t with b = a R^T + t (the placement's rotation and translation)."""
    ca, cb = a.mean(0), b.mean(0)
    u, _, vt = np.linalg.svd((a - ca).T @ (b - cb))
    d = np.sign(np.linalg.det(vt.T @ u.T))
    r = vt.T @ np.diag([1, 1, d]) @ u.T
    return r, cb - ca @ r.T


def test_nucleus_equilibrium_matches_reference():
    print("case nucleus equilibrium, reference units", flush=True)
    g = zfish()
    from zebraflow_amd._abi import ZfMembraneConfig
    import ctypes as C
    eng = Engine(reference_config(16, 16, 16, periodic=(1, 0, 0)))
    eng.upload_lattice(geometry.tube((16, 16, 16), axis=0, r=5.0, periodic=True))
    mc = ZfMembraneConfig()
    eng.lib.zf_membrane_config_reference(C.byref(mc))
    mc.max_cells = 1
    eng.membrane_setup(mc, g["tris"], g["nodes"])
    eng.nucleus_setup(g["nuc_edge"], g["nuc_nodes"])
    eq = eng.nucleus_equilibrium()
    for k in ("lmax", "plength", "kp"):
        np.testing.assert_allclose(eq[k], g["nuc_" + k].astype(np.float64), rtol=2e-6)
    eng.close()


def setup(ncell=2, seed=5, list_interval=1, store=STORE_F64):
    g = zfish()
    xyz, tris, xn0, ned = g["nodes"], g["tris"], g["nuc_nodes"], g["nuc_edge"]
    b = geometry.tube((48, 30, 30), axis=0, r=12.0, periodic=True)
    cells = geometry.place_cells(b, xyz, 6.7, 1.0, ncell, seed)
    assert len(cells) == ncell
    # strained nuclei (3 % larger than the rest network, with noise), pushed
    # towards one face of their cell
    shift = np.array([0.05, 0.0, 0.15])
    noise = 0.01 * np.random.default_rng(seed).standard_normal(xn0.shape)
    nuclei = np.stack([(1.03 * xn0 + noise + shift) @ r.T + t
                       for r, t in (rigid_fit(xyz, c) for c in cells)])
    mp, edge, eq, tae = lattice_params(xyz, tris)
    cfg = lattice_config(*b.shape, 0.8, storage=store, periodic=(1, 0, 0))
    eng = Engine(cfg)
    eng.upload_lattice(b)
    eng.membrane_setup(membrane_config(dict(mp, es=5e-3), ncell), tris, xyz)
    eng.nucleus_setup(ned, xn0, list_interval=list_interval)
    eng.upload_cells(cells)
    eng.upload_nuclei(nuclei)
    eng.ibm_setup(ibm_config())
    neq = po.nucleus_equilibrium(xn0, ned, mp, 0.1 * 5e-3)
    return b, g, cells, nuclei, eng, (mp, edge, eq, tae), neq


def test_nucleus_forces_and_contact_match_oracle():
    print("case nucleus forces and contact, 2 zebrafish cells", flush=True)
    b, g, cells, nuclei, eng, (mp, edge, eq, tae), neq = setup()
    eng.membrane_forces()
    eng.nucleus_forces()
    _, _, fn = eng.nuclei()
    ref = po.nucleus_forces(nuclei, g["nuc_edge"], neq, mp)
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(fn, ref, rtol=0, atol=1e-12 * scale)
    eng.upload_repulsion(np.zeros_like(cells))
    lost = eng.nucleus_contact(1)
    assert lost == 0
    v = eng.repulsion()
    gm = po.membrane_forces(cells, g["tris"], edge, eq, tae, mp)
    vr, lost_o = po.nucleus_contact(nuclei, cells, g["tris"], np.ones(len(cells), np.int32),
                                    gm["tri_centroid"], gm["tri_normal"])
    assert lost_o == 0
    s = np.abs(vr).max()
    assert s > 1e-8  # the nucleus touches the membrane
    np.testing.assert_allclose(v, vr, rtol=0, atol=1e-12 * s)
    eng.close()


def test_step_with_nuclei_matches_composed_oracle():
    """Three RBC iterations of zf_step (LBM step, membrane and nucleus
    forces, cell list, interpolation of membrane and nucleus nodes, cell and
    wall repulsion, nucleus contact, spreading of both node sets, position
    update) against the same oracle calls."""
    print("case zf_step with nuclei, 2 cells, 3 iterations", flush=True)
    b, g, cells, nuclei, eng, (mp, edge, eq, tae), neq = setup()
    tris = g["tris"]
    wall = geometry.wall_mesh(b)
    eng.repulsion_setup(repulsion_config())
    eng.wall_setup(*wall)
    rng = np.random.default_rng(4)
    domid, dl = po.dom_setup(b)
    n = len(dl)
    f0 = po.feq(1 + 0.001 * rng.standard_normal(n), 0.01 * rng.standard_normal((3, n)))
    eng.upload_populations(f0)
    o = po.LbmOracle(b, dict(wall_mode=1, visc_scale=0.3, force_dt=1.0), periodic=(1, 0, 0))
    o.set_populations(f0)
    p = po.ibm_params(b.shape)
    r = po.rep_params()
    nwt, windex = po.wall_tables(p, wall[0], wall[1])
    x, xn = cells.copy(), nuclei.copy()
    status = np.ones(len(cells), np.int32)
    nm = x.shape[1] * len(x)
    for it in range(1, 4):
        eng.full_step(it)
        o.step(it, rbc_step=1)
        gm = po.membrane_forces(x, tris, edge, eq, tae, mp)
        fn = po.nucleus_forces(xn, g["nuc_edge"], neq, mp)
        bins = po.cell_list(p, b, domid, gm["tri_centroid"], status, r.bin_cutoff)
        u = macro_of(o.populations)
        v = po.ibm_interpolate(p, b, domid, u, x)
        vn = po.ibm_interpolate(p, b, domid, u, xn)
        vrep = po.cell_repulsion(p, r, b, domid, tris, status, x, gm["node_normal"],
                                 gm["tri_centroid"], gm["tri_normal"], bins)
        vrep += po.wall_repulsion(p, r, windex, *wall, nwt, status, x)[0]
        vrep, _ = po.nucleus_contact(xn, x, tris, status, gm["tri_centroid"], gm["tri_normal"],
                                     vrep=vrep)
        pts = np.concatenate([x.reshape(-1, 3), xn.reshape(-1, 3)])
        ffl = np.concatenate([(0.5 * vrep + gm["force"]).reshape(-1, 3), fn.reshape(-1, 3)])
        po.ibm_spread(p, b, domid, n, pts, ffl, force=o.force)
        x, status, removed = po.advance(p, x, v + vrep, status)
        assert removed == 0
        xn = xn + vn * p.dt_advance
    xg, _, st = eng.cells()
    xng, _, _ = eng.nuclei()
    assert (st > 0).all()
    assert np.abs(xn - nuclei).max() > 1e-5  # the nuclei moved
    np.testing.assert_allclose(xg, x, rtol=0, atol=1e-8)
    np.testing.assert_allclose(xng, xn, rtol=0, atol=1e-8)
    fg = eng.force()
    scale = np.abs(o.force).max()
    assert scale > 0
    np.testing.assert_allclose(fg, o.force, rtol=0, atol=2e-6 * scale)
    np.testing.assert_allclose(eng.populations(), o.populations, rtol=0, atol=1e-9)
    assert nm > 0
    eng.close()


def test_checkpoints_with_nuclei_continue_bit_identically(tmp_path):
    """A device slot (zf_checkpoint_save/restore) and a restart file
    (zfh_save_state/load_state) taken between two list rebuilds: the runs
    continued from them equal the uninterrupted run bit for bit, nuclei and
    their contact lists included."""
    print("case checkpoints with nuclei, 2 cells, lists every 3 iterations", flush=True)

    def make():
        b, g, cells, nuclei, eng, *_ = setup(list_interval=3)
        eng.repulsion_setup(repulsion_config())
        eng.wall_setup(*geometry.wall_mesh(b))
        eng.init_equilibrium(1.0, (0.01, 0.0, 0.0))
        return eng, len(cells), cells.shape[1]

    def state(eng):
        x, _, st = eng.cells()
        xn, vn, fn = eng.nuclei()
        lst, cnt, valid = eng.nucleus_lists()
        return [eng.populations_raw(), eng.force(), x, st, xn, vn, fn, cnt, valid,
                np.concatenate([r[:c] for r, c in zip(lst, cnt)])]

    eng, ncell, nn = make()
    for it in range(1, 5):  # lists built at 1 (first call) and 3
        eng.full_step(it)
    eng.checkpoint_save(0)
    path = tmp_path / "ZFstate-000000004.bin"
    eng.save_state(path, 4, ncell, nn)
    for it in range(5, 8):  # rebuilt at 6
        eng.full_step(it)
    ref = state(eng)
    assert ref[8] == 1 and ref[7].sum() > 0
    eng.checkpoint_restore(0)
    for it in range(5, 8):
        eng.full_step(it)
    for a, b_ in zip(state(eng), ref):
        np.testing.assert_array_equal(a, b_)
    eng.close()
    eng2, _, _ = make()
    assert eng2.load_state(path, ncell, nn) == 4
    for it in range(5, 8):
        eng2.full_step(it)
    for a, b_ in zip(state(eng2), ref):
        np.testing.assert_array_equal(a, b_)
    eng2.close()
