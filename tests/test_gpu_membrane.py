"""GPU membrane (a7-a12) vs the oracle, through the C ABI.

Topology and the reference-mesh equilibrium parameters come from the host
side of libzfgpu.so and must equal the oracle's (bit-exact topology; FP64
parameters to 1e-12).  Forces, normals and the FP64 cell reductions of the
device kernels must match the FP64 oracle to 1e-10 relative.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from tests.test_oracle_mesh import deformed, lattice_params
from zebraflow_amd.engine import Engine, lattice_config, membrane_config

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "mesh_zfish.npz")


def engine():
    eng = Engine(lattice_config(8, 8, 8, 0.8))
    return eng


def test_reference_mesh_topology_and_equilibrium():
    g = np.load(GOLD)
    print("case zebrafish mesh", flush=True)
    eng = engine()
    mp = po.reference_membrane_params()
    eng.membrane_setup(membrane_config(dict(mp, es=3e-6), 1), g["tris"], g["nodes"])
    edge, nbn, nbt = eng.mesh_topology()
    np.testing.assert_array_equal(edge, g["edge"])
    np.testing.assert_array_equal(nbn, g["nbr_node"])
    np.testing.assert_array_equal(nbt, g["nbr_tri"])
    eq = eng.mesh_equilibrium()
    ora = po.mesh_equilibrium(g["nodes"], g["edge"], mp, es=3e-6)
    # the dihedral sine is sqrt(1 - cos^2): FP64 rounding near 0 is ~1e-11 absolute
    for k in ora:
        np.testing.assert_allclose(eq[k], ora[k], rtol=1e-10, atol=1e-10 if "0" in k else 0)
    np.testing.assert_allclose(eq["lmax"], g["lmax"], rtol=2e-6)
    eng.close()


@pytest.mark.parametrize("ncell,blend", [(1, 0.0), (7, 0.0), (3, 0.02)])
def test_forces_match_oracle(ncell, blend):
    """blend > 0: the smooth bending cut-off over |cos dihedral| in
    [0.97, 0.99] (its formula is pinned on the oracle)."""
    xyz, x1, tris = deformed(amp=0.1)
    mp, edge, eq, tae = lattice_params(xyz, tris, cos_blend=blend)
    if blend:
        p = x1[tris - 1]
        av = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        n = av / np.linalg.norm(av, axis=1)[:, None]
        ac = np.abs((n[edge[:, 4] - 1] * n[edge[:, 5] - 1]).sum(1))
        assert ((ac > 0.99 - blend) & (ac <= 0.99)).sum() > 50  # hinges inside the fade
    rng = np.random.default_rng(ncell)
    cells = np.stack([x1 * (1 + 0.02 * rng.standard_normal()) + rng.uniform(-50, 50, 3)
                      for _ in range(ncell)])
    print(f"case lattice-unit RBCs ncell {ncell} cos_blend {blend}", flush=True)
    eng = engine()
    eng.membrane_setup(membrane_config(dict(mp, es=5e-3), ncell + 1), tris, xyz)
    eng.upload_cells(cells)
    eng.membrane_forces()
    got = eng.membrane()
    ref = po.membrane_forces(cells, tris, edge, eq, tae, mp)
    scale = np.abs(ref["force"]).max()
    np.testing.assert_allclose(got["force"], ref["force"], rtol=0, atol=1e-10 * scale)
    np.testing.assert_allclose(got["node_normal"], ref["node_normal"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(got["cell_area"], ref["cell_area"], rtol=1e-13)
    np.testing.assert_allclose(got["cell_vol"], ref["cell_vol"], rtol=1e-13)
    np.testing.assert_allclose(got["cell_centroid"], ref["cell_centroid"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(got["guard_fail"], ref["guard_fail"])
    eng.close()


@pytest.mark.parametrize("amp", [0.1, 0.3])
def test_fp32_node_forces_match_oracle(amp):
    """ZF_FLAG_F32_CELLS: the node forces in FP32 arithmetic (positions
    relative to each node) against the FP64 oracle, to 1e-5 of the largest
    force (FP32's 6e-8 rounding amplified by the cancellation of the spring,
    bending and area terms at each node); the triangle geometry and the cell
    reductions stay FP64 and match as tightly as without the flag."""
    from zebraflow_amd.engine import FLAG_F32_CELLS, STORE_F32DEV
    xyz, x1, tris = deformed(amp=amp)
    mp, edge, eq, tae = lattice_params(xyz, tris)
    rng = np.random.default_rng(3)
    cells = np.stack([x1 * (1 + 0.02 * rng.standard_normal()) + rng.uniform(100, 400, 3)
                      for _ in range(4)])
    print(f"case FP32 node forces, amp {amp}", flush=True)
    eng = Engine(lattice_config(8, 8, 8, 0.8, storage=STORE_F32DEV, flags=FLAG_F32_CELLS))
    eng.membrane_setup(membrane_config(dict(mp, es=5e-3), 4), tris, xyz)
    eng.upload_cells(cells)
    eng.membrane_forces()
    got = eng.membrane()
    ref = po.membrane_forces(cells, tris, edge, eq, tae, mp)
    scale = np.abs(ref["force"]).max()
    err = np.abs(got["force"] - ref["force"]).max() / scale
    print(f"FP32 node force error {err:.3e} of the largest force", flush=True)
    assert 0 < err <= 1e-5  # FP32, not FP64, arithmetic ran
    np.testing.assert_allclose(got["cell_area"], ref["cell_area"], rtol=1e-13)
    np.testing.assert_allclose(got["cell_vol"], ref["cell_vol"], rtol=1e-13)
    np.testing.assert_array_equal(got["guard_fail"], ref["guard_fail"])
    eng.close()


def test_guards_drop_terms_like_oracle():
    """A crushed triangle trips the force guards; the device drops the same
    terms and counts the same failures as the oracle."""
    xyz, x1, tris = deformed(amp=0.05)
    mp, edge, eq, tae = lattice_params(xyz, tris, force_limit=2e-3)
    x2 = x1.copy()
    t = tris[10] - 1
    x2[t[0]] = x2[t[1]] + 1e-3 * (x2[t[0]] - x2[t[1]])
    eng = engine()
    eng.membrane_setup(membrane_config(dict(mp, es=5e-3), 1), tris, xyz)
    eng.upload_cells(x2[None])
    eng.membrane_forces()
    got = eng.membrane()
    ref = po.membrane_forces(x2[None], tris, edge, eq, tae, mp)
    assert ref["guard_fail"][0] > 0
    np.testing.assert_array_equal(got["guard_fail"], ref["guard_fail"])
    scale = np.abs(ref["force"]).max()
    np.testing.assert_allclose(got["force"], ref["force"], rtol=0, atol=1e-9 * scale)
    eng.close()
