"""GPU IBM coupling (a13, a16, a17) and a whole membrane iteration vs the
oracle, through the C ABI.

Tolerances: interpolation reads FP32 macro velocities (rel. 1e-6 of the
largest node velocity); spreading accumulates FP32 weights and forces
(rel. 2e-6 of the largest spread force); the position update is FP64
(1e-12).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from tests.test_oracle_mesh import lattice_params
from zebraflow_amd.engine import STORE_F32DEV, STORE_F64, Engine, lattice_config, membrane_config
from zebraflow_amd._abi import ZfIbmConfig

pytestmark = pytest.mark.gpu


def ibm_config(**kw):
    ic = ZfIbmConfig()
    d = dict(x0=0.0, y0=0.0, z0=0.0, dx=1.0, vel_scale=1.0, force_scale=1.0, gamma_p=0.5,
             dt_advance=1.0, vmax=1.0, rbc_interval=1)
    d.update(kw)
    for k, v in d.items():
        setattr(ic, k, v)
    return ic


def setup(store=STORE_F32DEV, ncell=3, seed=1):
    b = geometry.tube((48, 30, 30), axis=0, r=12.0, periodic=True)
    cfg = lattice_config(*b.shape, 0.8, storage=store, periodic=(1, 0, 0))
    xyz, tris = geometry.rbc_mesh(radius=3.0)
    cells = geometry.place_cells(b, xyz, 3.0, 1.0, ncell, seed)
    assert len(cells) == ncell
    mp, edge, eq, tae = lattice_params(xyz, tris)
    eng = Engine(cfg)
    eng.upload_lattice(b)
    eng.membrane_setup(membrane_config(dict(mp, es=5e-3), ncell), tris, xyz)
    eng.upload_cells(cells)
    eng.ibm_setup(ibm_config())
    rng = np.random.default_rng(seed)
    domid, dl = po.dom_setup(b)
    n = len(dl)
    rho = 1 + 0.001 * rng.standard_normal(n)
    u = 0.02 * rng.standard_normal((3, n))
    f0 = po.feq(rho, u)
    eng.upload_populations(f0)
    return b, domid, dl, eng, cells, f0, (xyz, tris, mp, edge, eq, tae)


def macro_of(f):
    d = f.sum(0)
    return [(po.E[:, a] @ f) / d for a in range(3)]


def test_interpolation_matches_oracle():
    b, domid, dl, eng, cells, f0, _ = setup()
    print("case interpolation, 3 cells", flush=True)
    eng.ibm_interpolate()
    _, v, st = eng.cells()
    p = po.ibm_params(b.shape)
    ref = po.ibm_interpolate(p, b, domid, macro_of(f0), cells)
    assert (st > 0).all()
    scale = np.abs(ref).max()
    assert scale > 1e-3
    np.testing.assert_allclose(v, ref, rtol=0, atol=1e-6 * scale)
    eng.close()


def test_spreading_and_advance_match_oracle():
    b, domid, dl, eng, cells, f0, _ = setup(ncell=4, seed=2)
    print("case spreading, 4 cells", flush=True)
    rng = np.random.default_rng(5)
    F = 1e-3 * rng.standard_normal(cells.shape)
    vrep = 1e-3 * rng.standard_normal(cells.shape)
    eng.ibm_interpolate()
    _, v0, _ = eng.cells()
    eng.upload_membrane_force(F)
    eng.upload_repulsion(vrep)
    eng.ibm_spread()
    force = eng.force()
    p = po.ibm_params(b.shape)
    ref = po.ibm_spread(p, b, domid, len(dl), cells, 0.5 * vrep + F)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(force, ref, rtol=0, atol=2e-6 * scale)
    # total spread force equals the total node force (partition of unity is
    # exact only at lattice points; the cells sit in the bulk, so compare loosely)
    np.testing.assert_allclose(force.sum(1), (0.5 * vrep + F).reshape(-1, 3).sum(0), rtol=0.05,
                               atol=0.05 * np.abs(F).sum() / len(F.reshape(-1, 3)))
    _, v1, _ = eng.cells()
    np.testing.assert_allclose(v1, v0 + vrep, rtol=0, atol=1e-15)
    removed = eng.advance_cells()
    x2, _, st = eng.cells()
    assert removed == 0 and (st > 0).all()
    np.testing.assert_allclose(x2, cells + v1 * 1.0, rtol=0, atol=1e-12)
    eng.close()


def test_deterministic_spreading():
    """Two spreads of the same state give bit-identical force fields."""
    b, domid, dl, eng, cells, f0, _ = setup(ncell=4, seed=3)
    rng = np.random.default_rng(9)
    eng.upload_membrane_force(1e-3 * rng.standard_normal(cells.shape))
    eng.ibm_spread()
    a = eng.force()
    eng.ibm_spread()
    c = eng.force()
    np.testing.assert_array_equal(a, c)
    eng.close()
