"""Transposes of the IBM coupling (zf_ibm_vjp) through the C ABI, against the
FP64 oracle's interpolation and spreading (oracle/zf_oracle_ibm.c).

Spreading is linear in the node forces, so <lam_F, S(X) G> = <lam_G, G>
holds to rounding (1e-12).  The population and position transposes are
checked against central differences of the oracle (h = 1e-6, agreement
1e-6 relative).  Grid with an offset origin and DX = 0.5, and velocity /
force scales other than 1, so that every scale factor is exercised.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from tests.test_gpu_ibm import ibm_config
from tests.test_oracle_mesh import lattice_params
from zebraflow_amd.engine import STORE_F64, Engine, lattice_config, membrane_config

pytestmark = pytest.mark.gpu

ORIGIN, DX, VS, FS = np.array([-1.0, 2.0, 0.5]), 0.5, 1.3, 0.7


def setup(ncell=3, seed=4):
    b = geometry.tube((40, 26, 26), axis=0, r=10.0, periodic=True)
    cfg = lattice_config(*b.shape, 0.8, storage=STORE_F64, periodic=(1, 0, 0))
    xyz, tris = geometry.rbc_mesh(radius=3.0)
    cells = ORIGIN + DX * geometry.place_cells(b, xyz, 3.0, 1.0, ncell, seed)
    mp, _, _, _ = lattice_params(xyz, tris)
    eng = Engine(cfg)
    eng.upload_lattice(b)
    eng.membrane_setup(membrane_config(dict(mp, es=5e-3), ncell), tris, xyz)
    eng.upload_cells(cells)
    eng.ibm_setup(ibm_config(x0=ORIGIN[0], y0=ORIGIN[1], z0=ORIGIN[2], dx=DX, vel_scale=VS,
                             force_scale=FS))
    rng = np.random.default_rng(seed)
    domid, dl = po.dom_setup(b)
    n = len(dl)
    f0 = po.feq(1 + 0.001 * rng.standard_normal(n), 0.02 * rng.standard_normal((3, n)))
    eng.upload_populations(f0)
    p = po.ibm_params(b.shape, x0=ORIGIN[0], y0=ORIGIN[1], z0=ORIGIN[2], dx=DX, vel_scale=VS,
                      force_scale=FS)
    return b, domid, n, eng, cells, f0, p, rng


def macro_of(f):
    d = f.sum(0)
    return [(po.E[:, a] @ f) / d for a in range(3)]


def test_spread_transpose_is_adjoint():
    b, domid, n, eng, cells, f0, p, rng = setup()
    print("case spread transpose, 3 cells", flush=True)
    G = rng.standard_normal(cells.shape)
    lamF = rng.standard_normal((3, n))
    lg, _, _ = eng.ibm_vjp(lam_force=lamF, node_force=G, want_f=False)
    F = po.ibm_spread(p, b, domid, n, cells, G)
    lhs, rhs = (lamF * F).sum(), (lg * G).sum()
    assert abs(rhs) > 1e-3
    np.testing.assert_allclose(rhs, lhs, rtol=1e-12)
    eng.close()


def test_interpolation_transpose_in_populations():
    b, domid, n, eng, cells, f0, p, rng = setup()
    print("case interpolation transpose, populations", flush=True)
    lamV = rng.standard_normal(cells.shape)
    _, _, lam = eng.ibm_vjp(lam_vel=lamV)
    J = lambda f: (lamV * po.ibm_interpolate(p, b, domid, macro_of(f), cells)).sum()
    for trial in range(2):
        d = rng.standard_normal(f0.shape)
        h = 1e-6
        fd = (J(f0 + h * d) - J(f0 - h * d)) / (2 * h)
        np.testing.assert_allclose((lam * d).sum(), fd, rtol=1e-6)
    eng.close()


def test_position_transpose():
    b, domid, n, eng, cells, f0, p, rng = setup()
    print("case position transpose", flush=True)
    G = 1e-3 * rng.standard_normal(cells.shape)
    lamF = rng.standard_normal((3, n))
    lamV = rng.standard_normal(cells.shape)
    lg, lx, lam = eng.ibm_vjp(lam_force=lamF, lam_vel=lamV, node_force=G)
    _, lx_s, _ = eng.ibm_vjp(lam_force=lamF, node_force=G, want_f=False)
    _, lx_i, lam_i = eng.ibm_vjp(lam_vel=lamV)
    np.testing.assert_allclose(lx, lx_s + lx_i, rtol=1e-12, atol=1e-15 * np.abs(lx).max())
    np.testing.assert_array_equal(lam, lam_i)
    u = macro_of(f0)

    def J(x):
        return ((lamF * po.ibm_spread(p, b, domid, n, x, G)).sum() +
                (lamV * po.ibm_interpolate(p, b, domid, u, x)).sum())

    for trial in range(3):
        d = rng.standard_normal(cells.shape)
        h = 1e-6
        fd = (J(cells + h * d) - J(cells - h * d)) / (2 * h)
        np.testing.assert_allclose((lx * d).sum(), fd, rtol=1e-6)
    # one node at a time: both terms matter on their own
    for a in range(3):
        d = np.zeros_like(cells)
        d[1, 17, a] = 1.0
        h = 1e-6
        fd = (J(cells + h * d) - J(cells - h * d)) / (2 * h)
        np.testing.assert_allclose(lx[1, 17, a], fd, rtol=1e-6, atol=1e-9)
    eng.close()


def test_cube_support_transpose():
    """ZF_IBM_CUBE_SUPPORT (no spherical cut-off): the position transpose
    against central differences of the device's own FP64 interpolation and
    spreading with that support."""
    b, domid, n, eng, cells, f0, p, rng = setup()
    print("case cube-support transpose", flush=True)
    eng.ibm_setup(ibm_config(x0=ORIGIN[0], y0=ORIGIN[1], z0=ORIGIN[2], dx=DX, vel_scale=VS,
                             force_scale=FS, flags=1))
    G = 1e-3 * rng.standard_normal(cells.shape)
    lamF = rng.standard_normal((3, n))
    lamV = rng.standard_normal(cells.shape)
    _, lx, _ = eng.ibm_vjp(lam_force=lamF, lam_vel=lamV, node_force=G)

    def J(x):
        eng.upload_cells(x)
        eng.ibm_interpolate()
        _, v, _ = eng.cells()
        eng.upload_membrane_force(G)
        eng.upload_repulsion(np.zeros_like(G))
        eng.ibm_spread()
        return (lamF * eng.force()).sum() + (lamV * v).sum()

    for trial in range(2):
        d = rng.standard_normal(cells.shape)
        h = 1e-6
        fd = (J(cells + h * d) - J(cells - h * d)) / (2 * h)
        np.testing.assert_allclose((lx * d).sum(), fd, rtol=1e-6)
    eng.close()
