"""GPU wall shear stress and wall pressure against the oracle, at the nodes
of the lumen wall mesh of a driven tube, through the C ABI.  The operator
reads the same populations on both sides (the oracle gets the GPU's FP64
macro fields), so the results agree to FP64 rounding."""
import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from zebraflow_amd.engine import (B_PRESSURE_OUTLET, B_VELOCITY_INLET, STORE_F32DEV, STORE_F64,
                                  Engine, lattice_config)

pytestmark = pytest.mark.gpu


def node_normals(xyz, tri, tnorm):
    """Area-weighted average of the incident triangle normals, unit length."""
    p = xyz[tri - 1]
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    acc = np.zeros_like(xyz)
    for s in range(3):
        np.add.at(acc, tri[:, s] - 1, area[:, None] * tnorm)
    return acc / np.linalg.norm(acc, axis=1, keepdims=True)


@pytest.mark.parametrize("store", [STORE_F64, STORE_F32DEV])
def test_wss_matches_oracle(store):
    print(f"case wss, store {store}", flush=True)
    b = geometry.tube((40, 24, 24), axis=0, r=9.0, inlet=B_VELOCITY_INLET,
                      outlet=B_PRESSURE_OUTLET)
    cfg = lattice_config(*b.shape, 0.8, storage=store, inlet_u0=0.02, inlet_amp=0.3,
                         inlet_period=200.0, inlet_dir=(1, 0, 0))
    eng = Engine(cfg)
    eng.upload_lattice(b)
    eng.run(0, 300)
    wx, wt, wn = geometry.wall_mesh(b)
    nrm = node_normals(wx, wt, wn)
    mu = 0.1
    wss, press, valid = eng.wss(wx, nrm, mu, pgauge=1.0)
    d, u, v, w = eng.macro()
    domid, _ = po.dom_setup(b)
    p = po.ibm_params(b.shape)
    ref, pref, vref = po.wss(p, b, domid, (d, u, v, w), wx, nrm, mu, cfg.pressure_scale, 1.0)
    np.testing.assert_array_equal(valid, vref)
    scale = np.abs(ref).max()
    assert scale > 1e-4
    np.testing.assert_allclose(wss, ref, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(press, pref, rtol=0, atol=1e-10 * np.abs(pref).max())
    # the shear points against the flow on the tube wall (tau = -mu u_t / DX
    # with u_t along +x away from the inlet/outlet planes)
    mid = (wx[:, 0] > 8) & (wx[:, 0] < 31)
    assert (wss[mid, 0] < 0).mean() > 0.95
    eng.close()
