"""The allocation-free WSS entry point (zf_wss_noalloc) against zf_wss,
through the C ABI: two calls in a row, the second with fewer points, each
bit-identical to zf_wss on the same state; and the size limit is refused."""
import numpy as np
import pytest

from zebraflow_amd import geometry
from zebraflow_amd.engine import (B_PRESSURE_OUTLET, B_VELOCITY_INLET, STORE_F32DEV, Engine,
                                  WSS_NOALLOC_MAX_POINTS, ZfError, lattice_config)

pytestmark = pytest.mark.gpu


def test_wss_noalloc_matches_wss_on_two_calls():
    print("case wss_noalloc: tube 40x24x24, f32dev, two calls", flush=True)
    b = geometry.tube((40, 24, 24), axis=0, r=9.0, inlet=B_VELOCITY_INLET,
                      outlet=B_PRESSURE_OUTLET)
    cfg = lattice_config(*b.shape, 0.8, storage=STORE_F32DEV, inlet_u0=0.02, inlet_amp=0.3,
                         inlet_period=200.0, inlet_dir=(1, 0, 0))
    eng = Engine(cfg)
    eng.upload_lattice(b)
    eng.run(0, 200)
    wx, wt, wn = geometry.wall_mesh(b)
    # per-node normals: mean of the incident triangle normals
    nrm = np.zeros_like(wx)
    for s in range(3):
        np.add.at(nrm, wt[:, s] - 1, wn)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    n1 = len(wx)
    assert n1 > 1000
    ref1 = eng.wss(wx, nrm, 0.1, pgauge=1.0)
    got1 = eng.wss(wx, nrm, 0.1, pgauge=1.0, noalloc=True)
    for a, r in zip(got1, ref1):
        np.testing.assert_array_equal(a, r)
    assert np.abs(ref1[0]).max() > 1e-4
    # second call: fewer points in another order, on a later state
    eng.run(200, 50)
    sel = np.random.default_rng(7).permutation(n1)[: n1 // 3]
    x2, m2 = wx[sel], nrm[sel]
    ref2 = eng.wss(x2, m2, 0.2, pgauge=0.5)
    got2 = eng.wss(x2, m2, 0.2, pgauge=0.5, noalloc=True)
    for a, r in zip(got2, ref2):
        np.testing.assert_array_equal(a, r)
    # beyond the scratch's limit the call is refused, not run
    lim = WSS_NOALLOC_MAX_POINTS
    big = np.zeros((lim + 1, 3))
    with pytest.raises(ZfError):
        eng.wss(big, big, 0.1, noalloc=True)
    eng.close()
