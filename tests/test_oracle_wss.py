"""Oracle wall shear stress and wall pressure (PostProWSSandPressure.c:267-369)
pinned by closed forms and planted cases.

- At a lattice-aligned probe the cosine kernel's weights are 8, 4, 2, 1 /64
  on the 3x3x3 neighbourhood: a partition of unity with zero first moment, so
  a linear shear field is reproduced exactly and the WSS has a closed form.
- Velocity along the normal has no tangential part.
- The post-processor's window clips i to [3, IM] but j, k to [1, ...]: a
  velocity planted at i = 2 is not seen, the same at j = 2 is.
- Nodes outside 0.11 < B < 4997 (walls, B = 4998) are not averaged; pressure
  nodes 1001-1009 are.
- The wall pressure reads the nearest listed node within 1.42 DX (brute
  force over the box), p = (D - 1) * scale - Pgauge.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def box_case(shape=(14, 14, 14), b_val=1):
    b = np.full(shape, b_val, np.int32)
    domid, dl = po.dom_setup(b)
    n = len(dl)
    return b, domid, dl, n


def macro_from(fn, dl):
    """D, U, V, W per listed node from a function of 0-based positions."""
    pos = (dl - 1).astype(float)
    d, u, v, w = fn(pos)
    return [np.broadcast_to(np.asarray(a, float), (len(dl),)).copy() for a in (d, u, v, w)]


def test_linear_shear_closed_form():
    b, domid, dl, n = box_case()
    p = po.ibm_params(b.shape)
    gam, zw = 0.013, 4.0
    macro = macro_from(lambda x: (1.0, gam * (x[:, 2] - zw) + 0.002 * x[:, 1], 0.004 * x[:, 0], 0.0), dl)
    # wall points one lattice unit below lattice nodes, normals +z (and
    # tilted normals, whose probes are not lattice-aligned, only for the
    # projection check below)
    pts = np.array([[6.0, 7.0, 5.0], [7.0, 6.0, 6.0], [5.0, 5.0, 7.0]])
    nrm = np.tile([0.0, 0.0, 1.0], (3, 1))
    mu = 0.7
    wss, _, _ = po.wss(p, b, domid, macro, pts, nrm, mu)
    xp = pts + nrm
    u = np.stack([gam * (xp[:, 2] - zw) + 0.002 * xp[:, 1], 0.004 * xp[:, 0], 0 * xp[:, 0]], 1)
    expect = -mu * (u - (u * nrm).sum(1, keepdims=True) * nrm)
    np.testing.assert_allclose(wss, expect, rtol=1e-13, atol=1e-16)


def test_normal_flow_has_no_shear_and_tangential_projection():
    b, domid, dl, n = box_case()
    p = po.ibm_params(b.shape)
    rng = np.random.default_rng(3)
    nv = rng.standard_normal((5, 3))
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    pts = 6.0 + rng.uniform(-0.5, 0.5, (5, 3))
    for k in range(5):
        c = 0.03 * nv[k]
        macro = macro_from(lambda x: (1.0, c[0], c[1], c[2]), dl)
        wss, _, _ = po.wss(p, b, domid, macro, pts[k:k + 1], nv[k:k + 1], 1.0)
        assert np.abs(wss).max() < 1e-15
    # any velocity: the result is orthogonal to n
    macro = macro_from(lambda x: (1.0, 0.01 * x[:, 1], -0.02 * x[:, 2], 0.003 * x[:, 0]), dl)
    wss, _, _ = po.wss(p, b, domid, macro, pts, nv, 1.0)
    assert np.abs((wss * nv).sum(1)).max() < 1e-15
    assert np.abs(wss).max() > 1e-3


def test_window_clips_i_to_three_but_j_to_one():
    b, domid, dl, n = box_case()
    p = po.ibm_params(b.shape)
    nrm = np.array([[0.0, 0.0, 1.0]])
    for axis, seen in ((0, False), (1, True)):
        # velocity only on the plane of 1-based index 2 along `axis`
        macro = macro_from(lambda x: (1.0, np.where(x[:, axis] == 1.0, 0.05, 0.0), 0.0, 0.0), dl)
        pt = np.array([[6.0, 6.0, 5.0]])
        pt[0, axis] = 2.2  # probe 1.2 from that plane: inside the 2 DX support
        wss, _, _ = po.wss(p, b, domid, macro, pt, nrm, 1.0)
        assert (np.abs(wss[0, 0]) > 0) == seen


def test_fluid_code_filter():
    b, domid, dl, n = box_case()
    p = po.ibm_params(b.shape)
    pt, nrm = np.array([[6.0, 6.0, 5.0]]), np.array([[0.0, 0.0, 1.0]])
    macro = macro_from(lambda x: (1.0, 0.05, 0.0, 0.0), dl)
    ref, _, _ = po.wss(p, b, domid, macro, pt, nrm, 1.0)
    for code, counted in ((1001, True), (2001, True), (4998, False), (0, False)):
        b2 = b.copy()
        b2[6, 6, 6] = code  # the probe's own node (weight 8/64)
        wss, _, _ = po.wss(p, b2, domid, macro, pt, nrm, 1.0)
        expect = ref[0, 0] if counted else ref[0, 0] * (1 - 8 / 64)
        np.testing.assert_allclose(wss[0, 0], expect, rtol=1e-14)


def test_wall_pressure_nearest_listed_node():
    b, domid, dl, n = box_case()
    b[5:9, 5:9, 5:9] = -1  # unlisted block: its nodes cannot be chosen
    domid, dl = po.dom_setup(b)
    p = po.ibm_params(b.shape)
    macro = macro_from(lambda x: (1.0 + 1e-4 * (x[:, 0] + 10 * x[:, 1] + 100 * x[:, 2]), 0, 0, 0), dl)
    rng = np.random.default_rng(7)
    pts = rng.uniform(2.0, 11.0, (200, 3))
    nrm = np.tile([1.0, 0.0, 0.0], (200, 1))
    scale, pg = 10000.0 / 3, 120.0
    _, press, valid = po.wss(p, b, domid, macro, pts, nrm, 1.0, scale, pg)
    pos = (dl - 1).astype(float)
    for m in range(len(pts)):
        d = np.linalg.norm(pos - pts[m], axis=1)
        k = int(np.argmin(d))
        if d[k] < 1.42:
            assert valid[m] == 1
            np.testing.assert_allclose(press[m], (macro[0][k] - 1) * scale - pg, rtol=1e-12)
        else:
            assert valid[m] == 0
    assert valid.sum() > 100 and (valid == 0).sum() > 0
