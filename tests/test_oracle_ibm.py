"""Oracle IBM coupling (a13, a16, a17) pinned by what the mathematics fixes.

No reference output exists for these loops.  Pins:
  - partition of unity: at a lattice point the spherical-cut-off cosine
    kernel sums to exactly one, so a uniform field interpolates exactly and a
    point force spreads its full value (a wrong window, index or weight fails);
  - brute force over every lattice node on a tiny box checks the
    (int)((x-X0)/DX + 1.01) window and its clipping for arbitrary positions;
  - interpolation and spreading are adjoint, <I u, F> = <u, S F>, away from
    walls (same node set, same weights);
  - the position update is x + v dt and removes a cell with a node above vmax.
"""
import numpy as np

from oracle import pyoracle as po


def box(shape, fluid=1):
    b = np.full(shape, fluid, np.int32)
    domid, domlist = po.dom_setup(b)
    return b, domid, domlist


def test_partition_of_unity_at_lattice_points():
    b, domid, dl = box((10, 11, 12))
    p = po.ibm_params(b.shape, x0=-1.0, y0=29.0, z0=-27.0, dx=0.5)
    n = len(dl)
    u = [np.full(n, 0.3), np.full(n, -0.2), np.full(n, 0.05)]
    pts = np.array([[p.x0 + 4 * 0.5, p.y0 + 5 * 0.5, p.z0 + 6 * 0.5]])
    v = po.ibm_interpolate(p, b, domid, u, pts)
    np.testing.assert_allclose(v[0], [0.3, -0.2, 0.05], rtol=1e-14)
    f = po.ibm_spread(p, b, domid, n, pts, np.array([[1.0, 2.0, -3.0]]))
    np.testing.assert_allclose(f.sum(1), [1.0, 2.0, -3.0], rtol=1e-14)


def brute_interp(b, domid, u, x, dx):
    im, jm, km = b.shape
    out = np.zeros(3)
    for i in range(im):
        for j in range(jm):
            for k in range(km):
                if not (0 < b[i, j, k] < 4998):
                    continue
                d = np.array([i * dx, j * dx, k * dx]) - x
                if np.sqrt((d * d).sum()) > 2 * dx:
                    continue
                w = np.prod(1 + np.cos(np.pi * np.abs(d) / (2 * dx))) / 64
                o = domid[i, j, k] - 1
                out += w * np.array([u[0][o], u[1][o], u[2][o]])
    return out


def test_window_matches_brute_force():
    rng = np.random.default_rng(7)
    b, domid, dl = box((7, 8, 6))
    b[3, 4, 2] = 0          # a wall node is skipped by interpolation
    b[1, 1, 1] = 4998       # and so is B == 4998
    n = len(dl)
    u = [rng.standard_normal(n) for _ in range(3)]
    p = po.ibm_params(b.shape)
    pts = rng.uniform([-0.5, -0.5, -0.5], [6.5, 7.5, 5.5], size=(40, 3))
    v = po.ibm_interpolate(p, b, domid, u, pts)
    for x, vv in zip(pts, v):
        np.testing.assert_allclose(vv, brute_interp(b, domid, u, x, 1.0), rtol=1e-12, atol=1e-15)


def test_interpolation_and_spreading_are_adjoint():
    rng = np.random.default_rng(3)
    b, domid, dl = box((16, 16, 16))
    n = len(dl)
    p = po.ibm_params(b.shape, x0=0.0, dx=1.0)
    x = rng.uniform(4, 11, size=(50, 3))
    u = rng.standard_normal((3, n))
    F = rng.standard_normal((50, 3))
    Iu = po.ibm_interpolate(p, b, domid, u, x)
    SF = po.ibm_spread(p, b, domid, n, x, F)
    np.testing.assert_allclose((Iu * F).sum(), (u * SF).sum(), rtol=1e-12)


def test_advance_and_removal():
    p = po.ibm_params((4, 4, 4), dt_advance=25.0, vmax=5e-3)
    x = np.zeros((2, 3, 3))
    v = np.full((2, 3, 3), 1e-4)
    v[1, 2, 0] = 6e-3
    st = np.array([1, 1], np.int32)
    x2, st2, removed = po.advance(p, x.copy(), v, st)
    np.testing.assert_allclose(x2[0], 25.0 * 1e-4)
    assert removed == 1 and st2[1] == -5000 and st2[0] == 1
