"""Oracle of the cells on a parent lattice (oracle/zf_oracle_wrap.c and the
per-cell rules of oracle/pyoracle.py), which restate the reference's
periodic parent window (ZebraflowWT.c:5898-6000, :7541-8688, :10199-10733,
:11179-11917, :12196-12403).  Pinned by:
  - interpolation and spreading: a field periodic along k with period
    L = (KM-1) DX, so that the window's kmap is invisible, equals the plain
    (non-periodic) interpolation and spreading of the ROI oracle on the box
    tiled three times along k (nodes of crossing cells at zpmap);
  - the kmap's asymmetry (plane KM read through k = KM, plane 1 through
    k = 1 only): a field on plane 1 alone is seen from the entry side and
    not from the exit side nor through zpmap;
  - the wall layer seen through zpmap contributing nothing;
  - spreading as the exact transpose of the interpolation;
  - the cell list against the nearest listed node over the box and its
    periodic images (brute force);
  - the cell-cell repulsion of cells across the wrap equal to the ROI
    oracle's on the same cells moved to the middle of a long box, with the
    triangles binned on plane KM removed (the parent's repulsion window
    reads plane 1 for k = KM, :10231-10234);
  - the position update's wrap, flags and status rules by hand-computed
    cases, and the cloning's slot choice and unwrapping.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from tests.test_oracle_mesh import lattice_params

KM, DX, Z0 = 21, 0.5, -10.0      # window [Z0, ZM] = [-10, 0], L = 10
L = (KM - 1) * DX


def parent_box(periodic=True):
    b = geometry.tube((12, 12, KM), axis=2, r=4.5, periodic=periodic)
    return b


def params(shape, z0=Z0):
    return po.ibm_params(shape, x0=0.0, y0=0.0, z0=z0, dx=DX)


def field(shape, z0, nplanes):
    """A smooth velocity field whose z dependence has period L."""
    im, jm, _ = shape
    i, j, k = np.meshgrid(np.arange(im), np.arange(jm), np.arange(nplanes), indexing="ij")
    z = z0 + k * DX
    ph = 2 * np.pi * (z - Z0) / L
    return [0.01 * np.sin(ph) * (1 + 0.1 * i), 0.02 * np.cos(ph) + 0.001 * j,
            0.03 + 0.005 * np.sin(2 * ph) * np.cos(0.3 * i)]


def on_list(b, f):
    domid, dl = po.dom_setup(b)
    idx = dl - 1
    return domid, [a[idx[:, 0], idx[:, 1], idx[:, 2]] for a in f]


def tiled(b):
    """The box repeated along k: planes 1..KM-1 of three periods, then the
    last plane, origin Z0 - L."""
    core = b[:, :, :KM - 1]
    return np.concatenate([core, core, core, b[:, :, -1:]], axis=2)


def wrapz(x):
    """The reference's storage: every node wrapped into [Z0, ZM)."""
    x = np.array(x, dtype=np.float64, copy=True)
    z = x[..., 2]
    z[z >= 0.0] -= L
    z[z < Z0] += L
    return x


def cell_nodes(seed, centre, r=1.5, n=40):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d *= (r * rng.random(n) ** (1 / 3) / np.linalg.norm(d, axis=1))[:, None]
    return np.asarray(centre) + d


def test_interpolation_periodic_field_equals_tiled_box():
    b = parent_box()
    p = params(b.shape)
    w = po.wrap_of(p, 1)
    domid, u = on_list(b, field(b.shape, Z0, KM))
    bt = tiled(b)
    pt = params(bt.shape, z0=Z0 - L)
    domid_t, ut = on_list(bt, field(bt.shape, Z0 - L, bt.shape[2]))
    # status 2 nodes near both ends and in the middle; a status 4 cell with
    # nodes on both sides of the wrap (wrapped into the window)
    c2 = wrapz(np.stack([cell_nodes(1, (3.0, 3.0, Z0 + 0.6)), cell_nodes(2, (2.5, 3.2, -0.4)),
                         cell_nodes(3, (3.1, 2.9, -5.0))]))
    c4 = cell_nodes(4, (3.0, 3.0, 0.2))          # continuous, across ZM = 0
    c4w = c4.copy()
    c4w[c4w[:, 2] >= 0.0, 2] -= L                # as the reference stores it
    x = np.concatenate([c2, c4w[None]])
    st = np.array([2, 2, 2, 4], np.int32)
    v = po.wrap_interpolate(p, w, b, domid, u, x, st)
    # expected: plain interpolation at the same places in the tiled box,
    # the crossing cell's nodes at zpmap (its continuous coordinates)
    xe = np.concatenate([c2, c4[None]])
    ve = po.ibm_interpolate(pt, bt, domid_t, ut, xe)
    assert np.abs(ve).max() > 1e-3
    np.testing.assert_allclose(v, ve, rtol=0, atol=1e-15)


def test_interpolation_kmap_reads_plane_km_from_the_exit_side():
    b = np.ones((8, 8, KM), np.int32)
    p = params(b.shape)
    w = po.wrap_of(p, 1)
    domid, dl = po.dom_setup(b)
    u = np.zeros((3, len(dl)))
    u[0][dl[:, 2] == 1] = 1.0                     # plane 1 only
    x = np.array([[[2.0, 2.0, Z0 + 0.3]], [[2.0, 2.0, -0.3]], [[2.0, 2.0, Z0 + 0.3]]])
    st = np.array([2, 2, 4], np.int32)
    v = po.wrap_interpolate(p, w, b, domid, u, x, st)
    assert v[0, 0, 0] > 0.1                       # entry side reads plane 1
    assert v[1, 0, 0] == 0.0                      # exit side reads plane KM
    assert v[2, 0, 0] == 0.0                      # zpmap: plane KM as well
    u[0][:] = 0
    u[0][dl[:, 2] == KM] = 1.0                    # plane KM only
    v = po.wrap_interpolate(p, w, b, domid, u, x, st)
    assert v[0, 0, 0] == 0.0 and v[1, 0, 0] > 0.1 and v[2, 0, 0] > 0.1


def test_wall_layer_seen_through_zpmap_adds_nothing():
    b = parent_box()
    p = params(b.shape)
    w = po.wrap_of(p, 1)
    domid, u = on_list(b, field(b.shape, Z0, KM))
    windex = np.where(b <= 0, 1, 0).astype(np.int32)
    uvoid = np.zeros((3,) + b.shape)
    uvoid[0][b <= 0] = 0.5
    # nodes near the wall (r = 4.5 around (5.5, 5.5) lu = (2.75, 2.75))
    near = np.array([[2.75 + 1.9, 2.75, Z0 + 0.4], [2.75 + 1.9, 2.75, -4.0]])
    x = np.stack([near, near])
    st = np.array([2, 4], np.int32)
    v = po.wrap_interpolate(p, w, b, domid, u, x, st, windex=windex, uvoid=uvoid)
    v0 = po.wrap_interpolate(p, w, b, domid, u, x, st)
    assert (v[0] - v0[0])[:, 0].min() > 1e-3      # status 2: the layer's terms
    assert v[1, 0, 0] == v0[1, 0, 0]              # status 4, node behind the middle: none
    assert v[1, 1, 0] > v0[1, 1, 0]               # status 4, node past the middle: terms


def test_spreading_is_the_transpose_of_interpolation():
    b = np.where(parent_box() > 0, 1, -1).astype(np.int32)
    p = params(b.shape)
    w = po.wrap_of(p, 1)
    domid, dl = po.dom_setup(b)
    n = len(dl)
    rng = np.random.default_rng(7)
    x = wrapz(np.stack([cell_nodes(11, (3.0, 3.0, Z0 + 0.2)), cell_nodes(12, (3.0, 2.8, -0.2)),
                        cell_nodes(13, (2.8, 3.0, -5.0))]))
    st = np.array([2, 2, 2], np.int32)
    u = 0.01 * rng.standard_normal((3, n))
    F = rng.standard_normal(x.shape)
    v = po.wrap_interpolate(p, w, b, domid, u, x, st)
    s = po.wrap_spread(p, w, b, domid, n, x, F, st)
    lhs, rhs = (v * F).sum(), (u * s).sum()
    assert abs(lhs) > 1e-3
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_spreading_periodic_equals_tiled_box_folded():
    b = parent_box()
    p = params(b.shape)
    w = po.wrap_of(p, 1)
    domid, dl = po.dom_setup(b)
    bt = tiled(b)
    pt = params(bt.shape, z0=Z0 - L)
    domid_t, dlt = po.dom_setup(bt)
    x = wrapz(np.stack([cell_nodes(21, (3.0, 3.0, Z0 + 0.3)), cell_nodes(22, (3.0, 3.0, -0.3))]))
    st = np.array([2, 2], np.int32)
    F = np.random.default_rng(5).standard_normal(x.shape)
    s = po.wrap_spread(p, w, b, domid, len(dl), x, F, st)
    st_ = po.ibm_spread(pt, bt, domid_t, len(dlt), x, F)
    # fold the tiled box's planes onto the window through the kmap of the
    # spreading: plane k' of the tiled box is k' - (KM-1) of the middle copy
    folded = np.zeros_like(s)
    kt = dlt[:, 2] - (KM - 1)
    kk = np.where(kt < 1, kt + (KM - 1), np.where(kt > KM, kt - (KM - 1), kt))
    # planes seen from the exit side keep KM, from the entry side keep 1:
    # those are the physical copies the middle tiles hold
    key = {tuple(r): q for q, r in enumerate(dl)}
    for q, (i, j, _) in enumerate(dlt):
        folded[:, key[(i, j, kk[q])]] += st_[:, q]
    assert np.abs(s).max() > 0.05
    np.testing.assert_allclose(s, folded, rtol=0, atol=1e-14)


def test_cell_list_nearest_listed_node_over_periodic_images():
    b = parent_box()
    p = params(b.shape)
    w = po.wrap_of(p, 1)
    domid, dl = po.dom_setup(b)
    rng = np.random.default_rng(9)
    # centroids inside the lumen, some near both ends; a crossing cell's
    # centroids past the exit (the reference's unwrapped TRI_CENTROID)
    tc2 = np.column_stack([2.75 + 1.5 * rng.uniform(-1, 1, 60), 2.75 + 1.5 * rng.uniform(-1, 1, 60),
                           rng.uniform(Z0, 0.0, 60)])
    tc4 = tc2.copy()
    tc4[:, 2] = rng.uniform(-0.4, 0.8, 60)
    st = np.array([2, 4], np.int32)
    bins = po.wrap_cell_list(p, w, b, domid, np.stack([tc2, tc4]), st)
    pos = np.column_stack([(dl[:, 0] - 1) * DX, (dl[:, 1] - 1) * DX, Z0 + (dl[:, 2] - 1) * DX])
    for c, tc in enumerate((tc2, tc4)):
        for t, q in enumerate(tc):
            z = q[2] - L if (st[c] != 2 and q[2] >= 0.0) else q[2]
            best = np.inf
            for sh in (-L, 0.0, L):
                d = np.linalg.norm(pos + [0, 0, sh] - [q[0], q[1], z], axis=1)
                # planes 1 and KM lie L apart: the place they share is plane 1
                # seen from below and plane KM seen from above, never their
                # images (an image of plane KM at -L is plane 1's place)
                if sh < 0:
                    d[dl[:, 2] == KM] = np.inf
                if sh > 0:
                    d[dl[:, 2] == 1] = np.inf
                k = int(np.argmin(d))
                if d[k] < best:
                    best, arg = d[k], k
            assert bins[c, t] == domid[tuple(dl[arg] - 1)], (c, t)


def test_cell_repulsion_across_the_wrap_equals_roi_oracle_in_the_middle():
    b = parent_box()
    p = params(b.shape)
    w = po.wrap_of(p, 1)
    domid, dl = po.dom_setup(b)
    xyz, tris = geometry.rbc_mesh(radius=1.4, level=2)
    mp, edge, eq, tae = lattice_params(xyz, tris)
    r = po.rep_params(cutoff=0.5, search=1.0)
    # continuous coordinates: cell A across ZM, cell B just behind it
    a = xyz * [1.0, 1.0, 0.35] + [2.75, 2.75, -0.1]
    bc = xyz * [1.0, 1.0, 0.35] + [2.75, 2.75, -1.3]
    cont = np.stack([a, bc])
    gm = po.membrane_forces(cont, tris, edge, eq, tae, mp)
    # the reference's storage: every node wrapped into the window
    wr = cont.copy()
    wr[..., 2] = np.where(wr[..., 2] >= 0.0, wr[..., 2] - L, wr[..., 2])
    st = np.array([4, 2], np.int32)
    # TRI_CENTROID of the split cell is unwrapped (continuous); B is inside
    bins = po.wrap_cell_list(p, w, b, domid, gm["tri_centroid"], st)
    v = po.wrap_cell_repulsion(p, w, r, b, domid, tris, st, wr, gm["node_normal"],
                               gm["tri_centroid"], gm["tri_normal"], bins)
    assert np.abs(v).max() > 1e-8
    # the same cells moved by -L/2 into the middle of the box: no wrap
    shift = np.array([0, 0, -L / 2])
    mid = cont + shift
    gm2 = po.membrane_forces(mid, tris, edge, eq, tae, mp)
    pm = params(b.shape)
    bins2 = po.cell_list(pm, b, domid, gm2["tri_centroid"], np.array([1, 1], np.int32), 5.0)
    # the parent's window never reads plane KM: its triangles drop out
    on_km = dl[np.maximum(bins, 1) - 1][..., 2] == KM
    bins2[on_km & (bins > 0)] = 0
    ve = po.cell_repulsion(pm, r, b, domid, tris, np.array([1, 1], np.int32), mid,
                           gm2["node_normal"], gm2["tri_centroid"], gm2["tri_normal"], bins2)
    np.testing.assert_allclose(v, ve, rtol=0, atol=1e-12 * np.abs(ve).max())


def test_advance_wraps_flags_and_returns_to_status_2():
    p = params((12, 12, KM))
    w = po.wrap_of(p, 1)
    base = cell_nodes(31, (3.0, 3.0, 0.0), r=1.0, n=20)
    # cell 0: status 2 moving across the exit, REGRESSFLAG 0 -> flagged;
    # cell 1: the same with REGRESSFLAG 1 -> wraps, no flag;
    # cell 2: status 4 wrapped well inside -> back to 2
    x = np.stack([base - [0, 0, 1.2], base - [0, 0, 1.2], base - [0, 0, 5.0]])
    x = np.ascontiguousarray(x)
    v = np.zeros_like(x)
    v[..., 2] = 0.05
    st = np.array([2, 2, 4], np.int32)
    reg = np.array([0, 1, 0], np.int32)
    x0 = x.copy()
    flags = po.wrap_advance(w, 25.0, x, v, st, reg)
    moved = x0[..., 2] + 1.25
    exp = np.where(moved >= 0.0, moved - L, moved)
    np.testing.assert_array_equal(x[..., 2], exp)
    assert list(flags) == [1, 0, 0]
    assert list(st) == [2, 2, 2] and list(reg) == [0, 1, 1]
    # a status 4 cell still across the wrap stays 4
    x2 = np.ascontiguousarray(base[None] - [0, 0, 0.5])
    st2, reg2 = np.array([4], np.int32), np.array([0], np.int32)
    po.wrap_advance(w, 25.0, x2, np.zeros_like(x2), st2, reg2)
    assert st2[0] == 4 and reg2[0] == 1


def test_clone_takes_the_first_free_slot_unwrapped():
    p = params((12, 12, KM))
    w = po.wrap_of(p, 1)
    base = cell_nodes(41, (3.0, 3.0, -0.3), r=1.0, n=20)
    px = base.copy()
    px[px[:, 2] >= 0.0, 2] -= L                   # wrapped nodes near Z0
    pxyz = np.stack([px, px + [0, 0, -3.0]])
    pvel = np.full_like(pxyz, 0.01)
    pst, preg = np.array([2, 2], np.int32), np.array([0, 0], np.int32)
    flags = np.array([1, 0], np.int32)
    mst = np.array([1, -5000, -5000], np.int32)
    mxyz = np.zeros((3, 20, 3))
    mvel = np.zeros_like(mxyz)
    pairs = po.clone_cells(w, 188, pst, preg, flags, pxyz, pvel, mst, mxyz, mvel)
    assert pairs == [(0, 1)]
    assert list(mst) == [1, -189, -5000] and list(pst) == [4, 2] and list(preg) == [1, 0]
    np.testing.assert_allclose(mxyz[1], base, atol=1e-14)   # straddles the exit again
    assert mxyz[1][:, 2].min() < 0.0 < mxyz[1][:, 2].max()
    # no free slot: the parent still turns to 4
    mst2 = np.array([1, 1, 1], np.int32)
    pst2, preg2 = np.array([2], np.int32), np.array([0], np.int32)
    assert po.clone_cells(w, 0, pst2, preg2, np.array([1], np.int32), pxyz[:1], pvel[:1], mst2,
                          mxyz, mvel) == []
    assert pst2[0] == 4 and preg2[0] == 1


def test_clone_band_and_unwrap_rules_for_both_directions():
    p = params((12, 12, KM))
    wp, wm = po.wrap_of(p, 1), po.wrap_of(p, -1)
    z = np.array([-1.0, 0.0, 2.4, 2.6, 9.0])
    assert list(po.clone_band(wp, z, 2.5)) == [True, True, True, False, False]
    z = np.array([Z0 + 1.0, Z0, Z0 - 2.4, Z0 - 2.6])
    assert list(po.clone_band(wm, z, 2.5)) == [True, True, True, False]
    zz = np.array([[Z0 + 0.5, -0.5, -9.0, -1.0]])
    st = np.array([4], np.int32)
    half = L / 2  # the split span of a window of length L (the reference's 30 for 60)
    u, cut = po.wrap_unwrap(wp, zz, st, zz.min(1), zz.max(1), split=half)
    assert cut[0] and list(u[0]) == [Z0 + 0.5 + L, -0.5, -9.0 + L, -1.0]
    u, cut = po.wrap_unwrap(wm, zz, st, zz.min(1), zz.max(1), split=half)
    assert cut[0] and list(u[0]) == [Z0 + 0.5, -0.5 - L, -9.0, -1.0 - L]
    u, cut = po.wrap_unwrap(wp, zz, np.array([2], np.int32), zz.min(1), zz.max(1), split=half)
    assert not cut[0]
