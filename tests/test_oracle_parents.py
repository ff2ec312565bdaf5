"""Oracle of the parent (feeder) lattices' lists against the reference's own
LBM_PARENTDOMAIN_SETUP / LBM_PARENTBC_SETUP (ZebraflowWT.c:515-1350), run by
tests/golden/make_parent_goldens.py on the three parent boundary files the
reference ships.  The overlap geometry is the reference's lattice origins
(ZebraflowWT.c:85-112, DX 0.5): parent 1 (DA) meets the main lattice at its
top plane k = KM_1 = 121, main k = 3; parents 2 and 3 (PCV) at k = 1, main
k = 720; the offsets are (X0_p - X0)/DX etc."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "parents_ref.npz")
# p -> (k_over, offset of the parent index space in the main lattice's)
GEOMETRY = {1: (121, (143, 362, -118)), 2: (1, (126, 406, 719)), 3: (1, (118, 361, 719))}


def test_overlap_offsets_follow_the_reference_origins():
    # (int)((X_p - X0)/DX + 1.01) of the reference with its origins, in float
    X0, Y0, Z0, DX = -1.0, 29.0, -27.0, 0.5
    origins = {1: (70.5, 210.0, -86.0, 121), 2: (62.0, 232.0, 332.5, 121), 3: (58.0, 209.5, 332.5, 121)}
    for p, (x0, y0, z0, km) in origins.items():
        k_over, off = GEOMETRY[p]
        zo = z0 + (k_over - 1) * DX
        main = [int(np.float32((np.float32(x0) - np.float32(X0)) / np.float32(DX) + np.float32(1.01))),
                int(np.float32((np.float32(y0) - np.float32(Y0)) / np.float32(DX) + np.float32(1.01))),
                int(np.float32((np.float32(zo) - np.float32(Z0)) / np.float32(DX) + np.float32(1.01)))]
        assert main == [1 + off[0], 1 + off[1], k_over + off[2]], (p, main)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_parent_lists_match_reference(p):
    g = np.load(GOLD)
    b = g[f"b{p}"].astype(np.int32)
    # the reference's codes: walls 0, fluid 1, copy 10+p at the overlap plane,
    # upstream 20+p at the far end
    assert set(np.unique(b).tolist()) == {-1, 0, 1, 10 + p, 20 + p}
    k_over, off = GEOMETRY[p]
    assert (b[:, :, k_over - 1] == 10 + p).any()
    domid, domlist = po.dom_setup(b)
    assert len(domlist) == int(g[f"dommax{p}"])
    bl, bn, vb, nfail = po.parent_bc_setup(b, p, k_over, off)
    np.testing.assert_array_equal(bl, g[f"bclist{p}"])
    np.testing.assert_array_equal(bn, g[f"bcnlist{p}"])
    np.testing.assert_array_equal(vb, g[f"vbcn{p}"])
    assert nfail == 0


def test_parent_tiers_differ_from_the_main_rule():
    # a parent's BC list takes codes >= 10 and its neighbour tiers stop at
    # B < 10: the main lattice's rule (>= 1000, < 1000) gives other lists on
    # the same box, so the goldens above pin the parent thresholds
    g = np.load(GOLD)
    b = g["b1"].astype(np.int32)
    bl_main, _, _ = po.bc_setup(b)
    assert len(bl_main) < len(g["bclist1"])


def _pulled_moments(b, rho, periodic_k):
    """Moments after the pull of rest-equilibrium populations w_c rho(x)
    (numpy, over the box): D = sum_c w_c rho(x - e_c), U = sum_c c w_c
    rho(x - e_c) / D, at every box node whose pull sources are in the box."""
    im, jm, km = b.shape
    D = np.zeros(b.shape)
    J = np.zeros((3,) + b.shape)
    for c in range(19):
        e = po.E[c]
        src = np.roll(rho, shift=(e[0], e[1], e[2]), axis=(0, 1, 2))  # rho(x - e_c)
        D += po.W[c] * src
        for a in range(3):
            J[a] += e[a] * po.W[c] * src
    if not periodic_k:
        D[:, :, 0] = D[:, :, -1] = np.nan  # sources outside the box
    return D, J / D


def test_parent_boundary_pass_against_brute_force():
    """One paired step (main lattice, then parent 1; ZebraflowWT.c:6101-6626)
    from rest-equilibrium populations with different density fields on the
    two lattices.  The parent's copy nodes must hold feq(D, U) of the main
    node over their column at the overlap plane, its upstream nodes
    feq(D_p at their neighbour, U of that main node): both recomputed here by
    brute force over the box, independently of the oracle's lists."""
    from zebraflow_amd import geometry
    from zebraflow_amd.engine import WALL_EQUILIBRIUM, lattice_config
    from tests.test_gpu_lbm import oracle_params
    bm = geometry.tube((16, 16, 24), axis=2, r=6.5, periodic=True)
    kp, k_main = 12, 10
    plane = bm[:, :, 0]
    bp = np.repeat(plane[:, :, None], kp, axis=2).astype(np.int32)
    bp[bp > 0] = 1
    bp[:, :, kp - 1][plane > 0] = 11
    bp[:, :, 0][plane > 0] = 21
    off = (0, 0, k_main - kp)
    cm = lattice_config(*bm.shape, 0.8, periodic=(0, 0, 1), wall_mode=WALL_EQUILIBRIUM)
    cp = lattice_config(*bp.shape, 0.8, wall_mode=WALL_EQUILIBRIUM)
    om = po.LbmOracle(bm, oracle_params(cm))
    op = po.LbmOracle(bp, oracle_params(cp), parent=(1, kp, off))
    ii, jj, kk = np.meshgrid(np.arange(1, 17), np.arange(1, 17), np.arange(1, 25), indexing="ij")
    rho_m = 1 + 0.01 * np.sin(ii) + 0.02 * np.cos(0.7 * jj) + 0.03 * kk / 24
    ii, jj, kk = np.meshgrid(np.arange(1, 17), np.arange(1, 17), np.arange(1, kp + 1), indexing="ij")
    rho_p = 1 - 0.02 * np.cos(ii) + 0.01 * np.sin(1.3 * jj) + 0.02 * kk / kp
    for o, rho in ((om, rho_m), (op, rho_p)):
        d = o.domlist - 1
        r = rho[d[:, 0], d[:, 1], d[:, 2]]
        o.set_populations(po.W[:, None] * r[None, :])
    po.step_with_parents(om, [op], 1)
    Dm, Um = _pulled_moments(bm, rho_m, True)
    Dp, _ = _pulled_moments(bp, rho_p, False)
    bl, bn = op.bclist, op.bcnlist
    checked = 0
    for row, nb in zip(bl, bn):
        i, j, k, code = row
        if code not in (11, 21):
            continue
        q = op.domid[i - 1, j - 1, k - 1] - 1
        v = (i - 1, j - 1, k_main - 1)  # main node over the column
        u = Um[:, v[0], v[1], v[2]]
        if code == 11:
            dbc = Dm[v]
        else:
            # the upstream neighbour: the first interior node among c = 1..18
            nbr = next((i + e[0], j + e[1], k + e[2]) for e in po.E[:18]
                       if 1 <= k + e[2] <= kp and 0 < bp[i + e[0] - 1, j + e[1] - 1, k + e[2] - 1] < 10)
            dbc = Dp[nbr[0] - 1, nbr[1] - 1, nbr[2] - 1]
        want = po.feq(np.array([dbc]), u[:, None])[:, 0]
        np.testing.assert_allclose(op.populations[:, q], want, rtol=0, atol=1e-15)
        checked += 1
    assert checked == 2 * int((plane > 0).sum())


def test_phase_split_equals_the_whole_step():
    """zfo_lbm_phase (0, 1, 2 in turn) is zfo_lbm_step, bit for bit."""
    from zebraflow_amd import geometry
    from zebraflow_amd.engine import lattice_config, reference_config
    from tests.test_gpu_lbm import oracle_params
    for cfg, b in ((reference_config(20, 18, 18), geometry.tube((20, 18, 18), axis=0, r=6.0,
                                                                 inlet=1001, outlet=1003)),
                   (lattice_config(16, 16, 24, 0.8, periodic=(0, 0, 1), body_force=(0, 0, 1e-4)),
                    geometry.tube((16, 16, 24), axis=2, r=6.5, periodic=True))):
        a, c = po.LbmOracle(b, oracle_params(cfg)), po.LbmOracle(b, oracle_params(cfg))
        rng = np.random.default_rng(7)
        f0 = po.feq(1 + 0.002 * rng.standard_normal(a.n), 0.01 * rng.standard_normal((3, a.n)))
        a.set_populations(f0)
        c.set_populations(f0)
        for it in range(1, 6):
            a.step(it)
            for ph in (0, 1, 2):
                c.phase(ph, it)
        np.testing.assert_array_equal(a.populations, c.populations)
        np.testing.assert_array_equal(a.U, c.U)
