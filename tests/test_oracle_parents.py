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
