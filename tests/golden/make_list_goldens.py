"""Write golden vectors for the index lists (a1 DOMLIST/DOMID, a3 BCLIST/BCNLIST).

The expected outputs come from the reference's own LBM_DOM_SETUP
(ZebraflowWT.c:1481-1502) and LBM_BC_SETUP (ZebraflowWT.c:1506-1589), compiled
from the reference source by ``oracle/Makefile`` (target ``ref``) and driven by
``oracle/ref_harness.c``.  Run in the development container only:

    make -C oracle ref && python tests/golden/make_list_goldens.py

Each fixture ``lists_<name>.npz`` holds the input B field (int8/int16) and the
reference's DOMLIST, BCLIST and BCNLIST (1-based, as the reference stores them).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")

E = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [-1, 1, 0], [1, -1, 0], [1, 0, 1], [-1, 0, -1],
              [-1, 0, 1], [1, 0, -1], [0, 1, 1], [0, -1, -1], [0, -1, 1], [0, 1, -1]])


def wall_shell(fluid):
    """Non-fluid nodes with a D3Q19 neighbour in `fluid` (no wrap)."""
    im, jm, km = fluid.shape
    pad = np.zeros((im + 2, jm + 2, km + 2), bool)
    pad[1:-1, 1:-1, 1:-1] = fluid
    sh = np.zeros_like(fluid)
    for e in E:
        sh |= pad[1 + e[0]:1 + e[0] + im, 1 + e[1]:1 + e[1] + jm, 1 + e[2]:1 + e[2] + km]
    return sh & ~fluid


def tube_field(im, jm, km, r, inlet_code, outlet_code):
    i, j, k = np.meshgrid(np.arange(im), np.arange(jm), np.arange(km), indexing="ij")
    cj, ck = (jm - 1) / 2.0, (km - 1) / 2.0
    fluid = (j - cj) ** 2 + (k - ck) ** 2 <= r * r
    b = np.full((im, jm, km), -1, np.int32)
    b[wall_shell(fluid)] = 0
    b[fluid] = 1
    b[0][fluid[0]] = inlet_code
    b[-1][fluid[-1]] = outlet_code
    return b


def pair_field(im, jm, km):
    i, j, k = np.meshgrid(np.arange(im), np.arange(jm), np.arange(km), indexing="ij")
    r = 4.0
    c1 = (j - 7.0 - 1.5 * np.sin(2 * np.pi * i / im)) ** 2 + (k - 11.0) ** 2 <= r * r
    c2 = (j - 17.0) ** 2 + (k - 11.0 - 1.5 * np.cos(2 * np.pi * i / im)) ** 2 <= r * r
    conn = np.zeros_like(c1)
    for x0 in (10, 25):
        conn |= ((i - x0) ** 2 + (k - 11.0) ** 2 <= 2.2 ** 2) & (j >= 7) & (j <= 17)
    fluid = c1 | c2 | conn
    b = np.full((im, jm, km), -1, np.int32)
    b[wall_shell(fluid)] = 0
    b[fluid] = 1
    b[0][c1[0]] = 2001
    b[-1][c1[-1]] = 3001
    b[0][c2[0]] = 3001
    b[fluid & (i == 20) & (j > 10) & (j < 14)] = 2   # a tau-class change
    return b


def bc_has_neighbour(b):
    """True where a BC node (B==0 or >=1000) finds some neighbour tier."""
    im, jm, km = b.shape
    ok = np.zeros(b.shape, bool)
    pad = np.full((im + 2, jm + 2, km + 2), -1, np.int32)
    pad[1:-1, 1:-1, 1:-1] = b
    for e in E:
        nb = pad[1 + e[0]:1 + e[0] + im, 1 + e[1]:1 + e[1] + jm, 1 + e[2]:1 + e[2] + km]
        ok |= ((nb > 0) & (nb < 1000)) | (nb > 1000) | (nb == 0)
    return ok


def random_field(shape, seed):
    rng = np.random.default_rng(seed)
    vals = np.array([-1, 0, 1, 2, 11, 1000, 1001, 2001, 3001, 4998])
    prob = np.array([0.30, 0.25, 0.25, 0.05, 0.02, 0.02, 0.03, 0.03, 0.03, 0.02])
    b = rng.choice(vals, size=shape, p=prob).astype(np.int32)
    # the reference leaves stale neighbour indices for a BC node with no
    # neighbour (ZebraflowWT.c:1575-1581); drop such nodes to keep the goldens
    # free of that undefined case.
    while True:
        isbc = (b == 0) | (b >= 1000)
        bad = isbc & ~bc_has_neighbour(b)
        if not bad.any():
            return b
        b[bad] = -1


def run_ref(b):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array(b.shape, np.int32).tofile(f)
            b.astype(np.int32).tofile(f)
        subprocess.run([HARNESS, "lists", fin, fout], check=True,
                       stdout=subprocess.DEVNULL)
        raw = np.fromfile(fout, np.int32)
    n = raw[0]
    domlist = raw[1:1 + 3 * n].reshape(n, 3)
    o = 1 + 3 * n
    m = raw[o]
    bclist = raw[o + 1:o + 1 + 4 * m].reshape(m, 4)
    bcnlist = raw[o + 1 + 4 * m:o + 1 + 7 * m].reshape(m, 3)
    return domlist, bclist, bcnlist


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -C oracle ref")
    cases = {
        "tube": tube_field(20, 18, 18, 5.0, 1001, 2001),
        "tube_outlet": tube_field(12, 15, 13, 4.2, 2001, 3001),
        "pair": pair_field(36, 24, 22),
        "random_a": random_field((14, 12, 10), 20261203),
        "random_b": random_field((9, 17, 11), 20261204),
    }
    for name, b in cases.items():
        domlist, bclist, bcnlist = run_ref(b)
        out = os.path.join(HERE, f"lists_{name}.npz")
        np.savez_compressed(out, b=b.astype(np.int16), domlist=domlist,
                            bclist=bclist, bcnlist=bcnlist)
        print(f"{name}: box {b.shape} DOMMAX {len(domlist)} BCMAX {len(bclist)} -> {out}")


if __name__ == "__main__":
    main()
