"""Brute-force checks of the geometric facts the device-side binning relies
on (no GPU):

- spreading bins (zf_ibm.hip, spread_bin and kSpreadSubs): a membrane node
  binned by (tile, octant) lies in one of the 64 bins a tile gathers from
  whenever it is within the kernel's reach of that tile (2 lattice units
  along every axis of some lattice node of the tile), so the gather misses
  no contribution;
- the cell list's nearest-node search (zf_cells.hip, k_tri_bin): every
  lattice point outside the 4^3 block of points ci-1..ci+2 around the unit
  cell holding a centroid is at least 2 dx from it, so a listed block
  point nearer than that is the nearest in the whole window.
"""
import itertools

import numpy as np

TILE = 4


def spread_bin(u):
    """(tile, octant) of a 0-based lattice coordinate triple u, as
    spread_bin: the octant bit is set in the upper half (>= 2) of the tile
    along that axis."""
    t = np.floor(u / TILE).astype(int)
    h = (u - TILE * t >= 2.0).astype(int)
    return t, h


def gathered_bins(tile):
    """The bins a tile gathers from: the 27 neighbouring tiles, and in each
    the octants in the halves facing the tile (both halves along an axis
    where the neighbour is level with it) -- the device table."""
    out = set()
    for d in itertools.product((-1, 0, 1), repeat=3):
        for o in itertools.product((0, 1), repeat=3):
            if all(dd == 0 or (dd < 0 and hh == 1) or (dd > 0 and hh == 0) for dd, hh in zip(d, o)):
                out.add((tuple(np.array(tile) + np.array(d)), o))
    return out


def test_gathered_bins_are_64():
    assert len(gathered_bins((5, 5, 5))) == 64


def test_every_node_in_reach_of_a_tile_is_in_its_gathered_bins():
    rng = np.random.default_rng(20261204)
    u = rng.uniform(8.0, 24.0, (4000, 3))
    # nodes exactly on half and tile boundaries too
    u[:500] = np.round(u[:500] * 2) / 2
    for p in u:
        t, h = spread_bin(p)
        key = (tuple(t), tuple(h))
        # every tile with a lattice node within 2 along every axis of p
        lo = np.ceil(p - 2.0).astype(int)
        hi = np.floor(p + 2.0).astype(int)
        tiles = {tuple(v // TILE for v in x)
                 for x in itertools.product(*[range(a, b + 1) for a, b in zip(lo, hi)])}
        for tile in tiles:
            assert key in gathered_bins(tile), (p, tile)


def test_points_outside_the_block_are_at_least_two_dx_away():
    rng = np.random.default_rng(7)
    for x in rng.uniform(3.0, 9.0, (2000, 3)):
        c = np.floor(x).astype(int)  # 0-based lower corner of the unit cell
        block = [range(ci - 1, ci + 3) for ci in c]
        for p in itertools.product(*[range(ci - 4, ci + 6) for ci in c]):
            if all(pi in b for pi, b in zip(p, block)):
                continue
            assert np.linalg.norm(np.array(p) - x) >= 2.0 - 1e-12


def test_lexicographic_least_equals_the_in_order_strict_scan():
    """The wall search (k_wall_repulsion) keeps the least (distance, visit
    order) of the points of the listed wall nodes, evaluated in any order,
    with a wall node repeated along the walk listed again; the reference
    scans the walk in order, keeping a point only when strictly nearer and
    skipping a wall node equal to the one just seen.  Both pick the same
    triangle, ties of equal distances included."""
    rng = np.random.default_rng(11)
    for _ in range(300):
        nwall = 12
        # each wall node: up to 12 triangles of 4 points, distances drawn
        # from a small set so that ties are frequent
        tris = {w: [(int(rng.integers(1, 50)), rng.choice([0.5, 0.75, 1.0, 1.25, 2.0], 4))
                    for _ in range(int(rng.integers(1, 13)))] for w in range(1, nwall + 1)}
        walk = rng.integers(0, nwall + 1, 60)  # W_index entries along the window (0: none)
        search = 1.5
        # reference: in order, strict <, consecutive repeats skipped
        mind, win, last = search, 0, 0
        for wn in walk:
            if wn <= 0 or wn == last:
                continue
            last = wn
            for tw, ds in tris[wn]:
                for d in ds:
                    if d < mind:
                        mind, win = d, tw
        # lexicographic least over every listed occurrence, shuffled
        cands = []
        for pos, wn in enumerate(walk):
            if wn <= 0:
                continue
            for l, (tw, ds) in enumerate(tris[wn]):
                for sv, d in enumerate(ds):
                    if d < search:
                        cands.append((d, pos * 64 + 4 * (l + 1) + sv, tw))
        rng.shuffle(cands)
        best = min(cands, key=lambda c: (c[0], c[1])) if cands else None
        assert (best[2] if best else 0) == win
