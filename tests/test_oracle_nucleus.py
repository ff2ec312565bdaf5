"""Oracle of the nucleus network (ZebraflowWT.c:1903-1922 data, :2003-2050
and :2203-2221 equilibrium parameters, :7388-7466 spring forces,
:9519-9730 nucleus-membrane contact), pinned by:
  - the reference's own parameters for its zebrafish nucleus (goldens from
    CGSM_FILAMENT_PARAMETERS / CGSM_EQUILIBRIUM_METRICS run by
    oracle/ref_harness.c on NucleusNodes/Edges.dat);
  - the spring forces being the negative gradient of the WLC + POW energy
    (central differences) and summing to zero per cell;
  - the contact against a brute force over all triangles whose 19 probe
    points are written as barycentric weights of the triangle's vertices.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mesh_zfish.npz")
ES_REF = 3e-6  # Es, ZebraflowWT.c:128


def zfish():
    g = np.load(GOLD)
    return g


def test_nucleus_equilibrium_matches_reference():
    g = zfish()
    assert g["nuc_nodes"].shape == (121, 3) and g["nuc_edge"].shape == (544, 2)
    eq = po.nucleus_equilibrium(g["nuc_nodes"], g["nuc_edge"], po.reference_membrane_params(),
                                0.1 * ES_REF)
    for k in ("lmax", "plength", "kp"):
        np.testing.assert_allclose(eq[k], g["nuc_" + k].astype(np.float64), rtol=2e-6)
    # the network's modulus is the membrane's times 0.1: with Es itself kp
    # comes out about ten times larger (a dropped factor fails above)
    eq1 = po.nucleus_equilibrium(g["nuc_nodes"], g["nuc_edge"], po.reference_membrane_params(), ES_REF)
    assert np.median(eq1["kp"] / eq["kp"]) > 5


def spring_energy(x, edge, eq, mp):
    """WLC + POW (m = 2) filament energies whose negative gradients the
    reference's spring forces are (no clamp active)."""
    kT, L = mp["kT"], mp["len_scale"]
    d = x[edge[:, 0] - 1] - x[edge[:, 1] - 1]
    l = np.linalg.norm(d, axis=1) * L
    xr = l / eq["lmax"]
    u_wlc = kT * eq["lmax"] / (4 * eq["plength"]) * (3 * xr ** 2 - 2 * xr ** 3) / (1 - xr)
    u_pow = eq["kp"] / l
    return (u_wlc + u_pow).sum()


def test_nucleus_forces_are_energy_gradients_and_balance():
    g = zfish()
    xn0 = g["nuc_nodes"]
    mp = dict(po.reference_membrane_params(), len_scale=1.0, kT=0.0036 * 5e-3 * 0.3 ** 2)
    eq = po.nucleus_equilibrium(xn0, g["nuc_edge"], mp, 5e-4)
    rng = np.random.default_rng(3)
    xn = xn0 * 1.04 + 0.01 * rng.standard_normal(xn0.shape)
    F = po.nucleus_forces(np.stack([xn, xn + 5.0]), g["nuc_edge"], eq, mp)
    np.testing.assert_allclose(F[1], F[0], rtol=1e-12, atol=1e-12 * np.abs(F).max())
    f = F[0]
    scale = np.abs(f).max()
    assert scale > 0
    np.testing.assert_allclose(f.sum(0), 0, atol=1e-12 * scale * len(f))
    for node in rng.choice(len(xn), 10, replace=False):
        for ax in range(3):
            h = 1e-6
            xp, xm = xn.copy(), xn.copy()
            xp[node, ax] += h
            xm[node, ax] -= h
            gr = (spring_energy(xp, g["nuc_edge"], eq, mp) - spring_energy(xm, g["nuc_edge"], eq, mp)) / (2 * h)
            assert abs(f[node, ax] + gr) <= 1e-6 * scale + 1e-5 * abs(gr), (node, ax, f[node, ax], -gr)
    # at the rest network every filament sits at its equilibrium ratio, where
    # WLC and POW balance: no force
    f0 = po.nucleus_forces(xn0[None], g["nuc_edge"], eq, mp)[0]
    assert np.abs(f0).max() < 1e-9 * scale


# the 19 probe points (ZebraflowWT.c:9662-9699) as barycentric weights x 12
# of (p1, p2, p3): the centroid, the corner sub-triangle centroids, the
# midpoints of edge-midpoint pairs, then points halfway to those centroids
PROBES12 = np.array([[4, 4, 4], [8, 2, 2], [2, 8, 2], [2, 2, 8], [6, 3, 3], [3, 6, 3], [3, 3, 6],
                     [10, 1, 1], [7, 4, 1], [7, 1, 4], [4, 7, 1], [1, 10, 1], [1, 7, 4],
                     [4, 1, 7], [1, 4, 7], [1, 1, 10], [5, 5, 2], [2, 5, 5], [5, 2, 5]]) / 12.0


def brute_contact(xn, x, tris, list_r, near_r, contact, k):
    v = np.zeros_like(x)
    p = x[tris - 1]
    cen = p.mean(1)
    av = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    nrm = av / np.linalg.norm(av, axis=1, keepdims=True)
    probes = np.einsum("sk,tkd->tsd", PROBES12, p)
    for q in xn:
        dc = np.linalg.norm(cen - q, axis=1)
        for t in np.nonzero((dc < list_r) & (dc < near_r))[0]:
            d = np.linalg.norm(probes[t] - q, axis=1)
            rf = ((contact - d[d <= contact]) / contact * k).sum()
            for n in tris[t] - 1:
                v[n] += rf * nrm[t] / 3
    return v, nrm, cen


@pytest.mark.parametrize("shift", [0.15, -0.12])
def test_nucleus_contact_against_brute_force(shift):
    g = zfish()
    x, tris, xn0 = g["nodes"], g["tris"], g["nuc_nodes"]
    xn = xn0 + np.array([0.05, 0.0, shift])  # the nucleus pushed against one face
    vb, nrm, cen = brute_contact(xn, x, tris, 0.5, 1.0, 0.2, 4.1666667e-07)
    v, lost = po.nucleus_contact(xn[None], x[None], tris, np.ones(1, np.int32), cen[None],
                                 nrm[None], list_cap=1004)
    assert lost == 0
    scale = np.abs(vb).max()
    assert scale > 1e-8  # contacts exist
    np.testing.assert_allclose(v[0], vb, rtol=0, atol=1e-12 * scale)
    # the membrane is pushed outwards, on the face the nucleus was moved to
    moved = np.linalg.norm(v[0], axis=1) > 0
    assert (np.sign(x[moved, 2]) == np.sign(shift)).all()
    assert (v[0][moved, 2] * np.sign(shift) > 0).all()
    # a full list drops entries and reports them
    _, lost = po.nucleus_contact(xn[None], x[None], tris, np.ones(1, np.int32), cen[None],
                                 nrm[None], list_cap=1)
    assert lost > 0
