"""Transpose of the membrane forces (zf_membrane_vjp) through the C ABI.

For J = <lam, F(X)> on deformed, displaced cells: dJ/dX against central
differences (h = 1e-7 lu) of the FP64 oracle's forces
(oracle/zf_oracle_membrane.c) and of the device forward, 1e-6 relative;
the anti-sliver terms are switched on over the whole mesh in one case so
that their derivative is exercised.  dJ/d ln es and dJ/d ln kbend against
central differences (h = 1e-4 relative) of the lattice-unit set-up with the
modulus changed, 1e-8 relative.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.test_oracle_mesh import deformed, lattice_params
from zebraflow_amd.engine import Engine, lattice_config, lattice_membrane_config, membrane_config

pytestmark = pytest.mark.gpu
ES, KB = 5e-3, 2e-3


def cells(x1, ncell, seed):
    rng = np.random.default_rng(seed)
    return np.stack([x1 * (1 + 0.03 * rng.standard_normal()) + rng.uniform(-20, 20, 3) +
                     0.04 * rng.standard_normal(x1.shape) for _ in range(ncell)])


def forces(eng, X):
    eng.upload_cells(X)
    eng.membrane_forces()
    m = eng.membrane()
    assert (m["guard_fail"] == 0).all()
    return m["force"]


@pytest.mark.parametrize("sliver,blend", [(False, False), (True, False), (False, True)])
def test_position_gradient(sliver, blend):
    print(f"case membrane vjp positions, sliver terms {sliver}, smooth cut-off {blend}", flush=True)
    xyz, x1, tris = deformed(amp=0.1)
    # The force is discontinuous where a term switches (bending off above
    # |cos dihedral| = cos_cut, the anti-sliver thresholds); with 3840 edges
    # some dihedral lies within 1e-6 of the reference's 0.99.  The
    # thresholds are put in gaps of these cells' values (|cos| 0.985165 ..
    # 0.985216; node-to-centroid 0.4931..0.4956 and 0.0758..0.0876) so that
    # no difference below crosses one.
    # With the smooth cut-off (bending faded to zero over |cos| 0.97..0.99)
    # the bending term has no jump and the reference's 0.99 is kept.
    over = dict(cos_cut=0.99, cos_blend=0.02) if blend else dict(cos_cut=0.98519)
    if sliver:
        over.update(sliver_len=0.4943, sliver_len_min=0.08)
    mp, edge, eq, tae = lattice_params(xyz, tris, **over)
    X = cells(x1, 3, 7)
    eng = Engine(lattice_config(8, 8, 8, 0.8))
    eng.membrane_setup(membrane_config(dict(mp, es=ES), 3), tris, xyz)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(X.shape)
    forces(eng, X)
    gx, _ = eng.membrane_vjp(lam)
    J_dev = lambda x: (lam * forces(eng, x)).sum()
    J_ora = lambda x: (lam * po.membrane_forces(x, tris, edge, eq, tae, mp)["force"]).sum()
    h = 1e-7
    for J in (J_ora, J_dev):
        d = rng.standard_normal(X.shape)
        fd = (J(X + h * d) - J(X - h * d)) / (2 * h)
        np.testing.assert_allclose((gx * d).sum(), fd, rtol=1e-6)
    for c, n in ((0, 0), (2, 311)):
        for a in range(3):
            d = np.zeros_like(X)
            d[c, n, a] = 1.0
            fd = (J_dev(X + h * d) - J_dev(X - h * d)) / (2 * h)
            np.testing.assert_allclose(gx[c, n, a], fd, rtol=1e-5, atol=1e-7 * np.abs(gx).max())
    eng.close()


def test_modulus_gradients():
    print("case membrane vjp moduli", flush=True)
    xyz, x1, tris = deformed(amp=0.1)
    X = cells(x1, 2, 9)
    lam = np.random.default_rng(4).standard_normal(X.shape)
    eng = Engine(lattice_config(8, 8, 8, 0.8))

    def J(es, kb):
        eng.membrane_setup(lattice_membrane_config(xyz, tris, 2, es=es, kb=kb), tris, xyz)
        return (lam * forces(eng, X)).sum()

    J(ES, KB)
    _, dp = eng.membrane_vjp(lam)
    h = 1e-4
    fd_es = (J(ES * (1 + h), KB) - J(ES * (1 - h), KB)) / (2 * h)
    fd_kb = (J(ES, KB * (1 + h)) - J(ES, KB * (1 - h))) / (2 * h)
    print("adjoint", dp, "fd", fd_es, fd_kb, flush=True)
    assert abs(dp[1]) > 1e-3 * abs(dp[0])
    np.testing.assert_allclose(dp, [fd_es, fd_kb], rtol=1e-8)
    eng.close()
