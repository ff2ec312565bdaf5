"""Parent (feeder) lattices on the GPU: their lists against the reference's
own (LBM_PARENTBC_SETUP, ZebraflowWT.c:1015-1350, through the golden
vectors of the three parent lattices it ships), and the coupled step of a
main lattice with a parent (ZebraflowWT.c:6101-6626: the parent's stream +
collide :6148-6286 and boundary pass :6436-6594) against the oracle."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from tests.test_gpu_lbm import oracle_params
from zebraflow_amd.engine import (STORE_F32DEV, STORE_F64, WALL_EQUILIBRIUM, WALL_HALFWAY_BB, Engine,
                                  ZfError, lattice_config)
from tests.test_oracle_parents import GEOMETRY, GOLD

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p", [1, 2, 3])
def test_parent_lists_match_reference(p):
    g = np.load(GOLD)
    b = g[f"b{p}"].astype(np.int32)
    print(f"case parent {p} lists box {b.shape}", flush=True)
    k_over, off = GEOMETRY[p]
    eng = Engine(lattice_config(*b.shape, 0.8, storage=STORE_F32DEV))
    eng.set_parent(p, k_over, off)
    eng.upload_lattice(b)
    dl, bl, bn = eng.lists()
    vb = eng.parent_lists()
    eng.close()
    assert len(dl) == int(g[f"dommax{p}"])
    np.testing.assert_array_equal(bl, g[f"bclist{p}"])
    np.testing.assert_array_equal(bn, g[f"bcnlist{p}"])
    np.testing.assert_array_equal(vb, g[f"vbcn{p}"])


def parent_of(bm, kp, p=1):
    """A parent lattice extruded from the main lattice's first plane, as the
    reference builds its parents (ZebraflowWT.c:590-606): fluid 1, the far
    end 20+p, the overlap plane 10+p (k = kp for parent 1, k = 1 else)."""
    plane = bm[:, :, 0]
    b = np.repeat(plane[:, :, None], kp, axis=2).astype(np.int32)
    fl = b > 0
    b[fl] = 1
    over, far = (kp - 1, 0) if p == 1 else (0, kp - 1)
    b[:, :, over][plane > 0] = 10 + p
    b[:, :, far][plane > 0] = 20 + p
    return b


def run_coupled(store, wall, p, steps=40):
    bm = geometry.tube((16, 16, 24), axis=2, r=6.5, periodic=True)
    kp = 12
    bp = parent_of(bm, kp, p)
    k_over = kp if p == 1 else 1
    off = (0, 0, 10 - k_over)  # overlap plane over main k = 10
    cm = lattice_config(*bm.shape, 0.8, storage=store, periodic=(0, 0, 1), body_force=(0, 0, 1e-4),
                        wall_mode=wall)
    cp = lattice_config(*bp.shape, 0.8, storage=store, wall_mode=wall)
    print(f"case coupled store {store} wall {wall} parent {p} main {bm.shape} parent {bp.shape}",
          flush=True)
    om = po.LbmOracle(bm, oracle_params(cm))
    op = po.LbmOracle(bp, oracle_params(cp), parent=(p, k_over, off))
    em, ep = Engine(cm), Engine(cp)
    em.upload_lattice(bm)
    ep.set_parent(p, k_over, off)
    ep.upload_lattice(bp)
    assert em.dommax == om.n and ep.dommax == op.n
    ep.parent_link(em)
    rng = np.random.default_rng(20261204)
    for o, e in ((om, em), (op, ep)):
        rho = 1 + 0.002 * rng.standard_normal(o.n)
        u = 0.01 * rng.standard_normal((3, o.n))
        f0 = po.feq(rho, u) * (1 + 0.01 * rng.standard_normal((19, o.n)))
        o.set_populations(f0)
        e.upload_populations(f0)
    force = 2e-6 * rng.standard_normal((3, om.n))
    om.force[...] = force
    em.upload_force(force)
    for it in range(1, steps + 1):
        rbc = 1 if it == steps // 2 else 0
        po.step_with_parents(om, [op], it, rbc_step=rbc)
        em.step(it, rbc_step=rbc)
        ep.step(it)
    out = (em.populations(), ep.populations(), ep.macro())
    ep.close()
    em.close()
    return om, op, out


@pytest.mark.parametrize("wall", [WALL_EQUILIBRIUM, WALL_HALFWAY_BB])
@pytest.mark.parametrize("p", [1, 2])
def test_coupled_step_f64_matches_oracle(wall, p):
    om, op, (fm, fp, mp) = run_coupled(STORE_F64, wall, p)
    # the coupled nodes did change: the upstream and copy planes carry the
    # main lattice's flow (about 0.01 after the forced start)
    bc = op.bclist
    coupled = (bc[:, 3] == 10 + p) | (bc[:, 3] == 20 + p)
    assert coupled.sum() > 50
    np.testing.assert_allclose(fm, om.populations, rtol=0, atol=1e-13)
    # the parent's coupled nodes take moments recovered from stored
    # populations (sum f' c = rho u + DT F / 2): equal to rounding
    np.testing.assert_allclose(fp, op.populations, rtol=0, atol=1e-13)
    np.testing.assert_allclose(mp[0], op.D, rtol=0, atol=1e-13)
    np.testing.assert_allclose(mp[3], op.W, rtol=0, atol=1e-13)
    assert np.abs(op.W).max() > 1e-4


def test_coupled_step_f32dev_matches_oracle():
    om, op, (fm, fp, mp) = run_coupled(STORE_F32DEV, WALL_EQUILIBRIUM, 1)
    np.testing.assert_allclose(fm, om.populations, rtol=0, atol=2e-7)
    np.testing.assert_allclose(fp, op.populations, rtol=0, atol=2e-7)


def test_coupled_step_f16dev_matches_oracle():
    # FP16 deviations: the tolerance of the plain FP16 form (velocity within
    # 1e-2 relative L2, density within 2e-3)
    from zebraflow_amd.engine import STORE_F16DEV
    om, op, (fm, fp, mp) = run_coupled(STORE_F16DEV, WALL_EQUILIBRIUM, 1)
    uo = np.stack([op.U, op.V, op.W])
    assert np.linalg.norm(mp[1:4] - uo) / np.linalg.norm(uo) <= 1e-2
    np.testing.assert_allclose(mp[0], op.D, rtol=0, atol=2e-3)


def test_link_rejects_a_parent_outside_the_main_box():
    bm = geometry.tube((16, 16, 24), axis=2, r=6.5, periodic=True)
    bp = parent_of(bm, 12)
    em = Engine(lattice_config(*bm.shape, 0.8, periodic=(0, 0, 1)))
    em.upload_lattice(bm)
    ep = Engine(lattice_config(*bp.shape, 0.8))
    ep.set_parent(1, 12, (0, 0, 20))  # overlap plane over main k = 32 > 24
    ep.upload_lattice(bp)
    with pytest.raises(ZfError, match="outside the main box"):
        ep.parent_link(em)
    ep.close()
    em.close()


def test_two_parents_share_one_main_lattice():
    """Parents 1 and 2 on one main lattice (the reference runs three): their
    export slots are shared, and both match the oracle's paired step."""
    bm = geometry.tube((16, 16, 24), axis=2, r=6.5, periodic=True)
    kp = 12
    cm = lattice_config(*bm.shape, 0.8, storage=STORE_F64, periodic=(0, 0, 1), body_force=(0, 0, 1e-4),
                        wall_mode=WALL_EQUILIBRIUM)
    om = po.LbmOracle(bm, oracle_params(cm))
    em = Engine(cm)
    em.upload_lattice(bm)
    parents = []
    # parent 1 over main plane 10 (its k = 12), parent 2 over main plane 10 too
    # (its k = 1): the two share every exported node
    for p, k_over in ((1, kp), (2, 1)):
        bp = parent_of(bm, kp, p)
        off = (0, 0, 10 - k_over)
        cp = lattice_config(*bp.shape, 0.8, storage=STORE_F64, wall_mode=WALL_EQUILIBRIUM)
        op = po.LbmOracle(bp, oracle_params(cp), parent=(p, k_over, off))
        ep = Engine(cp)
        ep.set_parent(p, k_over, off)
        ep.upload_lattice(bp)
        ep.parent_link(em)
        parents.append((op, ep))
    print("case two parents on one main lattice", flush=True)
    rng = np.random.default_rng(20261205)
    for o, e in [(om, em)] + parents:
        rho = 1 + 0.002 * rng.standard_normal(o.n)
        u = 0.01 * rng.standard_normal((3, o.n))
        f0 = po.feq(rho, u)
        o.set_populations(f0)
        e.upload_populations(f0)
    for it in range(1, 31):
        po.step_with_parents(om, [op for op, _ in parents], it)
        em.step(it)
        for _, ep in parents:
            ep.step(it)
    np.testing.assert_allclose(em.populations(), om.populations, rtol=0, atol=1e-13)
    for op, ep in parents:
        np.testing.assert_allclose(ep.populations(), op.populations, rtol=0, atol=1e-13)
    # destroying the main lattice first unlinks its parents, which then step
    # on their own
    em.close()
    for _, ep in parents:
        ep.step(31)
        assert np.isfinite(ep.populations()).all()
        ep.close()
