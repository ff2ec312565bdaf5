"""Cells on a parent lattice and their cloning into the main lattice, on the
GPU through the C ABI (zf_cell_wrap_setup, zf_step_linked), against the
composed FP64 oracle: the coupled LBM step of the main lattice and its
parent (oracle/zf_oracle_lbm.inc), then per RBC iteration the parent's
stages with its periodic window (oracle/zf_oracle_wrap.c: cell list,
interpolation, wall and cell-cell repulsion, spreading; membrane forces on
the unwrapped positions), the main lattice's stages with its clones
(oracle/zf_oracle_*.c), the main lattice's update (clone band nodes with
the parent cell's velocity, release), the parent's wrapped update and the
cloning (oracle/pyoracle.py), as ZebraflowWT.c:5646-12403 orders them.

The cases: a main vessel (16x16x40, periodic, forced) fed by parent 1
(16x16x41, window [-38, 2], L = 40 lu, cells leaving through its upper
end, copy plane over main plane 3) or by a parent of the other direction
(window [37, 77], cells leaving through its lower end into main plane 38,
as the reference's parents 2 and 3); three parent cells (one about to
leave through the exit, one in the middle, one at the entry), one main
cell and three free main slots; the flow carries the first parent cell
across the exit, where it is cloned, its clone released into the main
lattice, and the parent cell wraps round to re-enter the window."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.test_gpu_lbm import oracle_params
from tests.test_gpu_parent import parent_of
from zebraflow_amd import geometry
from zebraflow_amd.engine import (STORE_F64, Engine, ZfError, lattice_config, lattice_ibm_config,
                                  lattice_membrane_config, lattice_repulsion_config)

pytestmark = pytest.mark.gpu

KP = 41                            # parent planes (window L = 40 lu, DX 1)
DT_ADV, BAND, SPLIT = 25.0, 5.0, 20.0


class Case:
    """Parent p = 1: exit at its upper end over main plane 3; p = 2: exit at
    its lower end over main plane 38 (main origin 0)."""

    def __init__(self, p):
        self.p = p
        self.dir = 1 if p == 1 else -1
        self.k_over = KP if p == 1 else 1
        kc = 3 if p == 1 else 38
        self.off = (0, 0, kc - self.k_over)
        self.z0 = float(kc - KP) if p == 1 else float(kc - 1)
        self.zm = self.z0 + (KP - 1)
        self.exit = self.zm if p == 1 else self.z0
        self.roi = (self.exit, 1e3) if p == 1 else (-1e3, self.exit)
        self.id_base = 0 if p == 1 else 188


def macro_of(f):
    d = f.sum(0)
    return [(po.E[:, a] @ f) / d for a in range(3)]


def nucleus_rest():
    """The reference's 121-node nucleus network, scaled into the test cells."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mesh_zfish.npz"))
    return 0.55 * g["nuc_nodes"], g["nuc_edge"]


def setup(cs, nuc=False):
    bm = geometry.tube((16, 16, 40), axis=2, r=6.5, periodic=True)
    bp = parent_of(bm, KP, cs.p)
    xyz, tris = geometry.rbc_mesh(radius=2.6, level=2)
    cm = lattice_config(*bm.shape, 0.8, storage=STORE_F64, periodic=(0, 0, 1),
                        body_force=(0, 0, 2e-5 * cs.dir))
    cp = lattice_config(*bp.shape, 0.8, storage=STORE_F64)
    em, ep = Engine(cm), Engine(cp)
    em.upload_lattice(bm)
    ep.set_parent(cs.p, cs.k_over, cs.off)
    ep.upload_lattice(bp)
    ep.parent_link(em)
    om = po.LbmOracle(bm, oracle_params(cm))
    op = po.LbmOracle(bp, oracle_params(cp), parent=(cs.p, cs.k_over, cs.off))
    u0 = np.zeros((3, 1))
    u0[2] = 0.02 * cs.dir
    for o, e in ((om, em), (op, ep)):
        f0 = po.feq(np.ones(o.n), np.repeat(u0, o.n, axis=1))
        o.set_populations(f0)
        e.upload_populations(f0)
    # cells: parent slots 0-2, main slot 0 live and 1-3 free (far away)
    c = np.array([7.5, 7.5, 0.0])
    ent = cs.z0 + 3.0 if cs.dir > 0 else cs.zm - 3.0  # the entry end
    pcells = np.stack([xyz + c + [0, 0, cs.exit - 3.3 * cs.dir],
                       xyz + c + [0.3, -0.2, 0.5 * (cs.z0 + cs.zm)], xyz + c + [-0.2, 0.3, ent]])
    mcells = np.stack([xyz + c + [0, 0, 20.0]] + [xyz + [100.0 + 20 * k, 0, 0] for k in range(3)])
    nuclei = None
    if nuc:
        xn0, ned = nucleus_rest()
        nuclei = [np.stack([xn0 + cc.mean(0) for cc in cells]) for cells in (pcells, mcells)]
    for e, cells, z0, k in ((ep, pcells, cs.z0, 0), (em, mcells, 0.0, 1)):
        e.membrane_setup(lattice_membrane_config(xyz, tris, len(cells)), tris, xyz)
        if nuc:
            e.nucleus_setup(ned, xn0, list_interval=1)
        e.upload_cells(cells)
        if nuc:
            e.upload_nuclei(nuclei[k])
        ic = lattice_ibm_config(rbc_interval=1)
        ic.z0 = z0
        ic.dt_advance = DT_ADV
        e.ibm_setup(ic)
        e.repulsion_setup(lattice_repulsion_config())
    em.wall_setup(*geometry.wall_mesh(bm))
    ep.wall_setup(*geometry.wall_mesh(bp, z0=cs.z0))
    em.upload_cell_status(np.array([1, -5000, -5000, -5000], np.int32))
    ep.cell_wrap_setup(em, cs.p, split_span=SPLIT, clone_band=BAND, roi_lo=cs.roi[0],
                       roi_hi=cs.roi[1])
    return bm, bp, xyz, tris, em, ep, om, op, pcells, mcells, nuclei


class Oracle:
    """The composed FP64 oracle of one RBC iteration of the pair."""

    def __init__(self, cs, bm, bp, xyz, tris, om, op, pcells, mcells, nuclei=None):
        from tests.test_oracle_mesh import lattice_params
        self.cs = cs
        self.om, self.op, self.tris = om, op, tris
        self.mp, self.edge, self.eq, self.tae = lattice_params(xyz, tris)
        rc = lattice_repulsion_config()
        self.r = po.rep_params(bin_cutoff=rc.bin_cutoff, k_cell=rc.k_cell, cutoff=rc.cutoff,
                               search=rc.search, k_wall=rc.k_wall, wall_d0=rc.wall_d0,
                               wall_dmax=rc.wall_dmax, wall_search=rc.wall_search)
        self.pm = po.ibm_params(bm.shape, dt_advance=DT_ADV, vmax=1.0)
        self.pp = po.ibm_params(bp.shape, z0=cs.z0, dt_advance=DT_ADV, vmax=1.0)
        self.w = po.wrap_of(self.pp, cs.dir)
        self.bm, self.bp = bm, bp
        self.dm, _ = po.dom_setup(bm)
        self.dp, _ = po.dom_setup(bp)
        self.wall_m = geometry.wall_mesh(bm)
        self.wall_p = geometry.wall_mesh(bp, z0=cs.z0)
        self.nwt_m, self.wi_m = po.wall_tables(self.pm, *self.wall_m[:2])
        self.nwt_p, self.wi_p = po.wall_tables(self.pp, *self.wall_p[:2])
        self.xp, self.xm = pcells.copy(), mcells.copy()
        self.sp = np.full(len(pcells), 2, np.int32)
        self.rp = np.zeros(len(pcells), np.int32)
        self.sm = np.array([1, -5000, -5000, -5000], np.int32)
        self.clones = 0
        self.xnp = self.xnm = None
        if nuclei is not None:
            xn0, self.ned = nucleus_rest()
            self.neq = po.nucleus_equilibrium(xn0, self.ned, self.mp, 0.1 * 5e-3)
            self.xnp, self.xnm = nuclei[0].copy(), nuclei[1].copy()

    def forces(self, x):
        return po.membrane_forces(x, self.tris, self.edge, self.eq, self.tae, self.mp)

    def iteration(self, it):
        gamma = 0.5
        po.step_with_parents(self.om, [self.op], it, rbc_step=1, parent_rbc_step=1)
        # parent: unwrapped membrane stage, REGRESSFLAG, cell list, interpolation,
        # wall and cell-cell repulsion, spreading
        w, pp, xp, sp, xnp = self.w, self.pp, self.xp, self.sp, self.xnp
        nuc = xnp is not None
        zmin, zmax = po.wrap_span(xp, xnp, sp)
        zu, _ = po.wrap_unwrap(w, xp[..., 2], sp, zmin, zmax, split=SPLIT)
        xu = xp.copy()
        xu[..., 2] = zu
        gp = self.forces(xu)
        if nuc:
            xnu = xnp.copy()
            xnu[..., 2] = po.wrap_unwrap(w, xnp[..., 2], sp, zmin, zmax, split=SPLIT)[0]
            fnp = po.nucleus_forces(xnu, self.ned, self.neq, self.mp)
        po.wrap_regress_reset(w, sp, self.rp, zmin, zmax, gp["cell_centroid"][:, 2], split=SPLIT)
        binp = po.wrap_cell_list(pp, w, self.bp, self.dp, gp["tri_centroid"], sp)
        vp = po.wrap_interpolate(pp, w, self.bp, self.dp, macro_of(self.op.populations), xp, sp)
        vrp, _ = po.wrap_wall_repulsion(pp, w, self.r, self.wi_p, self.wall_p[0], self.wall_p[1],
                                        self.wall_p[2], self.nwt_p, sp, xp)
        if nuc:
            vnp = po.wrap_interpolate(pp, w, self.bp, self.dp, macro_of(self.op.populations), xnp, sp)
            po.nucleus_contact(xnu, xu, self.tris, sp, gp["tri_centroid"], gp["tri_normal"],
                               vrep=vrp)
        po.wrap_cell_repulsion(pp, w, self.r, self.bp, self.dp, self.tris, sp, xp, gp["node_normal"],
                               gp["tri_centroid"], gp["tri_normal"], binp, vrep=vrp)
        vp += vrp
        po.wrap_spread(pp, w, self.bp, self.dp, self.op.n, xp, gamma * vrp + gp["force"], sp,
                       force=self.op.force)
        if nuc:
            po.wrap_spread(pp, w, self.bp, self.dp, self.op.n, xnp, fnp, sp, force=self.op.force)
        # main: clones take part as live cells
        xm, sm = self.xm, self.sm
        clone = (sm < 0) & (sm > -4998)
        s1 = np.where(clone, 1, sm).astype(np.int32)
        gm = self.forces(xm)
        binm = po.cell_list(self.pm, self.bm, self.dm, gm["tri_centroid"], s1, self.r.bin_cutoff)
        vm = po.ibm_interpolate(self.pm, self.bm, self.dm, macro_of(self.om.populations), xm)
        vrm = po.wall_repulsion(self.pm, self.r, self.wi_m, *self.wall_m, self.nwt_m, s1, xm)[0]
        xnm = self.xnm
        if nuc:
            fnm = po.nucleus_forces(xnm, self.ned, self.neq, self.mp)
            vnm = po.ibm_interpolate(self.pm, self.bm, self.dm, macro_of(self.om.populations), xnm)
            po.nucleus_contact(xnm, xm, self.tris, s1, gm["tri_centroid"], gm["tri_normal"],
                               vrep=vrm)
        vrm += po.cell_repulsion(self.pm, self.r, self.bm, self.dm, self.tris, s1, xm,
                                 gm["node_normal"], gm["tri_centroid"], gm["tri_normal"], binm)
        vm += vrm
        ffl = gamma * vrm + gm["force"]
        keep = np.repeat((s1 > -4998)[:, None], xm.shape[1], axis=1)
        for c in np.nonzero(clone)[0]:
            keep[c] &= ~po.clone_band(w, xm[c, :, 2], BAND)
        po.ibm_spread(self.pm, self.bm, self.dm, self.om.n, xm[keep], ffl[keep], force=self.om.force)
        if nuc:
            keepn = np.repeat((s1 > -4998)[:, None], xnm.shape[1], axis=1)
            for c in np.nonzero(clone)[0]:
                keepn[c] &= ~po.clone_band(w, xnm[c, :, 2], BAND)
            po.ibm_spread(self.pm, self.bm, self.dm, self.om.n, xnm[keepn], fnm[keepn],
                          force=self.om.force)
        # main update: released cells, then the clones (band with the parent)
        before = sm.copy()
        po.advance(self.pm, xm, vm, sm)
        if nuc:
            for c in np.nonzero((before >= 0) & (sm >= 0))[0]:
                xnm[c] += vnm[c] * DT_ADV
        lo, hi = self.cs.roi
        for c in np.nonzero(clone)[0]:
            s = -sm[c] - 1 - self.cs.id_base
            band = po.clone_band(w, xm[c, :, 2], BAND)
            vm[c][band] = vp[s][band]
            xm[c] += vm[c] * DT_ADV
            if nuc:
                bn = po.clone_band(w, xnm[c, :, 2], BAND)
                vnm[c][bn] = vnp[s][bn]
                xnm[c] += vnm[c] * DT_ADV
            if xm[c, :, 2].min() > lo and xm[c, :, 2].max() < hi:
                sm[c] = 1
        # parent update with the wrap, then the cloning
        flags = po.wrap_advance(w, DT_ADV, xp, vp, sp, self.rp, xyzn=xnp,
                                veln=vnp if nuc else None)
        self.clones += len(po.clone_cells(w, self.cs.id_base, sp, self.rp, flags, xp, vp, sm, xm,
                                          vm, pxyzn=xnp, pveln=vnp if nuc else None, mxyzn=xnm,
                                          mveln=vnm if nuc else None))


@pytest.mark.parametrize("p,nuc", [(1, False), (2, False), (1, True)])
def test_parent_cells_clone_into_the_main_lattice(p, nuc):
    print(f"case parent {p} cells and cloning (nuclei {nuc}), 3 parent cells, 1 + 3 free main "
          "slots", flush=True)
    cs = Case(p)
    bm, bp, xyz, tris, em, ep, om, op, pcells, mcells, nuclei = setup(cs, nuc)
    o = Oracle(cs, bm, bp, xyz, tris, om, op, pcells, mcells, nuclei)
    crossed = cloned = False
    for it in range(1, 31):
        em.step_linked(it)
        o.iteration(it)
        xp, _, sp = ep.cells()
        xm, _, sm = em.cells()
        reg, _ = ep.cell_flags()
        np.testing.assert_array_equal(sp, o.sp, err_msg=f"parent status, iteration {it}")
        np.testing.assert_array_equal(sm, o.sm, err_msg=f"main status, iteration {it}")
        np.testing.assert_array_equal(reg, o.rp, err_msg=f"REGRESSFLAG, iteration {it}")
        live = o.sm > -4998
        np.testing.assert_allclose(xp, o.xp, rtol=0, atol=1e-8, err_msg=f"parent, iteration {it}")
        np.testing.assert_allclose(xm[live], o.xm[live], rtol=0, atol=1e-8,
                                   err_msg=f"main, iteration {it}")
        if nuc:
            np.testing.assert_allclose(ep.nuclei()[0], o.xnp, rtol=0, atol=1e-8,
                                       err_msg=f"parent nuclei, iteration {it}")
            np.testing.assert_allclose(em.nuclei()[0][live], o.xnm[live], rtol=0, atol=1e-8,
                                       err_msg=f"main nuclei, iteration {it}")
        crossed |= bool(sp[0] == 4)
        cloned |= bool(sm[1] == -(cs.id_base + 1))
    assert o.clones == 1                       # parent cell 0 was cloned once
    assert cloned and sm[1] == 1               # its clone was released into the ROI
    assert crossed                             # the parent cell crossed the wrap
    np.testing.assert_allclose(em.populations(), om.populations, rtol=0, atol=1e-11)
    np.testing.assert_allclose(ep.populations(), op.populations, rtol=0, atol=1e-11)
    np.testing.assert_allclose(em.force(), om.force, rtol=0, atol=1e-12)
    np.testing.assert_allclose(ep.force(), op.force, rtol=0, atol=1e-12)
    em.close()
    ep.close()


def test_cell_wrap_setup_needs_the_link():
    bm = geometry.tube((16, 16, 40), axis=2, r=6.5, periodic=True)
    bp = parent_of(bm, KP, 1)
    xyz, tris = geometry.rbc_mesh(radius=2.6, level=2)
    em = Engine(lattice_config(*bm.shape, 0.8, periodic=(0, 0, 1)))
    ep = Engine(lattice_config(*bp.shape, 0.8))
    em.upload_lattice(bm)
    ep.set_parent(1, KP, Case(1).off)
    ep.upload_lattice(bp)
    for e in (em, ep):
        e.membrane_setup(lattice_membrane_config(xyz, tris, 2), tris, xyz)
        e.ibm_setup(lattice_ibm_config())
    with pytest.raises(ZfError, match="zf_parent_link"):
        ep.cell_wrap_setup(em, 1)
    ep.parent_link(em)
    ep.cell_wrap_setup(em, 1)
    em.close()
    ep.close()
