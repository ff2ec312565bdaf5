"""Benchmark of the Zebraflow LBM time step on one MI355X.

One "step" is one LBM iteration (stream + collide + boundary) over the whole
fluid-node list of the workload, inputs resident in HBM.  Prints ONE JSON
line (fluid MLUPS of the FP32-deviation form) with:
  - fp64_form: the plain FP64-on-GPU form on the same workload and ratio;
  - error: relative L2 velocity error of the FP32-deviation form against the
    FP64 form after `--err-steps` steps (the FP64 form agrees with the FP64
    oracle to 1e-13 in tests/test_gpu_lbm.py);
  - roofline: the stream-collide kernel's algorithmic bytes per launch over
    its average launch time (HIP events on the launch stream);
  - cpu_baseline: the reference-structured OpenMP CPU step (oracle FP32) on a
    bounded sample of the same workload, rank 0 only.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
# the reference's OpenMP environment (OMPenvironment.txt): no thread binding;
# set before any OpenMP runtime starts (the host and oracle libraries)
os.environ.setdefault("OMP_PROC_BIND", "false")

HBM_PEAK_GBS = 8000.0
STORE_NAMES = {"f64": 0, "f32dev": 1, "f16dev": 2}
STORE_BYTES = {"f64": 8, "f32dev": 4, "f16dev": 2}

# Workloads (box, generator args).  vessel_pair_15: the vessel-pair topology
# (DA + PCV, sinusoidal centrelines of amplitude 8, 6 connectors) with radii
# scaled so the fluid fraction is 15 % of the 512x128x128 box (about 1.26 M
# fluid nodes); vessel_pair_r10: the same with radius 10 / 4 (3.9 % fluid).
WORKLOADS = {
    "vessel_pair_15": dict(box=(512, 128, 128), kind="pair", r_main=19.8, r_conn=7.9, amp=8.0,
                           n_conn=6, seed=20261205),
    "vessel_pair_r10": dict(box=(512, 128, 128), kind="pair", r_main=10.0, r_conn=4.0, amp=8.0,
                            n_conn=6, seed=20261205),
    "tube_cfg1": dict(box=(32, 32, 128), kind="tube", r=14.0),
    "tree_cfg5": dict(box=(1024, 256, 256), kind="tree", r_root=24.0, ratio=0.79, gens=4,
                      a_lo=30.0, a_hi=45.0, seed=20261207),
    # the same tree with the root radius scaled to 42.4 for 12 % fluid
    # (8.12 M fluid nodes); radius 24 gives 4.1 %
    "tree_cfg5_12": dict(box=(1024, 256, 256), kind="tree", r_root=42.4, ratio=0.79, gens=4,
                         a_lo=30.0, a_hi=45.0, seed=20261207),
}


def make_geometry(name):
    from zebraflow_amd import geometry as g
    w = WORKLOADS[name]
    if w["kind"] == "pair":
        return g.vessel_pair(w["box"], w["r_main"], w["r_conn"], w["amp"], w["n_conn"], w["seed"])
    if w["kind"] == "tube":
        return g.tube(w["box"], axis=2, r=w["r"], periodic=True)
    return g.tree(w["box"], w["r_root"], w["ratio"], w["gens"], w["a_lo"], w["a_hi"], w["seed"])


FLAGS = 0  # zf_config.flags of the measured forms (--flags)

# HBM bytes per stream-collide launch on vessel_pair_15 (default flags), from
# two rocprofv3 --pmc runs of `bench.py --probe` (FETCH_SIZE, then WRITE_SIZE,
# KiB per dispatch, averaged over the 20 launches of each form), each scaled by
# what the counter reads for the probe's 256 MiB copy kernel: FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0 (profiles/stream_collide_traffic_vessel_pair_15.md).
TRAFFIC = {"f32dev": 203603931, "f64": 399333437, "f16dev": 106048845}


def engine_for(name, b, store, extra_flags=0):
    from zebraflow_amd.engine import Engine, lattice_config
    w = WORKLOADS[name]
    flags = FLAGS | extra_flags
    if w["kind"] == "tube":
        # cfg1: tau 0.8, body force giving a peak velocity of 0.02 (g = 4 nu u / R^2)
        nu = 0.3 / 3
        cfg = lattice_config(*b.shape, 0.8, storage=STORE_NAMES[store], periodic=(0, 0, 1),
                             body_force=(0, 0, 4 * nu * 0.02 / w["r"] ** 2), flags=flags)
    else:
        # pulsatile inlet u(t) = 0.03 (1 + 0.5 sin(2 pi t / 4000)) along +i
        cfg = lattice_config(*b.shape, 0.8, storage=STORE_NAMES[store], inlet_u0=0.03,
                             inlet_amp=0.5, inlet_period=4000.0, inlet_dir=(1, 0, 0), flags=flags)
    eng = Engine(cfg)
    eng.upload_lattice(b)
    return eng


def time_form(name, b, store, steps, warmup):
    eng = engine_for(name, b, store)
    eng.run(0, warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.record(0)
    eng.run(warmup, steps)
    eng.record(1)
    eng.sync()
    wall = time.perf_counter() - t0
    ms = eng.elapsed_ms(0, 1)
    return eng, ms, wall


def kernel_roofline(eng, store, n_int, ms, steps, workload):
    """The timed region is `steps` back-to-back launches of the one fused
    stream-collide kernel on the engine's stream, so its HIP-event time over
    `steps` is the kernel's average launch duration (dispatch gaps included).
    A second pass brackets every launch with its own event pair (adds the
    event overhead; reported for comparison)."""
    avg_ms = ms / steps
    algo = 2 * 19 * STORE_BYTES[store] * n_int
    achieved = algo / (avg_ms * 1e-3) / 1e9
    t = eng.time_kernels(10 ** 6, min(steps, 500))
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=TRAFFIC.get(store) if workload == "vessel_pair_15" and FLAGS == 0 else None,
                kernel="k_stream_collide", avg_us=round(avg_ms * 1e3, 3),
                algorithmic_bytes_per_launch=algo,
                per_launch_event_us=round(t[0] / t[2] * 1e3, 3),
                note="boundary nodes are updated inside the same launch")


def node_normals(wx, wt, wn):
    """Area-weighted unit normals of the wall-mesh nodes (NaN rows for nodes
    whose triangles all have zero area)."""
    p = wx[wt - 1]
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    acc = np.zeros_like(wx)
    for k in range(3):
        np.add.at(acc, wt[:, k] - 1, area[:, None] * wn)
    with np.errstate(invalid="ignore", divide="ignore"):
        return acc / np.linalg.norm(acc, axis=1, keepdims=True)


def form_errors(name, b, steps):
    """The three forms from the same equilibrium start: rel. L2 error of the
    FP32- and FP16-deviation forms' velocity and wall shear stress (zf_wss at
    every lumen wall-mesh node, mu/DX = 1) against the FP64 form."""
    from zebraflow_amd import geometry as g
    wx, wt, wn = g.wall_mesh(b)
    nrm = node_normals(wx, wt, wn)
    ok = np.isfinite(nrm).all(axis=1)  # nodes of degenerate wall triangles only: no normal
    wx, nrm = wx[ok], nrm[ok]
    out = {}
    for store in ("f64", "f32dev", "f16dev"):
        eng = engine_for(name, b, store)
        eng.run(0, steps)
        out[store] = (eng.macro()[1:4], eng.wss(wx, nrm, 1.0)[0])
        eng.close()
    rel = lambda a, r: float(np.linalg.norm(a - r) / np.linalg.norm(r))
    (u64, w64), (u32, w32), (u16, w16) = out["f64"], out["f32dev"], out["f16dev"]
    return dict(vel_rel_l2_vs_fp64=rel(u32, u64), wss_rel_l2_vs_fp64=rel(w32, w64),
                fp16_vel_rel_l2_vs_fp64=rel(u16, u64), fp16_wss_rel_l2_vs_fp64=rel(w16, w64),
                wall_points=len(wx), steps=steps)


def cpu_baseline(name, b, seconds):
    """Reference-structured OpenMP step (oracle, FP32 storage) on the same
    workload; as many whole steps as fit in about `seconds`."""
    from oracle import pyoracle as po
    w = WORKLOADS[name]
    if w["kind"] == "tube":
        params = dict(wall_mode=1, visc_scale=0.3, force_dt=1.0, periodic=(0, 0, 1),
                      body_force=(0, 0, 4 * 0.1 * 0.02 / w["r"] ** 2))
    else:
        params = dict(wall_mode=1, visc_scale=0.3, force_dt=1.0, inlet_u0=0.03, inlet_amp=0.5,
                      inlet_period=4000.0, inlet_dir=(1, 0, 0), outlet_rho=1.0)
    o = po.LbmOracle(b, params, dtype=np.float32)
    n_int = int(((o.bnode() > 0) & (o.bnode() < 10)).sum())
    threads = po.max_threads()
    o.step(1)  # warm-up (first touch)
    t0 = po.wtime()
    n = 0
    while True:
        o.step(2 + n)
        n += 1
        el = po.wtime() - t0
        if el >= seconds or n >= 2000:
            break
    return dict(value=round(n_int * n / el / 1e6, 3), unit="MLUPS", cores=threads,
                kind="port",
                sample=f"{n} steps of the full {name} lattice ({n_int} fluid nodes), FP32 "
                       f"reference loop structure (F/F_OLD/FEQ arrays, DOMID lookups, "
                       f"schedule(static,5000)), {el:.1f} s")


def beyond_llc_leg(store, steps, warmup, name="tree_cfg5_12"):
    """The same stream kernel beyond the last-level cache: the 12 %-fluid
    tree (8.12 M fluid nodes, 684 MB per FP32 population buffer), timed as
    the headline (HIP events around back-to-back steps)."""
    bt = make_geometry(name)
    eng, ms, _ = time_form(name, bt, store, steps, warmup)
    n = eng.n_interior
    info = eng.info()
    eng.close()
    avg_ms = ms / steps
    algo = 2 * 19 * STORE_BYTES[store] * n
    ach = algo / (avg_ms * 1e-3) / 1e9
    return dict(workload=name, fluid_nodes=n, steps=steps, avg_us=round(avg_ms * 1e3, 3),
                mlups=round(n / (avg_ms * 1e-3) / 1e6, 1), achieved=round(ach, 1),
                frac=round(ach / HBM_PEAK_GBS, 4),
                population_footprint_bytes=2 * 19 * STORE_BYTES[store] * int(info["stride"]))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_rbc_baseline(name, b, ncell, seconds):
    """The RBC iteration on the CPU, reference-structured OpenMP (oracle
    restatements of ZebraflowWT.c:6800-7535 membrane forces, one thread per
    cell; :5851-5900 cell list; :8693-9144 interpolation and wall
    repulsion, :9913-10197 cell-cell repulsion, parallel over nodes;
    :10739-11173 spreading with OpenMP atomics; :11975-11989 position
    update), plus the FP32 LBM step of the same iteration, on the dense-cell
    workload of the GPU leg (same placement); whole iterations until about
    `seconds`."""
    from oracle import pyoracle as po
    from zebraflow_amd import geometry as g
    from zebraflow_amd.engine import lattice_membrane_config
    xyz, tris = g.rbc_mesh(radius=4.0)
    cells = g.place_cells_dense(b, xyz, 0.6, ncell, 20261206)
    wx, wt, wn = g.wall_mesh(b)
    mc = lattice_membrane_config(xyz, tris, len(cells))
    mp = {f: getattr(mc, f) for f, _ in po.MembParams._fields_}
    edge, _, _ = po.mesh_topology(len(xyz), tris)
    eq = po.mesh_equilibrium(xyz, edge, mp, mc.es)
    tae = po.tri_areas(xyz, tris, 1.0)
    p = po.ibm_params(b.shape, vmax=1.0)  # lattice units, as lattice_ibm_config
    r = po.rep_params(bin_cutoff=2.0, cutoff=1.0, search=2.0)
    nwt, windex = po.wall_tables(p, wx, wt)
    params = dict(wall_mode=1, visc_scale=0.3, force_dt=1.0, inlet_u0=0.03, inlet_amp=0.5,
                  inlet_period=4000.0, inlet_dir=(1, 0, 0), outlet_rho=1.0)
    o = po.LbmOracle(b, params, dtype=np.float32)
    domid = o.domid
    x = cells.copy()
    status = np.ones(len(cells), np.int32)
    threads = po.max_threads()
    t0 = po.wtime()
    n = 0
    while True:
        o.step(1 + n, rbc_step=1)
        gm = po.membrane_forces(x, tris, edge, eq, tae, mp)
        bins = po.cell_list(p, b, domid, gm["tri_centroid"], status, r.bin_cutoff)
        v = po.ibm_interpolate(p, b, domid, (o.U, o.V, o.W), x)
        vrep = po.cell_repulsion(p, r, b, domid, tris, status, x, gm["node_normal"],
                                 gm["tri_centroid"], gm["tri_normal"], bins)
        vrep += po.wall_repulsion(p, r, windex, wx, wt, wn, nwt, status, x)[0]
        frc = np.zeros((3, o.n), np.float32)
        f64 = po.ibm_spread(p, b, domid, o.n, x, 0.5 * vrep + gm["force"], parallel=True)
        frc[...] = f64
        o.force[...] = frc
        x, status, _ = po.advance(p, x, v + vrep, status)
        n += 1
        el = po.wtime() - t0
        if el >= seconds or n >= 50:
            break
    nn = len(xyz)
    return dict(value=round(len(cells) * nn * n / el, 1), unit="RBC node-updates/s",
                cores=threads, kind="port",
                sample=f"{n} RBC iterations (membrane/IBM/repulsion and the FP32 LBM step) of "
                       f"{len(cells)} cells ({len(cells) * nn} membrane nodes) on the full "
                       f"{name} lattice, {el:.1f} s",
                ms_per_iteration=round(el / n * 1e3, 2))


def cells_leg(name, b, ncell, steps, warmup):
    """Dense-cell load on the same lattice: `ncell` Evans-Fung cells (642
    nodes, radius 4 lu) placed at random without overlap, membrane shear
    modulus 5e-3 and bending modulus 2e-3 lu, cell-cell (cutoff 1 lu, 2-lu
    cell list) and cell-wall repulsion against the lumen mesh, membrane/IBM
    work on every step.  Times whole zf_step iterations of the FP32 form;
    the FP64 form runs the same iterations for the membrane-shape error
    (node positions, relative to the mean node displacement)."""
    from zebraflow_amd import geometry as g
    from zebraflow_amd.engine import (lattice_ibm_config, lattice_membrane_config,
                                      lattice_repulsion_config)
    xyz, tris = g.rbc_mesh(radius=4.0)
    cells = g.place_cells_dense(b, xyz, 0.6, ncell, 20261206)
    wall = g.wall_mesh(b)
    from zebraflow_amd.engine import FLAG_F32_CELLS
    res = {}
    for store, fl in (("f32dev", 0), ("f64", 0), ("f32cells", FLAG_F32_CELLS)):
        eng = engine_for(name, b, "f32dev" if store == "f32cells" else store, fl)
        eng.membrane_setup(lattice_membrane_config(xyz, tris, len(cells)), tris, xyz)
        eng.upload_cells(cells)
        eng.ibm_setup(lattice_ibm_config(rbc_interval=1))
        eng.repulsion_setup(lattice_repulsion_config())
        eng.wall_setup(*wall)
        for it in range(1, warmup + 1):
            eng.full_step(it)
        eng.sync()
        eng.record(0)
        for it in range(warmup + 1, warmup + steps + 1):
            eng.full_step(it)
        eng.record(1)
        eng.sync()
        x, _, st = eng.cells()
        res[store] = (eng.elapsed_ms(0, 1), x, st, eng.n_interior)
        eng.close()
    ms, x32, st, n_int = res["f32dev"]
    x64 = res["f64"][1]
    ms_c, x32c = res["f32cells"][0], res["f32cells"][1]
    disp = np.linalg.norm(x64 - cells, axis=-1).mean()
    nn = len(xyz)
    p = xyz[tris - 1]
    vol = (0.5 * np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]) * p.mean(1)).sum() / 3
    return dict(cells=len(cells), membrane_nodes=len(cells) * nn,
                haematocrit=round(len(cells) * vol / n_int, 4), steps=steps, warmup=warmup,
                ms_per_step=round(ms / steps, 4),
                rbc_node_updates_per_s=round(len(cells) * nn * steps / (ms * 1e-3), 1),
                fluid_mlups=round(n_int * steps / (ms * 1e-3) / 1e6, 1),
                cells_live_after=int((st > 0).sum()), rbc_interval=1,
                shape_error_vs_fp64=dict(
                    max_node_dev=float(np.abs(x32 - x64).max()), mean_displacement=float(disp),
                    rel=float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64 - cells))),
                fp32_node_forces=dict(
                    flag="ZF_FLAG_F32_CELLS", ms_per_step=round(ms_c / steps, 4),
                    max_node_dev_vs_fp64=float(np.abs(x32c - x64).max()),
                    rel_vs_fp64=float(np.linalg.norm(x32c - x64) / np.linalg.norm(x64 - cells))))


def parent_leg(steps=500, warmup=50):
    """A main lattice with one parent (feeder) lattice in the layout of the
    reference's DA parent: the main lattice a straight vessel (96x96x512,
    radius 40, periodic along k, driven by a body force), the parent its
    cross-section extruded over 121 planes with the upstream code 21 at
    k = 1 and the copy code 11 at k = 121, which lies on main plane k = 3.
    FP32-deviation storage.  Times main-only iterations and coupled ones
    (main step, then parent step with its boundary pass) by the host clock
    around whole loops, the parent's stream kernel by HIP events, and checks
    that the parent's copy plane carries the main plane's axial velocity."""
    from zebraflow_amd import geometry as g
    from zebraflow_amd.engine import Engine, lattice_config
    bm = g.tube((96, 96, 512), axis=2, r=40.0, periodic=True)
    kp = 121
    plane = bm[:, :, 0]
    bp = np.repeat(plane[:, :, None], kp, axis=2).astype(np.int32)
    bp[bp > 0] = 1
    bp[:, :, kp - 1][plane > 0] = 11
    bp[:, :, 0][plane > 0] = 21
    cm = lattice_config(*bm.shape, 0.8, storage=STORE_NAMES["f32dev"], periodic=(0, 0, 1),
                        body_force=(0, 0, 2e-6), flags=FLAGS)
    cp = lattice_config(*bp.shape, 0.8, storage=STORE_NAMES["f32dev"], flags=FLAGS)
    em, ep = Engine(cm), Engine(cp)
    em.upload_lattice(bm)
    ep.set_parent(1, kp, (0, 0, 3 - kp))
    ep.upload_lattice(bp)
    em.init_equilibrium()
    ep.init_equilibrium()

    def loop(it0, n, coupled):
        em.sync()
        ep.sync()
        t = time.perf_counter()
        for it in range(it0, it0 + n):
            em.step(it)
            if coupled:
                ep.step(it)
        em.sync()
        ep.sync()
        return (time.perf_counter() - t) * 1e3

    # the main lattice alone, before the link (no export of coupled moments)
    loop(1, warmup, False)
    ms_main = loop(warmup + 1, steps, False)
    em.init_equilibrium()
    ep.parent_link(em)
    loop(1, warmup, True)
    ms_pair = loop(warmup + 1, steps, True)
    # the parent's own stream kernel (its boundary pass runs after it)
    pk_ms, _, pk_n = ep.time_kernels(warmup + 2 * steps, 100)
    pk_us = pk_ms * 1e3 / pk_n
    loop(warmup + 2 * steps + 101, 2000, True)  # let the flow reach the parent
    dm, um, vm, wm = em.macro()
    dp, up, vp, wp = ep.macro()
    dlm, dlp = em.lists()[0], ep.lists()[0]
    n_main, n_par = em.n_interior, ep.n_interior
    ep.close()
    em.close()
    sel_m = (dlm[:, 2] == 3) & (bm[dlm[:, 0] - 1, dlm[:, 1] - 1, 2] > 0)
    sel_p = (dlp[:, 2] == kp) & (bp[dlp[:, 0] - 1, dlp[:, 1] - 1, kp - 1] == 11)
    km = {tuple(x[:2]): w for x, w in zip(dlm[sel_m], wm[sel_m])}
    diff = [abs(w - km[tuple(x[:2])]) for x, w in zip(dlp[sel_p], wp[sel_p])]
    return dict(main_fluid_nodes=n_main, parent_fluid_nodes=n_par, coupled_nodes=int(sel_p.sum() * 2),
                steps=steps, main_ms_per_step=round(ms_main / steps, 4),
                coupled_ms_per_iteration=round(ms_pair / steps, 4),
                parent_kernel_us=round(pk_us, 2),
                parent_mlups=round(n_par / pk_us, 1),
                coupling_overhead_ms=round((ms_pair - ms_main) / steps - pk_us * 1e-3, 4),
                coupling_note="main timed unlinked; the overhead is the export of the coupled "
                              "moments plus the parent's boundary pass (and launch gaps)",
                copy_plane_w_max_abs_diff=float(max(diff)), main_plane_w_max=float(max(km.values())),
                timing="host clock around whole loops (includes the Python call per step)")


def mixed_precision_leg(steps=3000, ncell=6000, interval=1, chunk=500):
    """The large mixed-precision run: the 4-generation tree (1024x256x256,
    12 % fluid) with as many Evans-Fung cells (642 nodes, radius 4) as the
    dense placer fits up to `ncell`, membrane/IBM/repulsion every `interval`
    iterations with `interval`-step position updates, `steps` iterations in
    the FP16, FP32 and FP64 forms from the same start, advanced in lockstep.
    Every `chunk` steps: velocity, wall shear stress (every lumen wall-mesh
    node, mu/DX = 1) and cell shape (node positions relative to the mean node
    displacement, cells live in every form) of the FP16 and FP32 forms
    against the FP64 form and of FP16 against FP32, so that the growth of
    the differences along the run is recorded."""
    from zebraflow_amd import geometry as g
    from zebraflow_amd.engine import (lattice_ibm_config, lattice_membrane_config,
                                      lattice_repulsion_config)
    name = "tree_cfg5_12"
    b = make_geometry(name)
    xyz, tris = g.rbc_mesh(radius=4.0)
    cells = g.place_cells_dense(b, xyz, 0.6, ncell, 20261208)
    wx, wt, wn = g.wall_mesh(b)
    nrm = node_normals(wx, wt, wn)
    ok = np.isfinite(nrm).all(axis=1)  # nodes of degenerate wall triangles only: no normal
    px, pn = wx[ok], nrm[ok]
    print(f"mixed precision: {len(cells)} cells placed, {len(px)} wall points", flush=True)
    forms = ("f16dev", "f32dev", "f64")
    engs = {}
    for store in forms:
        eng = engine_for(name, b, store)
        eng.membrane_setup(lattice_membrane_config(xyz, tris, len(cells)), tris, xyz)
        eng.upload_cells(cells)
        ic = lattice_ibm_config(rbc_interval=interval)
        ic.dt_advance = float(interval)
        eng.ibm_setup(ic)
        eng.repulsion_setup(lattice_repulsion_config())
        eng.wall_setup(wx, wt, wn)
        eng.sync()
        engs[store] = eng
    ms = {k: 0.0 for k in forms}

    def err(res, a, ref, live):
        ea, er = res[a], res[ref]
        dx = np.linalg.norm(res["f64"]["x"][live] - cells[live])
        return dict(vel_rel_l2=float(np.linalg.norm(ea["u"] - er["u"]) / np.linalg.norm(er["u"])),
                    wss_rel_l2=float(np.linalg.norm(ea["wss"] - er["wss"]) / np.linalg.norm(er["wss"])),
                    shape_rel=float(np.linalg.norm(ea["x"][live] - er["x"][live]) / dx))

    growth = []
    for c0 in range(0, steps, chunk):
        n = min(chunk, steps - c0)
        res = {}
        for store, eng in engs.items():
            eng.record(0)
            eng.full_run(c0, n)
            eng.record(1)
            eng.sync()
            ms[store] += eng.elapsed_ms(0, 1)
            x, _, st = eng.cells()
            _, u, v, w = eng.macro()
            res[store] = dict(u=np.stack([u, v, w]), wss=eng.wss(px, pn, 1.0)[0], x=x, st=st)
        live = (res["f16dev"]["st"] > 0) & (res["f32dev"]["st"] > 0) & (res["f64"]["st"] > 0)
        umax = float(np.sqrt((res["f64"]["u"] ** 2).sum(0)).max())
        row = dict(step=c0 + n, cells_live=int(live.sum()), max_u_f64=round(umax, 5),
                   f16_vs_f64=err(res, "f16dev", "f64", live), f32_vs_f64=err(res, "f32dev", "f64", live),
                   f16_vs_f32=err(res, "f16dev", "f32dev", live))
        growth.append(row)
        print(f"mixed precision: {c0 + n} steps, {row['cells_live']} cells live in all forms, "
              f"max |u| {umax:.4f}, vel f16/f64 {row['f16_vs_f64']['vel_rel_l2']:.2e} "
              f"f32/f64 {row['f32_vs_f64']['vel_rel_l2']:.2e}", flush=True)
    n_int = engs["f64"].n_interior
    for eng in engs.values():
        eng.close()
    nn = len(xyz)
    last = growth[-1]
    return dict(workload=f"{name}: 1024x256x256 tree, {n_int} fluid nodes, "
                         f"{len(cells)} cells of {nn} nodes",
                cells=len(cells), membrane_nodes=len(cells) * nn, steps=steps, rbc_interval=interval,
                cells_live_all_forms=last["cells_live"], wall_points=len(px),
                ms={k: round(v, 1) for k, v in ms.items()},
                ms_per_step={k: round(v / steps, 3) for k, v in ms.items()},
                rbc_node_updates_per_s={k: round(len(cells) * nn * steps / interval / (v * 1e-3), 1)
                                        for k, v in ms.items()},
                f16_vs_f64=last["f16_vs_f64"], f32_vs_f64=last["f32_vs_f64"],
                f16_vs_f32=last["f16_vs_f32"], growth=growth)


def adjoint_leg(steps=3000, t0=1000, checkpoint=100, ncell=32, interval=10, fd_h=(1e-4, 1e-6),
                fd_h_smooth=(1e-3, 1e-4)):
    """Adjoint cost and gradients on a 128x64x64 straight vessel of radius 26
    (velocity inlet, pressure outlet, pulsatile inflow u0 (1 + 0.5 sin 2 pi
    t/1000), u0 = 0.005) with `ncell` RBCs (642-node Evans-Fung meshes of
    radius 4, shear modulus 5e-3, bending modulus 2e-3 lu), membrane/IBM every
    `interval` iterations with the positions advanced by `interval` lattice
    steps (every 50, the reference's stagger, runs unstable here within 800
    iterations: the explicit update overshoots once membrane forces reach
    ~0.02 lu): J = time-averaged |WSS| over the wall patch 48 <= x <= 80 for
    iterations t0..steps, populations checkpointed every `checkpoint`
    iterations.  Gradients w.r.t. inlet amplitude, es and kb in FP32 and FP64,
    and central differences of the FP64 forward pass for each relative step
    in fd_h, for the model as the reference has it and for its smooth
    variant (IBM kernel on the whole 4^3 cube instead of the sphere, bending
    faded to zero by a smoothstep over |cos dihedral| 0.97..0.99 instead of
    the reference's cut at 0.99, no anti-sliver threshold), whose forward map
    is C1, with the 642-node cells and with 162-node cells (radius 3)."""
    from zebraflow_amd import geometry as g
    from zebraflow_amd.engine import (B_PRESSURE_OUTLET, B_VELOCITY_INLET, Engine,
                                      lattice_config, lattice_ibm_config, lattice_membrane_config)
    es0, kb0, amp0, u0 = 5e-3, 2e-3, 0.5, 0.005
    b = g.tube((128, 64, 64), axis=0, r=26.0, inlet=B_VELOCITY_INLET, outlet=B_PRESSURE_OUTLET)
    wx, wt, wn = g.wall_mesh(b)
    p = wx[wt - 1]
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    acc = np.zeros_like(wx)
    for k in range(3):
        np.add.at(acc, wt[:, k] - 1, area[:, None] * wn)
    nrm = acc / np.linalg.norm(acc, axis=1, keepdims=True)
    patch = (wx[:, 0] >= 48) & (wx[:, 0] <= 80)
    # cells start in the upstream half (x < 60), so that none reaches the
    # outlet within the pass (the peak velocity moves a cell ~45 lu)
    bp = b.copy()
    bp[60:] = -1
    meshes = {}
    for name, radius, level in (("642", 4.0, 3), ("162", 3.0, 2)):
        xyz, tris = g.rbc_mesh(radius=radius, level=level)
        cells = g.place_cells(bp, xyz, radius, 1.0, ncell, 20261207)
        assert len(cells) == ncell
        meshes[name] = (xyz, tris, cells)

    def build(store, k, es=es0, kb=kb0, smooth=False, mesh="642"):
        xyz, tris, cells = meshes[mesh]
        cfg = lattice_config(*b.shape, 0.8, storage=STORE_NAMES[store], inlet_u0=u0,
                             inlet_amp=amp0, inlet_period=1000.0, inlet_dir=(1, 0, 0))
        eng = Engine(cfg)
        eng.upload_lattice(b)
        mc = lattice_membrane_config(xyz, tris, ncell, es=es, kb=kb)
        ic = lattice_ibm_config(rbc_interval=interval)
        ic.dt_advance = float(interval)
        if smooth:
            # bending faded out over |cos| 0.97..0.99 instead of cut at 0.99;
            # ic.flags 1 = ZF_IBM_CUBE_SUPPORT
            mc.cos_cut, mc.cos_blend, mc.sliver_len, ic.flags = 0.99, 0.02, 0.0, 1
        eng.membrane_setup(mc, tris, xyz)
        eng.upload_cells(cells)
        eng.ibm_setup(ic)
        eng.adjoint_setup(wx[patch], nrm[patch], 0.1, t0, steps, steps, k, cells=True)
        eng.zero_force = np.zeros((3, eng.dommax))
        eng.cells0 = cells
        return eng

    def J_at(eng, amp):
        eng.init_equilibrium()
        eng.upload_force(eng.zero_force)
        eng.upload_cells(eng.cells0)
        eng.set_forcing(u0, amp)
        return eng.adjoint_forward(0, steps)

    def gradients(eng):
        grad = eng.adjoint_backward()
        dp, _ = eng.adjoint_cell_gradients()
        return np.array([grad[0], dp[0] / es0, dp[1] / kb0])  # d/d amp, es, kb

    def fd_check(eng, smooth, mesh):
        out = {}
        for h in (fd_h_smooth if smooth else fd_h):
            def J_mod(es, kb):
                e = build("f64", steps, es, kb, smooth, mesh)
                v = J_at(e, amp0)
                e.close()
                return v
            out[h] = np.array([
                (J_at(eng, amp0 * (1 + h)) - J_at(eng, amp0 * (1 - h))) / (2 * h * amp0),
                (J_mod(es0 * (1 + h), kb0) - J_mod(es0 * (1 - h), kb0)) / (2 * h * es0),
                (J_mod(es0, kb0 * (1 + h)) - J_mod(es0, kb0 * (1 - h))) / (2 * h * kb0)])
        return out

    rel = lambda a, b: [float(x) for x in np.abs(a - b) / np.abs(b)]
    res = {}
    for store, k, smooth, mesh in (("f32dev", checkpoint, False, "642"), ("f32dev", 1, False, "642"),
                                   ("f64", checkpoint, False, "642"), ("f64", checkpoint, True, "642"),
                                   ("f64", checkpoint, True, "162"), ("f32dev", checkpoint, True, "162"),
                                   ("f32dev", checkpoint, True, "642")):
        eng = build(store, k, smooth=smooth, mesh=mesh)
        eng.full_run(0, 2 * interval)  # warm-up of every kernel
        eng.init_equilibrium()
        eng.upload_force(eng.zero_force)
        eng.upload_cells(eng.cells0)
        eng.sync()
        eng.record(0)
        eng.full_run(0, steps)
        eng.record(1)
        eng.sync()
        plain_ms = eng.elapsed_ms(0, 1)
        J = J_at(eng, amp0)
        guard = int(eng.membrane()["guard_fail"].sum())
        gr = gradients(eng)
        fwd_ms, bwd_ms = eng.adjoint_timing()
        r = dict(J=J, g=gr, plain_ms=plain_ms, fwd_ms=fwd_ms, bwd_ms=bwd_ms, n=eng.n_interior,
                 guard=guard)
        if store == "f64":
            r["fd"] = fd_check(eng, smooth, mesh)
        res[(store, k, smooth, mesh)] = r
        eng.close()
        print(f"adjoint leg: {store} checkpoint {k} smooth {smooth} mesh {mesh} done", flush=True)
    r32, r32all = res[("f32dev", checkpoint, False, "642")], res[("f32dev", 1, False, "642")]
    r64 = res[("f64", checkpoint, False, "642")]
    smooth = {m: res[("f64", checkpoint, True, m)] for m in ("642", "162")}
    fdd = lambda r: {f"h={h:g}": [float(x) for x in v] for h, v in r["fd"].items()}
    fde = lambda r: {f"h={h:g}": rel(r["g"], v) for h, v in r["fd"].items()}
    return dict(workload=f"straight vessel 128x64x64, r=26, {ncell} RBCs of 642 nodes",
                fluid_nodes=r32["n"], steps=steps, rbc_interval=interval,
                objective_window=[t0, steps], wall_points=int(patch.sum()),
                forward_ms=round(r32["plain_ms"], 2), checkpoint=checkpoint,
                forward_recorded_ms=round(r32["fwd_ms"], 2), backward_ms=round(r32["bwd_ms"], 2),
                adjoint_over_forward=round(r32["bwd_ms"] / r32["plain_ms"], 3),
                adjoint_over_recorded_forward=round(r32["bwd_ms"] / r32["fwd_ms"], 3),
                every_state_stored=dict(forward_recorded_ms=round(r32all["fwd_ms"], 2),
                                        backward_ms=round(r32all["bwd_ms"], 2),
                                        adjoint_over_forward=round(r32all["bwd_ms"] / r32all["plain_ms"], 3)),
                gradient_names=["dJ/d inlet_amp", "dJ/d es", "dJ/d kb"],
                grad_fp32=[float(x) for x in r32["g"]], grad_fp64=[float(x) for x in r64["g"]],
                grad_rel_err_fp32_vs_fp64=rel(r32["g"], r64["g"]),
                fd_fp64=fdd(r64), grad_rel_err_fp64_vs_fd=fde(r64),
                smooth_variant={f"cells_{m}": dict(grad_fp64=[float(x) for x in rs["g"]], fd_fp64=fdd(rs),
                                                   grad_rel_err_vs_fd=fde(rs), J=rs["J"])
                                for m, rs in smooth.items()},
                smooth_variant_fp32_vs_fp64_cells_162=rel(res[("f32dev", checkpoint, True, "162")]["g"],
                                                          smooth["162"]["g"]),
                smooth_variant_fp32_vs_fp64_cells_642=rel(res[("f32dev", checkpoint, True, "642")]["g"],
                                                          smooth["642"]["g"]),
                J=[r32["J"], r64["J"]])


def probe(name, b):
    """A copy of known size (256 MiB read + 256 MiB written per launch, three
    launches) to calibrate the HBM counters, then 20 steps of each form."""
    eng = engine_for(name, b, "f32dev")
    eng.copy_bandwidth(256 << 20, 3)
    eng.close()
    print("probe: copies done", flush=True)
    for store in ("f32dev", "f64", "f16dev"):
        eng = engine_for(name, b, store)
        eng.run(0, 20)
        eng.sync()
        print(f"probe: {store} 20 steps, {eng.n_interior} fluid nodes", flush=True)
        eng.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="vessel_pair_15", choices=sorted(WORKLOADS))
    ap.add_argument("--store", default="f32dev", choices=sorted(STORE_NAMES))
    ap.add_argument("--flags", type=int, default=0, help="zf_config.flags (ZF_FLAG_*)")
    ap.add_argument("--err-steps", type=int, default=10000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip FP64 form and error legs")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg and print its JSON object")
    ap.add_argument("--cells", type=int, default=1200,
                    help="cells of the dense-cell leg (0 skips it)")
    ap.add_argument("--cell-steps", type=int, default=200)
    ap.add_argument("--no-parents", action="store_true", help="skip the parent-lattice leg")
    ap.add_argument("--parents-only", action="store_true",
                    help="only the parent-lattice coupling leg (for profiling)")
    ap.add_argument("--cells-only", action="store_true",
                    help="only the dense-cell leg (for profiling its kernels)")
    ap.add_argument("--mixed-steps", type=int, default=3000)
    ap.add_argument("--mixed-cells", type=int, default=6000)
    ap.add_argument("--mixed-precision-only", action="store_true",
                    help="run only the large mixed-precision leg (tree, FP16/FP32/FP64 with "
                         "cells) and print its JSON object")
    ap.add_argument("--probe", action="store_true",
                    help="counter-calibration run for rocprofv3 --pmc: three 256 MiB copies, "
                         "then 20 steps of each storage form; prints no JSON line")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine runs the time step on a single MI355X "
              "(multi-GPU slab decomposition is a design note, not built); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    global FLAGS
    FLAGS = args.flags
    b = make_geometry(args.workload)
    if args.probe:
        probe(args.workload, b)
        return
    if args.mixed_precision_only:
        print(json.dumps(dict(mixed_precision=mixed_precision_leg(
            steps=args.mixed_steps, ncell=args.mixed_cells, chunk=min(500, args.mixed_steps)))), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps(dict(adjoint=adjoint_leg())), flush=True)
        return
    if args.parents_only:
        print(json.dumps(dict(parent_coupling=parent_leg())), flush=True)
        return
    if args.cells_only:
        print(json.dumps(dict(dense_cells=cells_leg(args.workload, b, args.cells, args.cell_steps,
                                                     max(1, args.cell_steps // 10)))), flush=True)
        return
    n_fluid = int(((b > 0) & (b < 10)).sum())
    eng, ms, wall = time_form(args.workload, b, args.store, args.steps, args.warmup)
    n_int = eng.n_interior
    assert n_int == n_fluid
    mlups = n_int * args.steps / (ms * 1e-3) / 1e6
    roof = kernel_roofline(eng, args.store, n_int, ms, args.steps, args.workload)
    # measured device-copy ceiling (2 GiB each way: far beyond the 256 MB
    # last-level cache, which inflates copies of one launch's size)
    roof["copy_ceiling"] = round(eng.copy_bandwidth(2 << 30, 20), 1)
    info = eng.info()
    eng.close()
    # cache residency of the measured workload: both population buffers
    # against MI355X's 256 MB last-level cache (Infinity Cache)
    foot = 2 * 19 * STORE_BYTES[args.store] * int(info["stride"])
    roof["population_footprint_bytes"] = foot
    roof["llc_bytes"] = 256 << 20
    roof["llc_resident"] = foot <= (256 << 20)

    extra = {}
    if not args.no_extra:
        e64, ms64, _ = time_form(args.workload, b, "f64", args.steps, args.warmup)
        mlups64 = n_int * args.steps / (ms64 * 1e-3) / 1e6
        roof64 = kernel_roofline(e64, "f64", n_int, ms64, args.steps, args.workload)
        e64.close()
        extra["fp64_form"] = dict(value=round(mlups64, 1), unit="MLUPS",
                                  ms_per_step=round(ms64 / args.steps, 5),
                                  kernel_avg_us=roof64["avg_us"], kernel_frac=roof64["frac"])
        extra["speedup_vs_fp64_form"] = round(mlups / mlups64, 3)
        if roof["llc_resident"]:
            roof["beyond_llc"] = beyond_llc_leg(args.store, 200, 20)
        extra["error"] = form_errors(args.workload, b, args.err_steps)
        if not args.no_adjoint:
            extra["adjoint"] = adjoint_leg()
        if args.cells > 0:
            extra["dense_cells"] = cells_leg(args.workload, b, args.cells, args.cell_steps,
                                             max(1, args.cell_steps // 10))
        if not args.no_parents:
            extra["parent_coupling"] = parent_leg()

    line = {
        "metric": "fluid MLUPS (FP32-deviation D3Q19 step) + % HBM roofline vs FP64 form",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.store != "f64" else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "box": list(b.shape), "fluid_nodes": n_int,
                   "listed_nodes": info["dommax"], "storage": args.store,
                   "fluid_fraction": round(n_int / b.size, 4), "tau": 0.8,
                   "wall": "halfway bounce-back",
                   "inlet": "u(t)=0.03(1+0.5 sin 2 pi t/4000)" if "pair" in args.workload else None,
                   "chords": info["n_chords"], "flags": args.flags},
        "wall_clock_s": round(wall, 4),
        "roofline": roof,
    }
    line.update(extra)
    if not args.no_cpu:
        cb = cpu_baseline(args.workload, b, args.cpu_seconds)
        cb.update(cpu_model=cpu_model(), omp_proc_bind=os.environ.get("OMP_PROC_BIND"))
        if args.cells > 0 and not args.no_extra:
            cb["rbc"] = cpu_rbc_baseline(args.workload, b, args.cells, args.cpu_seconds)
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
