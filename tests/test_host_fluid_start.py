"""The reference's fluid start-state input (ZebraflowWT.c:2626-3387 for the
main lattice, :3394-3620 for the parents) read by libzfhost.so, on the
start file the reference ships for its PCV2 parent lattice
(tests/golden/PCV2parentdomainStart.csv.gz, a copy of the reference's own
data file: 96,074 rows in OUTPUT()'s 13-column layout).

  - the file's rows are the parent's listed nodes in DOMID_3 order
    (LBM_PARENTDOMAIN_SETUP, ZebraflowWT.c:744-758: lexicographic over
    B_index_3 >= 0) and their B values are the parent's B field from
    PCVparent2Boundary.csv (the reference-written golden);
  - zfh_read_fluid_start puts each row's D, U, V, W and force at that list
    position, whole or split into pieces with only the first holding the
    header (as FLUIDdomainStart.aa .. .aj are), equal to a numpy parse;
  - on the GPU, the populations at equilibrium with that state
    (zfh_init_from_macro, ZebraflowWT.c:5253-5262) step as the oracle does.
"""
import gzip
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import load_host_library

GOLD = os.path.join(os.path.dirname(__file__), "golden")
# PCV parent 2 (the reference's lattice 3): box and origin, ZebraflowWT.c:106-112
IM3, JM3, KM3, ORIGIN3, DX = 42, 27, 121, (58.0, 209.5, 332.5), 0.5


def start_text():
    with gzip.open(os.path.join(GOLD, "PCV2parentdomainStart.csv.gz"), "rt") as f:
        return f.read()


def parse(text):
    rows = np.loadtxt(text.splitlines()[1:], delimiter=",", dtype=np.float64)
    xyz = rows[:, :3].astype(np.float32).astype(np.float64)
    ijk = np.array([[int((x - o) / DX + 1.01) for x, o in zip(r, ORIGIN3)] for r in xyz])
    return rows, ijk


def b3():
    return np.load(os.path.join(GOLD, "parents_ref.npz"))["b3"].astype(np.int32)


def read_start(paths, b):
    import ctypes as C
    n = int((b >= 0).sum())
    out = np.zeros((7, n))
    arr = (C.c_char_p * len(paths))(*[p.encode() for p in paths])
    got = load_host_library().zfh_read_fluid_start(
        len(paths), C.cast(arr, C.c_void_p), b.ctypes.data, *b.shape, *ORIGIN3, DX, n,
        out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data, out[3].ctypes.data,
        out[4:].ctypes.data)
    return got, out


def test_start_rows_follow_the_parent_domain_order():
    rows, ijk = parse(start_text())
    b = b3()
    assert b.shape == (IM3, JM3, KM3)
    _, domlist = po.dom_setup(b)
    assert len(rows) == len(domlist) == 96074
    np.testing.assert_array_equal(ijk, domlist)
    np.testing.assert_array_equal(rows[:, 3].astype(int), b[ijk[:, 0] - 1, ijk[:, 1] - 1, ijk[:, 2] - 1])
    # a realistic state: flow along the parent's axis, density above 1
    assert np.abs(rows[:, 6]).max() > 5e-4 and rows[:, 7].min() > 1.0


def test_reader_whole_and_split_equals_numpy(tmp_path):
    text = start_text()
    rows, ijk = parse(text)
    b = b3()
    whole = tmp_path / "start.csv"
    whole.write_text(text)
    lines = text.splitlines(keepends=True)
    cuts = [0, 30001, 61234, len(lines)]
    pieces = []
    for a, (lo, hi) in enumerate(zip(cuts[:-1], cuts[1:])):
        p = tmp_path / f"FLUIDdomainStart.a{'abc'[a]}"
        p.write_text("".join(lines[lo:hi]))
        pieces.append(str(p))
    got1, out1 = read_start([str(whole)], b)
    got3, out3 = read_start(pieces, b)
    assert got1 == got3 == len(rows)
    np.testing.assert_array_equal(out1, out3)
    # columns U, V, W, D, Bx, By, Bz in list order (rows are in list order)
    want = rows[:, [7, 4, 5, 6, 8, 9, 10]].astype(np.float32).astype(np.float64).T
    np.testing.assert_array_equal(out1, want)


def test_reader_rejects_rows_off_the_lattice(tmp_path):
    text = start_text()
    b = b3()
    i, j, k = po.dom_setup(b)[1][100]
    b[i - 1, j - 1, k - 1] = -1  # unlist a node that the file has a row for
    p = tmp_path / "start.csv"
    p.write_text(text)
    got, _ = read_start([str(p)], b)
    assert got < 0


@pytest.mark.gpu
def test_gpu_step_from_the_shipped_start_state_matches_oracle():
    from zebraflow_amd.engine import STORE_F32DEV, STORE_F64, Engine, reference_config
    from tests.test_gpu_lbm import oracle_params
    print("case PCV2 parent start state, 10 steps", flush=True)
    rows, _ = parse(start_text())
    b = b3()
    d, u, v, w = rows[:, 7], rows[:, 4], rows[:, 5], rows[:, 6]
    force = rows[:, 8:11].T.copy()
    for store, tol in ((STORE_F64, 1e-13), (STORE_F32DEV, 2e-7)):
        cfg = reference_config(IM3, JM3, KM3, storage=store)
        eng = Engine(cfg)
        eng.upload_lattice(b)
        assert eng.dommax == len(rows)
        mac = [np.ascontiguousarray(a, np.float64) for a in (d, u, v, w, force)]
        assert load_host_library().zfh_init_from_macro(eng.h, *[a.ctypes.data for a in mac]) == 0
        ora = po.LbmOracle(b, oracle_params(cfg))
        ora.set_populations(po.feq(d, np.stack([u, v, w])))
        ora.force[...] = force
        np.testing.assert_allclose(eng.populations(), ora.populations, rtol=0, atol=tol)
        for it in range(1, 11):
            eng.step(it)
            ora.step(it)
        np.testing.assert_allclose(eng.populations(), ora.populations, rtol=0, atol=tol)
        _, ug, vg, wg = eng.macro()
        assert np.abs(wg).max() > 5e-4
        eng.close()


@pytest.mark.gpu
def test_host_program_starts_from_split_start_files(tmp_path):
    """zf_host with fluid_start = two pieces of a FLUIDdomain CSV (the second
    without header) writes, 6 steps later, the FLUIDdomain file the engine
    writes after the same start and steps through the C ABI."""
    import filecmp
    import subprocess
    from zebraflow_amd import PKG, geometry
    from zebraflow_amd.engine import B_PRESSURE_OUTLET, B_VELOCITY_INLET, STORE_F64, Engine, lattice_config
    print("case zf_host fluid_start", flush=True)
    h = load_host_library()
    b = geometry.tube((30, 18, 18), axis=0, r=6.5, inlet=B_VELOCITY_INLET, outlet=B_PRESSURE_OUTLET)
    im, jm, km = b.shape
    cfg = lattice_config(im, jm, km, 0.8, storage=STORE_F64, inlet_u0=0.02, inlet_amp=0.3,
                         inlet_period=40.0, inlet_dir=(1, 0, 0))
    eng = Engine(cfg)
    eng.upload_lattice(b)
    eng.run(0, 60)
    d, u, v, w = [np.ascontiguousarray(a) for a in eng.macro()]
    n = eng.dommax
    force = np.zeros((3, n))
    full = tmp_path / "start.csv"
    bc = np.ascontiguousarray(b, np.int32)
    assert h.zfh_write_fluid_csv(str(full).encode(), bc.ctypes.data, im, jm, km, 0.0, 0.0, 0.0, 1.0,
                                 d.ctypes.data, u.ctypes.data, v.ctypes.data, w.ctypes.data,
                                 force.ctypes.data, None, None, n) == 0
    lines = full.read_text().splitlines(keepends=True)
    (tmp_path / "FLUIDdomainStart.aa").write_text("".join(lines[:2000]))
    (tmp_path / "FLUIDdomainStart.ab").write_text("".join(lines[2000:]))
    bpath = tmp_path / "FLUIDdomainBindex.csv"
    assert h.zfh_write_bindex_csv(str(bpath).encode(), bc.ctypes.data, im, jm, km, 0.0, 0.0, 0.0, 1.0) == 0
    out = tmp_path / "out"
    case = tmp_path / "case.cfg"
    case.write_text(f"""box = {im} {jm} {km}
grid = 0 0 0 1
bindex = {bpath}
units = lattice
tau = 0.8
storage = f64
inlet = 0.02 0.3 40 1 0 0
fluid_start = {tmp_path / "FLUIDdomainStart.aa"} {tmp_path / "FLUIDdomainStart.ab"}
steps = 6
output_every = 6
output_dir = {out}
""")
    r = subprocess.run([os.path.join(PKG, "zf_host"), str(case)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout[-800:], flush=True)
    assert r.returncode == 0 and "fluid start state from 2 file(s)" in r.stdout
    # the same start through the C ABI
    e2 = Engine(cfg)
    e2.upload_lattice(b)
    got, mac = read_start_box([str(tmp_path / "FLUIDdomainStart.aa"),
                               str(tmp_path / "FLUIDdomainStart.ab")], bc)
    assert got == n
    mac = [np.ascontiguousarray(a) for a in (mac[0], mac[1], mac[2], mac[3], mac[4:])]
    assert h.zfh_init_from_macro(e2.h, *[a.ctypes.data for a in mac]) == 0
    e2.run(0, 6)
    d2, u2, v2, w2 = [np.ascontiguousarray(a) for a in e2.macro()]
    ref = tmp_path / "ref.csv"
    f2 = np.ascontiguousarray(e2.force())
    assert h.zfh_write_fluid_csv(str(ref).encode(), bc.ctypes.data, im, jm, km, 0.0, 0.0, 0.0, 1.0,
                                 d2.ctypes.data, u2.ctypes.data, v2.ctypes.data, w2.ctypes.data,
                                 f2.ctypes.data, None, None, n) == 0
    e2.close()
    eng.close()
    produced = out / f"FLUIDdomain-{6:09d}.csv"
    # zf_host writes the EulNeigh/Windex columns from its cell list and wall map
    # (none here: zeros), as the engine-side file above does
    assert filecmp.cmp(produced, ref, shallow=False)


def read_start_box(paths, b):
    import ctypes as C
    n = int((b >= 0).sum())
    out = np.zeros((7, n))
    arr = (C.c_char_p * len(paths))(*[p.encode() for p in paths])
    im, jm, km = b.shape
    got = load_host_library().zfh_read_fluid_start(
        len(paths), C.cast(arr, C.c_void_p), b.ctypes.data, im, jm, km, 0.0, 0.0, 0.0, 1.0, n,
        out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data, out[3].ctypes.data,
        out[4:].ctypes.data)
    return got, out
