"""Exact state checkpoints: device slots (zf_checkpoint_save/restore) and the
host program's restart files (checkpoint_every / restart keys of zf_host),
which hold the populations as stored words.

A tube with a pulsatile inlet and two cells, membrane/IBM every second
iteration: saving at iteration 6, running on, restoring and running the
same iterations again reproduces populations, force field and cell
positions bit for bit; a zf_host run restarted from its iteration-6 file
writes the same iteration-12 files as the uninterrupted run.
"""
import filecmp
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.test_host_program import write_case
from zebraflow_amd import geometry
from zebraflow_amd import PKG
from zebraflow_amd.engine import (B_PRESSURE_OUTLET, B_VELOCITY_INLET, STORE_F16DEV, STORE_F32DEV,
                                  STORE_F64, Engine, lattice_config, lattice_ibm_config,
                                  lattice_membrane_config, lattice_repulsion_config)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("store", [STORE_F64, STORE_F32DEV])
def test_device_slots_restore_bit_identically(store):
    print(f"case checkpoint slots, storage {store}", flush=True)
    b = geometry.tube((30, 20, 20), axis=0, r=8.0, inlet=B_VELOCITY_INLET, outlet=B_PRESSURE_OUTLET)
    xyz, tris = geometry.rbc_mesh(radius=3.0, level=2)
    cells = geometry.place_cells(b, xyz, 3.0, 1.0, 2, 5)
    eng = Engine(lattice_config(*b.shape, 0.8, storage=store, inlet_u0=0.02, inlet_amp=0.3,
                                inlet_period=40.0, inlet_dir=(1, 0, 0)))
    eng.upload_lattice(b)
    eng.membrane_setup(lattice_membrane_config(xyz, tris, 2), tris, xyz)
    eng.upload_cells(cells)
    eng.ibm_setup(lattice_ibm_config(rbc_interval=2))
    eng.repulsion_setup(lattice_repulsion_config())
    eng.wall_setup(*geometry.wall_mesh(b))
    eng.init_equilibrium()
    eng.full_run(0, 6)
    eng.checkpoint_save(3)
    eng.full_run(6, 7)
    a = (eng.populations(), eng.force(), eng.cells()[0])
    eng.checkpoint_restore(3)
    eng.full_run(6, 7)
    b2 = (eng.populations(), eng.force(), eng.cells()[0])
    for x, y in zip(a, b2):
        np.testing.assert_array_equal(x, y)
    assert np.abs(a[1]).max() > 0  # the force field was spread
    with pytest.raises(Exception):
        eng.checkpoint_restore(4)  # an empty slot is an error, not a no-op
    eng.close()


def test_host_restart_file_continues_bit_identically(tmp_path):
    print("case zf_host restart", flush=True)
    b = geometry.tube((30, 20, 20), axis=0, r=8.0, inlet=B_VELOCITY_INLET, outlet=B_PRESSURE_OUTLET)
    xyz, tris = geometry.rbc_mesh(radius=3.0, level=2)
    cells = geometry.place_cells(b, xyz, 3.0, 1.0, 2, 5)
    exe = os.path.join(PKG, "zf_host")
    d1 = tmp_path / "full"
    d1.mkdir()
    cfg = write_case(str(d1), b, cells, xyz, tris, 12, 6, extra="checkpoint_every = 6")
    r = subprocess.run([exe, cfg], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0
    state = d1 / "out" / f"ZFstate-{6:09d}.bin"
    assert state.exists()
    d2 = tmp_path / "restart"
    d2.mkdir()
    cfg2 = write_case(str(d2), b, cells, xyz, tris, 6, 6, extra=f"restart = {state}")
    r = subprocess.run([exe, cfg2], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0 and "restarted from" in r.stdout
    for name in (f"FLUIDdomain-{12:09d}.csv", f"AllCellMemb-{12:09d}.csv"):
        assert filecmp.cmp(d1 / "out" / name, d2 / "out" / name, shallow=False), name


@pytest.mark.parametrize("store", [STORE_F32DEV, STORE_F16DEV, STORE_F64])
def test_restart_file_keeps_every_stored_bit(tmp_path, store):
    """A restart file holds the populations as stored: on a state whose FP32
    deviations f - w sit near 1e-11 (where f = w + g in double no longer
    holds g: the double layout would round about a third of them), the
    reloaded words are the saved ones bit for bit, and the run continues
    bit-identically."""
    print(f"case restart raw words, store {store}", flush=True)
    b = geometry.tube((20, 14, 14), axis=0, r=5.0, inlet=B_VELOCITY_INLET, outlet=B_PRESSURE_OUTLET)
    cfg = lattice_config(*b.shape, 0.8, storage=store, inlet_u0=0.01, inlet_amp=0.2,
                         inlet_period=50.0, inlet_dir=(1, 0, 0))
    eng = Engine(cfg)
    eng.upload_lattice(b)
    n = eng.dommax
    rng = np.random.default_rng(9)
    w = np.array(po.W)[:, None]
    g = (10.0 ** rng.uniform(-14, -9, (19, n))) * rng.choice([-1.0, 1.0], (19, n))
    eng.upload_populations(w + g)
    eng.run(0, 3)
    raw0 = eng.populations_raw()
    if store == STORE_F32DEV:
        # the double round trip f = w + g, g' = f - w changes many of these words
        back = ((raw0.astype(np.float64) + w) - w).astype(np.float32)
        assert (back != raw0).mean() > 0.05
    path = tmp_path / "state.bin"
    eng.save_state(path, 3)
    eng.run(3, 5)
    after = eng.populations_raw()
    e2 = Engine(cfg)
    e2.upload_lattice(b)
    assert e2.load_state(path) == 3
    np.testing.assert_array_equal(e2.populations_raw().view(np.uint8), raw0.view(np.uint8))
    e2.run(3, 5)
    np.testing.assert_array_equal(e2.populations_raw().view(np.uint8), after.view(np.uint8))
    eng.close()
    e2.close()
