"""Device-built index lists are bit-exact with the reference (goldens) and
with the oracle on the benchmark geometries (runs through libzfgpu.so)."""
import glob
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from zebraflow_amd.engine import Engine, ZfError, lattice_config, reference_config

pytestmark = pytest.mark.gpu

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "lists_*.npz")))


@pytest.mark.parametrize("path", GOLD, ids=lambda p: os.path.basename(p))
def test_device_lists_match_reference_goldens(path):
    g = np.load(path)
    b = g["b"].astype(np.int32)
    print(f"case {os.path.basename(path)} box {b.shape}", flush=True)
    eng = Engine(reference_config(*b.shape))
    # Random fixtures hold interior nodes next to B<0 nodes, which the stream
    # tables reject (the reference would pull DOMID=-1); the lists stay valid.
    try:
        eng.upload_lattice(b)
    except ZfError as e:
        assert "outside the list" in str(e)
        eng.dommax = int((b >= 0).sum())
        eng.bcmax = int(((b == 0) | (b >= 1000)).sum())
    dl, bl, bn = eng.lists()
    np.testing.assert_array_equal(dl, g["domlist"])
    np.testing.assert_array_equal(bl, g["bclist"])
    np.testing.assert_array_equal(bn, g["bcnlist"])
    domid = eng.domid()
    ref_domid, _ = po.dom_setup(b)
    np.testing.assert_array_equal(domid, ref_domid)
    eng.close()


@pytest.mark.parametrize("case", ["tube_periodic", "vessel_pair", "tree"])
def test_device_lists_match_oracle(case):
    if case == "tube_periodic":
        b = geometry.tube((32, 32, 48), axis=2, r=14.0, periodic=True)
        per = (0, 0, 1)
    elif case == "vessel_pair":
        b = geometry.vessel_pair((160, 64, 64), 8.0, 3.0, 6.0, 4, 11)
        per = (0, 0, 0)
    else:
        b = geometry.tree((160, 96, 96), 12.0, 0.79, 3, 30.0, 45.0, 5)
        per = (0, 0, 0)
    print(f"case {case} box {b.shape}", flush=True)
    eng = Engine(lattice_config(*b.shape, 0.8, periodic=per))
    eng.upload_lattice(b)
    dl, bl, bn = eng.lists()
    domid, dl2 = po.dom_setup(b)
    bl2, bn2, _ = po.bc_setup(b, per)
    np.testing.assert_array_equal(dl, dl2)
    np.testing.assert_array_equal(bl, bl2)
    np.testing.assert_array_equal(bn, bn2)
    np.testing.assert_array_equal(eng.domid(), domid)
    np.testing.assert_array_equal(eng.neighbours(), po.neighbours(domid, dl2, per))
    info = eng.info()
    assert info["n_interior"] == ((b > 0) & (b < 10)).sum()
    eng.close()
