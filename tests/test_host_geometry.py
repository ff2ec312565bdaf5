"""Host-side geometry and reference-format I/O (CPU)."""
import numpy as np

from zebraflow_amd import geometry
from zebraflow_amd import load_host_library
from oracle import pyoracle as po


def test_tube_layers_and_codes():
    b = geometry.tube((12, 20, 20), axis=0, r=6.0, inlet=2001, outlet=3001)
    assert set(np.unique(b)) == {-1, 0, 1, 2001, 3001}
    assert (b[0] == 2001).sum() == (b[-1] == 3001).sum() > 0
    # every fluid node's 18 neighbours are listed (B >= 0): the wall layer closes
    fl = np.argwhere(b == 1)
    for e in po.E[:18]:
        x = fl + e
        inside = ((x >= 0) & (x < b.shape)).all(1)
        assert (b[tuple(x[inside].T)] >= 0).all()


def test_vessel_pair_is_closed_and_connected():
    b = geometry.vessel_pair((128, 64, 48), 6.0, 3.0, 4.0, 3, 7)
    assert (b == 2001).sum() > 0 and (b == 3001).sum() > 0
    fl = np.argwhere((b > 0) & (b < 10))
    for e in po.E[:18]:
        x = fl + e
        assert ((x >= 0) & (x < b.shape)).all()
        assert (b[tuple(x.T)] >= 0).all()


def test_bindex_csv_roundtrip(tmp_path):
    b = geometry.tube((8, 12, 12), axis=2, r=4.0, periodic=True)
    p = str(tmp_path / "FLUIDdomainBindex.csv")
    lib = load_host_library()
    assert lib.zfh_write_bindex_csv(p.encode(), b.ctypes.data, 8, 12, 12, -1.0, 29.0, -27.0, 0.5) == 0
    r = np.zeros_like(b)
    n = lib.zfh_read_bindex_csv(p.encode(), r.ctypes.data, 8, 12, 12, -1.0, 29.0, -27.0, 0.5)
    assert n == (b >= 0).sum()
    np.testing.assert_array_equal(r, b)
