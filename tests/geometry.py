"""Small test geometries from the product's C generators (libzfhost.so)."""
import numpy as np

from zebraflow_amd import load_host_library


def tube(shape, axis, r, periodic=False, inlet=0, outlet=0):
    im, jm, km = shape
    b = np.zeros(shape, np.int32)
    n = load_host_library().zfh_geom_tube(b.ctypes.data, im, jm, km, axis, r, int(periodic),
                                          inlet, outlet)
    assert n > 0
    return b


def vessel_pair(shape, r_main, r_conn, amp, n_conn, seed):
    im, jm, km = shape
    b = np.zeros(shape, np.int32)
    n = load_host_library().zfh_geom_vessel_pair(b.ctypes.data, im, jm, km, r_main, r_conn, amp,
                                                 n_conn, seed)
    assert n > 0
    return b


def tree(shape, r_root, ratio, generations, a_lo, a_hi, seed):
    im, jm, km = shape
    b = np.zeros(shape, np.int32)
    n = load_host_library().zfh_geom_tree(b.ctypes.data, im, jm, km, r_root, ratio, generations,
                                          a_lo, a_hi, seed)
    assert n > 0
    return b
