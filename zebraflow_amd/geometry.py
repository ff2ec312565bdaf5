"""Python binding of the product's C geometry generators (libzfhost.so): tubes,
vessel pairs, branching trees, RBC meshes, cell placement and wall meshes
in the layouts the C ABI takes."""
import numpy as np

from . import load_host_library


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


def rbc_mesh(radius=4.0, level=3, biconcave=True):
    """Evans-Fung RBC mesh: (xyz [nn,3], tris [nt,3] 1-based)."""
    import ctypes as C
    cap_v, cap_t = 10 * 4 ** level + 2, 20 * 4 ** level
    xyz = np.zeros((cap_v, 3))
    tri = np.zeros((cap_t, 3), np.int32)
    nt = C.c_int32()
    nv = load_host_library().zfh_mesh_icosphere(level, radius, int(biconcave), xyz.ctypes.data,
                                                tri.ctypes.data, cap_v, cap_t, C.byref(nt))
    assert nv == cap_v and nt.value == cap_t
    return xyz, tri


def place_cells(b, ref, r_cell, gap, ncell, seed, max_tries=200000):
    im, jm, km = b.shape
    b = np.ascontiguousarray(b, np.int32)
    ref = np.ascontiguousarray(ref, np.float64)
    out = np.zeros((ncell, len(ref), 3))
    n = load_host_library().zfh_place_cells(b.ctypes.data, im, jm, km, ref.ctypes.data, len(ref),
                                            r_cell, gap, ncell, seed, max_tries, out.ctypes.data)
    return out[:n]


def place_cells_dense(b, ref, gap, ncell, seed, max_tries=2000000):
    im, jm, km = b.shape
    b = np.ascontiguousarray(b, np.int32)
    ref = np.ascontiguousarray(ref, np.float64)
    out = np.zeros((ncell, len(ref), 3))
    n = load_host_library().zfh_place_cells_dense(b.ctypes.data, im, jm, km, ref.ctypes.data,
                                                  len(ref), gap, ncell, seed, max_tries,
                                                  out.ctypes.data)
    assert n >= 0
    return out[:n]


def wall_mesh(b, x0=0.0, y0=0.0, z0=0.0, dx=1.0):
    """Surface-nets wall mesh: (xyz [nv,3], tri [nt,3] 1-based, normal [nt,3])."""
    import ctypes as C
    lib = load_host_library()
    im, jm, km = b.shape
    b = np.ascontiguousarray(b, np.int32)
    nv, nt = C.c_int64(), C.c_int64()
    assert lib.zfh_wall_mesh(b.ctypes.data, im, jm, km, x0, y0, z0, dx, None, 0, None, None, 0,
                             C.byref(nv), C.byref(nt)) == 0
    xyz = np.zeros((nv.value, 3))
    tri = np.zeros((nt.value, 3), np.int32)
    nrm = np.zeros((nt.value, 3))
    assert lib.zfh_wall_mesh(b.ctypes.data, im, jm, km, x0, y0, z0, dx, xyz.ctypes.data, nv.value,
                             tri.ctypes.data, nrm.ctypes.data, nt.value, C.byref(nv),
                             C.byref(nt)) == 0
    return xyz, tri, nrm
