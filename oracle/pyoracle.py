"""ctypes bindings of liboracle.so, the CPU restatement (TEST INFRASTRUCTURE).

Only tests/, the smoke entry point and bench.py's CPU-baseline leg import
this module.  See zf_oracle.h for the functions and the reference lines they
restate.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


class LbmParams(C.Structure):
    _fields_ = [
        ("im", C.c_int), ("jm", C.c_int), ("km", C.c_int), ("periodic", C.c_int * 3),
        ("wall_mode", C.c_int), ("visc_scale", C.c_double), ("force_dt", C.c_double),
        ("body_force", C.c_double * 3), ("inlet_u0", C.c_double), ("inlet_amp", C.c_double),
        ("inlet_period", C.c_double), ("inlet_dir", C.c_double * 3), ("outlet_rho", C.c_double),
        ("t_scale", C.c_double), ("pressure_scale", C.c_double),
    ]


class Lattice(C.Structure):
    _fields_ = [
        ("im", C.c_int), ("jm", C.c_int), ("km", C.c_int), ("dommax", C.c_int), ("bcmax", C.c_int),
        ("b_index", C.c_void_p), ("domid", C.c_void_p), ("domlist", C.c_void_p),
        ("bclist", C.c_void_p), ("bcnlist", C.c_void_p),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: make -C oracle")
        _lib = C.CDLL(LIB)
        _lib.zfo_dom_setup.restype = C.c_int
        _lib.zfo_bc_setup.restype = C.c_int
        _lib.zfo_parent_bc_setup.restype = C.c_int
        _lib.zfo_feq.restype = C.c_double
        _lib.zfo_feq.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        _lib.zfo_bc_pressure.restype = C.c_float
        _lib.zfo_bc_pressure.argtypes = [C.c_int, C.c_float]
        _lib.zfo_wtime.restype = C.c_double
        _lib.zfo_max_threads.restype = C.c_int
        _lib.zfo_nucleus_contact.restype = C.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


E = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [-1, 1, 0], [1, -1, 0], [1, 0, 1], [-1, 0, -1],
              [-1, 0, 1], [1, 0, -1], [0, 1, 1], [0, -1, -1], [0, -1, 1], [0, 1, -1],
              [0, 0, 0]])
W = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])


def dom_setup(b):
    b = np.ascontiguousarray(b, np.int32)
    im, jm, km = b.shape
    domid = np.zeros(b.shape, np.int32)
    cap = int((b >= 0).sum())
    domlist = np.zeros((max(cap, 1), 3), np.int32)
    n = lib().zfo_dom_setup(_p(b), im, jm, km, _p(domid), _p(domlist), cap)
    assert n == cap
    return domid, domlist[:n]


def bc_setup(b, periodic=(0, 0, 0)):
    b = np.ascontiguousarray(b, np.int32)
    im, jm, km = b.shape
    cap = int(((b == 0) | (b >= 1000)).sum())
    bl = np.zeros((max(cap, 1), 4), np.int32)
    bn = np.zeros((max(cap, 1), 3), np.int32)
    per = (C.c_int * 3)(*periodic)
    nf = C.c_int()
    n = lib().zfo_bc_setup(_p(b), im, jm, km, per, _p(bl), _p(bn), cap, C.byref(nf))
    assert n == cap
    return bl[:n], bn[:n], nf.value


def parent_bc_setup(b, p, k_over, off):
    """BC list, neighbour list and upstream velocity sources of parent lattice p."""
    b = np.ascontiguousarray(b, np.int32)
    im, jm, km = b.shape
    cap = int(((b == 0) | (b >= 10)).sum())
    bl = np.zeros((max(cap, 1), 4), np.int32)
    bn = np.zeros((max(cap, 1), 3), np.int32)
    vb = np.zeros((max(cap, 1), 3), np.int32)
    o = (C.c_int * 3)(*off)
    nf = C.c_int()
    n = lib().zfo_parent_bc_setup(_p(b), im, jm, km, p, k_over, o, _p(bl), _p(bn), _p(vb), cap,
                                  C.byref(nf))
    assert n == cap
    return bl[:n], bn[:n], vb[:n], nf.value


def neighbours(domid, domlist, periodic=(0, 0, 0)):
    im, jm, km = domid.shape
    n = len(domlist)
    nbr = np.zeros((18, n), np.int32)
    per = (C.c_int * 3)(*periodic)
    lib().zfo_neighbours(_p(np.ascontiguousarray(domid)), _p(np.ascontiguousarray(domlist)), n,
                         im, jm, km, per, _p(nbr))
    return nbr


def feq(rho, u):
    """feq_c(rho, u) for rho[n], u[3, n] -> [19, n] (ZebraflowWT.c:6139, C_LBM = 1)."""
    rho = np.asarray(rho, np.float64)
    u = np.asarray(u, np.float64)
    cu = E.astype(np.float64) @ u
    usq = (u * u).sum(axis=0)
    return W[:, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)


class LbmOracle:
    """Reference-structured LBM (F, F_OLD, FEQ, D, U, V, W, BODYFORCE)."""

    def __init__(self, b, params, dtype=np.float64, periodic=None, parent=None):
        """parent: (p, k_over, offset) for a parent lattice (its BC lists and
        upstream velocity sources from parent_bc_setup)."""
        self.b = np.ascontiguousarray(b, np.int32)
        self.im, self.jm, self.km = self.b.shape
        per = periodic if periodic is not None else tuple(params.get("periodic", (0, 0, 0)))
        self.domid, self.domlist = dom_setup(self.b)
        self.parent = parent
        if parent is None:
            self.bclist, self.bcnlist, self.n_fail = bc_setup(self.b, per)
            self.vbcn = None
        else:
            self.bclist, self.bcnlist, self.vbcn, self.n_fail = parent_bc_setup(self.b, *parent)
        self.n = len(self.domlist)
        self.dtype = np.dtype(dtype)
        p = LbmParams()
        p.im, p.jm, p.km = self.im, self.jm, self.km
        for i, v in enumerate(per):
            p.periodic[i] = v
        for k, v in params.items():
            if k == "periodic":
                continue
            if k in ("body_force", "inlet_dir"):
                arr = getattr(p, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(p, k, v)
        self.params = p
        lat = Lattice()
        lat.im, lat.jm, lat.km, lat.dommax, lat.bcmax = self.im, self.jm, self.km, self.n, len(self.bclist)
        self._keep = [np.ascontiguousarray(self.domid), np.ascontiguousarray(self.domlist),
                      np.ascontiguousarray(self.bclist), np.ascontiguousarray(self.bcnlist)]
        lat.b_index = _p(self.b)
        lat.domid, lat.domlist, lat.bclist, lat.bcnlist = [_p(a) for a in self._keep]
        self.lat = lat
        n = self.n
        self.F = np.zeros((19, n), self.dtype)
        self.F_OLD = np.zeros((19, n), self.dtype)
        self.FEQ = np.zeros((19, n), self.dtype)
        self.D = np.ones(n, self.dtype)
        self.U = np.zeros(n, self.dtype)
        self.V = np.zeros(n, self.dtype)
        self.W = np.zeros(n, self.dtype)
        self.force = np.zeros((3, n), self.dtype)
        self.set_populations(feq(np.ones(n), np.zeros((3, n))))

    def set_populations(self, f):
        self.F[...] = f
        self.F_OLD[...] = f
        d = self.F_OLD.astype(np.float64).sum(0)
        self.D[...] = d
        self.U[...] = (E[:, 0] @ self.F_OLD) / d
        self.V[...] = (E[:, 1] @ self.F_OLD) / d
        self.W[...] = (E[:, 2] @ self.F_OLD) / d

    def step(self, it, rbc_step=0, nthreads=0):
        fn = lib().zfo_lbm_step_f64 if self.dtype == np.float64 else lib().zfo_lbm_step_f32
        fn(C.byref(self.lat), C.byref(self.params), _p(self.F), _p(self.F_OLD), _p(self.FEQ),
           _p(self.D), _p(self.U), _p(self.V), _p(self.W), _p(self.force), int(it), int(rbc_step),
           int(nthreads))

    def phase(self, ph, it, rbc_step=0, nthreads=0):
        assert self.dtype == np.float64
        lib().zfo_lbm_phase_f64(C.byref(self.lat), C.byref(self.params), int(ph), _p(self.F),
                                _p(self.F_OLD), _p(self.FEQ), _p(self.D), _p(self.U), _p(self.V),
                                _p(self.W), _p(self.force), int(it), int(rbc_step), int(nthreads))

    @property
    def populations(self):
        return self.F_OLD

    def bnode(self):
        d = self.domlist
        return self.b[d[:, 0] - 1, d[:, 1] - 1, d[:, 2] - 1]


def step_with_parents(main, parents, it, rbc_step=0, nthreads=0, parent_rbc_step=0):
    """One iteration of a main lattice and its parent lattices (FP64) in the
    reference's order (ZebraflowWT.c:6101-6626): every stream + collide, the
    main boundary pass, the parents' boundary passes, every macro + copy
    (rbc_step, parent_rbc_step: the force field is spread again after this
    step, on the main lattice and on the parents carrying cells)."""
    main.phase(0, it, nthreads=nthreads)
    for p in parents:
        p.phase(0, it, nthreads=nthreads)
    main.phase(1, it, nthreads=nthreads)
    for p in parents:
        vb = np.ascontiguousarray(p.vbcn)
        lib().zfo_parent_boundary_f64(C.byref(p.lat), C.byref(p.params), int(p.parent[0]), _p(vb),
                                      _p(p.F), _p(p.D), C.byref(main.lat), _p(main.D), _p(main.U),
                                      _p(main.V), _p(main.W), int(nthreads))
    main.phase(2, it, rbc_step=rbc_step, nthreads=nthreads)
    for p in parents:
        p.phase(2, it, rbc_step=parent_rbc_step, nthreads=nthreads)


def wtime():
    return lib().zfo_wtime()


def max_threads():
    return lib().zfo_max_threads()


# ---- membrane (zf_oracle_mesh.c) ------------------------------------------
class MembParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "len_scale", "kT", "x0", "kbend", "kd", "ka", "kv", "area0", "vol0", "x_max", "x_floor",
        "cos_cut", "cos_blend", "area_lo", "area_hi", "force_limit", "vol_limit", "sliver_len",
        "sliver_len_min", "sliver_f1", "sliver_f2", "relax", "relax2")]


def reference_membrane_params():
    """SI constants of ZebraflowWT.c:34-36, :124-134, :5654-5655."""
    return dict(len_scale=1e-6, kT=300 * 1.3806488e-23, x0=0.3154574132, kbend=1.3856406e-19,
                kd=3.6e-5, ka=7.7e-5, kv=55.0, area0=144.916e-12, vol0=97.5671e-18, x_max=0.95,
                x_floor=0.02435, cos_cut=0.99, cos_blend=0.0, area_lo=0.1, area_hi=10.0, force_limit=1e-10,
                vol_limit=1e-10, sliver_len=0.05, sliver_len_min=0.001, sliver_f1=1e-14,
                sliver_f2=1e-11, relax=1.0, relax2=1.0)


def memb_params(d):
    p = MembParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def mesh_topology(nn, tris, nlist=10):
    tris = np.ascontiguousarray(tris, np.int32)
    nt = len(tris)
    cap = 3 * nt
    edge = np.zeros((cap, 6), np.int32)
    nbn = np.zeros((nn, nlist), np.int32)
    nbt = np.zeros((nn, nlist), np.int32)
    ne = lib().zfo_mesh_topology(nn, nt, _p(tris), _p(edge), cap, _p(nbn), _p(nbt), nlist)
    assert ne > 0
    return edge[:ne].copy(), nbn, nbt


def mesh_equilibrium(xyz, edge, mp, es, passes=51):
    xyz = np.ascontiguousarray(xyz, np.float64)
    edge = np.ascontiguousarray(edge, np.int32)
    ne = len(edge)
    out = [np.zeros(ne) for _ in range(6)]
    lib().zfo_mesh_equilibrium(len(xyz), ne, _p(xyz), _p(edge), C.byref(memb_params(mp)),
                               C.c_double(es), passes, *[_p(a) for a in out])
    return dict(zip(("lmax", "plength", "kp", "theta0", "sin0", "cos0"), out))


def tri_areas(xyz, tris, len_scale):
    p = xyz[tris - 1]
    a = 0.5 * np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]) * len_scale ** 2
    return np.linalg.norm(a, axis=1)


def membrane_forces(xyz_cells, tris, edge, eq, tri_area_eq, mp):
    """xyz_cells: [ncell, nn, 3].  Returns a dict of per-cell outputs."""
    x = np.ascontiguousarray(xyz_cells, np.float64)
    nc, nn, _ = x.shape
    nt, ne = len(tris), len(edge)
    out = dict(force=np.zeros((nc, nn, 3)), node_normal=np.zeros((nc, nn, 3)),
               tri_area=np.zeros((nc, nt)), tri_normal=np.zeros((nc, nt, 3)),
               tri_centroid=np.zeros((nc, nt, 3)), cell_area=np.zeros(nc), cell_vol=np.zeros(nc),
               cell_centroid=np.zeros((nc, 3)), guard_fail=np.zeros(nc, np.int32))
    t = np.ascontiguousarray(tris, np.int32)
    e = np.ascontiguousarray(edge, np.int32)
    arrs = [np.ascontiguousarray(eq[k], np.float64) for k in ("lmax", "plength", "kp", "sin0", "cos0")]
    tae = np.ascontiguousarray(tri_area_eq, np.float64)
    lib().zfo_membrane_forces(nc, nn, nt, ne, _p(t), _p(e), *[_p(a) for a in arrs], _p(tae),
                              C.byref(memb_params(mp)), _p(x),
                              *[_p(out[k]) for k in ("force", "node_normal", "tri_area",
                                                     "tri_normal", "tri_centroid", "cell_area",
                                                     "cell_vol", "cell_centroid", "guard_fail")])
    return out


def nucleus_equilibrium(xyz, edge, mp, es, passes=51):
    """LMAX_N, Plength_N, kp_N of the rest network (es: the network's
    shear modulus, 0.1 of the membrane's in the reference)."""
    x = np.ascontiguousarray(xyz, np.float64)
    e = np.ascontiguousarray(edge, np.int32)
    ne = len(e)
    out = [np.zeros(ne) for _ in range(3)]
    lib().zfo_nucleus_equilibrium(len(x), ne, _p(x), _p(e), C.byref(memb_params(mp)),
                                  C.c_double(es), passes, *[_p(a) for a in out])
    return dict(zip(("lmax", "plength", "kp"), out))


def nucleus_forces(xyzn_cells, edge, eq, mp):
    x = np.ascontiguousarray(xyzn_cells, np.float64)
    nc, nnuc, _ = x.shape
    e = np.ascontiguousarray(edge, np.int32)
    out = np.zeros_like(x)
    arrs = [np.ascontiguousarray(eq[k], np.float64) for k in ("lmax", "plength", "kp")]
    lib().zfo_nucleus_forces(nc, nnuc, len(e), _p(e), *[_p(a) for a in arrs],
                             C.byref(memb_params(mp)), _p(x), _p(out))
    return out


def nucleus_contact(xyzn_cells, xyz_cells, tris, status, tri_centroid, tri_normal, list_r=0.5,
                    near_r=1.0, contact=0.2, k=4.1666667e-07, list_cap=64, vrep=None):
    """Adds the contact's pushes to vrep [ncell, nn, 3] (zeros if None);
    returns (vrep, entries dropped from full lists)."""
    xn = np.ascontiguousarray(xyzn_cells, np.float64)
    x = np.ascontiguousarray(xyz_cells, np.float64)
    nc, nnuc, _ = xn.shape
    nn = x.shape[1]
    t = np.ascontiguousarray(tris, np.int32)
    v = np.zeros_like(x) if vrep is None else vrep
    lost = lib().zfo_nucleus_contact(nc, nnuc, nn, len(t), _p(t),
                                     _p(np.ascontiguousarray(status, np.int32)), _p(xn), _p(x),
                                     _p(np.ascontiguousarray(tri_centroid, np.float64)),
                                     _p(np.ascontiguousarray(tri_normal, np.float64)),
                                     C.c_double(list_r), C.c_double(near_r), C.c_double(contact),
                                     C.c_double(k), int(list_cap), _p(v))
    return v, lost


# ---- immersed boundary (zf_oracle_ibm.c) ----------------------------------
class IbmParams(C.Structure):
    _fields_ = [("im", C.c_int), ("jm", C.c_int), ("km", C.c_int)] + [
        (n, C.c_double) for n in ("x0", "y0", "z0", "dx", "vel_scale", "force_scale",
                                  "dt_advance", "vmax")]


def ibm_params(shape, **kw):
    p = IbmParams()
    p.im, p.jm, p.km = shape
    d = dict(x0=0.0, y0=0.0, z0=0.0, dx=1.0, vel_scale=1.0, force_scale=1.0, dt_advance=1.0,
             vmax=1e30)
    d.update(kw)
    for k, v in d.items():
        setattr(p, k, v)
    return p


def ibm_interpolate(p, b, domid, uvw, xyz, windex=None, uvoid=None):
    """Interpolated node velocities; with windex and uvoid ([3, im, jm, km])
    the wall layer's Uvoid terms are added (zfo_ibm_interpolate_wall)."""
    b = np.ascontiguousarray(b, np.int32)
    domid = np.ascontiguousarray(domid, np.int32)
    u = [np.ascontiguousarray(a, np.float64) for a in uvw]
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    v = np.zeros_like(x)
    if uvoid is None:
        lib().zfo_ibm_interpolate(C.byref(p), _p(b), _p(domid), _p(u[0]), _p(u[1]), _p(u[2]),
                                  C.c_int64(len(x)), _p(x), _p(v))
    else:
        wi = np.ascontiguousarray(windex, np.int32)
        uv = np.ascontiguousarray(uvoid, np.float64)
        lib().zfo_ibm_interpolate_wall(C.byref(p), _p(b), _p(domid), _p(u[0]), _p(u[1]), _p(u[2]),
                                       _p(wi), _p(uv), C.c_int64(len(x)), _p(x), _p(v))
    return v.reshape(np.shape(xyz))


def windex_dist(p, wall_xyz):
    wx = np.ascontiguousarray(wall_xyz, np.float64)
    windex = np.zeros((p.im, p.jm, p.km), np.int32)
    wdist = np.zeros((p.im, p.jm, p.km), np.float32)
    lib().zfo_windex_dist(C.byref(p), len(wx), _p(wx), _p(windex), _p(wdist))
    return windex, wdist


def uvoid(p, b, windex, wdist, wall_xyz, wall_tri, wall_normal, wall_node_tri, speed=1e-3,
          d0=0.7, d1=3.0, k_lo=3, k_hi=2, select=1, point=(78.5, 214.5, 392.5)):
    out = np.zeros((3, p.im, p.jm, p.km))
    sp = np.ascontiguousarray(point, np.float64)
    lib().zfo_uvoid(C.byref(p), _p(np.ascontiguousarray(b, np.int32)),
                    _p(np.ascontiguousarray(windex, np.int32)),
                    _p(np.ascontiguousarray(wdist, np.float32)),
                    _p(np.ascontiguousarray(wall_xyz, np.float64)),
                    _p(np.ascontiguousarray(wall_tri, np.int32)),
                    _p(np.ascontiguousarray(wall_normal, np.float64)),
                    _p(np.ascontiguousarray(wall_node_tri, np.int32)), C.c_double(speed),
                    C.c_double(d0), C.c_double(d1), int(k_lo), int(k_hi), int(select), _p(sp),
                    _p(out))
    return out


def ibm_spread(p, b, domid, dommax, xyz, ffluid, force=None, parallel=False):
    """Spreading onto force [3, dommax] (added); parallel: the reference's
    OpenMP-atomic structure (zfo_ibm_spread_omp, order of sums free)."""
    b = np.ascontiguousarray(b, np.int32)
    domid = np.ascontiguousarray(domid, np.int32)
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(ffluid, np.float64).reshape(-1, 3)
    out = np.zeros((3, dommax)) if force is None else force
    fn = lib().zfo_ibm_spread_omp if parallel else lib().zfo_ibm_spread
    fn(C.byref(p), _p(b), _p(domid), C.c_int64(dommax), C.c_int64(len(x)), _p(x), _p(f), _p(out))
    return out


def wss(p, b, domid, macro, xyz, nrm, mu_over_dx, pressure_scale=10000.0 / 3, pgauge=0.0):
    """Wall shear stress [n,3], wall pressure [n], valid [n] at wall points;
    macro = (D, U, V, W) per listed node."""
    x = np.ascontiguousarray(xyz, np.float64)
    nv = np.ascontiguousarray(nrm, np.float64)
    n = len(x)
    out, pr, va = np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.int32)
    m = [np.ascontiguousarray(a, np.float64) for a in macro]
    lib().zfo_wss(C.byref(p), _p(np.ascontiguousarray(b, np.int32)),
                  _p(np.ascontiguousarray(domid, np.int32)), *[_p(a) for a in m], C.c_int64(n),
                  _p(x), _p(nv), C.c_double(mu_over_dx), C.c_double(pressure_scale),
                  C.c_double(pgauge), _p(out), _p(pr), _p(va))
    return out, pr, va


def advance(p, xyz, vel, status):
    x = np.ascontiguousarray(xyz, np.float64)
    v = np.ascontiguousarray(vel, np.float64)
    st = np.ascontiguousarray(status, np.int32)
    nc, nn, _ = x.shape
    removed = lib().zfo_advance(C.byref(p), nc, nn, _p(x), _p(v), _p(st))
    return x, st, removed


# ---- cell list and repulsion (zf_oracle_cells.c) ---------------------------
class RepParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("bin_cutoff", "k_cell", "cutoff", "search", "k_wall",
                                          "wall_d0", "wall_dmax", "wall_search")]


def rep_params(**kw):
    d = dict(bin_cutoff=1.0, k_cell=8.333e-7, cutoff=1.0, search=2.0, k_wall=1e-4, wall_d0=0.25,
             wall_dmax=1.0, wall_search=1.5)
    d.update(kw)
    r = RepParams()
    for k, v in d.items():
        setattr(r, k, v)
    return r


def cell_list(p, b, domid, tri_centroid, status, bin_cutoff):
    tc = np.ascontiguousarray(tri_centroid, np.float64)
    nc, nt, _ = tc.shape
    out = np.zeros((nc, nt), np.int32)
    lib().zfo_cell_list(C.byref(p), _p(np.ascontiguousarray(b, np.int32)),
                        _p(np.ascontiguousarray(domid, np.int32)), nc, nt,
                        _p(np.ascontiguousarray(status, np.int32)), _p(tc), C.c_double(bin_cutoff),
                        _p(out))
    return out


def cell_repulsion(p, r, b, domid, tris, status, xyz, node_normal, tri_centroid, tri_normal, bins):
    x = np.ascontiguousarray(xyz, np.float64)
    nc, nn, _ = x.shape
    nt = len(tris)
    v = np.zeros_like(x)
    lib().zfo_cell_repulsion(C.byref(p), C.byref(r), _p(np.ascontiguousarray(b, np.int32)),
                             _p(np.ascontiguousarray(domid, np.int32)), nc, nn, nt,
                             _p(np.ascontiguousarray(tris, np.int32)),
                             _p(np.ascontiguousarray(status, np.int32)), _p(x),
                             _p(np.ascontiguousarray(node_normal, np.float64)),
                             _p(np.ascontiguousarray(tri_centroid, np.float64)),
                             _p(np.ascontiguousarray(tri_normal, np.float64)),
                             _p(np.ascontiguousarray(bins, np.int32)), _p(v))
    return v


def wall_tables(p, wall_xyz, wall_tri):
    wx = np.ascontiguousarray(wall_xyz, np.float64)
    wt = np.ascontiguousarray(wall_tri, np.int32)
    nwt = np.zeros((len(wx), 13), np.int32)
    lib().zfo_wall_node_tri(len(wx), len(wt), _p(wt), _p(nwt))
    windex = np.zeros((p.im, p.jm, p.km), np.int32)
    lib().zfo_windex(C.byref(p), len(wx), _p(wx), _p(windex))
    return nwt, windex


def wall_repulsion(p, r, windex, wall_xyz, wall_tri, wall_normal, wall_node_tri, status, xyz):
    x = np.ascontiguousarray(xyz, np.float64)
    nc, nn, _ = x.shape
    v = np.zeros_like(x)
    near = np.zeros((nc, nn), np.int32)
    wx = np.ascontiguousarray(wall_xyz, np.float64)
    wt = np.ascontiguousarray(wall_tri, np.int32)
    lib().zfo_wall_repulsion(C.byref(p), C.byref(r), _p(np.ascontiguousarray(windex, np.int32)),
                             len(wx), _p(wx), len(wt), _p(wt),
                             _p(np.ascontiguousarray(wall_normal, np.float64)),
                             _p(np.ascontiguousarray(wall_node_tri, np.int32)), nc, nn,
                             _p(np.ascontiguousarray(status, np.int32)), _p(x), _p(v), _p(near))
    return v, near


# ---- cells on a parent lattice (zf_oracle_wrap.c) ---------------------------
class Wrap(C.Structure):
    _fields_ = [("dir", C.c_int), ("z0", C.c_double), ("zm", C.c_double)]


def wrap_of(p, direction):
    """The parent's cell window along k: Z0_p = z0, ZM_p = z0 + (km-1) dx."""
    w = Wrap()
    w.dir, w.z0, w.zm = direction, p.z0, p.z0 + (p.km - 1) * p.dx
    return w


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def wrap_interpolate(p, w, b, domid, uvw, xyz, status, windex=None, uvoid=None):
    x = _f64(xyz)
    nc, nn, _ = x.shape
    v = np.zeros_like(x)
    u = [_f64(a) for a in uvw]
    wi = _i32(windex) if uvoid is not None else None
    uv = _f64(uvoid) if uvoid is not None else None
    lib().zfo_wrap_interpolate(C.byref(p), C.byref(w), _p(_i32(b)), _p(_i32(domid)), _p(u[0]),
                               _p(u[1]), _p(u[2]), _p(wi) if wi is not None else None,
                               _p(uv) if uv is not None else None, nc, nn, _p(_i32(status)),
                               _p(x), _p(v))
    return v


def wrap_wall_repulsion(p, w, r, windex, wall_xyz, wall_tri, wall_normal, wall_node_tri, status,
                        xyz, vrep=None):
    x = _f64(xyz)
    nc, nn, _ = x.shape
    v = np.zeros_like(x) if vrep is None else vrep
    near = np.zeros((nc, nn), np.int32)
    lib().zfo_wrap_wall_repulsion(C.byref(p), C.byref(w), C.byref(r), _p(_i32(windex)),
                                  _p(_f64(wall_xyz)), _p(_i32(wall_tri)), _p(_f64(wall_normal)),
                                  _p(_i32(wall_node_tri)), nc, nn, _p(_i32(status)), _p(x), _p(v),
                                  _p(near))
    return v, near


def wrap_cell_list(p, w, b, domid, tri_centroid, status):
    tc = _f64(tri_centroid)
    nc, nt, _ = tc.shape
    out = np.zeros((nc, nt), np.int32)
    lib().zfo_wrap_cell_list(C.byref(p), C.byref(w), _p(_i32(b)), _p(_i32(domid)), nc, nt,
                             _p(_i32(status)), _p(tc), _p(out))
    return out


def wrap_cell_repulsion(p, w, r, b, domid, tris, status, xyz, node_normal, tri_centroid,
                        tri_normal, bins, vrep=None):
    x = _f64(xyz)
    nc, nn, _ = x.shape
    v = np.zeros_like(x) if vrep is None else vrep
    lib().zfo_wrap_cell_repulsion(C.byref(p), C.byref(w), C.byref(r), _p(_i32(b)),
                                  _p(_i32(domid)), nc, nn, len(tris), _p(_i32(tris)),
                                  _p(_i32(status)), _p(x), _p(_f64(node_normal)),
                                  _p(_f64(tri_centroid)), _p(_f64(tri_normal)), _p(_i32(bins)),
                                  _p(v))
    return v


def wrap_spread(p, w, b, domid, dommax, xyz, ffluid, status, force=None):
    x = _f64(xyz)
    nc, nn, _ = x.shape
    out = np.zeros((3, dommax)) if force is None else force
    lib().zfo_wrap_spread(C.byref(p), C.byref(w), _p(_i32(b)), _p(_i32(domid)),
                          C.c_int64(dommax), nc, nn, _p(_i32(status)), _p(x),
                          _p(_f64(ffluid)), _p(out))
    return out


def wrap_advance(w, dt, xyz, vel, status, regress, xyzn=None, veln=None):
    """Moves and wraps in place; returns the clone flags."""
    nc, nn, _ = xyz.shape
    nnuc = 0 if xyzn is None else xyzn.shape[1]
    flags = np.zeros(nc, np.int32)
    assert xyz.flags.c_contiguous and status.dtype == np.int32 and regress.dtype == np.int32
    lib().zfo_wrap_advance(C.byref(w), C.c_double(dt), nc, nn, nnuc, _p(xyz), _p(_f64(vel)),
                           _p(xyzn) if xyzn is not None else None,
                           _p(_f64(veln)) if veln is not None else None, _p(status),
                           _p(regress), _p(flags))
    return flags


def wrap_span(xyz, xyzn, status):
    """CAPS_MAX/MIN along z over membrane and nucleus nodes (ZebraflowWT.c:
    5709-5741), per cell."""
    z = xyz[..., 2]
    if xyzn is not None:
        z = np.concatenate([z, xyzn[..., 2]], axis=1)
    return z.min(axis=1), z.max(axis=1)


def wrap_unwrap(w, z, status, zmin, zmax, split=30.0):
    """The coordinates the membrane stages use (:5769-5797, :6738-6756): a
    cell with status > 3 spanning more than `split` has its nodes behind the
    window's middle moved by L past its exit."""
    L, mid = w.zm - w.z0, 0.5 * (w.z0 + w.zm)
    out = np.array(z, dtype=np.float64, copy=True)
    cut = (status > 3) & (zmax - zmin > split)
    for c in np.nonzero(cut)[0]:
        zc = out[c]
        if w.dir > 0:
            zc[zc < mid] += L
        else:
            zc[zc > mid] -= L
    return out, cut


def wrap_regress_reset(w, status, regress, zmin, zmax, centroid_z, split=30.0):
    """REGRESSFLAG back to 0 for a cell not across the wrap whose centroid
    lies in the window's middle half (:6762-6778)."""
    L = w.zm - w.z0
    ok = (status > -5000) & (zmax - zmin < split) & (centroid_z > w.z0 + 0.25 * L) & \
        (centroid_z < w.z0 + 0.75 * L)
    regress[ok] = 0


def clone_cells(w, id_base, pstatus, pregress, cloneflag, pxyz, pvel, mstatus, mxyz, mvel,
                pxyzn=None, pveln=None, mxyzn=None, mveln=None):
    """The serial cloning of :12345-12403 for one parent: each flagged cell
    (in slot order) goes to the first free main slot (status -5000) with
    status -(id_base + slot + 1), its nodes behind the window's middle moved
    by L past the exit; the parent cell becomes status 4 with REGRESSFLAG 1
    whether or not a slot was free.  Returns the (parent, main) slot pairs."""
    L, mid = w.zm - w.z0, 0.5 * (w.z0 + w.zm)
    pairs = []
    for s in np.nonzero(cloneflag)[0]:
        free = np.nonzero(mstatus == -5000)[0]
        if len(free):
            m = free[0]
            mstatus[m] = -(id_base + s + 1)
            for src, dst, vs, vd in ((pxyz, mxyz, pvel, mvel), (pxyzn, mxyzn, pveln, mveln)):
                if src is None:
                    continue
                x = src[s].copy()
                if w.dir > 0:
                    x[x[:, 2] < mid, 2] += L
                else:
                    x[x[:, 2] > mid, 2] -= L
                dst[m] = x
                vd[m] = vs[s]
            pairs.append((s, m))
        pstatus[s] = 4
        pregress[s] = 1
    return pairs


def clone_band(w, z, band):
    """Clone nodes that move with the parent cell (:11933-11941): within
    band (5 DX) past the exit plane or behind it."""
    if w.dir > 0:
        return ~(z > w.zm + band)
    return ~(z < w.z0 - band)
