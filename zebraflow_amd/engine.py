"""Thin numpy wrapper over the C ABI of libzfgpu.so (tests and benchmarks).

Every method calls exactly one zf_* entry point; errors raise ZfError with
the library's message.  Arrays follow include/zf.h: per-node arrays in the
reference's list order p = DOMID-1, populations SoA [19, dommax].
"""
import ctypes as C

import numpy as np

from . import load_library
from ._abi import ZfConfig, ZfLatticeInfo

STORE_F64, STORE_F32DEV, STORE_F16DEV = 0, 1, 2
WALL_EQUILIBRIUM, WALL_HALFWAY_BB = 0, 1
B_VELOCITY_INLET, B_PRESSURE_OUTLET = 2001, 3001
WSS_NOALLOC_MAX_POINTS = 196608  # ZF_WSS_NOALLOC_MAX_POINTS of include/zf.h
# zf_config.flags (include/zf.h)
FLAG_LINEAR_BLOCKS, FLAG_NT_POPULATIONS, FLAG_F64_ARITH, FLAG_F32_CELLS = 1, 2, 4, 8
FLAG_SINGLE_STEPS, FLAG_PAIRED_STEPS = 16, 32


def flag_cached_dirs(k):
    """ZF_FLAG_CACHED_DIRS(k): directions 0..k-1 cached, the rest non-temporal."""
    return (k + 1) << 8


class ZfError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def reference_config(im, jm, km, **kw):
    cfg = ZfConfig()
    load_library().zf_config_reference(C.byref(cfg), im, jm, km)
    return _apply(cfg, kw)


def lattice_config(im, jm, km, tau, **kw):
    cfg = ZfConfig()
    load_library().zf_config_lattice(C.byref(cfg), im, jm, km, tau)
    return _apply(cfg, kw)


def _apply(cfg, kw):
    for k, v in kw.items():
        if k in ("periodic", "body_force", "inlet_dir"):
            arr = getattr(cfg, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(cfg, k, v)
    return cfg


class Engine:
    def __init__(self, cfg):
        self.lib = load_library()
        self.cfg = cfg
        h = C.c_void_p()
        rc = self.lib.zf_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise ZfError(f"zf_create failed ({rc})")
        self.h = h
        self.dommax = self.bcmax = self.n_interior = 0

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.zf_last_error(self.h).decode()
            raise ZfError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if self.h:
            self.lib.zf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- lists --------------------------------------------------------------
    def upload_lattice(self, b):
        b = np.ascontiguousarray(b, dtype=np.int32)
        assert b.shape == (self.cfg.im, self.cfg.jm, self.cfg.km), b.shape
        nd, nb, ni = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.zf_upload_lattice(self.h, _ptr(b), C.byref(nd), C.byref(nb),
                                               C.byref(ni)), "zf_upload_lattice")
        self.dommax, self.bcmax, self.n_interior = nd.value, nb.value, ni.value
        return self.dommax, self.bcmax, self.n_interior

    def set_parent(self, p, k_over, offset):
        """Make this context parent lattice p (before upload_lattice)."""
        off = np.ascontiguousarray(offset, np.int32)
        self._check(self.lib.zf_set_parent(self.h, p, k_over, _ptr(off)), "zf_set_parent")

    def parent_lists(self):
        vb = np.zeros((self.bcmax, 3), np.int32)
        self._check(self.lib.zf_download_parent_lists(self.h, _ptr(vb)), "zf_download_parent_lists")
        return vb

    def parent_link(self, main):
        """Couple this parent lattice's boundary to the main lattice `main`."""
        self._check(self.lib.zf_parent_link(self.h, main.h), "zf_parent_link")

    def cell_wrap_setup(self, main, parent_index=1, **kw):
        """zf_cell_wrap_setup: this parent lattice's cells in their periodic
        window, cloning into `main`; the reference configuration of its
        parent `parent_index` updated by kw."""
        from ._abi import ZfCellWrapConfig
        wc = ZfCellWrapConfig()
        self.lib.zf_cell_wrap_config_reference(C.byref(wc), parent_index)
        for k, v in kw.items():
            setattr(wc, k, v)
        self._check(self.lib.zf_cell_wrap_setup(self.h, main.h, C.byref(wc)), "zf_cell_wrap_setup")

    def step_linked(self, it):
        """zf_step_linked of this main lattice and its linked parents."""
        self._check(self.lib.zf_step_linked(self.h, it), "zf_step_linked")

    def cell_flags(self):
        reg = np.zeros(self.ncell, np.int32)
        clo = np.zeros(self.ncell, np.int32)
        self._check(self.lib.zf_download_cell_flags(self.h, _ptr(reg), _ptr(clo)),
                    "zf_download_cell_flags")
        return reg, clo

    def upload_cell_flags(self, regress):
        r = np.ascontiguousarray(regress, np.int32)
        self._check(self.lib.zf_upload_cell_flags(self.h, _ptr(r)), "zf_upload_cell_flags")

    def lists(self):
        dl = np.zeros((self.dommax, 3), np.int32)
        bl = np.zeros((self.bcmax, 4), np.int32)
        bn = np.zeros((self.bcmax, 3), np.int32)
        self._check(self.lib.zf_download_lists(self.h, _ptr(dl), _ptr(bl), _ptr(bn)),
                    "zf_download_lists")
        return dl, bl, bn

    def domid(self):
        d = np.zeros((self.cfg.im, self.cfg.jm, self.cfg.km), np.int32)
        self._check(self.lib.zf_download_domid(self.h, _ptr(d)), "zf_download_domid")
        return d

    def neighbours(self):
        n = np.zeros((18, self.dommax), np.int32)
        self._check(self.lib.zf_download_neighbours(self.h, _ptr(n)), "zf_download_neighbours")
        return n

    def info(self):
        inf = ZfLatticeInfo()
        self._check(self.lib.zf_get_lattice_info(self.h, C.byref(inf)), "zf_get_lattice_info")
        return {f: getattr(inf, f) for f, _ in ZfLatticeInfo._fields_}

    # -- state ----------------------------------------------------------------
    def init_equilibrium(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        uu = np.asarray(u, np.float64)
        self._check(self.lib.zf_init_equilibrium(self.h, rho, _ptr(uu)), "zf_init_equilibrium")

    def upload_populations(self, f):
        f = np.ascontiguousarray(f, np.float64)
        assert f.shape == (19, self.dommax)
        self._check(self.lib.zf_upload_populations(self.h, _ptr(f)), "zf_upload_populations")

    def populations(self):
        f = np.zeros((19, self.dommax), np.float64)
        self._check(self.lib.zf_download_populations(self.h, _ptr(f)), "zf_download_populations")
        return f

    def populations_raw(self):
        """The populations as stored ([19, dommax] of float64, float32 or
        float16 by the storage form), without conversion."""
        dt = {STORE_F64: np.float64, STORE_F32DEV: np.float32}.get(self.cfg.storage, np.float16)
        f = np.zeros((19, self.dommax), dt)
        self._check(self.lib.zf_download_populations_raw(self.h, _ptr(f), f.nbytes),
                    "zf_download_populations_raw")
        return f

    def upload_populations_raw(self, f):
        f = np.ascontiguousarray(f)
        self._check(self.lib.zf_upload_populations_raw(self.h, _ptr(f), f.nbytes),
                    "zf_upload_populations_raw")

    def save_state(self, path, iter_, ncell=0, nn=0):
        """Restart file of libzfhost.so (zfh_save_state)."""
        from . import load_host_library
        self._check(load_host_library().zfh_save_state(self.h, str(path).encode(), iter_, ncell, nn),
                    "zfh_save_state")

    def load_state(self, path, ncell=0, nn=0):
        import ctypes as C
        from . import load_host_library
        it = C.c_int32()
        self._check(load_host_library().zfh_load_state(self.h, str(path).encode(), C.byref(it),
                                                       ncell, nn), "zfh_load_state")
        return it.value

    def upload_force(self, force):
        if force is None:
            self._check(self.lib.zf_upload_force(self.h, None), "zf_upload_force")
            return
        force = np.ascontiguousarray(force, np.float64)
        assert force.shape == (3, self.dommax)
        self._check(self.lib.zf_upload_force(self.h, _ptr(force)), "zf_upload_force")

    def force(self):
        f = np.zeros((3, self.dommax), np.float64)
        self._check(self.lib.zf_download_force(self.h, _ptr(f)), "zf_download_force")
        return f

    def macro(self):
        out = np.zeros((4, self.dommax), np.float64)
        self._check(self.lib.zf_download_macro(self.h, _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                               _ptr(out[3])), "zf_download_macro")
        return out

    def conservation_sums(self):
        out = np.zeros(4, np.float64)
        self._check(self.lib.zf_conservation_sums(self.h, _ptr(out)), "zf_conservation_sums")
        return out

    # -- stepping -------------------------------------------------------------
    def step(self, it, rbc_step=0):
        self._check(self.lib.zf_lbm_step(self.h, it, rbc_step), "zf_lbm_step")

    def run(self, iter0, nsteps):
        self._check(self.lib.zf_lbm_run(self.h, iter0, nsteps), "zf_lbm_run")

    def pair_info(self):
        """zf_pair_info: the two-step launch on this lattice."""
        out = np.zeros(6, np.int32)
        self._check(self.lib.zf_pair_info(self.h, _ptr(out)), "zf_pair_info")
        keys = ("used", "admitted", "xcds", "ghost_blocks", "ring_blocks", "workgroups")
        return dict(zip(keys, (int(v) for v in out)))

    def time_kernels(self, iter0, nsteps):
        out = np.zeros(3, np.float64)
        self._check(self.lib.zf_lbm_time_kernels(self.h, iter0, nsteps, _ptr(out)),
                    "zf_lbm_time_kernels")
        return out

    # -- membranes -----------------------------------------------------------
    def membrane_setup(self, mc, tris, rest_xyz):
        tris = np.ascontiguousarray(tris, np.int32)
        rest = np.ascontiguousarray(rest_xyz, np.float64)
        self._check(self.lib.zf_membrane_setup(self.h, C.byref(mc), len(rest), len(tris), _ptr(tris),
                                               _ptr(rest)), "zf_membrane_setup")
        self.nn, self.nt = len(rest), len(tris)
        ne = C.c_int32()
        self._check(self.lib.zf_download_mesh_topology(self.h, C.byref(ne), None, None, None),
                    "zf_download_mesh_topology")
        self.ne = ne.value

    def mesh_topology(self):
        edge = np.zeros((self.ne, 6), np.int32)
        nbn = np.zeros((self.nn, 10), np.int32)
        nbt = np.zeros((self.nn, 10), np.int32)
        self._check(self.lib.zf_download_mesh_topology(self.h, None, _ptr(edge), _ptr(nbn),
                                                       _ptr(nbt)), "zf_download_mesh_topology")
        return edge, nbn, nbt

    def mesh_equilibrium(self):
        out = [np.zeros(self.ne) for _ in range(6)]
        self._check(self.lib.zf_download_mesh_equilibrium(self.h, *[_ptr(a) for a in out]),
                    "zf_download_mesh_equilibrium")
        return dict(zip(("lmax", "plength", "kp", "theta0", "sin0", "cos0"), out))

    def upload_cells(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64)
        self.ncell = len(xyz)
        self._check(self.lib.zf_upload_cells(self.h, self.ncell, _ptr(xyz)), "zf_upload_cells")

    def cells(self):
        x = np.zeros((self.ncell, self.nn, 3))
        v = np.zeros((self.ncell, self.nn, 3))
        st = np.zeros(self.ncell, np.int32)
        self._check(self.lib.zf_download_cells(self.h, _ptr(x), _ptr(v), _ptr(st)), "zf_download_cells")
        return x, v, st

    def membrane_forces(self):
        self._check(self.lib.zf_membrane_forces(self.h), "zf_membrane_forces")

    def membrane(self):
        nc, nn = self.ncell, self.nn
        out = dict(force=np.zeros((nc, nn, 3)), node_normal=np.zeros((nc, nn, 3)),
                   cell_area=np.zeros(nc), cell_vol=np.zeros(nc), cell_centroid=np.zeros((nc, 3)),
                   guard_fail=np.zeros(nc, np.int32))
        self._check(self.lib.zf_download_membrane(
            self.h, *[_ptr(out[k]) for k in ("force", "node_normal", "cell_area", "cell_vol",
                                             "cell_centroid", "guard_fail")]), "zf_download_membrane")
        return out

    # -- immersed boundary ------------------------------------------------------
    def ibm_setup(self, ic):
        self._check(self.lib.zf_ibm_setup(self.h, C.byref(ic)), "zf_ibm_setup")

    def ibm_interpolate(self):
        self._check(self.lib.zf_ibm_interpolate(self.h), "zf_ibm_interpolate")

    def ibm_spread(self):
        self._check(self.lib.zf_ibm_spread(self.h), "zf_ibm_spread")

    def advance_cells(self):
        n = C.c_int32()
        self._check(self.lib.zf_advance_cells(self.h, C.byref(n)), "zf_advance_cells")
        return n.value

    def upload_membrane_force(self, f):
        f = np.ascontiguousarray(f, np.float64)
        self._check(self.lib.zf_upload_membrane_force(self.h, _ptr(f)), "zf_upload_membrane_force")

    # -- nucleus network ------------------------------------------------------------
    def nucleus_setup(self, edge, rest_xyz, **kw):
        """zf_nucleus_setup with the reference configuration updated by kw."""
        from ._abi import ZfNucleusConfig
        nc = ZfNucleusConfig()
        self.lib.zf_nucleus_config_reference(C.byref(nc))
        for k, v in kw.items():
            setattr(nc, k, v)
        e = np.ascontiguousarray(edge, np.int32)
        x = np.ascontiguousarray(rest_xyz, np.float64)
        self._check(self.lib.zf_nucleus_setup(self.h, C.byref(nc), len(x), len(e), _ptr(e), _ptr(x)),
                    "zf_nucleus_setup")
        self.nnuc, self.ne_n = len(x), len(e)

    def nucleus_equilibrium(self):
        out = [np.zeros(self.ne_n) for _ in range(3)]
        self._check(self.lib.zf_download_nucleus_equilibrium(self.h, *[_ptr(a) for a in out]),
                    "zf_download_nucleus_equilibrium")
        return dict(zip(("lmax", "plength", "kp"), out))

    def upload_nuclei(self, xyz):
        x = np.ascontiguousarray(xyz, np.float64)
        assert x.shape == (self.ncell, self.nnuc, 3)
        self._check(self.lib.zf_upload_nuclei(self.h, _ptr(x)), "zf_upload_nuclei")

    def nucleus_lists(self):
        """(list, count, valid) of zf_download_nucleus_lists."""
        n = self.ncell * self.nnuc
        nuc, cap = C.c_int32(), C.c_int32()
        self._check(self.lib.zf_get_nucleus_info(self.h, C.byref(nuc), C.byref(cap)),
                    "zf_get_nucleus_info")
        lst = np.zeros((n, cap.value), np.int32)
        cnt = np.zeros(n, np.int32)
        valid = C.c_int32()
        self._check(self.lib.zf_download_nucleus_lists(self.h, _ptr(lst), _ptr(cnt), C.byref(valid)),
                    "zf_download_nucleus_lists")
        return lst, cnt, valid.value

    def nuclei(self):
        out = [np.zeros((self.ncell, self.nnuc, 3)) for _ in range(3)]
        self._check(self.lib.zf_download_nuclei(self.h, *[_ptr(a) for a in out]), "zf_download_nuclei")
        return tuple(out)

    def nucleus_forces(self):
        self._check(self.lib.zf_nucleus_forces(self.h), "zf_nucleus_forces")

    def nucleus_contact(self, it):
        lost = C.c_int32()
        self._check(self.lib.zf_nucleus_contact(self.h, it, C.byref(lost)), "zf_nucleus_contact")
        return lost.value

    def upload_repulsion(self, v):
        v = np.ascontiguousarray(v, np.float64)
        self._check(self.lib.zf_upload_repulsion(self.h, _ptr(v)), "zf_upload_repulsion")

    def repulsion(self):
        v = np.zeros((self.ncell, self.nn, 3))
        self._check(self.lib.zf_download_repulsion(self.h, _ptr(v)), "zf_download_repulsion")
        return v

    def full_step(self, it):
        self._check(self.lib.zf_step(self.h, it), "zf_step")

    def checkpoint_save(self, slot=0):
        self._check(self.lib.zf_checkpoint_save(self.h, slot), "zf_checkpoint_save")

    def checkpoint_restore(self, slot=0):
        self._check(self.lib.zf_checkpoint_restore(self.h, slot), "zf_checkpoint_restore")

    def upload_cell_status(self, status):
        st = np.ascontiguousarray(status, np.int32)
        self._check(self.lib.zf_upload_cell_status(self.h, _ptr(st)), "zf_upload_cell_status")

    def full_run(self, iter0, nsteps):
        self._check(self.lib.zf_run(self.h, iter0, nsteps), "zf_run")

    def ibm_vjp(self, lam_force=None, lam_vel=None, node_force=None, want_f=True):
        """Transposes of spreading and interpolation at the current state:
        (lam_node_force [ncell, nn, 3], lam_pos [ncell, nn, 3], lam_f [19, dommax])."""
        nc, nn = self.ncell, self.nn
        arr = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
        lf, lv, g = arr(lam_force), arr(lam_vel), arr(node_force)
        lg = np.zeros((nc, nn, 3))
        lx = np.zeros((nc, nn, 3))
        lam = np.zeros((19, self.dommax)) if want_f else None
        self._check(self.lib.zf_ibm_vjp(self.h, *[None if a is None else _ptr(a)
                                                  for a in (lf, lv, g, lg, lx, lam)]), "zf_ibm_vjp")
        return lg, lx, lam

    def membrane_vjp(self, lam_force):
        """(dJ/dX [ncell, nn, 3], (dJ/d ln es, dJ/d ln kbend)) for J = <lam_force, F(X)>."""
        lf = np.ascontiguousarray(lam_force, np.float64)
        lx = np.zeros((self.ncell, self.nn, 3))
        dp = np.zeros(2)
        self._check(self.lib.zf_membrane_vjp(self.h, _ptr(lf), _ptr(lx), _ptr(dp)), "zf_membrane_vjp")
        return lx, dp

    # -- cell list and repulsion ------------------------------------------------
    def repulsion_setup(self, rc):
        self._check(self.lib.zf_repulsion_setup(self.h, C.byref(rc)), "zf_repulsion_setup")

    def wall_setup(self, xyz, tri, normal):
        xyz = np.ascontiguousarray(xyz, np.float64)
        tri = np.ascontiguousarray(tri, np.int32)
        normal = np.ascontiguousarray(normal, np.float64)
        self._check(self.lib.zf_wall_setup(self.h, len(xyz), _ptr(xyz), len(tri), _ptr(tri),
                                           _ptr(normal)), "zf_wall_setup")

    def windex(self):
        w = np.zeros((self.cfg.im, self.cfg.jm, self.cfg.km), np.int32)
        self._check(self.lib.zf_download_windex(self.h, _ptr(w)), "zf_download_windex")
        return w

    def cell_list(self):
        self._check(self.lib.zf_cell_list(self.h), "zf_cell_list")
        out = np.zeros((self.ncell, self.nt), np.int32)
        self._check(self.lib.zf_download_cell_list(self.h, _ptr(out)), "zf_download_cell_list")
        return out

    def repulse(self):
        self._check(self.lib.zf_repulsion(self.h), "zf_repulsion")

    def uvoid_setup(self, uc=None, **kw):
        """zf_uvoid_setup with the reference configuration updated by kw
        (speed, d0, d1, k_lo, k_hi, select, select_point); uc=False turns it off."""
        from ._abi import ZfUvoidConfig
        if uc is False:
            self._check(self.lib.zf_uvoid_setup(self.h, None), "zf_uvoid_setup")
            return
        if uc is None:
            uc = ZfUvoidConfig()
            self.lib.zf_uvoid_config_reference(C.byref(uc))
        for k, v in kw.items():
            if k == "select_point":
                for a in range(3):
                    uc.select_point[a] = v[a]
            else:
                setattr(uc, k, v)
        self._check(self.lib.zf_uvoid_setup(self.h, C.byref(uc)), "zf_uvoid_setup")

    def uvoid(self):
        out = np.zeros((3, self.cfg.im, self.cfg.jm, self.cfg.km))
        self._check(self.lib.zf_download_uvoid(self.h, _ptr(out)), "zf_download_uvoid")
        return out

    def wall_neighbour(self):
        out = np.zeros((self.ncell, self.nn), np.int32)
        self._check(self.lib.zf_download_wall_neighbour(self.h, _ptr(out)),
                    "zf_download_wall_neighbour")
        return out

    # -- wall shear stress ------------------------------------------------------
    def wss(self, xyz, normal, mu_over_dx, pgauge=0.0, noalloc=False):
        """(wss [n,3], pressure [n], valid [n]) at wall points; noalloc calls
        the allocation-free zf_wss_noalloc instead of zf_wss."""
        xyz = np.ascontiguousarray(xyz, np.float64)
        normal = np.ascontiguousarray(normal, np.float64)
        n = len(xyz)
        out, pr, va = np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.int32)
        fn = self.lib.zf_wss_noalloc if noalloc else self.lib.zf_wss
        name = "zf_wss_noalloc" if noalloc else "zf_wss"
        self._check(fn(self.h, n, _ptr(xyz), _ptr(normal), mu_over_dx, pgauge,
                       _ptr(out), _ptr(pr), _ptr(va)), name)
        return out, pr, va

    # -- adjoint ------------------------------------------------------------------
    def adjoint_setup(self, xyz, normal, mu_over_dx, t_begin, t_end, max_steps, checkpoint=100,
                      cells=False):
        from ._abi import ZfAdjointConfig
        ac = ZfAdjointConfig(checkpoint, max_steps, t_begin, t_end, mu_over_dx, int(cells), 0)
        xyz = np.ascontiguousarray(xyz, np.float64)
        normal = np.ascontiguousarray(normal, np.float64)
        self._check(self.lib.zf_adjoint_setup(self.h, C.byref(ac), len(xyz), _ptr(xyz), _ptr(normal)),
                    "zf_adjoint_setup")

    def adjoint_forward(self, iter0, nsteps):
        j = C.c_double()
        self._check(self.lib.zf_adjoint_forward(self.h, iter0, nsteps, C.byref(j)), "zf_adjoint_forward")
        return j.value

    def adjoint_backward(self):
        """dJ/d(inlet_amp, inlet_u0, body_force x, y, z)."""
        g = np.zeros(5)
        self._check(self.lib.zf_adjoint_backward(self.h, _ptr(g)), "zf_adjoint_backward")
        return g

    def adjoint_download(self):
        """(dJ/df0 [19, dommax], dJ/dforce [3, dommax])."""
        lam = np.zeros((19, self.dommax))
        gf = np.zeros((3, self.dommax))
        self._check(self.lib.zf_adjoint_download(self.h, _ptr(lam), _ptr(gf)), "zf_adjoint_download")
        return lam, gf

    def adjoint_cell_gradients(self):
        """((dJ/d ln es, dJ/d ln kbend), dJ/dX0 [ncell, nn, 3])."""
        dp = np.zeros(2)
        gx = np.zeros((self.ncell, self.nn, 3))
        self._check(self.lib.zf_adjoint_cell_gradients(self.h, _ptr(dp), _ptr(gx)),
                    "zf_adjoint_cell_gradients")
        return dp, gx

    def adjoint_timing(self):
        ms = np.zeros(2)
        self._check(self.lib.zf_adjoint_timing(self.h, _ptr(ms)), "zf_adjoint_timing")
        return ms

    def set_forcing(self, inlet_u0, inlet_amp, body_force=(0.0, 0.0, 0.0)):
        g = np.ascontiguousarray(body_force, np.float64)
        self._check(self.lib.zf_set_forcing(self.h, inlet_u0, inlet_amp, _ptr(g)), "zf_set_forcing")

    def copy_bandwidth(self, nbytes, reps=20):
        g = C.c_double()
        self._check(self.lib.zf_copy_bandwidth(self.h, int(nbytes), reps, C.byref(g)),
                    "zf_copy_bandwidth")
        return g.value

    def sync(self):
        self._check(self.lib.zf_sync(self.h), "zf_sync")

    def record(self, slot):
        self._check(self.lib.zf_event_record(self.h, slot), "zf_event_record")

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self._check(self.lib.zf_event_elapsed(self.h, a, b, C.byref(ms)), "zf_event_elapsed")
        return ms.value


def membrane_config(d, max_cells):
    """ZfMembraneConfig from a dict of the oracle's parameter names plus es."""
    from ._abi import ZfMembraneConfig
    mc = ZfMembraneConfig()
    load_library().zf_membrane_config_reference(C.byref(mc))
    for k, v in d.items():
        setattr(mc, k, v)
    mc.max_cells = max_cells
    return mc


def lattice_membrane_config(xyz, tris, max_cells, es=5e-3, kb=2e-3):
    """Membrane constants in lattice units for a rest mesh (xyz [nn,3], tris
    [nt,3] 1-based): zfh_membrane_config_lattice of the host library (the
    reference's moduli as ratios to the shear modulus es, ZebraflowWT.c:124-134;
    bending modulus kb; rest area and volume as the global targets)."""
    from . import load_host_library
    from ._abi import ZfMembraneConfig
    xyz = np.ascontiguousarray(xyz, np.float64)
    tris = np.ascontiguousarray(tris, np.int32)
    mc = ZfMembraneConfig()
    load_host_library().zfh_membrane_config_lattice(_ptr(xyz), len(xyz), _ptr(tris), len(tris),
                                                    es, kb, int(max_cells), C.byref(mc))
    return mc


def lattice_ibm_config(rbc_interval=1, gamma_p=0.5, vmax=1.0):
    """IBM constants in lattice units: x0 = 0, dx = 1, unit velocity and
    force scales, one lattice step per position update."""
    from ._abi import ZfIbmConfig
    ic = ZfIbmConfig()
    load_library().zf_ibm_config_reference(C.byref(ic))
    for k, v in dict(x0=0.0, y0=0.0, z0=0.0, dx=1.0, vel_scale=1.0, force_scale=1.0,
                     gamma_p=gamma_p, dt_advance=1.0, vmax=vmax).items():
        setattr(ic, k, v)
    ic.rbc_interval = rbc_interval
    return ic


def lattice_repulsion_config(cutoff=1.0, bin_cutoff=2.0):
    """Repulsion constants in lattice units: cell-cell range `cutoff` with
    the centroid search at twice that, triangles binned to their nearest
    node within `bin_cutoff`; the reference's speed scales and wall
    distances (ZebraflowWT.c:9000-9142, :9941-10052) read as lattice units."""
    from ._abi import ZfRepulsionConfig
    rc = ZfRepulsionConfig()
    load_library().zf_repulsion_config_reference(C.byref(rc))
    rc.cutoff = cutoff
    rc.search = 2 * cutoff
    rc.bin_cutoff = bin_cutoff
    return rc
