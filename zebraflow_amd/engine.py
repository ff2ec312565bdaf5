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

    def time_kernels(self, iter0, nsteps):
        out = np.zeros(3, np.float64)
        self._check(self.lib.zf_lbm_time_kernels(self.h, iter0, nsteps, _ptr(out)),
                    "zf_lbm_time_kernels")
        return out

    def sync(self):
        self._check(self.lib.zf_sync(self.h), "zf_sync")

    def record(self, slot):
        self._check(self.lib.zf_event_record(self.h, slot), "zf_event_record")

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self._check(self.lib.zf_event_elapsed(self.h, a, b, C.byref(ms)), "zf_event_elapsed")
        return ms.value
