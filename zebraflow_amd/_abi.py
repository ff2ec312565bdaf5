"""ctypes declarations mirroring include/zf.h (keep in the same order)."""
import ctypes as C

i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double
P = C.c_void_p


class ZfConfig(C.Structure):
    _fields_ = [
        ("im", i32), ("jm", i32), ("km", i32),
        ("periodic", i32 * 3),
        ("storage", i32), ("wall_mode", i32), ("device", i32), ("flags", i32),
        ("visc_scale", f64), ("force_dt", f64), ("body_force", f64 * 3),
        ("inlet_u0", f64), ("inlet_amp", f64), ("inlet_period", f64), ("inlet_dir", f64 * 3),
        ("outlet_rho", f64), ("t_scale", f64), ("pressure_scale", f64),
    ]


class ZfLatticeInfo(C.Structure):
    _fields_ = [
        ("dommax", i32), ("bcmax", i32), ("n_interior", i32), ("n_chords", i32),
        ("n_bc_active", i32), ("n_bb_links", i32), ("n_macro_nodes", i32), ("stride", i32),
        ("device_bytes", i64),
    ]


class ZfMembraneConfig(C.Structure):
    _fields_ = [(n, f64) for n in (
        "len_scale", "kT", "x0", "es", "kbend", "kd", "ka", "kv", "area0", "vol0", "x_max",
        "x_floor", "cos_cut", "area_lo", "area_hi", "cos_blend", "force_limit", "vol_limit", "sliver_len",
        "sliver_len_min", "sliver_f1", "sliver_f2", "relax", "relax2")] + [
        ("passes", i32), ("max_cells", i32)]


class ZfIbmConfig(C.Structure):
    _fields_ = [(n, f64) for n in (
        "x0", "y0", "z0", "dx", "vel_scale", "force_scale", "gamma_p", "dt_advance", "vmax")] + [
        ("rbc_interval", i32), ("flags", i32)]


class ZfRepulsionConfig(C.Structure):
    _fields_ = [(n, f64) for n in ("bin_cutoff", "k_cell", "cutoff", "search", "k_wall", "wall_d0",
                                   "wall_dmax", "wall_search")]


class ZfUvoidConfig(C.Structure):
    _fields_ = [("speed", f64), ("d0", f64), ("d1", f64), ("k_lo", i32), ("k_hi", i32),
                ("select", i32), ("reserved", i32), ("select_point", f64 * 3)]


class ZfNucleusConfig(C.Structure):
    _fields_ = [(n, f64) for n in ("es_factor", "list_radius", "near_radius", "contact",
                                   "k_contact")] + [("list_interval", i32), ("list_cap", i32)]


class ZfCellWrapConfig(C.Structure):
    _fields_ = [("direction", i32), ("id_base", i32)] + [
        (n, f64) for n in ("split_span", "clone_band", "roi_lo", "roi_hi")]


class ZfAdjointConfig(C.Structure):
    _fields_ = [("checkpoint_interval", i32), ("max_steps", i32), ("t_begin", i32), ("t_end", i32),
                ("mu_over_dx", f64), ("cells", i32), ("reserved", i32)]


# name -> (restype, argtypes)
FUNCS = {
    "zf_version": (C.c_char_p, []),
    "zf_config_reference": (None, [C.POINTER(ZfConfig), i32, i32, i32]),
    "zf_config_lattice": (None, [C.POINTER(ZfConfig), i32, i32, i32, f64]),
    "zf_create": (C.c_int, [C.POINTER(ZfConfig), C.POINTER(P)]),
    "zf_destroy": (C.c_int, [P]),
    "zf_last_error": (C.c_char_p, [P]),
    "zf_upload_lattice": (C.c_int, [P, P, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "zf_download_lists": (C.c_int, [P, P, P, P]),
    "zf_download_domid": (C.c_int, [P, P]),
    "zf_download_neighbours": (C.c_int, [P, P]),
    "zf_get_lattice_info": (C.c_int, [P, C.POINTER(ZfLatticeInfo)]),
    "zf_init_equilibrium": (C.c_int, [P, f64, P]),
    "zf_upload_populations": (C.c_int, [P, P]),
    "zf_download_populations": (C.c_int, [P, P]),
    "zf_download_populations_raw": (C.c_int, [P, P, i64]),
    "zf_upload_populations_raw": (C.c_int, [P, P, i64]),
    "zf_upload_force": (C.c_int, [P, P]),
    "zf_download_force": (C.c_int, [P, P]),
    "zf_lbm_step": (C.c_int, [P, i32, i32]),
    "zf_lbm_run": (C.c_int, [P, i32, i32]),
    "zf_pair_info": (C.c_int, [P, P]),
    "zf_download_macro": (C.c_int, [P, P, P, P, P]),
    "zf_conservation_sums": (C.c_int, [P, P]),
    "zf_lbm_time_kernels": (C.c_int, [P, i32, i32, P]),
    "zf_set_parent": (C.c_int, [P, i32, i32, P]),
    "zf_download_parent_lists": (C.c_int, [P, P]),
    "zf_parent_link": (C.c_int, [P, P]),
    "zf_cell_wrap_config_reference": (None, [P, i32]),
    "zf_cell_wrap_setup": (C.c_int, [P, P, P]),
    "zf_step_linked": (C.c_int, [P, i32]),
    "zf_download_cell_flags": (C.c_int, [P, P, P]),
    "zf_upload_cell_flags": (C.c_int, [P, P]),
    "zf_membrane_config_reference": (None, [C.POINTER(ZfMembraneConfig)]),
    "zf_membrane_setup": (C.c_int, [P, C.POINTER(ZfMembraneConfig), i32, i32, P, P]),
    "zf_download_mesh_topology": (C.c_int, [P, C.POINTER(i32), P, P, P]),
    "zf_download_mesh_equilibrium": (C.c_int, [P, P, P, P, P, P, P]),
    "zf_upload_cells": (C.c_int, [P, i32, P]),
    "zf_download_cells": (C.c_int, [P, P, P, P]),
    "zf_membrane_forces": (C.c_int, [P]),
    "zf_download_membrane": (C.c_int, [P, P, P, P, P, P, P]),
    "zf_ibm_config_reference": (None, [C.POINTER(ZfIbmConfig)]),
    "zf_ibm_setup": (C.c_int, [P, C.POINTER(ZfIbmConfig)]),
    "zf_ibm_interpolate": (C.c_int, [P]),
    "zf_ibm_spread": (C.c_int, [P]),
    "zf_advance_cells": (C.c_int, [P, C.POINTER(i32)]),
    "zf_upload_membrane_force": (C.c_int, [P, P]),
    "zf_upload_repulsion": (C.c_int, [P, P]),
    "zf_download_repulsion": (C.c_int, [P, P]),
    "zf_step": (C.c_int, [P, i32]),
    "zf_repulsion_config_reference": (None, [C.POINTER(ZfRepulsionConfig)]),
    "zf_repulsion_setup": (C.c_int, [P, C.POINTER(ZfRepulsionConfig)]),
    "zf_wall_setup": (C.c_int, [P, i32, P, i32, P, P]),
    "zf_download_windex": (C.c_int, [P, P]),
    "zf_cell_list": (C.c_int, [P]),
    "zf_download_cell_list": (C.c_int, [P, P]),
    "zf_repulsion": (C.c_int, [P]),
    "zf_download_wall_neighbour": (C.c_int, [P, P]),
    "zf_wss": (C.c_int, [P, i64, P, P, f64, f64, P, P, P]),
    "zf_nucleus_config_reference": (None, [C.POINTER(ZfNucleusConfig)]),
    "zf_nucleus_setup": (C.c_int, [P, C.POINTER(ZfNucleusConfig), i32, i32, P, P]),
    "zf_download_nucleus_equilibrium": (C.c_int, [P, P, P, P]),
    "zf_upload_nuclei": (C.c_int, [P, P]),
    "zf_download_nuclei": (C.c_int, [P, P, P, P]),
    "zf_nucleus_forces": (C.c_int, [P]),
    "zf_nucleus_contact": (C.c_int, [P, i32, C.POINTER(i32)]),
    "zf_get_nucleus_info": (C.c_int, [P, C.POINTER(i32), C.POINTER(i32)]),
    "zf_download_nucleus_lists": (C.c_int, [P, P, P, C.POINTER(i32)]),
    "zf_upload_nucleus_lists": (C.c_int, [P, P, P, i32]),
    "zf_uvoid_config_reference": (None, [C.POINTER(ZfUvoidConfig)]),
    "zf_uvoid_setup": (C.c_int, [P, C.POINTER(ZfUvoidConfig)]),
    "zf_download_uvoid": (C.c_int, [P, P]),
    "zf_wss_noalloc": (C.c_int, [P, i64, P, P, f64, f64, P, P, P]),
    "zf_adjoint_setup": (C.c_int, [P, C.POINTER(ZfAdjointConfig), i64, P, P]),
    "zf_adjoint_forward": (C.c_int, [P, i32, i32, C.POINTER(f64)]),
    "zf_adjoint_backward": (C.c_int, [P, P]),
    "zf_adjoint_download": (C.c_int, [P, P, P]),
    "zf_adjoint_timing": (C.c_int, [P, P]),
    "zf_set_forcing": (C.c_int, [P, f64, f64, P]),
    "zf_ibm_vjp": (C.c_int, [P, P, P, P, P, P, P]),
    "zf_membrane_vjp": (C.c_int, [P, P, P, P]),
    "zf_adjoint_cell_gradients": (C.c_int, [P, P, P]),
    "zf_run": (C.c_int, [P, i32, i32]),
    "zf_checkpoint_save": (C.c_int, [P, i32]),
    "zf_checkpoint_restore": (C.c_int, [P, i32]),
    "zf_upload_cell_status": (C.c_int, [P, P]),
    "zf_get_config": (C.c_int, [P, C.POINTER(ZfConfig)]),
    "zf_copy_bandwidth": (C.c_int, [P, i64, i32, C.POINTER(f64)]),
    "zf_sync": (C.c_int, [P]),
    "zf_event_record": (C.c_int, [P, i32]),
    "zf_event_elapsed": (C.c_int, [P, i32, i32, C.POINTER(C.c_float)]),
}

HOST_FUNCS = {
    "zfh_geom_tube": (i64, [P, i32, i32, i32, i32, f64, i32, i32, i32]),
    "zfh_geom_vessel_pair": (i64, [P, i32, i32, i32, f64, f64, f64, i32, C.c_uint64]),
    "zfh_geom_tree": (i64, [P, i32, i32, i32, f64, f64, i32, f64, f64, C.c_uint64]),
    "zfh_mesh_icosphere": (C.c_int, [i32, f64, i32, P, P, i32, i32, C.POINTER(i32)]),
    "zfh_place_cells": (C.c_int, [P, i32, i32, i32, P, i32, f64, f64, i32, C.c_uint64, i32, P]),
    "zfh_place_cells_dense": (C.c_int, [P, i32, i32, i32, P, i32, f64, i32, C.c_uint64, i32, P]),
    "zfh_wall_mesh": (C.c_int, [P, i32, i32, i32, f64, f64, f64, f64, P, i64, P, P, i64,
                                C.POINTER(i64), C.POINTER(i64)]),
    "zfh_write_bindex_csv":(C.c_int, [P, P, i32, i32, i32, f64, f64, f64, f64]),
    "zfh_read_bindex_csv": (i64, [P, P, i32, i32, i32, f64, f64, f64, f64]),
    "zfh_write_fluid_csv": (C.c_int, [P, P, i32, i32, i32, f64, f64, f64, f64, P, P, P, P, P, P, P,
                                      i64]),
    "zfh_read_points": (i64, [C.c_char_p, P, i64]),
    "zfh_read_rbc_triangles": (i64, [C.c_char_p, P, i64]),
    "zfh_read_wall_cells": (i64, [C.c_char_p, P, i64]),
    "zfh_read_wall_normals": (i64, [C.c_char_p, P, i64]),
    "zfh_write_points": (C.c_int, [C.c_char_p, P, i64]),
    "zfh_write_rbc_triangles": (C.c_int, [C.c_char_p, P, i64]),
    "zfh_write_wall_mesh": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, P, i64, P, P, i64]),
    "zfh_read_cells_csv": (C.c_int, [C.c_char_p, i32, P, P, i32]),
    "zfh_read_lines": (i64, [C.c_char_p, P, i64]),
    "zfh_read_nuclei_csv": (C.c_int, [C.c_char_p, i32, P, i32]),
    "zfh_write_nuclei_csv": (C.c_int, [C.c_char_p, i32, i32, P, P, P, P]),
    "zfh_write_cells_csv": (C.c_int, [C.c_char_p, i32, i32, P, P, P, P, P, P]),
    "zfh_append_volume_area": (C.c_int, [C.c_char_p, i32, i32, P, P, P, P]),
    "zfh_membrane_config_lattice": (None, [P, i32, P, i32, f64, f64, i32, P]),
    "zfh_save_state": (C.c_int, [P, C.c_char_p, i32, i32, i32]),
    "zfh_read_fluid_start": (i64, [i32, P, P, i32, i32, i32, f64, f64, f64, f64, i64, P, P, P, P, P]),
    "zfh_init_from_macro": (C.c_int, [P, P, P, P, P, P]),
    "zfh_load_state": (C.c_int, [P, C.c_char_p, C.POINTER(i32), i32, i32]),
}


def declare(lib):
    for name, (res, args) in FUNCS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def declare_host(lib):
    for name, (res, args) in HOST_FUNCS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
