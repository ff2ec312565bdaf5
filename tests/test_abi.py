"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from zebraflow_amd import HOST_LIB_PATH, LIB_PATH, load_host_library, load_library

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_gpu_library_exports_every_declared_entry_point():
    lib = load_library()
    names = declared(os.path.join(REPO, "include", "zf.h"), "zf_")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert b"gfx950" in lib.zf_version()


def test_host_library_exports_every_declared_entry_point():
    lib = load_host_library()
    names = declared(os.path.join(REPO, "zebraflow_amd", "host", "zf_host.h"), "zfh_")
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_gpu_library_has_gfx950_code_object():
    blob = open(LIB_PATH, "rb").read()
    assert b"gfx950" in blob
    assert os.path.exists(HOST_LIB_PATH)


def test_config_defaults_match_reference_constants():
    from zebraflow_amd.engine import reference_config, lattice_config
    c = reference_config(10, 11, 12)
    assert (c.im, c.jm, c.km) == (10, 11, 12)
    assert abs(c.visc_scale - 0.72) < 1e-12          # MU_SCALE*MU, ZebraflowWT.c:65-66
    assert c.force_dt == 0.5                         # DT, ZebraflowWT.c:41
    assert abs(c.pressure_scale - 10000.0 / 3) < 1e-9  # LBM_PRESSURE_SCALE, :69
    assert c.wall_mode == 0
    l = lattice_config(4, 4, 4, 0.8)
    assert abs(l.visc_scale - 0.3) < 1e-15 and l.force_dt == 1.0 and l.wall_mode == 1


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        return
    from zebraflow_amd.engine import Engine, ZfError, reference_config
    try:
        Engine(reference_config(4, 4, 4))
    except ZfError:
        return
    raise AssertionError("zf_create succeeded without a GPU")


def test_noalloc_limit_matches_header():
    from zebraflow_amd.engine import WSS_NOALLOC_MAX_POINTS
    text = open(os.path.join(REPO, "include", "zf.h")).read()
    m = re.search(r"#define ZF_WSS_NOALLOC_MAX_POINTS (\d+)", text)
    assert m and int(m.group(1)) == WSS_NOALLOC_MAX_POINTS
