"""zebraflow_amd: MI355X engine for the Zebraflow time step.

The product is ``libzfgpu.so`` (hand-written HIP kernels for gfx950 behind the
C ABI in ``include/zf.h``) and the C host program ``zf_host``.  This Python
package only loads the library with ctypes for tests and benchmarks
(:mod:`zebraflow_amd.engine`); it has no compute path of its own.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libzfgpu.so")
HOST_LIB_PATH = os.path.join(PKG, "libzfhost.so")

_lib = None
_host = None


def load_library():
    """Load libzfgpu.so; raise if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m zebraflow_amd.build")
        _lib = ctypes.CDLL(LIB_PATH)
        from . import _abi
        _abi.declare(_lib)
    return _lib


def load_host_library():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"{HOST_LIB_PATH} is missing: build it with python -m zebraflow_amd.build")
        load_library()  # the host library resolves zf_* from it
        _host = ctypes.CDLL(HOST_LIB_PATH)
        from . import _abi
        _abi.declare_host(_host)
    return _host
