"""Build the native parts in-tree (no JIT cache, nothing installed).

  libzfgpu.so   hipcc, gfx950: HIP kernels + the C ABI of include/zf.h
  libzfhost.so  gcc: reference-format I/O, synthetic geometry, the step loop
  zf_host       gcc: the C host program (drives libzfgpu.so)

Run ``python -m zebraflow_amd.build`` or call :func:`build`.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")

HIP_SOURCES = ["zf_core.hip", "zf_lists.hip", "zf_lbm.hip", "zf_membrane.hip", "zf_ibm.hip",
               "zf_cells.hip", "zf_wss.hip", "zf_adjoint.hip", "zf_adjoint_ibm.hip",
               "zf_adjoint_cells.hip", "zf_state.hip", "zf_parent.hip", "zf_nucleus.hip",
               "zf_wrap.hip", "zf_pipe.hip"]
HOST_SOURCES = ["zf_io.c", "zf_geom.c", "zf_mesh.c", "zf_wallmesh.c", "zf_loop.c", "zf_restart.c"]  # those present are built

# gfx950 only; packed-FP32 instructions are disabled for all device code.
HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-local-typedef", "-munsafe-fp-atomics"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "warning: unknown" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build_gpu(verbose=False, force=True):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    headers = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    headers.append(os.path.join(REPO, "include", "zf.h"))
    objs, cmds = [], []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        if force or _stale(o, [s] + headers):
            cmds.append([hipcc, "-c", *HIPCC_FLAGS, s, "-o", o])
        objs.append(o)
    # one hipcc per translation unit, a few at a time
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    lib = os.path.join(PKG, "libzfgpu.so")
    tmp = lib + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs], verbose)
    os.replace(tmp, lib)
    return lib


def build_host(verbose=False):
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if not srcs:
        return None
    inc = ["-I", os.path.join(REPO, "include"), "-I", HOST]
    lib = os.path.join(PKG, "libzfhost.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-Wall", *inc, "-o", lib + ".tmp",
          *srcs, "-L", PKG, "-lzfgpu", "-Wl,-rpath,$ORIGIN", "-lm"], verbose)
    os.replace(lib + ".tmp", lib)
    main = os.path.join(HOST, "zf_host.c")
    if os.path.exists(main):
        exe = os.path.join(PKG, "zf_host")
        _run(["gcc", "-O2", "-Wall", *inc, "-o", exe, main, *srcs, "-fopenmp",
              "-L", PKG, "-lzfgpu", "-Wl,-rpath,$ORIGIN", "-lm"], verbose)
    return lib


def build_oracle(verbose=False):
    """The oracle (test infrastructure): liboracle.so always, the reference
    harness only where the reference source exists (this container)."""
    odir = os.path.join(REPO, "oracle")
    _run(["make", "-C", odir, "-B", "liboracle.so"], verbose)
    if os.path.exists("/root/reference/ZebraflowWT.c"):
        _run(["make", "-C", odir, "ref"], verbose)


def build(verbose=False):
    build_gpu(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
