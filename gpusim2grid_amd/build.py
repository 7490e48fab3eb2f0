"""Build recipe of the native parts (in-tree, gfx950).

* ``libgpusim2grid_amd.so``: the HIP kernels, the host symbolic layer and the
  C ABI of ``include/gpusim2grid_amd.h`` (hipcc, --offload-arch=gfx950);
* ``_gpusim2grid*.so``: the pybind11 module with the reference's Python names,
  a C++ layer over the C ABI (g++).

Run ``python -m gpusim2grid_amd.build``.  Objects go to ``_build/`` and the
two shared objects next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libgpusim2grid_amd.so")
ARCH = os.environ.get("G2G_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed-FP32 instructions stay off: hipcc -O3 otherwise emits them from
# plain float code
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_lib(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, lang in (("kernels.hip", "hip"), ("capi.cpp", "hip"), ("plan.cpp", "c++"), ("mlp.hip", "hip"),
                      ("mlp_capi.cpp", "hip")):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        flags = list(HIP_FLAGS) if lang == "hip" else ["-O3", "-std=c++17", "-fPIC", "-Wall"]
        _run([HIPCC, "-x", lang, *flags, "-c", path, "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_module(verbose=False):
    import pybind11
    src = os.path.join(CSRC, "bindings.cpp")
    if not os.path.exists(src):
        return None
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_gpusim2grid" + ext)
    tmp = out + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    # host code only: the HIP runtime API for the sessions' device buffers
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-D__HIP_PLATFORM_AMD__",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{rocm}/include", src, "-o", tmp,
           f"-L{HERE}", "-lgpusim2grid_amd", f"-L{rocm}/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN",
           f"-Wl,-rpath,{rocm}/lib"]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build(verbose=False):
    build_lib(verbose)
    build_module(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
