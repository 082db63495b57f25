"""In-tree build of the native parts of qforte_amd.

* ``libqfe.so``  -- the HIP engine and its C ABI (``include/qfe.h``), compiled
  for gfx950 with ``hipcc``.
* ``_core*.so``  -- the C++ host core (qforte data model + Computer) with its
  pybind11 bindings, compiled with ``g++`` and linked against ``libqfe.so``
  through an ``$ORIGIN`` rpath.

Run ``python qforte_amd/_build.py`` (or ``__graft_entry__.build()``); it does
not import the package, which needs the libraries it builds.  Every call
recompiles.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build_tmp")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QFE_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # keep packed FP32 instructions out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Keep LLVM from merging the per-mask branches of the register-orbit
    # kernel into one block with a variable array index; that index forces the
    # amplitude array into scratch memory.
    "-mllvm", "-simplifycfg-sink-common=false",
    "-Wno-unused-result",
]
CXX = os.environ.get("CXX", "g++")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _pybind_includes():
    import pybind11

    return [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build(verbose: bool = True) -> None:
    os.makedirs(BUILD, exist_ok=True)
    hip_srcs = sorted(
        os.path.join(CSRC, "qfe", f) for f in os.listdir(os.path.join(CSRC, "qfe")) if f.endswith(".hip")
    )
    host_srcs = sorted(
        os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host")) if f.endswith(".cc")
    )
    jobs = []
    hip_objs = []
    for s in hip_srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        hip_objs.append(o)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
    host_objs = []
    for s in host_srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        host_objs.append(o)
        jobs.append([CXX, "-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden", *_pybind_includes(), "-c", s, "-o", o])
    workers = min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(_run, jobs))
    lib = os.path.join(PKG, "libqfe.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *hip_objs, "-o", tmp])
    os.replace(tmp, lib)
    core = os.path.join(PKG, "_core" + ext_suffix())
    tmp = core + ".tmp"
    _run([CXX, "-shared", "-fPIC", *host_objs, "-o", tmp, f"-L{PKG}", "-lqfe", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, core)
    shutil.rmtree(BUILD, ignore_errors=True)
    if verbose:
        print(f"built {lib} and {core}")


if __name__ == "__main__":
    build()
    sys.exit(0)
