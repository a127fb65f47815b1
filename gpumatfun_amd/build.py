"""Build ``libgpumatfun_amd.so`` in-tree with hipcc for gfx950.

Usage: ``python gpumatfun_amd/build.py`` (or ``gpumatfun_amd.build.build()``).
The packed-FP32 target feature is switched off for all device code: the
kernels are FP64, and hipcc must not emit packed FP32 instructions.  Host-only
sources (the sparse factorisation, which picks its dense kernel by the host
CPU's vector extensions at run time) are compiled by the host C++ compiler and
linked in.
"""

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("gm_solver.hip", "gm_batch.hip", "gm_adjoint.hip",
                                                      "gm_ubatch.hip", "gm_eigs.hip", "gm_bfab.hip",
                                                      "gm_setup.cpp")]
HOST_SOURCES = [os.path.join(HERE, "csrc", "gm_chol.cpp")]
HEADERS = [os.path.join(HERE, "csrc", "gm_kernels.h"), os.path.join(ROOT, "include", "gpumatfun.h")]
OUT = os.path.join(HERE, "libgpumatfun_amd.so")
ARCH = os.environ.get("GPUMATFUN_ARCH", "gfx950")


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def host_cxx():
    for cand in (shutil.which("g++"), shutil.which("c++")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("no host C++ compiler (g++) found")


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"{what} failed ({res.returncode})")


def build(verbose=False):
    tmp = OUT + ".tmp"
    with tempfile.TemporaryDirectory() as td:
        objs = []
        for src in HOST_SOURCES:
            obj = os.path.join(td, os.path.basename(src) + ".o")
            _run([host_cxx(), "-O3", "-std=c++17", "-fPIC", "-pthread", "-c", src, "-o", obj],
                 verbose, "host C++ compile")
            objs.append(obj)
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-result", "-Wno-unused-value", "-pthread",
               "-o", tmp] + SOURCES + ["-x", "none"] + objs
        _run(cmd, verbose, "hipcc")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
