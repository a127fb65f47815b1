"""Build ``libgpumatfun_amd.so`` in-tree with hipcc for gfx950.

Usage: ``python gpumatfun_amd/build.py`` (or ``gpumatfun_amd.build.build()``).
The packed-FP32 target feature is switched off for all device code: the
kernels are FP64, and hipcc must not emit packed FP32 instructions.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("gm_solver.hip", "gm_batch.hip", "gm_adjoint.hip",
                                                      "gm_ubatch.hip", "gm_eigs.hip", "gm_bfab.hip",
                                                      "gm_setup.cpp")]
HEADERS = [os.path.join(HERE, "csrc", "gm_kernels.h"), os.path.join(ROOT, "include", "gpumatfun.h")]
OUT = os.path.join(HERE, "libgpumatfun_amd.so")
ARCH = os.environ.get("GPUMATFUN_ARCH", "gfx950")


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result", "-Wno-unused-value",
           "-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
