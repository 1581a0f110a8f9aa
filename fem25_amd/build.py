"""Builds libfem25hip.so (gfx950) in-tree with hipcc.

    python -m fem25_amd.build [--force]

The library is host C++ (mesh numbering, tables) plus HIP kernels; the C ABI
is declared in include/fem25hip.h.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfem25hip.so")
SRCS = [os.path.join(HERE, "csrc", f) for f in ("host_setup.cpp", "capi.hip")]
DEPS = SRCS + [os.path.join(HERE, "csrc", f) for f in ("kernels.hip", "f25_internal.h")] + [
    os.path.join(ROOT, "include", "fem25hip.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FEM25_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
         # no packed FP32 instructions anywhere in device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # MFMA accumulators in ordinary VGPRs: without it the loop-carried
         # accumulators of k_apply are copied VGPR<->AGPR on every k step
         "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-Wno-unused-result"]


def stale() -> bool:
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    return any(os.path.getmtime(p) > t for p in DEPS)


def build(force: bool = False, verbose: bool = False) -> str:
    if not force and not stale():
        return LIB
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), *SRCS, "-o", tmp, "-L/opt/rocm/lib", "-lrocblas",
           "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libfem25hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
    print(LIB)
