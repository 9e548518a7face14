"""Builds libsigpde_hip.so in-tree with hipcc for gfx950 (no JIT, no torch extension).

Each source is compiled to an object with the common flags plus its own, then the objects are
linked into one shared library."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsigpde_hip.so")
OBJDIR = os.path.join(HERE, "_build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 (v_pk_*_f32) stays off: scalar FP32/FP64 VALU code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]

# source -> extra flags
SOURCES = {
    "capi.hip": [],
    "wavefront.hip": [],
    # MFMA results in VGPRs: the Gram producer writes its tiles to LDS straight from them (in AGPRs
    # they would be copied out every period)
    "fused_linear.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "fused_pairwise.hip": [],
    "static_increments.hip": [],
    "fused_point.hip": [],
    "fused_point_f32.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "fused_point_f64.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "fused_point_l1.hip": [],
    "rates.hip": [],
}


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)

    def compile_one(src: str) -> str:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, *SOURCES[src], *(extra or []), "-c", "-o", obj,
               os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, list(SOURCES)))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
