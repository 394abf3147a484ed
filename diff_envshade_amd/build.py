"""Build recipe for the HIP shader library (gfx950 only).

    python -m diff_envshade_amd.build        # compile lib/libdiffenvshade.so

The library is built in-tree so it travels with the source tree; it links only
libamdhip64 (no torch), exporting the C ABI of include/diff_envshade.h.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_NAME = "libdiffenvshade.so"
LIB_PATH = os.path.join(LIB_DIR, LIB_NAME)

SOURCES = ["envshade_kernels.hip", "point_order.hip", "envshade_capi.cpp", "relight_kernels.hip",
           "relight_capi.cpp", "mesh_kernels.hip", "mesh_capi.cpp", "envmap_kernels.hip",
           "envmap_capi.cpp"]
HEADERS = ["envshade_internal.h", os.path.join("..", "..", "include", "diff_envshade.h")]

ARCH = "gfx950"
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-fvisibility=hidden",
    # no packed-FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the shader library needs ROCm's hipcc")
    return path


def build(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 into lib/libdiffenvshade.so (always rebuilds)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *(os.path.join(CSRC, s) for s in SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
