"""Build recipe for the HIP shader library (gfx950 only).

    python -m diff_envshade_amd.build            # compile lib/libdiffenvshade.so
    python -m diff_envshade_amd.build --sanitize  # host code under ASan + UBSan

The library is built in-tree so it travels with the source tree; it links only
libamdhip64 (no torch), exporting the C ABI of include/diff_envshade.h.  Each
source is compiled to an object in parallel, then linked.

The sanitized build (lib/asan/libdiffenvshade.so) instruments the host code
only -- argument validation, workspace layout and dispatch -- with
AddressSanitizer and UndefinedBehaviorSanitizer; the device code is built as
usual.  It needs clang's ASan runtime preloaded into the process that loads it
(LD_PRELOAD=`asan_runtime()`), and is only used by the CPU tests.
"""
from __future__ import annotations

import concurrent.futures
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
SAN_LIB_PATH = os.path.join(LIB_DIR, "asan", LIB_NAME)

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
    "-fvisibility=hidden",
    # no packed-FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
# host-side sanitizers only (each flag scoped to the host pass)
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
_NOISE = "is not a recognized feature for this target"


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the shader library needs ROCm's hipcc")
    return path


def asan_runtime() -> str:
    """clang's shared ASan runtime (with UBSan) that a sanitized build needs."""
    out = subprocess.run([os.path.join(os.path.dirname(os.path.realpath(hipcc())), "..", "llvm", "bin",
                                       "clang"), "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True).stdout.strip()
    if not out or not os.path.isabs(out):
        llvm = "/opt/rocm/llvm/lib/clang"
        for ver in sorted(os.listdir(llvm)) if os.path.isdir(llvm) else []:
            cand = os.path.join(llvm, ver, "lib", "linux", "libclang_rt.asan-x86_64.so")
            if os.path.exists(cand):
                out = cand
    if not out or not os.path.exists(out):
        raise RuntimeError("clang's ASan runtime (libclang_rt.asan-x86_64.so) not found")
    return out


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if _NOISE not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)


def build(verbose: bool = False, sanitize: bool = False, jobs: int = 0) -> str:
    """Compile every HIP source for gfx950 and link lib/libdiffenvshade.so
    (always rebuilds); with sanitize, lib/asan/libdiffenvshade.so."""
    out = SAN_LIB_PATH if sanitize else LIB_PATH
    obj_dir = os.path.join(os.path.dirname(out), "obj")
    os.makedirs(obj_dir, exist_ok=True)
    flags = HIPCC_FLAGS + (SAN_FLAGS if sanitize else [])
    objs = [os.path.join(obj_dir, s + ".o") for s in SOURCES]
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1, 16)
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        futs = [pool.submit(_run, [hipcc(), *flags, "-c", "-o", o, os.path.join(CSRC, s)], verbose)
                for s, o in zip(SOURCES, objs)]
        for f in futs:
            f.result()
    tmp = out + ".tmp"
    _run([hipcc(), *flags, "-shared", "-o", tmp, *objs], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True, sanitize="--sanitize" in sys.argv[1:]))
