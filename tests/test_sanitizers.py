"""CPU sanitizer builds (AddressSanitizer + UndefinedBehaviorSanitizer).

* The FP64 oracle (`make -C oracle sanitize`): every shader variant and
  branch combination forward and backward on a vertex-indexed mesh problem,
  the mesh utilities, the bitfield decode and the env-map utilities run in a
  child process with gcc's libasan preloaded; any invalid access or undefined
  operation aborts it.  Its outputs equal the plain build's.
* The C ABI's host code (`python -m diff_envshade_amd.build --sanitize`:
  validation, workspace layout, dispatch instrumented; device code as usual):
  the CPU tests of the boundary (tests/test_capi.py, tests/test_dropin_host.py)
  run against it in a child process with clang's ASan runtime preloaded.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from _helpers import problem
from diff_envshade_amd import workloads as WL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _gcc_asan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                       text=True).stdout.strip()
    if not os.path.isabs(p) or not os.path.exists(p):
        pytest.skip("gcc's libasan is not installed")
    return p


def _inputs():
    d = problem(seed=3, n=77, H=8, W=16, vis="mesh", n_bright=2)
    d["roughness"][:5] = np.array([[0.0], [0.01], [1.0], [0.3], [0.05]], dtype=np.float32)
    d["view_dir"][5] = -d["normal"][5]  # back-facing
    rng = np.random.default_rng(4)
    m = WL.mesh_queries(seed=5, subdiv=1, n=33, K=6)
    inputs = dict(d)
    inputs.update({f"mesh/{k}": v for k, v in m.items()})
    inputs["mesh/grad_bary"] = rng.normal(size=(3, 33))
    tri = rng.normal(size=(21, 3, 3))
    inputs["uv/triangles"] = tri
    inputs["uv/query"] = tri.mean(axis=1) + 0.3 * rng.normal(size=(21, 3))
    inputs["uv/grad"] = rng.normal(size=(21, 3))
    q = rng.normal(size=(3, 50))
    q[:, 0] = [0.0, 0.0, 1.0]  # a pole
    inputs["env/query"] = q
    inputs["env/map"] = rng.lognormal(size=(3, 8, 16)).astype(np.float32)
    return inputs


def test_oracle_under_asan_and_ubsan(tmp_path):
    asan = _gcc_asan()
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True,
                   stdout=subprocess.DEVNULL)
    san = os.path.join(ROOT, "oracle", "_san", "liboracle_des.so")
    inputs = _inputs()
    np.savez(tmp_path / "in.npz", **inputs)
    res = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "_asan_oracle_child.py"),
                          san, str(tmp_path / "in.npz"), str(tmp_path / "out.npz")],
                         env=dict(ENV, LD_PRELOAD=asan), capture_output=True, text=True,
                         timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]
    assert "_san/liboracle_des.so" in res.stdout
    # the sanitized build computes what the plain one does
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _asan_oracle_child as child
    plain = child.run(inputs)
    with np.load(tmp_path / "out.npz") as f:
        assert sorted(f.files) == sorted(plain)
        for k in f.files:
            ref = np.asarray(plain[k], dtype=np.float64)
            got = np.asarray(f[k], dtype=np.float64)
            scale = max(np.abs(ref).max(), 1e-300) if ref.size else 1.0
            assert np.abs(got - ref).max(initial=0.0) <= 1e-12 * scale, k


def test_capi_host_code_under_asan_and_ubsan():
    from diff_envshade_amd.build import asan_runtime, build
    try:
        rt = asan_runtime()
    except RuntimeError as e:
        pytest.skip(str(e))
    san = build(sanitize=True)
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from diff_envshade_amd import _lib\n"
        "_lib.LIB_PATH = %r\n"
        "import ctypes, pytest\n"
        "assert hasattr(ctypes.CDLL(None), '__asan_init'), 'ASan runtime not loaded'\n"
        "rc = pytest.main(['-q', '-m', 'not gpu', '-p', 'no:cacheprovider', '--rootdir', %r,\n"
        "                  %r, %r])\n"
        "print('loaded', _lib.load()._name, flush=True)\n"
        "sys.exit(int(rc))\n"
    ) % (ROOT, san, ROOT, os.path.join(ROOT, "tests", "test_capi.py"),
         os.path.join(ROOT, "tests", "test_dropin_host.py"))
    res = subprocess.run([sys.executable, "-c", code], env=dict(ENV, LD_PRELOAD=rt), cwd=ROOT,
                         capture_output=True, text=True, timeout=900)
    assert res.returncode == 0, res.stdout[-4000:] + res.stderr[-4000:]
    assert "lib/asan/libdiffenvshade.so" in res.stdout
    assert " passed" in res.stdout and " failed" not in res.stdout
