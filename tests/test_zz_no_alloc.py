"""The backward entry point that allocates nothing
(des_bruteforce_specularvisibility_shader_backward2_noalloc) against the
workspace entry point it shares its kernels with: bitwise the same outputs on
two calls in a row, the second smaller, both through the C ABI; sizes beyond
the static scratch are rejected before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from _helpers import compare, gpu_inputs, oracle_run, problem

pytestmark = pytest.mark.gpu


def _args(t, P, H, W, outs):
    vis = t["visibility"]
    return [P, 3, H, W, t["face_vertex_list"].shape[1], vis.shape[1], vis.shape[0],
            t["grad_shading"].data_ptr(), t["grad_diffuse_shading"].data_ptr(),
            t["normal"].data_ptr(), t["view_dir"].data_ptr(), t["albedo"].data_ptr(),
            t["specular_albedo"].data_ptr(), t["roughness"].data_ptr(), t["envmap"].data_ptr(),
            t["light2obj_rotmat"].data_ptr(), t["face_vertex_list"].data_ptr(),
            t["nearest_triangle_id"].data_ptr(), t["barycentric_coord"].data_ptr(),
            vis.data_ptr(), 1] + [o.data_ptr() for o in outs]


def _outs(P, H, W):
    shapes = [(P, 3)] * 4 + [(P, 1), (3, H, W)]
    return [torch.full(s, float("nan"), device="cuda") for s in shapes]


def _both(d):
    from diff_envshade_amd import _lib
    L = _lib.load()
    t = gpu_inputs(d)
    P = d["normal"].shape[0]
    _, H, W = d["envmap"].shape
    stream = torch.cuda.current_stream().cuda_stream
    ref = _outs(P, H, W)
    nbytes = _lib.workspace_size(P, H, W)
    ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device="cuda")
    _lib.check(L.des_bruteforce_specularvisibility_shader_backward2(
        *_args(t, P, H, W, ref), ws.data_ptr(), nbytes, stream))
    got = _outs(P, H, W)
    print(f"noalloc backward: P={P} H={H} W={W}", flush=True)
    _lib.check(L.des_bruteforce_specularvisibility_shader_backward2_noalloc(
        *_args(t, P, H, W, got), stream))
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in ref], [x.cpu().numpy() for x in got]


def test_noalloc_backward_matches_the_workspace_entry_point_twice():
    # first call: the headline size (points sorted by normal, order in the
    # scratch); second call smaller, on a general vertex-indexed mesh and a
    # 32 x 64 map (also sorted), so it reads what the first left behind only
    # where it rewrites it first
    cases = [problem(seed=11, n=70000, H=16, W=32, vis="identity", n_bright=3),
             problem(seed=12, n=9000, H=32, W=64, vis="mesh")]
    for d in cases:
        ref, got = _both(d)
        for a, b in zip(ref, got):
            assert np.isfinite(b).all()
            assert np.array_equal(a, b)
    # the smaller call is also within the tolerance of the FP64 oracle
    d = cases[1]
    compare("vis", ([], got), ([], oracle_run("vis", d)[1]))


def test_noalloc_backward_rejects_problems_beyond_the_scratch():
    from diff_envshade_amd import _lib
    L = _lib.load()
    limit = L.des_static_scratch_bytes()
    assert limit == 32 << 20
    # the documented limits: 276 480 points at 16 x 32, 77 184 at 32 x 64
    for n, h, w in ((276480, 16, 32), (77184, 32, 64)):
        assert _lib.workspace_size(n, h, w) <= limit < _lib.workspace_size(n + 1, h, w)
    P, H, W = 400000, 16, 32
    assert _lib.workspace_size(P, H, W) > limit
    args = [P, 3, H, W, 1, 1, 16] + [4096] * 13 + [1] + [4096] * 6 + [None]
    rc = L.des_bruteforce_specularvisibility_shader_backward2_noalloc(*args)
    assert rc == -1
    assert b"static" in L.des_last_error()
