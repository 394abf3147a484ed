"""C-ABI library: loads, exports every symbol declared in include/*.h, and
rejects invalid arguments with the documented codes before touching a GPU
(CPU only: no call here reaches a kernel launch)."""
import ctypes
import glob
import os
import re

import pytest

from diff_envshade_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUMMY = 4096  # non-NULL placeholder pointer; never dereferenced on these paths


def declared_symbols():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        names += re.findall(r"DES_API\s+[\w\s\*]+?\b(des_\w+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_the_ten_shader_entry_points():
    names = declared_symbols()
    for v in ("diffuse_shader_forward", "diffuse_shader_backward", "specular_shader_forward",
              "specular_shader_backward", "specular_shader_clamp_forward",
              "specular_shader_clamp_backward", "specular_shader_forward2",
              "specular_shader_backward2", "specularvisibility_shader_forward2",
              "specularvisibility_shader_backward2"):
        assert f"des_bruteforce_{v}" in names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    assert L.des_abi_version() == 2


def test_library_links_no_torch():
    # the boundary library depends on the HIP runtime only
    with open(_lib.lib_path(), "rb") as f:
        blob = f.read()
    assert b"libamdhip64" in blob
    assert b"libtorch" not in blob and b"libc10" not in blob


def test_workspace_size():
    assert _lib.workspace_size(0, 16, 32) == 0
    n = _lib.workspace_size(70000, 16, 32)
    tiles = (70000 + 63) // 64
    assert n >= tiles * 3 * 512 * 4
    assert n % 256 == 0
    assert _lib.workspace_size(1, 16, 32) > 0


@pytest.mark.parametrize("P,H,W", [(1, 16, 32), (70000, 16, 32), (300000, 32, 64), (4096, 8, 8)])
def test_serving_charges_the_batch_workspace(P, H, W):
    # the server shades through the batch API: a request is charged its share
    # of the batch workspace (all of it minus the descriptor table)
    from diff_envshade_amd.batch import DesSet, _bind
    from diff_envshade_amd.serving import request_bytes
    L = _bind(_lib.load())
    s = DesSet()
    s.P, s.H, s.W = P, H, W
    ws, st = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert L.des_batch_plan(4, ctypes.byref(s), 1, ctypes.byref(ws), ctypes.byref(st)) == 0
    grads = P * 13 * 4 + 3 * H * W * 4
    outs = P * 3 * 4 * 2
    assert request_bytes(P, H * W, True) - grads - outs == ws.value - st.value
    assert request_bytes(P, H * W, False) - outs == ws.value - st.value


def test_batch_workspace_holds_the_order_of_large_sets():
    # a set of >= 2^24 point-texel pairs gets a point-order workspace (sorted at
    # prepare time); smaller sets do not
    from diff_envshade_amd.batch import DesSet, _bind
    L = _bind(_lib.load())

    def plan(*specs):
        arr = (DesSet * len(specs))()
        for a, (P, H, W) in zip(arr, specs):
            a.P, a.H, a.W = P, H, W
        ws, st = ctypes.c_size_t(0), ctypes.c_size_t(0)
        assert L.des_batch_plan(4, arr, len(specs), ctypes.byref(ws), ctypes.byref(st)) == 0
        return ws.value, st.value

    small, big = (20000, 16, 32), (70000, 16, 32)
    ws_s, st_s = plan(small)
    ws_b, st_b = plan(big)
    ws_sb, st_sb = plan(small, big)
    try:
        _lib.set_point_order(-1)
        ws_b0, st_b0 = plan(big)
    finally:
        _lib.set_point_order(0)
    assert ws_b0 - st_b0 < ws_b - st_b  # the order region
    assert ws_sb - st_sb == (ws_s - st_s) + (ws_b - st_b)


def test_forward_workspace_follows_point_order_policy():
    try:
        _lib.set_point_order(-1)
        assert _lib.forward_workspace_size(70000, 16, 32) == 0
        never = _lib.workspace_size(70000, 16, 32)
        _lib.set_point_order(1)
        fw = _lib.forward_workspace_size(100, 4, 8)
        assert fw >= 100 * 4 and fw % 256 == 0
        order = _lib.forward_workspace_size(70000, 16, 32)
        # the forward's also holds its tail slices' sums (70 tiles x 4 slices x 6 x 64 floats)
        fsum = (70 * 4 * 6 * 64 * 4 + 255) // 256 * 256
        assert _lib.workspace_size(70000, 16, 32) == never + order - fsum
        _lib.set_point_order(0)
        # automatic: both directions reorder from 2^24 pairs
        assert _lib.forward_workspace_size(70000, 16, 32) == order  # 35.8 M pairs
        assert _lib.workspace_size(70000, 16, 32) == never + order - fsum
        assert _lib.forward_workspace_size(32768, 16, 32) > 0  # 2^24 pairs
        assert _lib.forward_workspace_size(32767, 16, 32) == 0
        assert _lib.workspace_size(32767, 16, 32) < _lib.workspace_size(32768, 16, 32)
        assert _lib.forward_workspace_size(100, 4, 8) == 0
        with pytest.raises(RuntimeError):
            _lib.set_point_order(2)
    finally:
        _lib.set_point_order(0)


def test_forward_rejects_small_order_workspace():
    L = _lib.load()
    try:
        _lib.set_point_order(1)
        need = _lib.forward_workspace_size(10, 4, 8)
        rc = L.des_bruteforce_specularvisibility_shader_forward2(
            10, 3, 4, 8, 10, 10, 1, *([DUMMY] * 11), 1, DUMMY, DUMMY, DUMMY, need - 1, None)
        assert rc == -3
        rc = L.des_bruteforce_specularvisibility_shader_forward2(
            10, 3, 4, 8, 10, 10, 1, *([DUMMY] * 11), 1, DUMMY, DUMMY, DUMMY + 4, need, None)
        assert rc == -1
    finally:
        _lib.set_point_order(0)


def _vis_fwd(P=10, C=3, H=4, W=8, NF=10, NV=10, NS=1, ptr=DUMMY):
    L = _lib.load()
    return L.des_bruteforce_specularvisibility_shader_forward2(
        P, C, H, W, NF, NV, NS, *([ptr] * 11), 1, ptr, ptr, None, 0, None)


def test_channels_must_be_three():
    assert _vis_fwd(C=4) == -2
    assert "Channel number not supported!" in _lib.load().des_last_error().decode()


@pytest.mark.parametrize("kw,code", [({"P": -1}, -1), ({"H": 0}, -1), ({"W": -3}, -1),
                                     ({"NS": 0}, -1), ({"NF": 0}, -1), ({"NV": 0}, -1),
                                     ({"ptr": None}, -1)])
def test_invalid_arguments(kw, code):
    assert _vis_fwd(**kw) == code
    assert _lib.load().des_last_error().decode() != ""


def test_slots_must_cover_envmap():
    # 16 x 32 = 512 texels need 16 slots of 32 bits
    assert _vis_fwd(H=16, W=32, NS=15) == -1
    assert "needs 16" in _lib.load().des_last_error().decode()


def test_backward_rejects_small_workspace():
    L = _lib.load()
    P, H, W = 100, 4, 8
    need = _lib.workspace_size(P, H, W)
    rc = L.des_bruteforce_specular_shader_backward2(
        P, 3, H, W, *([DUMMY] * 9), 1, *([DUMMY] * 6), DUMMY, need - 1, None)
    assert rc == -3
    rc = L.des_bruteforce_specular_shader_backward2(
        P, 3, H, W, *([DUMMY] * 9), 1, *([DUMMY] * 6), None, need, None)
    assert rc == -3
    # misaligned workspace
    rc = L.des_bruteforce_specular_shader_backward2(
        P, 3, H, W, *([DUMMY] * 9), 1, *([DUMMY] * 6), DUMMY + 4, need, None)
    assert rc == -1


def test_tiling_override_validation():
    with pytest.raises(RuntimeError):
        _lib.set_tiling_override(3, 0)
    with pytest.raises(RuntimeError):
        _lib.set_tiling_override(0, 16)
    _lib.set_tiling_override(0, 0)


def test_error_raises_runtime_error_with_message():
    with pytest.raises(RuntimeError, match="Channel number not supported!"):
        _lib.check(_vis_fwd(C=1))


def test_timing_events_need_both():
    L = _lib.load()
    assert L.des_set_backward_timing_events(ctypes.c_void_p(DUMMY), None) == -1
    assert L.des_set_backward_timing_events(None, None) == 0
    assert L.des_set_forward_timing_events(ctypes.c_void_p(DUMMY), None) == -1
    assert L.des_set_forward_timing_events(None, None) == 0


# ---- mesh utilities ----------------------------------------------------------------

def test_header_declares_the_mesh_entry_points():
    names = declared_symbols()
    for n in ("get_nearest_mesh_points_forward", "get_nearest_mesh_points_backward",
              "get_uv_of_triangle_forward", "get_uv_of_triangle_backward",
              "extract_bitfield_forward"):
        assert f"des_{n}" in names


def _nearest_fwd(NV=10, NF=10, P=4, K=2, n_adj=30, ptr=DUMMY):
    return _lib.load().des_get_nearest_mesh_points_forward(NV, NF, P, K, ptr, ptr, n_adj, ptr,
                                                           ptr, ptr, ptr, ptr, ptr, None)


@pytest.mark.parametrize("kw", [{"NV": -1}, {"NF": -1}, {"P": -1}, {"K": -1}, {"n_adj": -1},
                                {"n_adj": 1 << 26}, {"P": 1 << 31}, {"ptr": None}])
def test_nearest_forward_rejects_invalid_arguments(kw):
    assert _nearest_fwd(**kw) == -1
    assert _lib.load().des_last_error().decode() != ""


def test_mesh_entry_points_with_no_points_do_nothing():
    L = _lib.load()
    assert _nearest_fwd(P=0, ptr=None) == 0
    assert L.des_get_nearest_mesh_points_backward(10, 10, 0, *([None] * 7)) == 0
    assert L.des_get_uv_of_triangle_forward(0, None, None, None, None) == 0
    assert L.des_get_uv_of_triangle_backward(0, None, None, None, None, None) == 0
    assert L.des_get_uv_of_triangle_forward(-1, None, None, None, None) == -1
    assert L.des_get_nearest_mesh_points_backward(10, 10, 5, DUMMY, DUMMY, DUMMY, None, DUMMY,
                                                  DUMMY, None) == -1


@pytest.mark.parametrize("args,ok", [
    ((2, 100, 64, 0, 100), True),     # whole range, N_bit = 32 * N_slot
    ((2, 100, 65, 0, 100), False),    # more bits than the slots hold
    ((2, 100, 10, 5, 5), True),       # empty range
    ((2, 100, 10, 6, 5), False),      # i_start > i_end
    ((2, 100, 10, 0, 101), False),    # beyond N
    ((2, 100, 10, -1, 3), False),
    ((2, 100, 0, 0, 100), True),      # no bits
])
def test_extract_bitfield_validation(args, ok):
    N_slot, N, N_bit, i0, i1 = args
    L = _lib.load()
    if ok and (i1 - i0) * N_bit > 0:
        return  # valid and non-empty: would launch a kernel
    rc = L.des_extract_bitfield_forward(DUMMY, N_slot, N, N_bit, i0, i1, DUMMY, None)
    assert rc == (0 if ok else -1), _lib.load().des_last_error()


def test_python_extract_bitfield_reports_invalid_ranges():
    import torch
    from diff_envshade_amd import _diff_envshade as de
    # the checks run before any device work, so a CPU tensor is rejected first
    with pytest.raises(RuntimeError, match="CUDA"):
        de.extract_bitfield_forward(torch.zeros(2, 4, dtype=torch.int32), 8, 0, 4)


# ---- env-map utilities ---------------------------------------------------------------

@pytest.mark.parametrize("args,code", [
    ((5, 8, 16, 4, 8, 10, 0.1), -2),   # C must be 3 or 4
    ((3, 0, 16, 4, 8, 10, 0.1), -1),
    ((3, 8, 16, -1, 8, 10, 0.1), -1),
    ((3, 8, 16, 4, 8, 0, 0.1), -1),    # at least one sample
    ((3, 8, 16, 4, 8, 10, float("nan")), -1),
    ((3, 8, 16, 0, 8, 10, 0.1), 0),    # empty target: nothing to do
])
def test_blur_validation(args, code):
    rc = _lib.load().des_blur_envmap_forward(*args, DUMMY, DUMMY, None)
    assert rc == code, _lib.load().des_last_error()


def test_sample_validation():
    L = _lib.load()
    assert L.des_sample_envmap_forward(-1, 3, 8, 16, DUMMY, DUMMY, DUMMY, None) == -1
    assert L.des_sample_envmap_forward(10, 0, 8, 16, DUMMY, DUMMY, DUMMY, None) == -1
    assert L.des_sample_envmap_forward(10, 3, 8, 16, None, DUMMY, DUMMY, None) == -1
    assert L.des_sample_envmap_forward(0, 3, 8, 16, None, None, None, None) == 0


def test_reuse_point_order_hook_validation():
    L = _lib.load()
    assert L.des_reuse_point_order_once(ctypes.c_void_p(DUMMY + 4), 10) == -1
    assert "256-byte aligned" in L.des_last_error().decode()
    assert L.des_reuse_point_order_once(ctypes.c_void_p(DUMMY), 0) == -1
    assert L.des_reuse_point_order_once(ctypes.c_void_p(DUMMY), 10) == 0
    # any call consumes the hook: a failing one here, before any launch
    assert _vis_fwd(C=4) == -2
    assert L.des_reuse_point_order_once(None, 0) == 0


def test_call_drops_the_hand_over_when_the_call_cannot_be_made(monkeypatch):
    # the entry point cannot be called (here: wrong argument count, as a
    # ctypes conversion error would): the helper disarms before re-raising
    hooks = []

    class Fake:
        def des_reuse_point_order_once(self, ptr, P):
            hooks.append((ptr, P))
            return 0

        def des_bruteforce_specular_shader_backward2(self, *args):
            raise ctypes.ArgumentError("argument 1: wrong type")

    monkeypatch.setattr(_lib, "load", lambda: Fake())
    with pytest.raises(ctypes.ArgumentError):
        _lib.call("des_bruteforce_specular_shader_backward2", "x", order=(DUMMY, 10))
    assert hooks == [(DUMMY, 10), (None, 0)]


def test_batch_launches_reject_an_unprepared_workspace():
    # des_batch_forward / _backward / _sort on a workspace that no
    # des_batch_prepare wrote return before any launch
    from diff_envshade_amd.batch import _bind
    L = _bind(_lib.load())
    assert L.des_batch_forward(4, 1, 1, 8, DUMMY, None) == -1
    assert b"not prepared" in L.des_last_error()
    assert L.des_batch_backward(4, 1, 1, 8, DUMMY, None) == -1
    assert b"not prepared" in L.des_last_error()
    assert L.des_batch_sort(DUMMY, None) == -1
    assert b"not prepared" in L.des_last_error()
