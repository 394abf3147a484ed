"""ctypes binding of the shader's C ABI (include/diff_envshade.h).

Loads the in-tree lib/libdiffenvshade.so built by diff_envshade_amd.build.
There is no fallback: if the library is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libdiffenvshade.so")

_lock = threading.Lock()
_lib = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_int = ctypes.c_int
_sz = ctypes.c_size_t

# argument kinds per entry point, in C order
_SIGS = {
    "des_bruteforce_diffuse_shader_forward":
        [_i64, _i32, _i32, _i32] + [_vp] * 4 + [_vp] + [_vp, _sz, _vp],
    "des_bruteforce_diffuse_shader_backward":
        [_i64, _i32, _i32, _i32] + [_vp] * 5 + [_vp] * 3 + [_vp, _sz, _vp],
    "des_bruteforce_specular_shader_forward":
        [_i64, _i32, _i32, _i32] + [_vp] * 7 + [_int, _int] + [_vp] + [_vp, _sz, _vp],
    "des_bruteforce_specular_shader_backward":
        [_i64, _i32, _i32, _i32] + [_vp] * 8 + [_int, _int] + [_vp] * 6 + [_vp, _sz, _vp],
    "des_bruteforce_specular_shader_clamp_forward":
        [_i64, _i32, _i32, _i32] + [_vp] * 7 + [_int, _int] + [_vp] + [_vp, _sz, _vp],
    "des_bruteforce_specular_shader_clamp_backward":
        [_i64, _i32, _i32, _i32] + [_vp] * 8 + [_int, _int] + [_vp] * 6 + [_vp, _sz, _vp],
    "des_bruteforce_specular_shader_forward2":
        [_i64, _i32, _i32, _i32] + [_vp] * 7 + [_int] + [_vp, _vp] + [_vp, _sz, _vp],
    "des_bruteforce_specular_shader_backward2":
        [_i64, _i32, _i32, _i32] + [_vp] * 9 + [_int] + [_vp] * 6 + [_vp, _sz, _vp],
    "des_bruteforce_specularvisibility_shader_forward2":
        [_i64, _i32, _i32, _i32, _i32, _i32, _i32] + [_vp] * 11 + [_int] + [_vp, _vp]
        + [_vp, _sz, _vp],
    "des_bruteforce_specularvisibility_shader_backward2":
        [_i64, _i32, _i32, _i32, _i32, _i32, _i32] + [_vp] * 13 + [_int] + [_vp] * 6
        + [_vp, _sz, _vp],
    "des_bruteforce_specularvisibility_shader_backward2_noalloc":
        [_i64, _i32, _i32, _i32, _i32, _i32, _i32] + [_vp] * 13 + [_int] + [_vp] * 6 + [_vp],
    "des_get_nearest_mesh_points_forward":
        [_i32, _i64, _i64, _i32, _vp, _vp, _i64] + [_vp] * 6 + [_vp],
    "des_get_nearest_mesh_points_backward":
        [_i32, _i64, _i64] + [_vp] * 6 + [_vp],
    "des_get_uv_of_triangle_forward": [_i64] + [_vp] * 3 + [_vp],
    "des_get_uv_of_triangle_backward": [_i64] + [_vp] * 4 + [_vp],
    "des_extract_bitfield_forward": [_vp, _i32, _i64, _i32, _i64, _i64, _vp, _vp],
    "des_sample_envmap_forward": [_i64, _i32, _i32, _i32, _vp, _vp, _vp, _vp],
    "des_blur_envmap_forward": [_i32] * 6 + [ctypes.c_float, _vp, _vp, _vp],
}

ERROR_NAMES = {
    -1: "DES_ERR_INVALID_ARGUMENT",
    -2: "DES_ERR_UNSUPPORTED_CHANNELS",
    -3: "DES_ERR_WORKSPACE_TOO_SMALL",
    -4: "DES_ERR_LAUNCH",
}


class ShaderError(RuntimeError):
    """A C-ABI call returned a negative status (raised as RuntimeError, like
    the reference's TORCH_CHECK failures)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def lib_path() -> str:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Load (once) and return the shader library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP shader library not found at {LIB_PATH}; build it with "
                "`python -m diff_envshade_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _int
        L.des_abi_version.argtypes = []
        L.des_abi_version.restype = _int
        L.des_last_error.argtypes = []
        L.des_last_error.restype = ctypes.c_char_p
        L.des_backward_workspace_size.argtypes = [_i64, _i32, _i32]
        L.des_backward_workspace_size.restype = _sz
        L.des_forward_workspace_size.argtypes = [_i64, _i32, _i32]
        L.des_forward_workspace_size.restype = _sz
        L.des_set_point_order.argtypes = [_i32]
        L.des_set_point_order.restype = _int
        L.des_set_tiling_override.argtypes = [_i32, _i32]
        L.des_set_tiling_override.restype = _int
        L.des_set_backward_timing_events.argtypes = [_vp, _vp]
        L.des_set_backward_timing_events.restype = _int
        L.des_set_forward_timing_events.argtypes = [_vp, _vp]
        L.des_set_forward_timing_events.restype = _int
        L.des_reuse_point_order_once.argtypes = [_vp, _i64]
        L.des_reuse_point_order_once.restype = _int
        L.des_static_scratch_bytes.argtypes = []
        L.des_static_scratch_bytes.restype = _sz
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().des_last_error().decode(errors="replace")
        raise ShaderError(rc, msg or ERROR_NAMES.get(rc, f"error {rc}"))


def call(name: str, *args, order=None) -> None:
    """Call entry point `name`; with order = (forward workspace pointer, P) the
    call takes its point order from that workspace (des_reuse_point_order_once).
    The hand-over is armed right before the call and dropped if the call
    cannot be made, so it never reaches a later call."""
    L = load()
    fn = getattr(L, name)
    if order is None:
        check(fn(*args))
        return
    check(L.des_reuse_point_order_once(*order))
    try:
        rc = fn(*args)
    except BaseException:
        L.des_reuse_point_order_once(None, 0)
        raise
    check(rc)


def workspace_size(P: int, H: int, W: int) -> int:
    return int(load().des_backward_workspace_size(P, H, W))


def forward_workspace_size(P: int, H: int, W: int) -> int:
    return int(load().des_forward_workspace_size(P, H, W))


def reuse_point_order_once(order_workspace_ptr, P: int) -> None:
    """The calling thread's next shading call for P points takes its point order
    from this forward workspace (see des_reuse_point_order_once)."""
    check(load().des_reuse_point_order_once(order_workspace_ptr, P))


def set_point_order(mode: int) -> None:
    """0 automatic, 1 always, -1 never (see des_set_point_order)."""
    check(load().des_set_point_order(mode))


def set_backward_timing_events(start, stop) -> None:
    """Record torch.cuda.Event / hipEvent_t handles around the next backward's
    main kernel (see des_set_backward_timing_events)."""
    h0 = start.cuda_event if hasattr(start, "cuda_event") else start
    h1 = stop.cuda_event if hasattr(stop, "cuda_event") else stop
    check(load().des_set_backward_timing_events(h0, h1))


def set_forward_timing_events(start, stop) -> None:
    """Record torch.cuda.Event / hipEvent_t handles around the next forward's
    main kernel (see des_set_forward_timing_events)."""
    h0 = start.cuda_event if hasattr(start, "cuda_event") else start
    h1 = stop.cuda_event if hasattr(stop, "cuda_event") else stop
    check(load().des_set_forward_timing_events(h0, h1))


def set_tiling_override(ntc_forward: int = 0, ntc_backward: int = 0) -> None:
    check(load().des_set_tiling_override(ntc_forward, ntc_backward))
