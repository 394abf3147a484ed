"""Many env maps lighting one point set, on the MFMA matrix cores.

    sh, ds = relight_forward(normal, view_dir, albedo, specular_albedo, roughness,
                             envmaps, light2obj_rotmat, visibility, precision="bf16x3")
    gE = relight_backward_envmap(normal, view_dir, albedo, specular_albedo, roughness,
                                 light2obj_rotmat, visibility, grad_shading,
                                 grad_diffuse_shading, precision="bf16x3")

envmaps: E x 3 x H x W; outputs and upstream gradients: P x E x 3 (map e of
point p at [p, e]); visibility: per-point words N_slot x P (int32, bit t%32 of
word [t/32][p] = texel t) or None for all visible.  Map e's results are those
of the drop-in visibility shader (bruteforce_specularvisibility_shader_forward2
/ _backward2 with one vertex per point) lit by envmaps[e], up to the chosen
operand precision:

  "bf16"   : weights and radiance rounded to bf16, one MFMA per product
  "fp16"   : the same in fp16; with specular on, roughness must be at least
             FP16_MIN_ROUGHNESS (0.0558) so every GGX weight fits fp16's 65504
             (ValueError otherwise)
  "bf16x3" : hi + lo bf16 split of both operands, three MFMAs per product

Accumulation is fp32 on the matrix cores; the cross-workgroup env-map-gradient
sum is FP64 in a fixed order.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _lib

PRECISIONS = {"bf16": 0, "fp16": 1, "bf16x3": 2}

# The GGX weight F G D / 4 of a pair is at most 2 / (pi r^4): F <= 1, the G
# term n.l / (n.l (1-k) + k) * 1/4 / (n.v (1-k) + k) <= 2 because
# k = (r+1)^2 / 8 >= 1/8, and D <= 1 / (pi alpha^2) with alpha^2 = r^4 because
# tD = alpha^2 (n.h)^2 + 1 - (n.h)^2 >= alpha^2.  fp16 holds |x| <= 65504, so
# every weight fits from this roughness up.
FP16_MIN_ROUGHNESS = (2.0 / (math.pi * 65504.0)) ** 0.25


def _check_range(precision, enable_specular, roughness):
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)} (got {precision!r})")
    if precision == "fp16" and enable_specular and roughness.numel() > 0:
        rmin = float(roughness.min())
        if not rmin >= FP16_MIN_ROUGHNESS:
            raise ValueError(
                f"fp16 cannot hold the GGX weights of roughness {rmin:.4g}: it needs roughness "
                f">= {FP16_MIN_ROUGHNESS:.4f}; use precision='bf16' or 'bf16x3'")

_vp, _i32, _i64, _int, _sz = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int,
                              ctypes.c_size_t)


def _bind():
    L = _lib.load()
    if not getattr(L, "_relight_bound", False):
        L.des_relight_workspace_size.argtypes = [_i64, _i32, _i32, _i32, _i32, _int]
        L.des_relight_workspace_size.restype = _sz
        L.des_relight_forward.argtypes = ([_i64, _i32, _i32, _i32, _i32] + [_vp] * 8
                                          + [_int, _i32, _vp, _vp, _vp, _sz, _vp])
        L.des_relight_forward.restype = _int
        L.des_relight_backward_envmap.argtypes = ([_i64, _i32, _i32, _i32, _i32] + [_vp] * 9
                                                  + [_int, _i32, _vp, _vp, _sz, _vp])
        L.des_relight_backward_envmap.restype = _int
        L._relight_bound = True
    return L


def _p(t):
    return t.data_ptr() if (t is not None and t.numel() > 0) else None


def _check(name, t, dtype=torch.float32):
    if t is None:
        return
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise RuntimeError(f"{name} must be a contiguous CUDA tensor")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype}")


def workspace_size(P, E, H, W, precision="bf16", with_envgrad=False) -> int:
    return int(_bind().des_relight_workspace_size(P, E, H, W, PRECISIONS[precision],
                                                   int(with_envgrad)))


def relight_forward(normal, view_dir, albedo, specular_albedo, roughness, envmaps,
                    light2obj_rotmat, visibility=None, enable_specular=True, precision="bf16",
                    workspace=None):
    L = _bind()
    for n, t in (("normal", normal), ("view_dir", view_dir), ("albedo", albedo),
                 ("specular_albedo", specular_albedo), ("roughness", roughness),
                 ("envmaps", envmaps), ("light2obj_rotmat", light2obj_rotmat)):
        _check(n, t)
    _check("visibility", visibility, torch.int32)
    _check_range(precision, enable_specular, roughness)
    P = normal.shape[0]
    E, C, H, W = envmaps.shape
    if C != 3:
        raise RuntimeError("Channel number not supported!")
    dev = normal.device
    sh = torch.empty((P, E, 3), dtype=torch.float32, device=dev)
    ds = torch.empty((P, E, 3), dtype=torch.float32, device=dev)
    need = workspace_size(P, E, H, W, precision)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(max(need, 256), dtype=torch.uint8, device=dev)
    ns = visibility.shape[0] if visibility is not None else 0
    with torch.cuda.device(dev):
        _lib.check(L.des_relight_forward(
            P, E, H, W, ns, _p(normal), _p(view_dir), _p(albedo), _p(specular_albedo),
            _p(roughness), _p(envmaps), _p(light2obj_rotmat), _p(visibility),
            int(bool(enable_specular)), PRECISIONS[precision], _p(sh), _p(ds), _p(workspace),
            workspace.numel(), torch.cuda.current_stream(dev).cuda_stream))
    return sh, ds


def relight_backward_envmap(normal, view_dir, albedo, specular_albedo, roughness, H, W,
                            light2obj_rotmat, visibility, grad_shading, grad_diffuse_shading,
                            enable_specular=True, precision="bf16", workspace=None):
    L = _bind()
    for n, t in (("normal", normal), ("view_dir", view_dir), ("albedo", albedo),
                 ("specular_albedo", specular_albedo), ("roughness", roughness),
                 ("light2obj_rotmat", light2obj_rotmat), ("grad_shading", grad_shading),
                 ("grad_diffuse_shading", grad_diffuse_shading)):
        _check(n, t)
    _check("visibility", visibility, torch.int32)
    _check_range(precision, enable_specular, roughness)
    P, E, _ = grad_shading.shape
    dev = normal.device
    gE = torch.empty((E, 3, H, W), dtype=torch.float32, device=dev)
    need = workspace_size(P, E, H, W, precision, with_envgrad=True)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(max(need, 256), dtype=torch.uint8, device=dev)
    ns = visibility.shape[0] if visibility is not None else 0
    with torch.cuda.device(dev):
        _lib.check(L.des_relight_backward_envmap(
            P, E, H, W, ns, _p(normal), _p(view_dir), _p(albedo), _p(specular_albedo),
            _p(roughness), _p(light2obj_rotmat), _p(visibility), _p(grad_shading),
            _p(grad_diffuse_shading), int(bool(enable_specular)), PRECISIONS[precision], _p(gE),
            _p(workspace), workspace.numel(), torch.cuda.current_stream(dev).cuda_stream))
    return gE


class Relight(torch.autograd.Function):
    """Autograd over env maps only: forward gives (shading, diffuse_shading)
    P x E x 3 for every map; backward gives grad_envmaps (the per-point
    gradients are not produced by this path; use the drop-in shader for them)."""

    @staticmethod
    def forward(ctx, envmaps, normal, view_dir, albedo, specular_albedo, roughness,
                light2obj_rotmat, visibility=None, enable_specular=True, precision="bf16"):
        sh, ds = relight_forward(normal, view_dir, albedo, specular_albedo, roughness, envmaps,
                                 light2obj_rotmat, visibility, enable_specular, precision)
        ctx.save_for_backward(normal, view_dir, albedo, specular_albedo, roughness,
                              light2obj_rotmat, visibility)
        ctx.meta = (envmaps.shape[0], envmaps.shape[2], envmaps.shape[3], bool(enable_specular),
                    precision)
        return sh, ds

    @staticmethod
    def backward(ctx, g_sh, g_ds):
        normal, view_dir, albedo, spec, rough, rot, vis = ctx.saved_tensors
        E, H, W, es, prec = ctx.meta
        shape = (normal.shape[0], E, 3)
        g_sh = normal.new_zeros(shape) if g_sh is None else g_sh.contiguous()
        g_ds = normal.new_zeros(shape) if g_ds is None else g_ds.contiguous()
        gE = relight_backward_envmap(normal, view_dir, albedo, spec, rough, H, W, rot, vis,
                                     g_sh, g_ds, es, prec)
        return gE, None, None, None, None, None, None, None, None, None
