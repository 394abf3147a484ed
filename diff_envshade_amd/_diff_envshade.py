"""Drop-in replacement for the reference's `_diff_envshade` extension module.

The ten shader functions keep the names, argument orders, return orders and
allocation semantics of bindings.cpp:6-19 / gsface_shader.h:3-118
(zjumsj/diff_envshade); each validates its tensors like the reference's
TORCH_CHECKs (bruteforce_shader.cu:20-24, raised as RuntimeError), allocates
its outputs on `normal`'s device and calls the HIP kernels through the C ABI
(include/diff_envshade.h) on the current stream of that device.

Differences from the reference, all deliberate: launches use the device's
current stream instead of the legacy default stream; tensor shapes and
devices are checked (the reference has `// TODO: shape check`); inputs must be
float32 (the reference admits half/double in CHECK_IS_FLOATING but then reads
them as float); grad_envmap is summed in a fixed order and is bitwise
reproducible.

The mesh utilities that produce the shader's mesh inputs (nearest mesh
points, closest point on a given triangle, visibility bitfield decoding;
bindings.cpp:27-33, etc.cu) and the env-map sampling and blur beside it
(bindings.cpp:21, 23) follow the same conventions.  The other etc.cu bindings
(camera rays, rotation conversion) are outside this library and raise
NotImplementedError.
"""
from __future__ import annotations

import weakref

import torch

from . import _lib

__all__ = [
    "bruteforce_diffuse_shader_forward",
    "bruteforce_diffuse_shader_backward",
    "bruteforce_specular_shader_forward",
    "bruteforce_specular_shader_backward",
    "bruteforce_specular_shader_clamp_forward",
    "bruteforce_specular_shader_clamp_backward",
    "bruteforce_specular_shader_forward2",
    "bruteforce_specular_shader_backward2",
    "bruteforce_specularvisibility_shader_forward2",
    "bruteforce_specularvisibility_shader_backward2",
    "get_nearest_mesh_points_forward",
    "get_nearest_mesh_points_backward",
    "get_uv_of_triangle_forward",
    "get_uv_of_triangle_backward",
    "extract_bitfield_forward",
    "sample_envmap_forward",
    "blur_envmap_forward",
]


# ---- checks mirroring bruteforce_shader.cu:20-24 ---------------------------------

def _check_cuda(x: torch.Tensor, name: str) -> None:
    if not isinstance(x, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if not x.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not x.is_contiguous():
        raise RuntimeError(f"{name} must be a contiguous tensor")


def _check_float(x: torch.Tensor, name: str) -> None:
    _check_cuda(x, name)
    if x.dtype != torch.float32:
        if x.dtype in (torch.float16, torch.float64, torch.bfloat16):
            raise RuntimeError(f"{name}: expected scalar type Float but found {x.dtype}")
        raise RuntimeError(f"{name} must be a floating tensor")


def _check_int(x: torch.Tensor, name: str) -> None:
    _check_cuda(x, name)
    if x.dtype != torch.int32:
        raise RuntimeError(f"{name} must be an int tensor")


def _check_long(x: torch.Tensor, name: str) -> None:
    _check_cuda(x, name)
    if x.dtype != torch.int64:
        raise RuntimeError(f"{name} must be an long tensor")


def _check_shape(x: torch.Tensor, name: str, shape: tuple) -> None:
    if tuple(x.shape) != tuple(shape):
        # roughness may also be given as a flat (P,) vector
        if not (len(shape) == 2 and shape[1] == 1 and tuple(x.shape) == (shape[0],)):
            raise RuntimeError(f"{name} must have shape {tuple(shape)}, got {tuple(x.shape)}")


def _same_device(dev: torch.device, tensors) -> None:
    for name, t in tensors:
        if t.device != dev:
            raise RuntimeError(f"{name} is on {t.device} but normal is on {dev}")


def _sizes(normal, albedo, envmap):
    if normal.dim() != 2 or normal.shape[1] != 3:
        raise RuntimeError(f"normal must have shape (P, 3), got {tuple(normal.shape)}")
    P = normal.shape[0]
    if albedo.dim() != 2 or albedo.shape[0] != P:
        raise RuntimeError(f"albedo must have shape (P, C), got {tuple(albedo.shape)}")
    C = albedo.shape[1]
    if C != 3:
        raise RuntimeError("Channel number not supported!")
    if envmap.dim() != 3 or envmap.shape[0] != C:
        raise RuntimeError(f"envmap must have shape (C, H, W), got {tuple(envmap.shape)}")
    return P, C, int(envmap.shape[1]), int(envmap.shape[2])


def _stream(dev: torch.device) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def _ptr(x):
    return x.data_ptr() if (x is not None and x.numel() > 0) else None


# The last forward's point order per device, for the backward that usually
# follows with the same normals: (weak reference to normal, its data pointer,
# its version counter, P, stream, the forward workspace holding the order).
_order_cache: dict = {}


def _workspace(P: int, H: int, W: int, dev: torch.device, normal: torch.Tensor):
    """Backward workspace, and the point order to hand over: when the device's
    last forward sorted these very normals -- the same tensor (alive, same
    data pointer, unchanged version counter) and P, on the current stream --
    (its workspace pointer, P), which the call passes to the library so that it
    takes that order instead of sorting again; else None."""
    n = _lib.workspace_size(P, H, W)
    if n == 0:
        return None, 0, None
    ws = torch.empty(n, dtype=torch.uint8, device=dev)
    order = None
    e = _order_cache.get(dev.index)
    if e is not None:
        ref, ptr, ver, p, stream, ows = e
        src = ref()
        if (src is not None and p == P and ptr == normal.data_ptr() == src.data_ptr()
                and ver == src._version == normal._version and stream == _stream(dev)):
            order = (ows.data_ptr(), P)
    return ws, n, order


def _fwd_workspace(P: int, H: int, W: int, dev: torch.device):
    """Point-order workspace of the forward entry points (None when the
    library would not reorder this problem)."""
    n = _lib.forward_workspace_size(P, H, W)
    if n == 0:
        return None, 0
    return torch.empty(n, dtype=torch.uint8, device=dev), n


def _remember_order(ws, P: int, dev: torch.device, normal: torch.Tensor) -> None:
    """After a forward call that sorted into `ws` has been enqueued: keep its
    order for the backward (the call raised otherwise, so the order is real)."""
    if ws is not None:
        _order_cache[dev.index] = (weakref.ref(normal), normal.data_ptr(), normal._version, P,
                                   _stream(dev), ws)


def _material_checks(P, C, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                     light2obj_rotmat):
    for name, t in (("normal", normal), ("view_dir", view_dir), ("albedo", albedo),
                    ("specular_albedo", specular_albedo), ("roughness", roughness),
                    ("envmap", envmap), ("light2obj_rotmat", light2obj_rotmat)):
        _check_float(t, name)
    _check_shape(view_dir, "view_dir", (P, 3))
    _check_shape(specular_albedo, "specular_albedo", (P, C))
    _check_shape(roughness, "roughness", (P, 1))
    _check_shape(light2obj_rotmat, "light2obj_rotmat", (3, 3))
    _same_device(normal.device, (("view_dir", view_dir), ("albedo", albedo),
                                 ("specular_albedo", specular_albedo), ("roughness", roughness),
                                 ("envmap", envmap), ("light2obj_rotmat", light2obj_rotmat)))


def _grad_checks(P, C, dev, grads):
    for name, g in grads:
        _check_float(g, name)
        _check_shape(g, name, (P, C))
        _same_device(dev, ((name, g),))


def _vis_checks(P, dev, face_vertex_list, nearest_triangle_id, barycentric_coord, visibility):
    _check_int(face_vertex_list, "face_vertex_list")
    _check_long(nearest_triangle_id, "nearest_triangle_id")
    _check_float(barycentric_coord, "barycentric_coord")
    _check_int(visibility, "visibility")
    if face_vertex_list.dim() != 2 or face_vertex_list.shape[0] != 3:
        raise RuntimeError(f"face_vertex_list must have shape (3, N_face), got "
                           f"{tuple(face_vertex_list.shape)}")
    _check_shape(nearest_triangle_id, "nearest_triangle_id", (P,))
    _check_shape(barycentric_coord, "barycentric_coord", (3, P))
    if visibility.dim() != 2:
        raise RuntimeError(f"visibility must have shape (N_slot, N_vertex), got "
                           f"{tuple(visibility.shape)}")
    _same_device(dev, (("face_vertex_list", face_vertex_list),
                       ("nearest_triangle_id", nearest_triangle_id),
                       ("barycentric_coord", barycentric_coord), ("visibility", visibility)))
    return (int(face_vertex_list.shape[1]), int(visibility.shape[1]), int(visibility.shape[0]))


# ---- diffuse (bindings.cpp:6-7, bruteforce_shader.cu:2687-2792) -------------------

def bruteforce_diffuse_shader_forward(normal, albedo, envmap, light2obj_rotmat):
    for name, t in (("normal", normal), ("albedo", albedo), ("envmap", envmap),
                    ("light2obj_rotmat", light2obj_rotmat)):
        _check_float(t, name)
    P, C, H, W = _sizes(normal, albedo, envmap)
    _check_shape(light2obj_rotmat, "light2obj_rotmat", (3, 3))
    dev = normal.device
    _same_device(dev, (("albedo", albedo), ("envmap", envmap),
                       ("light2obj_rotmat", light2obj_rotmat)))
    shading = torch.empty((P, C), dtype=normal.dtype, device=dev)
    with torch.cuda.device(dev):
        ws, nws = _fwd_workspace(P, H, W, dev)
        _lib.call("des_bruteforce_diffuse_shader_forward", P, C, H, W, _ptr(normal), _ptr(albedo),
                  _ptr(envmap), _ptr(light2obj_rotmat), _ptr(shading), _ptr(ws), nws,
                  _stream(dev))
        _remember_order(ws, P, dev, normal)
    return shading


def bruteforce_diffuse_shader_backward(grad_shading, normal, albedo, envmap, light2obj_rotmat):
    for name, t in (("normal", normal), ("albedo", albedo), ("envmap", envmap),
                    ("light2obj_rotmat", light2obj_rotmat)):
        _check_float(t, name)
    P, C, H, W = _sizes(normal, albedo, envmap)
    _check_shape(light2obj_rotmat, "light2obj_rotmat", (3, 3))
    dev = normal.device
    _same_device(dev, (("albedo", albedo), ("envmap", envmap),
                       ("light2obj_rotmat", light2obj_rotmat)))
    _grad_checks(P, C, dev, (("grad_shading", grad_shading),))
    grad_normal = torch.empty((P, 3), dtype=normal.dtype, device=dev)
    grad_albedo = torch.empty((P, C), dtype=normal.dtype, device=dev)
    grad_envmap = torch.empty((C, H, W), dtype=normal.dtype, device=dev)
    with torch.cuda.device(dev):
        ws, nws, order = _workspace(P, H, W, dev, normal)
        _lib.call("des_bruteforce_diffuse_shader_backward", P, C, H, W, _ptr(grad_shading),
                  _ptr(normal), _ptr(albedo), _ptr(envmap), _ptr(light2obj_rotmat),
                  _ptr(grad_normal), _ptr(grad_albedo), _ptr(grad_envmap), _ptr(ws), nws,
                  _stream(dev), order=order)
    return [grad_normal, grad_albedo, grad_envmap]


# ---- specular v1 and clamp (bindings.cpp:9-13, bruteforce_shader.cu:2796-3106) -----

def _spec_forward(entry, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                  light2obj_rotmat, enable_diffuse, enable_specular):
    _check_float(normal, "normal")
    _check_float(albedo, "albedo")
    _check_float(envmap, "envmap")
    P, C, H, W = _sizes(normal, albedo, envmap)
    _material_checks(P, C, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                     light2obj_rotmat)
    dev = normal.device
    shading = torch.empty((P, C), dtype=normal.dtype, device=dev)
    with torch.cuda.device(dev):
        ws, nws = _fwd_workspace(P, H, W, dev)
        _lib.call(entry, P, C, H, W, _ptr(normal), _ptr(view_dir), _ptr(albedo),
                  _ptr(specular_albedo), _ptr(roughness), _ptr(envmap), _ptr(light2obj_rotmat),
                  int(bool(enable_diffuse)), int(bool(enable_specular)), _ptr(shading),
                  _ptr(ws), nws, _stream(dev))
        _remember_order(ws, P, dev, normal)
    return shading


def _spec_backward(entry, grad_shading, normal, view_dir, albedo, specular_albedo, roughness,
                   envmap, light2obj_rotmat, enable_diffuse, enable_specular):
    _check_float(normal, "normal")
    _check_float(albedo, "albedo")
    _check_float(envmap, "envmap")
    P, C, H, W = _sizes(normal, albedo, envmap)
    _material_checks(P, C, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                     light2obj_rotmat)
    dev = normal.device
    _grad_checks(P, C, dev, (("grad_shading", grad_shading),))
    outs = _alloc_grads(P, C, H, W, normal)
    with torch.cuda.device(dev):
        ws, nws, order = _workspace(P, H, W, dev, normal)
        _lib.call(entry, P, C, H, W, _ptr(grad_shading), _ptr(normal), _ptr(view_dir),
                  _ptr(albedo), _ptr(specular_albedo), _ptr(roughness), _ptr(envmap),
                  _ptr(light2obj_rotmat), int(bool(enable_diffuse)), int(bool(enable_specular)),
                  *(_ptr(o) for o in outs), _ptr(ws), nws, _stream(dev), order=order)
    return outs


def _alloc_grads(P, C, H, W, normal):
    dev, dt = normal.device, normal.dtype
    return [torch.empty((P, 3), dtype=dt, device=dev),   # grad_normal
            torch.empty((P, 3), dtype=dt, device=dev),   # grad_view_dir
            torch.empty((P, C), dtype=dt, device=dev),   # grad_albedo
            torch.empty((P, C), dtype=dt, device=dev),   # grad_specular_albedo
            torch.empty((P, 1), dtype=dt, device=dev),   # grad_roughness
            torch.empty((C, H, W), dtype=dt, device=dev)]  # grad_envmap


def bruteforce_specular_shader_forward(normal, view_dir, albedo, specular_albedo, roughness,
                                       envmap, light2obj_rotmat, enable_diffuse, enable_specular):
    return _spec_forward("des_bruteforce_specular_shader_forward", normal, view_dir, albedo,
                         specular_albedo, roughness, envmap, light2obj_rotmat, enable_diffuse,
                         enable_specular)


def bruteforce_specular_shader_backward(grad_shading, normal, view_dir, albedo, specular_albedo,
                                        roughness, envmap, light2obj_rotmat, enable_diffuse,
                                        enable_specular):
    return _spec_backward("des_bruteforce_specular_shader_backward", grad_shading, normal,
                          view_dir, albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
                          enable_diffuse, enable_specular)


def bruteforce_specular_shader_clamp_forward(normal, view_dir, albedo, specular_albedo, roughness,
                                             envmap, light2obj_rotmat, enable_diffuse,
                                             enable_specular):
    return _spec_forward("des_bruteforce_specular_shader_clamp_forward", normal, view_dir,
                         albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
                         enable_diffuse, enable_specular)


def bruteforce_specular_shader_clamp_backward(grad_shading, normal, view_dir, albedo,
                                              specular_albedo, roughness, envmap,
                                              light2obj_rotmat, enable_diffuse, enable_specular):
    return _spec_backward("des_bruteforce_specular_shader_clamp_backward", grad_shading, normal,
                          view_dir, albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
                          enable_diffuse, enable_specular)


# ---- v2 (bindings.cpp:15-16, bruteforce_shader.cu:3108-3259) -----------------------

def bruteforce_specular_shader_forward2(normal, view_dir, albedo, specular_albedo, roughness,
                                        envmap, light2obj_rotmat, enable_specular):
    _check_float(normal, "normal")
    _check_float(albedo, "albedo")
    _check_float(envmap, "envmap")
    P, C, H, W = _sizes(normal, albedo, envmap)
    _material_checks(P, C, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                     light2obj_rotmat)
    dev = normal.device
    shading = torch.empty((P, C), dtype=normal.dtype, device=dev)
    diffuse_shading = torch.empty((P, C), dtype=normal.dtype, device=dev)
    with torch.cuda.device(dev):
        ws, nws = _fwd_workspace(P, H, W, dev)
        _lib.call("des_bruteforce_specular_shader_forward2", P, C, H, W, _ptr(normal),
                  _ptr(view_dir), _ptr(albedo), _ptr(specular_albedo), _ptr(roughness),
                  _ptr(envmap), _ptr(light2obj_rotmat), int(bool(enable_specular)),
                  _ptr(shading), _ptr(diffuse_shading), _ptr(ws), nws, _stream(dev))
        _remember_order(ws, P, dev, normal)
    return [shading, diffuse_shading]


def bruteforce_specular_shader_backward2(grad_shading, grad_diffuse_shading, normal, view_dir,
                                         albedo, specular_albedo, roughness, envmap,
                                         light2obj_rotmat, enable_specular):
    _check_float(normal, "normal")
    _check_float(albedo, "albedo")
    _check_float(envmap, "envmap")
    P, C, H, W = _sizes(normal, albedo, envmap)
    _material_checks(P, C, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                     light2obj_rotmat)
    dev = normal.device
    _grad_checks(P, C, dev, (("grad_shading", grad_shading),
                             ("grad_diffuse_shading", grad_diffuse_shading)))
    outs = _alloc_grads(P, C, H, W, normal)
    with torch.cuda.device(dev):
        ws, nws, order = _workspace(P, H, W, dev, normal)
        _lib.call("des_bruteforce_specular_shader_backward2", P, C, H, W, _ptr(grad_shading),
                  _ptr(grad_diffuse_shading), _ptr(normal), _ptr(view_dir), _ptr(albedo),
                  _ptr(specular_albedo), _ptr(roughness), _ptr(envmap), _ptr(light2obj_rotmat),
                  int(bool(enable_specular)), *(_ptr(o) for o in outs), _ptr(ws), nws,
                  _stream(dev), order=order)
    return outs


# ---- v2 + visibility (bindings.cpp:18-19, bruteforce_shader.cu:3261-3453) -----------

def bruteforce_specularvisibility_shader_forward2(normal, view_dir, albedo, specular_albedo,
                                                  roughness, envmap, light2obj_rotmat,
                                                  face_vertex_list, nearest_triangle_id,
                                                  barycentric_coord, visibility,
                                                  enable_specular):
    _check_float(normal, "normal")
    _check_float(albedo, "albedo")
    _check_float(envmap, "envmap")
    P, C, H, W = _sizes(normal, albedo, envmap)
    _material_checks(P, C, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                     light2obj_rotmat)
    dev = normal.device
    NF, NV, NS = _vis_checks(P, dev, face_vertex_list, nearest_triangle_id, barycentric_coord,
                             visibility)
    shading = torch.empty((P, C), dtype=normal.dtype, device=dev)
    diffuse_shading = torch.empty((P, C), dtype=normal.dtype, device=dev)
    with torch.cuda.device(dev):
        ws, nws = _fwd_workspace(P, H, W, dev)
        _lib.call("des_bruteforce_specularvisibility_shader_forward2", P, C, H, W, NF, NV, NS,
                  _ptr(normal), _ptr(view_dir), _ptr(albedo), _ptr(specular_albedo),
                  _ptr(roughness), _ptr(envmap), _ptr(light2obj_rotmat), _ptr(face_vertex_list),
                  _ptr(nearest_triangle_id), _ptr(barycentric_coord), _ptr(visibility),
                  int(bool(enable_specular)), _ptr(shading), _ptr(diffuse_shading), _ptr(ws),
                  nws, _stream(dev))
        _remember_order(ws, P, dev, normal)
    return [shading, diffuse_shading]


def bruteforce_specularvisibility_shader_backward2(grad_shading, grad_diffuse_shading, normal,
                                                   view_dir, albedo, specular_albedo, roughness,
                                                   envmap, light2obj_rotmat, face_vertex_list,
                                                   nearest_triangle_id, barycentric_coord,
                                                   visibility, enable_specular):
    _check_float(normal, "normal")
    _check_float(albedo, "albedo")
    _check_float(envmap, "envmap")
    P, C, H, W = _sizes(normal, albedo, envmap)
    _material_checks(P, C, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                     light2obj_rotmat)
    dev = normal.device
    _grad_checks(P, C, dev, (("grad_shading", grad_shading),
                             ("grad_diffuse_shading", grad_diffuse_shading)))
    NF, NV, NS = _vis_checks(P, dev, face_vertex_list, nearest_triangle_id, barycentric_coord,
                             visibility)
    outs = _alloc_grads(P, C, H, W, normal)
    with torch.cuda.device(dev):
        ws, nws, order = _workspace(P, H, W, dev, normal)
        _lib.call("des_bruteforce_specularvisibility_shader_backward2", P, C, H, W, NF, NV, NS,
                  _ptr(grad_shading), _ptr(grad_diffuse_shading), _ptr(normal), _ptr(view_dir),
                  _ptr(albedo), _ptr(specular_albedo), _ptr(roughness), _ptr(envmap),
                  _ptr(light2obj_rotmat), _ptr(face_vertex_list), _ptr(nearest_triangle_id),
                  _ptr(barycentric_coord), _ptr(visibility), int(bool(enable_specular)),
                  *(_ptr(o) for o in outs), _ptr(ws), nws, _stream(dev), order=order)
    return outs


# ---- mesh utilities (bindings.cpp:27-33, etc.cu) -------------------------------------

def _mesh_device(query_pos, name="query_pos"):
    _check_float(query_pos, name)
    return query_pos.device


def get_nearest_mesh_points_forward(adjacency_head, adjacency_list, face_vertex_list, vertex_pos,
                                    query_pos, idxs):
    """Nearest point on the faces around each query point's K nearest vertices
    (etc.cu:1277-1326): returns [nearest_triangle_id (P int64),
    barycentric_coord (3 x P)].  adjacency_head 2 x N_vertex (face counts,
    then their inclusive prefix sum), adjacency_list int32, face_vertex_list
    3 x N_face, vertex_pos 3 x N_vertex, query_pos 3 x P, idxs P x K int64."""
    dev = _mesh_device(query_pos)
    for t, n in ((adjacency_head, "adjacency_head"), (adjacency_list, "adjacency_list"),
                 (face_vertex_list, "face_vertex_list")):
        _check_int(t, n)
    _check_float(vertex_pos, "vertex_pos")
    _check_long(idxs, "idxs")
    if adjacency_head.dim() != 2 or adjacency_head.shape[0] != 2:
        raise RuntimeError(f"adjacency_head must have shape (2, N_vertex), got {tuple(adjacency_head.shape)}")
    NV = int(adjacency_head.shape[1])
    if face_vertex_list.dim() != 2 or face_vertex_list.shape[0] != 3:
        raise RuntimeError(f"face_vertex_list must have shape (3, N_face), got {tuple(face_vertex_list.shape)}")
    NF = int(face_vertex_list.shape[1])
    _check_shape(vertex_pos, "vertex_pos", (3, NV))
    if query_pos.dim() != 2 or query_pos.shape[0] != 3:
        raise RuntimeError(f"query_pos must have shape (3, P), got {tuple(query_pos.shape)}")
    P = int(query_pos.shape[1])
    if idxs.dim() != 2 or idxs.shape[0] != P:
        raise RuntimeError(f"idxs must have shape (P, K), got {tuple(idxs.shape)}")
    K = int(idxs.shape[1])
    _same_device(dev, [("adjacency_head", adjacency_head), ("adjacency_list", adjacency_list),
                       ("face_vertex_list", face_vertex_list), ("vertex_pos", vertex_pos),
                       ("idxs", idxs)])
    with torch.cuda.device(dev):
        tri = torch.empty(P, dtype=torch.int64, device=dev)
        bary = torch.empty(3, P, dtype=torch.float32, device=dev)
        _lib.call("des_get_nearest_mesh_points_forward", NV, NF, P, K, _ptr(adjacency_head),
                  _ptr(adjacency_list), int(adjacency_list.numel()), _ptr(face_vertex_list),
                  _ptr(vertex_pos), _ptr(query_pos), _ptr(idxs), _ptr(tri), _ptr(bary),
                  _stream(dev))
    return [tri, bary]


def get_nearest_mesh_points_backward(face_vertex_list, vertex_pos, query_pos, nearest_triangle_id,
                                     grad_barycentric_coord):
    """grad_query_pos (3 x P) of get_nearest_mesh_points_forward's barycentric
    output (etc.cu:1328-1365), from grad_barycentric_coord (3 x P)."""
    dev = _mesh_device(query_pos)
    _check_int(face_vertex_list, "face_vertex_list")
    _check_float(vertex_pos, "vertex_pos")
    _check_long(nearest_triangle_id, "nearest_triangle_id")
    _check_float(grad_barycentric_coord, "grad_barycentric_coord")
    if face_vertex_list.dim() != 2 or face_vertex_list.shape[0] != 3:
        raise RuntimeError(f"face_vertex_list must have shape (3, N_face), got {tuple(face_vertex_list.shape)}")
    if vertex_pos.dim() != 2 or vertex_pos.shape[0] != 3:
        raise RuntimeError(f"vertex_pos must have shape (3, N_vertex), got {tuple(vertex_pos.shape)}")
    if query_pos.dim() != 2 or query_pos.shape[0] != 3:
        raise RuntimeError(f"query_pos must have shape (3, P), got {tuple(query_pos.shape)}")
    NF, NV, P = int(face_vertex_list.shape[1]), int(vertex_pos.shape[1]), int(query_pos.shape[1])
    _check_shape(nearest_triangle_id, "nearest_triangle_id", (P,))
    _check_shape(grad_barycentric_coord, "grad_barycentric_coord", (3, P))
    _same_device(dev, [("face_vertex_list", face_vertex_list), ("vertex_pos", vertex_pos),
                       ("nearest_triangle_id", nearest_triangle_id),
                       ("grad_barycentric_coord", grad_barycentric_coord)])
    with torch.cuda.device(dev):
        g = torch.empty(3, P, dtype=torch.float32, device=dev)
        _lib.call("des_get_nearest_mesh_points_backward", NV, NF, P, _ptr(face_vertex_list),
                  _ptr(vertex_pos), _ptr(query_pos), _ptr(nearest_triangle_id),
                  _ptr(grad_barycentric_coord), _ptr(g), _stream(dev))
    return g


def _uv_checks(triangles, query_pos):
    dev = _mesh_device(query_pos)
    _check_float(triangles, "triangles")
    if query_pos.dim() != 2 or query_pos.shape[1] != 3:
        raise RuntimeError(f"query_pos must have shape (P, 3), got {tuple(query_pos.shape)}")
    P = int(query_pos.shape[0])
    _check_shape(triangles, "triangles", (P, 3, 3))
    _same_device(dev, [("triangles", triangles)])
    return dev, P


def get_uv_of_triangle_forward(triangles, query_pos):
    """Barycentric coordinates (P x 3: 1 - s - t, s, t) of the closest point of
    triangle p (triangles P x 3 x 3) to query point p (P x 3), etc.cu:1218-1244."""
    dev, P = _uv_checks(triangles, query_pos)
    with torch.cuda.device(dev):
        bary = torch.empty(P, 3, dtype=torch.float32, device=dev)
        _lib.call("des_get_uv_of_triangle_forward", P, _ptr(triangles), _ptr(query_pos),
                  _ptr(bary), _stream(dev))
    return bary


def get_uv_of_triangle_backward(triangles, query_pos, grad_barycentric_coord):
    """grad_query_pos (P x 3) of get_uv_of_triangle_forward, etc.cu:1246-1275."""
    dev, P = _uv_checks(triangles, query_pos)
    _check_float(grad_barycentric_coord, "grad_barycentric_coord")
    _check_shape(grad_barycentric_coord, "grad_barycentric_coord", (P, 3))
    _same_device(dev, [("grad_barycentric_coord", grad_barycentric_coord)])
    with torch.cuda.device(dev):
        g = torch.empty(P, 3, dtype=torch.float32, device=dev)
        _lib.call("des_get_uv_of_triangle_backward", P, _ptr(triangles), _ptr(query_pos),
                  _ptr(grad_barycentric_coord), _ptr(g), _stream(dev))
    return g


def extract_bitfield_forward(bitfield, N_bit, i_start, i_end):
    """(i_end - i_start) x N_bit floats, [b][k] = bit k % 32 of word
    [k / 32][i_start + b] of bitfield (N_slot x N int32), etc.cu:1044-1071."""
    _check_int(bitfield, "bitfield")
    if bitfield.dim() != 2:
        raise RuntimeError(f"bitfield must have shape (N_slot, N), got {tuple(bitfield.shape)}")
    dev = bitfield.device
    N_slot, N = int(bitfield.shape[0]), int(bitfield.shape[1])
    N_bit, i_start, i_end = int(N_bit), int(i_start), int(i_end)
    valid = 0 <= i_start <= i_end <= N and 0 <= N_bit <= 32 * N_slot
    with torch.cuda.device(dev):
        # on invalid ranges the library reports the error (no output is needed)
        out = torch.empty(i_end - i_start, N_bit, dtype=torch.float32, device=dev) if valid else None
        _lib.call("des_extract_bitfield_forward", _ptr(bitfield), N_slot, N, N_bit, i_start, i_end,
                  _ptr(out), _stream(dev))
    return out


# ---- env-map preprocessing (bindings.cpp:21, 23, etc.cu) --------------------------------

def _envmap_shape(envmap):
    _check_float(envmap, "envmap")
    if envmap.dim() != 3:
        raise RuntimeError(f"envmap must have shape (C, H, W), got {tuple(envmap.shape)}")
    return (int(x) for x in envmap.shape)


def sample_envmap_forward(query_point, envmap):
    """Bilinear lookup of envmap (C x H x W) along each direction of
    query_point (3 x P, normalised here), etc.cu:1145-1178: returns C x P (the
    reference allocates 3 x P whatever C is)."""
    C, H, W = _envmap_shape(envmap)
    _check_float(query_point, "query_point")
    if query_point.dim() != 2 or query_point.shape[0] != 3:
        raise RuntimeError(f"query_point must have shape (3, P), got {tuple(query_point.shape)}")
    dev = envmap.device
    _same_device(dev, [("query_point", query_point)])
    P = int(query_point.shape[1])
    with torch.cuda.device(dev):
        out = torch.empty(C, P, dtype=torch.float32, device=dev)
        _lib.call("des_sample_envmap_forward", P, C, H, W, _ptr(query_point), _ptr(envmap),
                  _ptr(out), _stream(dev))
    return out


def blur_envmap_forward(envmap, tar_H, tar_W, N_samples, alpha):
    """GGX blur of envmap (C x H x W, C = 3 or 4) onto a tar_H x tar_W map with
    N_samples directions per texel, etc.cu:1103-1143: returns C x tar_H x tar_W."""
    C, H, W = _envmap_shape(envmap)
    dev = envmap.device
    tar_H, tar_W, N_samples = int(tar_H), int(tar_W), int(N_samples)
    with torch.cuda.device(dev):
        out = torch.empty(C, max(tar_H, 0), max(tar_W, 0), dtype=torch.float32, device=dev)
        _lib.call("des_blur_envmap_forward", C, H, W, tar_H, tar_W, N_samples, float(alpha),
                  _ptr(envmap), _ptr(out), _stream(dev))
    return out


# ---- reference bindings outside this library (bindings.cpp:25, 32, etc.cu) -------------

def _not_here(name):
    def fn(*args, **kwargs):
        raise NotImplementedError(
            f"{name} (etc.cu of the reference) is not part of the MI355X shader library")
    fn.__name__ = name
    return fn


generate_camera_ray_forward = _not_here("generate_camera_ray_forward")
mt_format_conversion_forward = _not_here("mt_format_conversion_forward")
