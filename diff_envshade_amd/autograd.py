"""torch.autograd.Function wrappers over the drop-in shader bindings.

The reference ships only the raw forward/backward bindings; its callers wrap
them in their own autograd functions.  These classes are that wrapper, with the
contract fixed by the binding signatures and return orders
(gsface_shader.h:3-118, bruteforce_shader.cu:2791, 3164, 3452 of
zjumsj/diff_envshade):

  * gradients flow to normal, view_dir, albedo, specular_albedo, roughness and
    envmap; light2obj_rotmat and every integer / flag input get None (the
    reference computes and discards the light-direction gradient,
    bruteforce_shader.cu:2457);
  * a None upstream gradient is treated as zeros, and upstream gradients are
    made contiguous before the backward binding is called.

NearestMeshPoints and UVofTriangle wrap the mesh utilities the same way: their
barycentric outputs are differentiable in the query points.
"""
from __future__ import annotations

import torch

from . import _diff_envshade as _C


def _g(g, like):
    if g is None:
        return torch.zeros_like(like)
    return g.contiguous()


def _match(grad, inp):
    # roughness may be (P,) or (P, 1); the binding returns (P, 1)
    return grad.reshape(inp.shape) if grad is not None and grad.shape != inp.shape else grad


class DiffuseShader(torch.autograd.Function):
    """Lambert-only shading (bindings.cpp:6-7)."""

    @staticmethod
    def forward(ctx, normal, albedo, envmap, light2obj_rotmat):
        shading = _C.bruteforce_diffuse_shader_forward(normal, albedo, envmap, light2obj_rotmat)
        ctx.save_for_backward(normal, albedo, envmap, light2obj_rotmat)
        return shading

    @staticmethod
    def backward(ctx, grad_shading):
        normal, albedo, envmap, rot = ctx.saved_tensors
        gn, ga, gE = _C.bruteforce_diffuse_shader_backward(_g(grad_shading, normal), normal,
                                                           albedo, envmap, rot)
        return gn, ga, gE, None


class _SpecularBase(torch.autograd.Function):
    _fwd = None
    _bwd = None

    @classmethod
    def forward(cls, ctx, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                light2obj_rotmat, enable_diffuse=True, enable_specular=True):
        shading = cls._fwd(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                           light2obj_rotmat, bool(enable_diffuse), bool(enable_specular))
        ctx.save_for_backward(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                              light2obj_rotmat)
        ctx.flags = (bool(enable_diffuse), bool(enable_specular))
        return shading

    @classmethod
    def backward(cls, ctx, grad_shading):
        normal, view_dir, albedo, spec, rough, envmap, rot = ctx.saved_tensors
        gn, gv, ga, gs, gr, gE = cls._bwd(_g(grad_shading, normal), normal, view_dir, albedo,
                                          spec, rough, envmap, rot, *ctx.flags)
        return gn, gv, ga, gs, _match(gr, rough), gE, None, None, None


class SpecularShader(_SpecularBase):
    """Lambert and/or GGX (bindings.cpp:9-10)."""
    _fwd = staticmethod(_C.bruteforce_specular_shader_forward)
    _bwd = staticmethod(_C.bruteforce_specular_shader_backward)


class SpecularShaderClamp(_SpecularBase):
    """As SpecularShader with per-point back-face culling n.v > 0 (bindings.cpp:12-13)."""
    _fwd = staticmethod(_C.bruteforce_specular_shader_clamp_forward)
    _bwd = staticmethod(_C.bruteforce_specular_shader_clamp_backward)


class SpecularShader2(torch.autograd.Function):
    """Lambert + GGX with the extra white-albedo diffuse_shading output (bindings.cpp:15-16)."""

    @staticmethod
    def forward(ctx, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                light2obj_rotmat, enable_specular=True):
        shading, diffuse_shading = _C.bruteforce_specular_shader_forward2(
            normal, view_dir, albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
            bool(enable_specular))
        ctx.save_for_backward(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                              light2obj_rotmat)
        ctx.enable_specular = bool(enable_specular)
        return shading, diffuse_shading

    @staticmethod
    def backward(ctx, grad_shading, grad_diffuse_shading):
        normal, view_dir, albedo, spec, rough, envmap, rot = ctx.saved_tensors
        gn, gv, ga, gs, gr, gE = _C.bruteforce_specular_shader_backward2(
            _g(grad_shading, normal), _g(grad_diffuse_shading, normal), normal, view_dir, albedo,
            spec, rough, envmap, rot, ctx.enable_specular)
        return gn, gv, ga, gs, _match(gr, rough), gE, None, None


class SpecularVisibilityShader2(torch.autograd.Function):
    """SpecularShader2 times barycentric-interpolated per-vertex visibility
    bits (bindings.cpp:18-19)."""

    @staticmethod
    def forward(ctx, normal, view_dir, albedo, specular_albedo, roughness, envmap,
                light2obj_rotmat, face_vertex_list, nearest_triangle_id, barycentric_coord,
                visibility, enable_specular=True):
        shading, diffuse_shading = _C.bruteforce_specularvisibility_shader_forward2(
            normal, view_dir, albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
            face_vertex_list, nearest_triangle_id, barycentric_coord, visibility,
            bool(enable_specular))
        ctx.save_for_backward(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                              light2obj_rotmat, face_vertex_list, nearest_triangle_id,
                              barycentric_coord, visibility)
        ctx.enable_specular = bool(enable_specular)
        return shading, diffuse_shading

    @staticmethod
    def backward(ctx, grad_shading, grad_diffuse_shading):
        (normal, view_dir, albedo, spec, rough, envmap, rot, fvl, tri, bary,
         vis) = ctx.saved_tensors
        gn, gv, ga, gs, gr, gE = _C.bruteforce_specularvisibility_shader_backward2(
            _g(grad_shading, normal), _g(grad_diffuse_shading, normal), normal, view_dir, albedo,
            spec, rough, envmap, rot, fvl, tri, bary, vis, ctx.enable_specular)
        return (gn, gv, ga, gs, _match(gr, rough), gE, None, None, None, None, None, None)


def diffuse_shade(normal, albedo, envmap, light2obj_rotmat):
    return DiffuseShader.apply(normal, albedo, envmap, light2obj_rotmat)


def specular_shade(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                   light2obj_rotmat, enable_diffuse=True, enable_specular=True):
    return SpecularShader.apply(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                                light2obj_rotmat, enable_diffuse, enable_specular)


def specular_shade_clamp(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                         light2obj_rotmat, enable_diffuse=True, enable_specular=True):
    return SpecularShaderClamp.apply(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                                     light2obj_rotmat, enable_diffuse, enable_specular)


def specular_shade2(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                    light2obj_rotmat, enable_specular=True):
    return SpecularShader2.apply(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                                 light2obj_rotmat, enable_specular)


def specular_visibility_shade2(normal, view_dir, albedo, specular_albedo, roughness, envmap,
                               light2obj_rotmat, face_vertex_list, nearest_triangle_id,
                               barycentric_coord, visibility, enable_specular=True):
    return SpecularVisibilityShader2.apply(normal, view_dir, albedo, specular_albedo, roughness,
                                           envmap, light2obj_rotmat, face_vertex_list,
                                           nearest_triangle_id, barycentric_coord, visibility,
                                           enable_specular)


# ---- mesh utilities (etc.cu:707-920): gradients flow to the query points only ----

class NearestMeshPoints(torch.autograd.Function):
    """(nearest_triangle_id, barycentric_coord) of each query point
    (get_nearest_mesh_points_forward); the barycentrics are differentiable in
    query_pos (3 x P) through get_nearest_mesh_points_backward, with the chosen
    face held fixed.  The triangle ids are integers and carry no gradient."""

    @staticmethod
    def forward(ctx, adjacency_head, adjacency_list, face_vertex_list, vertex_pos, query_pos,
                idxs):
        tri, bary = _C.get_nearest_mesh_points_forward(adjacency_head, adjacency_list,
                                                       face_vertex_list, vertex_pos, query_pos,
                                                       idxs)
        ctx.save_for_backward(face_vertex_list, vertex_pos, query_pos, tri)
        ctx.mark_non_differentiable(tri)
        return tri, bary

    @staticmethod
    def backward(ctx, grad_tri, grad_bary):
        fvl, vpos, query, tri = ctx.saved_tensors
        gq = _C.get_nearest_mesh_points_backward(fvl, vpos, query, tri, _g(grad_bary, query))
        return None, None, None, None, gq, None


class UVofTriangle(torch.autograd.Function):
    """Barycentric coordinates (P x 3) of the closest point of each triangle
    (P x 3 x 3) to its query point (P x 3), get_uv_of_triangle_forward;
    differentiable in query_pos through get_uv_of_triangle_backward."""

    @staticmethod
    def forward(ctx, triangles, query_pos):
        bary = _C.get_uv_of_triangle_forward(triangles, query_pos)
        ctx.save_for_backward(triangles, query_pos)
        return bary

    @staticmethod
    def backward(ctx, grad_bary):
        triangles, query = ctx.saved_tensors
        return None, _C.get_uv_of_triangle_backward(triangles, query, _g(grad_bary, query))


def nearest_mesh_points(adjacency_head, adjacency_list, face_vertex_list, vertex_pos, query_pos,
                        idxs):
    return NearestMeshPoints.apply(adjacency_head, adjacency_list, face_vertex_list, vertex_pos,
                                   query_pos, idxs)


def uv_of_triangle(triangles, query_pos):
    return UVofTriangle.apply(triangles, query_pos)
