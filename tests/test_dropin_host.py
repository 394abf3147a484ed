"""Host-side behaviour of the drop-in module that needs no GPU: the reference's
TORCH_CHECK messages for non-CUDA tensors, the module surface, and the
non-shader bindings that are outside this library."""
import inspect

import numpy as np
import pytest
import torch

from diff_envshade_amd import _diff_envshade as C
from diff_envshade_amd import autograd as A


def _cpu(n=4):
    r = np.random.default_rng(0)
    f = lambda *s: torch.from_numpy(r.random(s).astype(np.float32))  # noqa: E731
    return f(n, 3), f(n, 3), f(n, 3), f(n, 3), f(n, 1), f(3, 2, 4), torch.eye(3)


def test_module_has_reference_names_and_argument_orders():
    # bindings.cpp:6-19 / gsface_shader.h:3-118
    expect = {
        "bruteforce_diffuse_shader_forward": ["normal", "albedo", "envmap", "light2obj_rotmat"],
        "bruteforce_diffuse_shader_backward": ["grad_shading", "normal", "albedo", "envmap",
                                               "light2obj_rotmat"],
        "bruteforce_specular_shader_forward": ["normal", "view_dir", "albedo", "specular_albedo",
                                               "roughness", "envmap", "light2obj_rotmat",
                                               "enable_diffuse", "enable_specular"],
        "bruteforce_specular_shader_backward": ["grad_shading", "normal", "view_dir", "albedo",
                                                "specular_albedo", "roughness", "envmap",
                                                "light2obj_rotmat", "enable_diffuse",
                                                "enable_specular"],
        "bruteforce_specular_shader_forward2": ["normal", "view_dir", "albedo", "specular_albedo",
                                                "roughness", "envmap", "light2obj_rotmat",
                                                "enable_specular"],
        "bruteforce_specular_shader_backward2": ["grad_shading", "grad_diffuse_shading", "normal",
                                                 "view_dir", "albedo", "specular_albedo",
                                                 "roughness", "envmap", "light2obj_rotmat",
                                                 "enable_specular"],
        "bruteforce_specularvisibility_shader_forward2": [
            "normal", "view_dir", "albedo", "specular_albedo", "roughness", "envmap",
            "light2obj_rotmat", "face_vertex_list", "nearest_triangle_id", "barycentric_coord",
            "visibility", "enable_specular"],
        "bruteforce_specularvisibility_shader_backward2": [
            "grad_shading", "grad_diffuse_shading", "normal", "view_dir", "albedo",
            "specular_albedo", "roughness", "envmap", "light2obj_rotmat", "face_vertex_list",
            "nearest_triangle_id", "barycentric_coord", "visibility", "enable_specular"],
    }
    for name, args in expect.items():
        assert list(inspect.signature(getattr(C, name)).parameters) == args, name
    clamp = inspect.signature(C.bruteforce_specular_shader_clamp_backward).parameters
    assert list(clamp) == expect["bruteforce_specular_shader_backward"]


def test_cpu_tensors_are_rejected_like_torch_check():
    n, v, a, s, r, e, m = _cpu()
    with pytest.raises(RuntimeError, match="normal must be a CUDA tensor"):
        C.bruteforce_specular_shader_forward2(n, v, a, s, r, e, m, True)
    with pytest.raises(RuntimeError, match="normal must be a CUDA tensor"):
        C.bruteforce_diffuse_shader_forward(n, a, e, m)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        C.bruteforce_specularvisibility_shader_backward2(n, n, n, v, a, s, r, e, m, None, None,
                                                         None, None, True)


@pytest.mark.parametrize("name", ["generate_camera_ray_forward", "mt_format_conversion_forward"])
def test_non_shader_bindings_are_outside_the_library(name):
    with pytest.raises(NotImplementedError):
        getattr(C, name)()


def test_mesh_bindings_check_their_tensors():
    # checked like the reference's CHECK_CUDA / CHECK_IS_* before any device work
    q = torch.zeros(3, 4)
    with pytest.raises(RuntimeError, match="CUDA"):
        C.get_nearest_mesh_points_forward(None, None, None, None, q, None)
    with pytest.raises(RuntimeError, match="CUDA"):
        C.get_nearest_mesh_points_backward(None, None, q, None, None)
    with pytest.raises(RuntimeError, match="CUDA"):
        C.get_uv_of_triangle_forward(torch.zeros(4, 3, 3), torch.zeros(4, 3))
    with pytest.raises(RuntimeError, match="CUDA"):
        C.get_uv_of_triangle_backward(torch.zeros(4, 3, 3), torch.zeros(4, 3), torch.zeros(4, 3))
    with pytest.raises(RuntimeError, match="tensor"):
        C.extract_bitfield_forward(None, 8, 0, 1)
    with pytest.raises(RuntimeError, match="envmap must be a CUDA tensor"):
        C.sample_envmap_forward(torch.zeros(3, 4), torch.zeros(3, 2, 4))
    with pytest.raises(RuntimeError, match="envmap must be a CUDA tensor"):
        C.blur_envmap_forward(torch.zeros(3, 2, 4), 2, 4, 8, 0.1)


def test_autograd_classes_exist():
    for cls in ("DiffuseShader", "SpecularShader", "SpecularShaderClamp", "SpecularShader2",
                "SpecularVisibilityShader2"):
        assert issubclass(getattr(A, cls), torch.autograd.Function)
