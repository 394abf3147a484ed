"""The torch.autograd.Function wrappers route to the drop-in bindings (GPU)."""
import numpy as np
import pytest
import torch

from _helpers import TOL, gpu_inputs, oracle_run, problem, relerr

pytestmark = pytest.mark.gpu


def _leaf(t):
    return t.clone().requires_grad_(True)


def test_visibility_autograd_matches_oracle():
    from diff_envshade_amd.autograd import specular_visibility_shade2
    d = problem(seed=11, n=400, H=8, W=16, vis="identity")
    t = gpu_inputs(d)
    leaves = [_leaf(t[k]) for k in ("normal", "view_dir", "albedo", "specular_albedo",
                                    "roughness", "envmap")]
    sh, ds = specular_visibility_shade2(*leaves, t["light2obj_rotmat"], t["face_vertex_list"],
                                        t["nearest_triangle_id"], t["barycentric_coord"],
                                        t["visibility"], True)
    loss = (sh * t["grad_shading"]).sum() + (ds * t["grad_diffuse_shading"]).sum()
    loss.backward()
    (rf, rb) = oracle_run("vis", d)
    assert relerr(sh.detach().cpu().numpy(), rf[0]) <= TOL
    assert relerr(ds.detach().cpu().numpy(), rf[1]) <= TOL
    for leaf, r in zip(leaves, rb):
        assert relerr(leaf.grad.cpu().numpy().reshape(r.shape), r) <= TOL


def test_none_upstream_grad_is_zero():
    from diff_envshade_amd.autograd import specular_shade2
    d = problem(seed=12, n=100, H=4, W=8)
    t = gpu_inputs(d)
    n = _leaf(t["normal"])
    sh, ds = specular_shade2(n, t["view_dir"], t["albedo"], t["specular_albedo"], t["roughness"],
                             t["envmap"], t["light2obj_rotmat"], True)
    # only shading feeds the loss: grad_diffuse_shading arrives as None
    (sh * t["grad_shading"]).sum().backward()
    d2 = dict(d)
    d2["grad_diffuse_shading"] = np.zeros_like(d["grad_diffuse_shading"])
    _, rb = oracle_run("v2", d2)
    assert relerr(n.grad.cpu().numpy(), rb[0]) <= TOL


@pytest.mark.parametrize("fn,variant", [("specular_shade", "specular"),
                                        ("specular_shade_clamp", "clamp")])
def test_v1_autograd(fn, variant):
    import diff_envshade_amd.autograd as A
    d = problem(seed=13, n=200, H=4, W=8)
    t = gpu_inputs(d)
    leaves = [_leaf(t[k]) for k in ("normal", "view_dir", "albedo", "specular_albedo",
                                    "roughness", "envmap")]
    sh = getattr(A, fn)(*leaves, t["light2obj_rotmat"], True, True)
    (sh * t["grad_shading"]).sum().backward()
    _, rb = oracle_run(variant, d)
    for leaf, r in zip(leaves, rb):
        assert relerr(leaf.grad.cpu().numpy().reshape(r.shape), r) <= TOL


def test_diffuse_autograd():
    from diff_envshade_amd.autograd import diffuse_shade
    d = problem(seed=14, n=200, H=4, W=8)
    t = gpu_inputs(d)
    leaves = [_leaf(t[k]) for k in ("normal", "albedo", "envmap")]
    sh = diffuse_shade(*leaves, t["light2obj_rotmat"])
    (sh * t["grad_shading"]).sum().backward()
    _, rb = oracle_run("diffuse", d)
    for leaf, r in zip(leaves, rb):
        assert relerr(leaf.grad.cpu().numpy(), r) <= TOL


def test_flat_roughness_shape():
    from diff_envshade_amd.autograd import specular_shade2
    d = problem(seed=15, n=64, H=4, W=8)
    t = gpu_inputs(d)
    r = _leaf(t["roughness"].reshape(-1))
    sh, ds = specular_shade2(t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"], r,
                             t["envmap"], t["light2obj_rotmat"], True)
    sh.sum().backward()
    assert r.grad.shape == r.shape
    assert torch.isfinite(r.grad).all()
