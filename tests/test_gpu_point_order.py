"""Point order by normal direction (GPU): sorting the points into
normal-coherent tiles, skipping texels that light no lane of a tile and
dealing the lit texel groups evenly to the tile's waves may change the
outputs only at fp32 rounding level (every sum is regrouped), and the
results stay bitwise reproducible.  Oracle parity of the sorted path is
covered by the parity suite, which runs every case with the order forced
on."""
import numpy as np
import pytest

from _helpers import TOL, compare, gpu_run, oracle_run, problem, relerr

pytestmark = pytest.mark.gpu


def _run(mode, variant, d, diff=True, spec=True):
    from diff_envshade_amd import _lib
    try:
        _lib.set_point_order(mode)
        return gpu_run(variant, d, diff, spec)
    finally:
        _lib.set_point_order(0)


@pytest.mark.parametrize("variant,diff,spec", [("vis", True, True), ("v2", True, True),
                                               ("clamp", True, True), ("specular", False, True),
                                               ("diffuse", True, False)])
def test_point_order_changes_outputs_only_by_rounding(variant, diff, spec):
    # every sum is regrouped; sums that cancel (grad_roughness) move by up to
    # ~2e-6 normwise, so the two orders are held to the 1e-5 accuracy bar they
    # each meet against FP64 in the parity suite
    d = problem(seed=21, n=20000, H=16, W=32, vis="identity" if variant == "vis" else "none")
    (f0, b0), (f1, b1) = _run(-1, variant, d, diff, spec), _run(1, variant, d, diff, spec)
    for a, b in zip(f0 + b0, f1 + b1):
        assert relerr(b, a) <= TOL


def test_sorted_outputs_are_bitwise_reproducible():
    d = problem(seed=22, n=30000, H=16, W=32, vis="identity")
    f0, b0 = _run(1, "vis", d)
    f1, b1 = _run(1, "vis", d)
    for a, b in zip(f0 + b0, f1 + b1):
        assert np.array_equal(a, b)


def test_coherent_normals_skip_half_the_map():
    # every normal within ~3 degrees of +y: the lower hemisphere of the map is
    # dark for whole tiles, so whole groups of texels take the skip path
    rng = np.random.default_rng(23)
    d = problem(seed=23, n=5000, H=16, W=32, rotation=False, vis="identity")
    n = np.array([0.0, 1.0, 0.0]) + 0.05 * rng.normal(size=(5000, 3))
    d["normal"] = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    from diff_envshade_amd.workloads import make_visibility
    d["visibility"] = make_visibility(rng, d["normal"], 16, 32)
    ref = oracle_run("vis", d)
    compare("vis", _run(1, "vis", d), ref)
    compare("vis", _run(-1, "vis", d), ref)


def test_degenerate_normals_are_ordered():
    # zero and non-finite-free degenerate normals get key 0; every point is
    # still shaded exactly once
    d = problem(seed=24, n=3000, H=8, W=16, vis="none")
    d["normal"][::7] = 0.0
    d["normal"][1::11] = np.array([0.0, 0.0, -1.0], dtype=np.float32)
    compare("v2", _run(1, "v2", d), oracle_run("v2", d))


@pytest.mark.parametrize("n", [4096 * 3, 4096 * 64 + 4000])
def test_order_block_boundaries(n):
    # 12288 points: three full 4096-point sort blocks (fused scan); 266144
    # points: 65 blocks, past the fused limit (separate per-key scan)
    d = problem(seed=25, n=n, H=8, W=16, vis="identity")
    (f0, b0), (f1, b1) = _run(-1, "vis", d), _run(1, "vis", d)
    for a, b in zip(f0 + b0, f1 + b1):
        assert relerr(b, a) <= TOL


def _vis_args(t):
    mat = [t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"], t["roughness"],
           t["envmap"], t["light2obj_rotmat"]]
    vis = [t["face_vertex_list"], t["nearest_triangle_id"], t["barycentric_coord"],
           t["visibility"]]
    return mat, vis


def _device(d):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0") for k, v in d.items()}


@pytest.fixture
def reuse_spy(monkeypatch):
    """Counts the order hand-overs the drop-in module makes."""
    from diff_envshade_amd import _diff_envshade as de
    from diff_envshade_amd import _lib
    calls = []
    real = _lib.call

    def spy(name, *args, order=None):
        if order is not None:
            calls.append(order)
        real(name, *args, order=order)
    monkeypatch.setattr(_lib, "call", spy)
    de._order_cache.clear()
    yield calls
    de._order_cache.clear()


def test_forward_order_serves_the_backward(reuse_spy):
    # 40,000 points x 512 texels (2^24 pairs and more): both directions sort;
    # a backward right after the forward takes the forward's order, and its
    # outputs are bitwise those of a backward that sorts by itself
    import torch
    from diff_envshade_amd import _diff_envshade as de
    t = _device(problem(seed=31, n=40000, H=16, W=32, vis="identity"))
    mat, vis = _vis_args(t)
    g = (t["grad_shading"], t["grad_diffuse_shading"])
    alone = de.bruteforce_specularvisibility_shader_backward2(*g, *mat, *vis, True)
    assert reuse_spy == []
    de.bruteforce_specularvisibility_shader_forward2(*mat, *vis, True)
    shared = de.bruteforce_specularvisibility_shader_backward2(*g, *mat, *vis, True)
    assert len(reuse_spy) == 1
    for a, b in zip(alone, shared):
        assert torch.equal(a, b)
    # an in-place change of the normals (version counter) stops the hand-over
    t["normal"].mul_(1.0)
    de.bruteforce_specularvisibility_shader_backward2(*g, *mat, *vis, True)
    assert len(reuse_spy) == 1
    # another variant's backward after the vis forward shares it too
    de.bruteforce_specularvisibility_shader_forward2(*mat, *vis, True)
    de.bruteforce_specular_shader_backward2(*g, *mat, True)
    assert len(reuse_spy) == 2


def test_an_order_of_other_normals_is_still_correct(reuse_spy):
    # the hand-over of an order sorted from different normals of the same P is
    # a permutation all the same: the outputs move by rounding only (the tiles
    # differ, so each point's sums are grouped differently)
    import torch
    from diff_envshade_amd import _diff_envshade as de
    d = problem(seed=32, n=40000, H=16, W=32, vis="identity")
    t = _device(d)
    mat, vis = _vis_args(t)
    g = (t["grad_shading"], t["grad_diffuse_shading"])
    fresh = de.bruteforce_specularvisibility_shader_backward2(*g, *mat, *vis, True)
    other = torch.roll(t["normal"], 977, dims=0).contiguous()
    de.bruteforce_specularvisibility_shader_forward2(other, *mat[1:], *vis, True)
    ows = de._order_cache[0][5]
    from diff_envshade_amd import _lib
    P, H, W = 40000, 16, 32

    def backward(order):
        outs = [torch.empty_like(x) for x in fresh]
        ws = torch.empty(_lib.workspace_size(P, H, W), dtype=torch.uint8, device="cuda:0")
        _lib.call("des_bruteforce_specularvisibility_shader_backward2", P, 3, H, W,
                  t["face_vertex_list"].shape[1], t["visibility"].shape[1],
                  t["visibility"].shape[0], *(x.data_ptr() for x in (*g, *mat, *vis)), 1,
                  *(o.data_ptr() for o in outs), ws.data_ptr(), ws.numel(),
                  torch.cuda.current_stream().cuda_stream, order=order)
        return outs
    outs = backward((ows.data_ptr(), P))
    for a, b in zip(fresh, outs):
        assert relerr(b.cpu().numpy(), a.cpu().numpy()) <= TOL
    # a hand-over for another number of points is ignored: the call sorts itself
    outs = backward((ows.data_ptr(), P - 1))
    for a, b in zip(fresh, outs):
        assert torch.equal(a, b)


def test_autograd_backward_takes_the_forward_order(reuse_spy):
    # through the torch.autograd.Function: the saved normals are the forward's
    # (same storage and version counter), so the backward takes its order; the
    # gradients are bitwise those of a backward that sorts by itself
    import torch
    from diff_envshade_amd import _diff_envshade as de
    from diff_envshade_amd.autograd import SpecularVisibilityShader2
    t = _device(problem(seed=33, n=40000, H=16, W=32, vis="identity"))
    names = ("normal", "view_dir", "albedo", "specular_albedo", "roughness", "envmap")
    leaves = {k: t[k].clone().requires_grad_(True) for k in names}
    sh, ds = SpecularVisibilityShader2.apply(
        *(leaves[k] for k in names), t["light2obj_rotmat"], t["face_vertex_list"],
        t["nearest_triangle_id"], t["barycentric_coord"], t["visibility"], True)
    ((sh * t["grad_shading"]).sum() + (ds * t["grad_diffuse_shading"]).sum()).backward()
    assert len(reuse_spy) == 1
    de._order_cache.clear()
    mat, vis = _vis_args(t)
    ref = de.bruteforce_specularvisibility_shader_backward2(
        t["grad_shading"], t["grad_diffuse_shading"], *mat, *vis, True)
    for k, r in zip(names, ref):
        assert torch.equal(leaves[k].grad.reshape(r.shape), r), k
