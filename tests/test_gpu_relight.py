"""Many-env-map relighting on MFMA (GPU): against the FP64 oracle per map and
against the FP32 drop-in shader, for every operand precision."""
import numpy as np
import pytest
import torch

from _helpers import gpu_inputs, oracle_run, relerr
from diff_envshade_amd import workloads as WL

pytestmark = pytest.mark.gpu

# normwise relative error vs FP64 allowed per precision (forward outputs and grad_envmap)
TOLS = {"bf16x3": 2e-5, "bf16": 1e-2, "fp16": 2e-3}


def relight_problem(seed, n, E, H, W, sigma=1.5):
    rng = np.random.default_rng(seed)
    d = WL.make_points(rng, n)
    d["light2obj_rotmat"] = WL.random_rotation(rng)
    d["visibility"] = WL.make_visibility(rng, d["normal"], H, W, d["light2obj_rotmat"])
    envs = np.stack([WL.make_envmap(rng, H, W, sigma) for _ in range(E)])
    gs = rng.normal(size=(n, E, 3)).astype(np.float32)
    gd = rng.normal(size=(n, E, 3)).astype(np.float32)
    return d, envs, gs, gd


def per_map_oracle(d, envs, gs, gd):
    n = d["normal"].shape[0]
    sh, ds, gE = [], [], []
    for e in range(envs.shape[0]):
        de = dict(d)
        de.update(WL.identity_mesh(n))
        de["envmap"] = envs[e]
        de["grad_shading"] = gs[:, e]
        de["grad_diffuse_shading"] = gd[:, e]
        (s, dd), b = oracle_run("vis", de)
        sh.append(s)
        ds.append(dd)
        gE.append(b[-1])
    return np.stack(sh, 1), np.stack(ds, 1), np.stack(gE)


@pytest.mark.parametrize("precision", ["bf16x3", "bf16", "fp16"])
def test_relight_matches_oracle(precision):
    from diff_envshade_amd.relight import relight_backward_envmap, relight_forward
    d, envs, gs, gd = relight_problem(0, n=1500, E=8, H=16, W=32)
    t = gpu_inputs(d)
    E = torch.from_numpy(envs).cuda()
    sh, ds = relight_forward(t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"],
                             t["roughness"], E, t["light2obj_rotmat"], t["visibility"], True,
                             precision)
    gE = relight_backward_envmap(t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"],
                                 t["roughness"], 16, 32, t["light2obj_rotmat"], t["visibility"],
                                 torch.from_numpy(gs).cuda(), torch.from_numpy(gd).cuda(), True,
                                 precision)
    torch.cuda.synchronize()
    rsh, rds, rgE = per_map_oracle(d, envs, gs, gd)
    errs = {"shading": relerr(sh.cpu().numpy(), rsh), "diffuse_shading": relerr(ds.cpu().numpy(), rds),
            "grad_envmap": relerr(gE.cpu().numpy(), rgE)}
    print(precision, errs)
    assert max(errs.values()) <= TOLS[precision], errs


def test_relight_many_maps_and_odd_sizes():
    # 70 maps -> two column groups; 1001 points and a 3 x 5 map exercise padding
    from diff_envshade_amd.relight import relight_backward_envmap, relight_forward
    d, envs, gs, gd = relight_problem(1, n=1001, E=70, H=3, W=5)
    t = gpu_inputs(d)
    E = torch.from_numpy(envs).cuda()
    sh, ds = relight_forward(t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"],
                             t["roughness"], E, t["light2obj_rotmat"], t["visibility"], True,
                             "bf16x3")
    gE = relight_backward_envmap(t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"],
                                 t["roughness"], 3, 5, t["light2obj_rotmat"], t["visibility"],
                                 torch.from_numpy(gs).cuda(), torch.from_numpy(gd).cuda(), True,
                                 "bf16x3")
    torch.cuda.synchronize()
    rsh, rds, rgE = per_map_oracle(d, envs, gs, gd)
    assert relerr(sh.cpu().numpy(), rsh) <= TOLS["bf16x3"]
    assert relerr(ds.cpu().numpy(), rds) <= TOLS["bf16x3"]
    assert relerr(gE.cpu().numpy(), rgE) <= TOLS["bf16x3"]


def test_relight_without_visibility_or_specular():
    from diff_envshade_amd.relight import relight_forward
    d, envs, gs, gd = relight_problem(2, n=700, E=4, H=8, W=16)
    t = gpu_inputs(d)
    E = torch.from_numpy(envs).cuda()
    sh, ds = relight_forward(t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"],
                             t["roughness"], E, t["light2obj_rotmat"], None, False, "bf16x3")
    torch.cuda.synchronize()
    for e in range(4):
        de = dict(d)
        de["envmap"] = envs[e]
        from oracle import oracle as O
        r_sh, r_ds = O.forward("v2", de, True, False)
        assert relerr(sh[:, e].cpu().numpy(), r_sh) <= TOLS["bf16x3"]
        assert relerr(ds[:, e].cpu().numpy(), r_ds) <= TOLS["bf16x3"]


def test_relight_envgrad_deterministic():
    from diff_envshade_amd.relight import relight_backward_envmap
    d, envs, gs, gd = relight_problem(3, n=3000, E=16, H=16, W=32)
    t = gpu_inputs(d)
    args = (t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"], t["roughness"], 16, 32,
            t["light2obj_rotmat"], t["visibility"], torch.from_numpy(gs).cuda(),
            torch.from_numpy(gd).cuda(), True, "bf16")
    a = relight_backward_envmap(*args).cpu().numpy()
    b = relight_backward_envmap(*args).cpu().numpy()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["bf16x3", "bf16"])
def test_relight_sharp_lobes_grazing_views(precision):
    # the shader's stress inputs (roughness down to 0, grazing and back-facing
    # views, bright texels) under four maps; fp16 is left out: its 65504 range
    # does not hold the GGX weights of lobes this sharp
    from _helpers import stress_problem
    from diff_envshade_amd.relight import relight_backward_envmap, relight_forward
    H, W, n, nE = 8, 16, 1500, 4
    d = stress_problem(seed=11, n=n, H=H, W=W)
    rng = np.random.default_rng(12)
    envs = np.stack([WL.make_envmap(rng, H, W, 1.0, n_bright=2) for _ in range(nE)])
    gs = rng.normal(size=(n, nE, 3)).astype(np.float32)
    gd = rng.normal(size=(n, nE, 3)).astype(np.float32)
    t = gpu_inputs(d)
    E = torch.from_numpy(envs).cuda()
    sh, ds = relight_forward(t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"],
                             t["roughness"], E, t["light2obj_rotmat"], t["visibility"], True,
                             precision)
    gE = relight_backward_envmap(t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"],
                                 t["roughness"], H, W, t["light2obj_rotmat"], t["visibility"],
                                 torch.from_numpy(gs).cuda(), torch.from_numpy(gd).cuda(), True,
                                 precision)
    torch.cuda.synchronize()
    rsh, rds, rgE = per_map_oracle(d, envs, gs, gd)
    errs = {"shading": relerr(sh.cpu().numpy(), rsh), "diffuse_shading": relerr(ds.cpu().numpy(), rds),
            "grad_envmap": relerr(gE.cpu().numpy(), rgE)}
    print(precision, errs)
    assert max(errs.values()) <= TOLS[precision], errs


def test_fp16_refuses_roughness_whose_weights_it_cannot_hold():
    # with specular on, fp16 needs roughness >= FP16_MIN_ROUGHNESS (the GGX
    # weight is at most 2 / (pi r^4)); bf16 / bf16x3, fp16 without specular
    # and fp16 at the bound itself run
    from diff_envshade_amd.relight import (FP16_MIN_ROUGHNESS, relight_backward_envmap,
                                           relight_forward)
    H, W = 8, 16
    d, envs, gs, gd = relight_problem(3, n=300, E=2, H=H, W=W)
    d["roughness"][7] = 0.03
    t = gpu_inputs(d)
    E = torch.from_numpy(envs).cuda()
    G = [torch.from_numpy(gs).cuda(), torch.from_numpy(gd).cuda()]
    mat = [t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"], t["roughness"]]

    def fwd(prec, spec=True):
        return relight_forward(*mat, E, t["light2obj_rotmat"], t["visibility"], spec, prec)

    def bwd(prec, spec=True):
        return relight_backward_envmap(*mat, H, W, t["light2obj_rotmat"], t["visibility"], *G,
                                       spec, prec)

    for call in (fwd, bwd):
        with pytest.raises(ValueError, match="fp16 cannot hold"):
            call("fp16")
        with pytest.raises(ValueError, match="precision must be one of"):
            call("fp8")
        for ok in (lambda: call("bf16"), lambda: call("bf16x3"), lambda: call("fp16", False)):
            out = ok()
            torch.cuda.synchronize()
            for x in (out if isinstance(out, tuple) else (out,)):
                assert torch.isfinite(x).all()
    t["roughness"][7] = FP16_MIN_ROUGHNESS
    mat[4] = t["roughness"]
    sh, ds = fwd("fp16")
    gE = bwd("fp16")
    torch.cuda.synchronize()
    assert torch.isfinite(sh).all() and torch.isfinite(ds).all() and torch.isfinite(gE).all()
