"""Pins for the FP64 oracle (CPU only).

The reference has no tests, golden vectors or buildable CPU path, so the
oracle is pinned by checks a plausible transcription mistake would fail:
closed forms of the lat-long geometry, textbook Cook-Torrance known answers in
angle form (tests/golden/kat_single_texel.json), one-hot env maps for row and
column orientation, visibility bit order and slot/vertex layout,
cross-variant identities, exact linearity in the env map (adjoint identity)
and FP64 central finite differences for every gradient.
"""
import json
import math
import os

import numpy as np
import pytest

from _helpers import BWD_NAMES, problem
from diff_envshade_amd import workloads as WL
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _d64(d):
    return {k: (v.astype(np.float64) if v.dtype == np.float32 else v) for k, v in d.items()}


# ---- lat-long geometry (bruteforce_shader.cu:124-147, config.h:11) ----------------

@pytest.mark.parametrize("H,W", [(1, 1), (2, 4), (16, 32), (7, 13), (64, 128)])
def test_solid_angles_sum_to_sphere(H, W):
    _, dw = O.texel_table(H, W)
    assert abs(dw.sum() - 4 * math.pi) < 1e-12
    assert np.all(dw > 0)


def test_solid_angle_is_latitude_band_over_W():
    H, W = 16, 32
    _, dw = O.texel_table(H, W)
    edges = np.linspace(math.pi / 2, -math.pi / 2, H + 1)  # top row first
    band = 2 * math.pi / W * (np.sin(edges[:-1]) - np.sin(edges[1:]))
    np.testing.assert_allclose(dw.reshape(H, W), np.repeat(band[:, None], W, axis=1), rtol=1e-13)


def test_texel_directions_orientation():
    H, W = 8, 16
    l, _ = O.texel_table(H, W)
    np.testing.assert_allclose(np.linalg.norm(l, axis=1), 1.0, rtol=1e-14)
    y = l[:, 1].reshape(H, W)
    # storage row 0 is the top of the map: +y, decreasing row by row
    assert np.all(y[0] > 0) and np.all(np.diff(y[:, 0]) < 0)
    assert np.allclose(y[:, 0], -y[::-1, 0])
    # column x points at longitude 2 pi (x + 1/2) / W from +x towards +z
    phi = np.arctan2(l[:, 2], l[:, 0]).reshape(H, W)[H // 2]
    expect = 2 * math.pi * (np.arange(W) + 0.5) / W
    np.testing.assert_allclose(np.mod(phi, 2 * math.pi), expect, atol=1e-13)
    # integral of the direction over the sphere vanishes
    _, dw = O.texel_table(H, W)
    assert np.abs((l * dw[:, None]).sum(axis=0)).max() < 1e-13


def test_rotation_is_row_major_M_times_d():
    rng = np.random.default_rng(0)
    M = WL.random_rotation(rng).astype(np.float64)
    l0, _ = O.texel_table(4, 8)
    l1, _ = O.texel_table(4, 8, M)
    np.testing.assert_allclose(l1, l0 @ M.T, atol=1e-15)


def test_workload_texel_dirs_match_oracle():
    rng = np.random.default_rng(1)
    M = WL.random_rotation(rng)
    a, da = WL.texel_dirs(8, 16, M)
    b, db = O.texel_table(8, 16, M)
    np.testing.assert_allclose(a, b, atol=1e-15)
    np.testing.assert_allclose(da, db, rtol=1e-13)


# ---- known answers (tests/golden/kat_single_texel.json) ----------------------------

def _kat_cases():
    with open(os.path.join(HERE, "golden", "kat_single_texel.json")) as f:
        return json.load(f)["cases"]


def kat_inputs(c):
    H, W = c["H"], c["W"]
    env = np.zeros((3, H * W))
    env[:, c["texel"]] = c["E"]
    return {
        "normal": np.array([c["normal"]]), "view_dir": np.array([c["view_dir"]]),
        "albedo": np.array([c["albedo"]]), "specular_albedo": np.array([c["specular_albedo"]]),
        "roughness": np.array([[c["roughness"]]]), "envmap": env.reshape(3, H, W),
        "light2obj_rotmat": np.array(c["light2obj_rotmat"]),
    }


@pytest.mark.parametrize("c", _kat_cases(), ids=lambda c: c["name"])
def test_known_answers(c):
    d = kat_inputs(c)
    e = c["expected"]
    sh, ds = O.forward("v2", d)
    np.testing.assert_allclose(sh[0], e["shading_v2"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(ds[0], e["diffuse_shading"], rtol=1e-12, atol=1e-15)
    sp, _ = O.forward("specular", d, enable_diffuse=False, enable_specular=True)
    np.testing.assert_allclose(sp[0], e["shading_specular_only"], rtol=1e-12, atol=1e-15)
    df = O.forward("diffuse", d)[0]
    np.testing.assert_allclose(df[0], np.array(c["albedo"]) * e["diffuse_shading"], rtol=1e-12,
                               atol=1e-15)
    # the clamp variant skips a point unless n.v > 0; the others shade a
    # view at or below the horizon with G = G1(n.l) / k (bs:161-171)
    cl, _ = O.forward("clamp", d, enable_diffuse=True, enable_specular=True)
    np.testing.assert_allclose(cl[0], e["shading_clamp"], rtol=1e-12, atol=1e-15)
    v1, _ = O.forward("specular", d, enable_diffuse=True, enable_specular=True)
    np.testing.assert_allclose(v1[0], e["shading_v2"], rtol=1e-12, atol=1e-15)


def test_roughness_one_gives_flat_ndf():
    # alpha^2 = 1: D = 1/pi for every half vector, so the specular lobe is
    # F G / (4 pi); check against the value at two different view angles
    c = [k for k in _kat_cases() if k["name"] == "rough_one"][0]
    d = kat_inputs(c)
    sp, _ = O.forward("specular", d, False, True)
    a = math.radians(60.0)
    nl = math.cos(a)
    v = np.array(c["view_dir"])
    nv = v[1]
    k = 0.5
    l = np.array([math.sin(a), nl, 0.0])
    h = (l + v) / np.linalg.norm(l + v)
    F = 0.04 + 0.96 * (1 - l @ h) ** 5
    G = nl / (nl * (1 - k) + k) / (nv * (1 - k) + k)
    dw = O.texel_table(c["H"], c["W"])[1][c["texel"]]
    expect = np.array(c["E"]) * np.array(c["specular_albedo"]) * F * G / (4 * math.pi) * dw
    np.testing.assert_allclose(sp[0], expect, rtol=1e-12)


def test_uniform_env_lambert_integrates_to_one():
    # E = 1, albedo = 1: shading -> (1/pi) * integral of max(cos, 0) = 1 as the map refines
    rng = np.random.default_rng(2)
    g = rng.normal(size=(16, 3))
    n = g / np.linalg.norm(g, axis=1, keepdims=True)
    for H, W, tol in ((32, 64, 3e-3), (128, 256, 2e-4)):
        d = {"normal": n, "albedo": np.ones((16, 3)), "envmap": np.ones((3, H, W)),
             "light2obj_rotmat": np.eye(3)}
        sh = O.forward("diffuse", d)[0]
        assert np.abs(sh - 1.0).max() < tol, (H, W, np.abs(sh - 1).max())


def test_top_row_lights_upward_normal_only():
    H, W = 8, 16
    env = np.zeros((3, H, W))
    env[:, 0, :] = 1.0  # storage row 0 = top of the map
    d = {"normal": np.array([[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]]), "albedo": np.ones((2, 3)),
         "envmap": env, "light2obj_rotmat": np.eye(3)}
    sh = O.forward("diffuse", d)[0]
    assert np.all(sh[0] > 0) and np.all(sh[1] == 0)


# ---- visibility decoding (bruteforce_shader.cu:2145-2151, 2180-2190) ----------------

def _one_hot_vis_problem(t, H=4, W=16, nvert=3):
    base = problem(seed=3, n=1, H=H, W=W, rotation=False)
    d = _d64(base)
    env = np.zeros((3, H * W))
    env[:, t] = [1.0, 2.0, 3.0]
    d["envmap"] = env.reshape(3, H, W)
    l, _ = O.texel_table(H, W)
    d["normal"] = l[t:t + 1].copy()  # face the lit texel
    d["view_dir"] = l[t:t + 1].copy()
    nslot = (H * W + 31) // 32
    d["visibility"] = np.zeros((nslot, nvert), dtype=np.uint32)
    d["face_vertex_list"] = np.array([[0], [1], [2]], dtype=np.int32)
    d["nearest_triangle_id"] = np.array([0], dtype=np.int64)
    d["barycentric_coord"] = np.array([[0.2], [0.3], [0.5]])
    return d


@pytest.mark.parametrize("t", [0, 5, 31, 32, 33, 63])
def test_visibility_bit_order_and_slot_layout(t):
    d = _one_hot_vis_problem(t)
    full = O.forward("v2", d)
    slot, bit = divmod(t, 32)
    for mask, weight in (((1, 0, 0), 0.2), ((0, 1, 0), 0.3), ((0, 0, 1), 0.5),
                         ((1, 1, 1), 1.0), ((0, 1, 1), 0.8)):
        vis = np.zeros_like(d["visibility"])
        for j, m in enumerate(mask):
            if m:
                vis[slot, j] = np.uint32(1) << np.uint32(bit)
        d["visibility"] = vis
        sh, ds = O.forward("vis", d)
        np.testing.assert_allclose(sh, weight * full[0], rtol=1e-14)
        np.testing.assert_allclose(ds, weight * full[1], rtol=1e-14)
    # every other bit set, this one clear: exactly dark
    vis = np.full_like(d["visibility"], 0xFFFFFFFF)
    vis[slot, :] &= ~np.uint32(np.uint32(1) << np.uint32(bit))
    d["visibility"] = vis
    sh, ds = O.forward("vis", d)
    assert np.all(sh == 0) and np.all(ds == 0)


def test_pack_bits_round_trip_and_order():
    rng = np.random.default_rng(4)
    bits = rng.random((5, 70)) < 0.5
    words = WL.pack_bits(bits)
    assert words.shape == (3, 5) and words.dtype == np.int32
    assert np.array_equal(WL.unpack_bits(words, 70), bits)
    one = np.zeros((1, 64), dtype=bool)
    one[0, 33] = True
    w = WL.pack_bits(one).view(np.uint32)
    assert w[0, 0] == 0 and w[1, 0] == 2


# ---- variant identities ----------------------------------------------------------

def test_variant_identities():
    d = _d64(problem(seed=5, n=200, H=8, W=16, vis="ones"))
    sh2, ds2 = O.forward("v2", d)
    sh1, _ = O.forward("specular", d, True, True)
    np.testing.assert_allclose(sh2, sh1, rtol=1e-13)
    dd = dict(d)
    dd["albedo"] = np.ones_like(d["albedo"])
    np.testing.assert_allclose(ds2, O.forward("diffuse", dd)[0], rtol=1e-13)
    np.testing.assert_allclose(O.forward("specular", d, True, False)[0],
                               O.forward("diffuse", d)[0], rtol=1e-13)
    shv, dsv = O.forward("vis", d)
    np.testing.assert_allclose(shv, sh2, rtol=1e-13)
    np.testing.assert_allclose(dsv, ds2, rtol=1e-13)
    # clamp: v1 where n.v > 0, zero elsewhere
    shc, _ = O.forward("clamp", d, True, True)
    front = (d["normal"] * d["view_dir"]).sum(1) > 0
    assert (~front).any() and front.any()
    np.testing.assert_allclose(shc[front], sh1[front], rtol=1e-13)
    assert np.all(shc[~front] == 0)
    gv = O.backward("vis", d, d["grad_shading"], d["grad_diffuse_shading"])
    g2 = O.backward("v2", d, d["grad_shading"], d["grad_diffuse_shading"])
    for k in g2:
        np.testing.assert_allclose(gv[k], g2[k], rtol=1e-12, atol=1e-300)


def test_disabled_branches_are_zero():
    d = _d64(problem(seed=6, n=50, H=4, W=8))
    g = O.backward("specular", d, d["grad_shading"], None, enable_diffuse=False,
                   enable_specular=True)
    assert np.all(g["albedo"] == 0)
    g = O.backward("specular", d, d["grad_shading"], None, enable_diffuse=True,
                   enable_specular=False)
    for k in ("view_dir", "specular_albedo", "roughness"):
        assert np.all(g[k] == 0)


# ---- gradients -----------------------------------------------------------------

def _loss(variant, d, diff, spec):
    sh, ds = O.forward(variant, d, diff, spec)
    L = (sh * d["grad_shading"]).sum()
    if ds is not None:
        L += (ds * d["grad_diffuse_shading"]).sum()
    return L


@pytest.mark.parametrize("variant,diff,spec", [("diffuse", True, False), ("specular", True, True),
                                               ("specular", False, True), ("clamp", True, True),
                                               ("v2", True, True), ("vis", True, True)])
def test_gradients_match_central_differences(variant, diff, spec):
    d = _d64(problem(seed=7, n=6, H=4, W=8, vis="mesh" if variant == "vis" else "none"))
    g = O.backward(variant, d, d["grad_shading"], d["grad_diffuse_shading"], diff, spec)
    h = 1e-6
    keys = BWD_NAMES[variant]
    for key in keys:
        x = d[key]
        fd = np.zeros_like(x)
        for idx in np.ndindex(x.shape):
            xp = x.copy()
            xp[idx] += h
            xm = x.copy()
            xm[idx] -= h
            fd[idx] = (_loss(variant, {**d, key: xp}, diff, spec)
                       - _loss(variant, {**d, key: xm}, diff, spec)) / (2 * h)
        err = np.abs(fd - g[key]).max() / max(np.abs(fd).max(), 1e-12)
        assert err < 1e-6, (variant, key, err)


def test_envmap_gradient_adjoint_identity():
    # shading is linear in E, so <gE, E'> = <gs, sh(E')> + <gd, ds(E')> exactly
    d = _d64(problem(seed=8, n=100, H=8, W=16, vis="identity"))
    g = O.backward("vis", d, d["grad_shading"], d["grad_diffuse_shading"])
    rng = np.random.default_rng(8)
    E2 = rng.normal(size=d["envmap"].shape)
    sh, ds = O.forward("vis", {**d, "envmap": E2})
    lhs = (g["envmap"] * E2).sum()
    rhs = (sh * d["grad_shading"]).sum() + (ds * d["grad_diffuse_shading"]).sum()
    assert abs(lhs - rhs) <= 1e-11 * max(abs(rhs), 1.0)


def test_zero_visibility_gives_zero():
    d = _d64(problem(seed=9, n=40, H=4, W=8, vis="identity"))
    d["visibility"] = np.zeros_like(d["visibility"])
    sh, ds = O.forward("vis", d)
    assert np.all(sh == 0) and np.all(ds == 0)
    g = O.backward("vis", d, d["grad_shading"], d["grad_diffuse_shading"])
    for v in g.values():
        assert np.all(v == 0)


def test_back_facing_views_shade_specular_except_in_the_clamp_variant():
    # the known answers with n.v <= 0 carry a non-zero specular term that the
    # non-clamp variants reproduce and the clamp variant does not
    cases = [c for c in _kat_cases() if np.dot(c["normal"], c["view_dir"]) <= 0]
    assert len(cases) == 3
    for c in cases:
        d = kat_inputs(c)
        assert min(c["expected"]["shading_specular_only"]) > 0
        sp, _ = O.forward("specular", d, enable_diffuse=False, enable_specular=True)
        assert (sp[0] > 0).all()
        cl, _ = O.forward("clamp", d, enable_diffuse=False, enable_specular=True)
        assert (cl[0] == 0).all()


# ---- the conditioning scale of grad_roughness --------------------------------------

def test_roughness_scale_bounds_the_one_hot_terms():
    # grad_roughness is linear in the env map, so with every texel but t set
    # to zero it is exactly texel t's term.  The scale the oracle reports sums
    # the magnitudes of the parts of every texel's term (each texel's parts
    # depend on its own radiance only), so it is additive over one-hot maps
    # and bounds the sum of the terms' magnitudes, and asking for it changes
    # no output
    d = _d64(problem(seed=8, n=60, H=4, W=8, vis="identity"))
    gs, gd = d["grad_shading"], d["grad_diffuse_shading"]
    g = O.backward("vis", d, gs, gd, roughness_abs=True)
    g0 = O.backward("vis", d, gs, gd)
    assert "roughness_abs" not in g0
    for k in g0:
        assert np.array_equal(g[k], g0[k]), k
    T = 32
    terms = np.zeros((60, T))
    scales = np.zeros((60, T))
    for t in range(T):
        e = np.zeros_like(d["envmap"])
        e[:, t // 8, t % 8] = d["envmap"][:, t // 8, t % 8]
        gt = O.backward("vis", dict(d, envmap=e), gs, gd, roughness_abs=True)
        terms[:, t] = gt["roughness"][:, 0]
        scales[:, t] = gt["roughness_abs"][:, 0]
    np.testing.assert_allclose(terms.sum(1), g["roughness"][:, 0], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(scales.sum(1), g["roughness_abs"][:, 0], rtol=1e-12, atol=0)
    assert np.all(scales >= np.abs(terms) * (1 - 1e-12))
    # the terms cancel across texels and their parts within texels
    assert np.any(np.abs(terms).sum(1) > 2 * np.abs(g["roughness"][:, 0]))
    assert np.any(scales > 2 * np.abs(terms))
    # scaling the radiance scales it; no specular lobe: no roughness terms
    g3 = O.backward("vis", dict(d, envmap=3.0 * d["envmap"]), gs, gd, roughness_abs=True)
    np.testing.assert_allclose(g3["roughness_abs"], 3.0 * g["roughness_abs"], rtol=1e-12)
    gn = O.backward("vis", d, gs, gd, enable_specular=False, roughness_abs=True)
    assert np.all(gn["roughness_abs"] == 0)
