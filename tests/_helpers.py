"""Shared helpers for the parity tests: run one variant through the HIP drop-in
and through the FP64 oracle on the same inputs, and compare.

Error metrics, per output tensor:
  * normwise relative error  err(x, ref) = max|x - ref| / max|ref|;
  * for the colors (shading, diffuse_shading: sums of non-negative terms) the
    elementwise relative error max |x - ref| / |ref| over the elements with
    |ref| >= 1e-3 max|ref|, so that dim points are held to the same bound;
  * for the gradients (signed sums, where elementwise relative error is
    unbounded near zero crossings) the per-point normwise error
    max_p |x_p - ref_p| / |ref_p| over the points (texels, for grad_envmap)
    whose gradient vector has |ref_p| >= 1e-3 max_p |ref_p|.
"""
from __future__ import annotations

import numpy as np

from diff_envshade_amd import workloads as WL
from oracle import oracle as O

# colors and every gradient must match the FP64 oracle to this normwise relative error
TOL = 1e-5
# the same bound for the elementwise error of the colors and the per-point
# error of the gradients (over the elements / points above 1e-3 of the largest).
# grad_roughness is a scalar per point summing terms of both signs (the GGX
# NDF's derivative in alpha^2 is negative near the lobe's peak and positive
# away from it), so its error is taken relative to the sum of the terms'
# magnitudes, which the oracle reports, rather than to the sum itself: where
# the terms cancel, no finite-precision sum holds the latter's digits.
TOL_LOCAL = 1e-5
LOCAL_FLOOR = 1e-3

FWD_NAMES = {"diffuse": ["shading"], "specular": ["shading"], "clamp": ["shading"],
             "v2": ["shading", "diffuse_shading"], "vis": ["shading", "diffuse_shading"]}
BWD_NAMES = {"diffuse": ["normal", "albedo", "envmap"]}
for _v in ("specular", "clamp", "v2", "vis"):
    BWD_NAMES[_v] = ["normal", "view_dir", "albedo", "specular_albedo", "roughness", "envmap"]


def relerr(x, ref) -> float:
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert x.shape == ref.shape, (x.shape, ref.shape)
    den = np.abs(ref).max() if ref.size else 0.0
    num = np.abs(x - ref).max() if ref.size else 0.0
    return float(num / den) if den > 0 else float(num)


def elemerr(x, ref, floor=LOCAL_FLOOR) -> float:
    """Elementwise relative error over |ref| >= floor * max|ref|."""
    x = np.asarray(x, dtype=np.float64).reshape(np.shape(ref))
    ref = np.asarray(ref, dtype=np.float64)
    if ref.size == 0:
        return 0.0
    top = np.abs(ref).max()
    if top == 0:
        return float(np.abs(x).max()) if x.size else 0.0
    m = np.abs(ref) >= floor * top
    return float((np.abs(x - ref)[m] / np.abs(ref)[m]).max())


def pointerr(x, ref, floor=LOCAL_FLOOR, texel_major=False, scale=None) -> float:
    """Per-point normwise error: rows of a P x k gradient (columns of a
    3 x H x W env-map gradient, one per texel) as vectors, over the rows with
    |ref_row| >= floor * max |ref_row|.  With scale (one value per row) the
    error of a row is taken relative to its scale instead of |ref_row|."""
    ref = np.asarray(ref, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64).reshape(ref.shape)
    if ref.size == 0:
        return 0.0
    if texel_major:
        ref = ref.reshape(ref.shape[0], -1).T
        x = x.reshape(x.shape[0], -1).T
    else:
        ref = ref.reshape(ref.shape[0], -1)
        x = x.reshape(x.shape[0], -1)
    rn = np.linalg.norm(ref, axis=1) if scale is None else np.asarray(scale, np.float64).ravel()
    top = rn.max()
    if top == 0:
        return float(np.abs(x).max())
    m = rn >= floor * top
    return float((np.linalg.norm(x - ref, axis=1)[m] / rn[m]).max())


def problem(seed=0, n=300, H=4, W=8, rotation=True, vis="none", p_vis=0.7, n_bright=0,
            sigma=1.0):
    """Random inputs. vis: 'none' | 'ones' | 'identity' | 'mesh'."""
    rng = np.random.default_rng(seed)
    d = WL.make_points(rng, n)
    d["envmap"] = WL.make_envmap(rng, H, W, sigma, n_bright=n_bright)
    d["light2obj_rotmat"] = WL.random_rotation(rng) if rotation else np.eye(3, dtype=np.float32)
    d.update(WL.upstream_grads(rng, n))
    if vis == "ones":
        d["visibility"] = np.full(((H * W + 31) // 32, n), -1, dtype=np.int32)
        d.update(WL.identity_mesh(n))
    elif vis == "identity":
        d["visibility"] = WL.make_visibility(rng, d["normal"], H, W, d["light2obj_rotmat"], p_vis)
        d.update(WL.identity_mesh(n))
    elif vis == "mesh":
        nv = max(1, n // 2)
        nf = max(1, n // 3)
        words = rng.integers(0, 2**32, size=((H * W + 31) // 32, nv), dtype=np.uint64)
        d["visibility"] = words.astype(np.uint32).view(np.int32)
        d.update(WL.random_mesh(rng, n, nv, nf))
    return d


def stress_problem(seed, n=1500, H=8, W=16):
    """Sharp and maximal lobes, grazing and back-facing views, bright texels:
    roughness in {0, 0.01, 0.03, 0.08, 0.3, 1} (alpha^2 down to 1e-8, so D
    peaks near 1 / (pi alpha^2)); view directions a third uniform on the
    sphere (half of them back-facing), a third grazing (n.v ~ 1e-3) and a
    third along the normal; three texels 50x brighter."""
    d = problem(seed=seed, n=n, H=H, W=W, vis="identity", n_bright=3)
    rng = np.random.default_rng(seed + 100)
    d["roughness"] = rng.choice([0.0, 0.01, 0.03, 0.08, 0.3, 1.0], size=(n, 1)).astype(np.float32)
    nrm = d["normal"].astype(np.float64)
    t = np.cross(nrm, rng.normal(size=(n, 3)))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    graze = t + 1e-3 * nrm
    graze /= np.linalg.norm(graze, axis=1, keepdims=True)
    kind = (np.arange(n) % 3)[:, None]
    d["view_dir"] = np.where(kind == 0, u, np.where(kind == 1, graze, nrm)).astype(np.float32)
    return d


class RefGrads(list):
    """The oracle's gradients in output order, with .scale: {name: per-point
    scale} for the outputs whose local error is taken against a scale."""
    scale: dict = {}


def oracle_run(variant, d, diff=True, spec=True):
    sh, ds = O.forward(variant, d, diff, spec)
    fwd = [sh] if ds is None else [sh, ds]
    g = O.backward(variant, d, d["grad_shading"], d.get("grad_diffuse_shading"), diff, spec,
                   roughness_abs=True)
    bwd = RefGrads(g[k] for k in BWD_NAMES[variant])
    bwd.scale = {"grad_roughness": g["roughness_abs"]} if "roughness" in BWD_NAMES[variant] else {}
    return fwd, bwd


def gpu_inputs(d, device="cuda"):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in d.items()}


def gpu_run(variant, d, diff=True, spec=True, device="cuda"):
    """Forward + backward through the drop-in bindings; returns numpy lists."""
    import torch
    from diff_envshade_amd import _diff_envshade as C
    t = gpu_inputs(d, device)
    mat = [t["normal"], t["view_dir"], t["albedo"], t["specular_albedo"], t["roughness"],
           t["envmap"], t["light2obj_rotmat"]]
    if variant == "diffuse":
        fwd = [C.bruteforce_diffuse_shader_forward(t["normal"], t["albedo"], t["envmap"],
                                                   t["light2obj_rotmat"])]
        bwd = C.bruteforce_diffuse_shader_backward(t["grad_shading"], t["normal"], t["albedo"],
                                                   t["envmap"], t["light2obj_rotmat"])
    elif variant == "specular":
        fwd = [C.bruteforce_specular_shader_forward(*mat, diff, spec)]
        bwd = C.bruteforce_specular_shader_backward(t["grad_shading"], *mat, diff, spec)
    elif variant == "clamp":
        fwd = [C.bruteforce_specular_shader_clamp_forward(*mat, diff, spec)]
        bwd = C.bruteforce_specular_shader_clamp_backward(t["grad_shading"], *mat, diff, spec)
    elif variant == "v2":
        fwd = C.bruteforce_specular_shader_forward2(*mat, spec)
        bwd = C.bruteforce_specular_shader_backward2(t["grad_shading"],
                                                     t["grad_diffuse_shading"], *mat, spec)
    elif variant == "vis":
        vis = [t["face_vertex_list"], t["nearest_triangle_id"], t["barycentric_coord"],
               t["visibility"]]
        fwd = C.bruteforce_specularvisibility_shader_forward2(*mat, *vis, spec)
        bwd = C.bruteforce_specularvisibility_shader_backward2(
            t["grad_shading"], t["grad_diffuse_shading"], *mat, *vis, spec)
    else:
        raise ValueError(variant)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in fwd], [x.cpu().numpy() for x in bwd]


def compare(variant, got, ref, tol=TOL, tol_local=TOL_LOCAL):
    """Assert every output within tol (normwise) and tol_local (elementwise for
    the colors, per point for the gradients); returns {name: normwise err}.
    tol_local=None skips the local check."""
    (gf, gb), (rf, rb) = got, ref
    scale = getattr(rb, "scale", {})
    errs, local = {}, {}
    for name, x, r in zip(FWD_NAMES[variant], gf, rf):
        errs[name] = relerr(x, r)
        local[name] = elemerr(x, r)
    for name, x, r in zip(BWD_NAMES[variant], gb, rb):
        errs["grad_" + name] = relerr(x.reshape(r.shape), r)
        local["grad_" + name] = pointerr(x, r, texel_major=(name == "envmap"),
                                         scale=scale.get("grad_" + name))
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"{variant}: outputs beyond {tol}: {bad} (all: {errs})"
    if tol_local is not None:
        bad = {k: v for k, v in local.items() if not v <= tol_local}
        assert not bad, (f"{variant}: local errors beyond {tol_local}: {bad} (all: {local}; "
                         f"normwise {errs})")
    return errs
