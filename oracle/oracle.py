"""ctypes wrapper around the FP64 oracle (oracle/des_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline.  The product
package never imports this module.

All float inputs are widened to float64 (exactly) before the call; outputs are
float64.  Argument meaning follows the reference bindings
(gsface_shader.h:3-118 of zjumsj/diff_envshade).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_des.so")

VARIANTS = {"diffuse": 0, "specular": 1, "clamp": 2, "v2": 3, "vis": 4}

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH) or any(
        os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, src))
        for src in ("des_oracle.c", "mesh_oracle.c", "envmap_oracle.c")
    ):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle_des.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        c_int, c_i64 = ctypes.c_int, ctypes.c_int64
        L.des_oracle_forward.restype = c_int
        L.des_oracle_forward.argtypes = [c_int, c_int, c_int, c_i64, c_int, c_int, c_int, c_int,
                                         dp, dp, dp, dp, dp, dp, dp, i32p, i64p, dp, u32p,
                                         dp, dp, c_int]
        L.des_oracle_backward.restype = c_int
        L.des_oracle_backward.argtypes = [c_int, c_int, c_int, c_i64, c_int, c_int, c_int, c_int,
                                          dp, dp, dp, dp, dp, dp, dp, dp, dp, i32p, i64p, dp,
                                          u32p, dp, dp, dp, dp, dp, dp, c_int]
        L.des_oracle_backward_ex.restype = c_int
        L.des_oracle_backward_ex.argtypes = L.des_oracle_backward.argtypes[:-1] + [dp, c_int]
        L.des_oracle_texel_table.restype = None
        L.des_oracle_texel_table.argtypes = [c_int, c_int, dp, dp, dp]
        L.des_oracle_delta_omega.restype = ctypes.c_double
        L.des_oracle_delta_omega.argtypes = [c_int, c_int, c_int]
        L.des_oracle_max_threads.restype = c_int
        L.des_oracle_max_threads.argtypes = []
        L.des_oracle_closest_point.restype = None
        L.des_oracle_closest_point.argtypes = [dp, dp, dp, dp, dp, dp]
        L.des_oracle_uv_of_triangle_forward.restype = c_int
        L.des_oracle_uv_of_triangle_forward.argtypes = [c_i64, dp, dp, dp]
        L.des_oracle_uv_of_triangle_backward.restype = c_int
        L.des_oracle_uv_of_triangle_backward.argtypes = [c_i64, dp, dp, dp, dp]
        L.des_oracle_nearest_forward.restype = c_int
        L.des_oracle_nearest_forward.argtypes = [c_int, c_i64, c_i64, c_int, i32p, i32p, c_i64,
                                                 i32p, dp, dp, i64p, i64p, dp]
        L.des_oracle_nearest_backward.restype = c_int
        L.des_oracle_nearest_backward.argtypes = [c_int, c_i64, c_i64, i32p, dp, dp, i64p, dp,
                                                  dp]
        L.des_oracle_extract_bitfield.restype = None
        L.des_oracle_extract_bitfield.argtypes = [u32p, c_i64, c_int, c_i64, c_i64, dp]
        fp = ctypes.POINTER(ctypes.c_float)
        L.des_oracle_philox.restype = None
        L.des_oracle_philox.argtypes = [u32p, u32p, u32p]
        L.des_oracle_sample_envmap.restype = None
        L.des_oracle_sample_envmap.argtypes = [c_i64, c_int, c_int, c_int, dp, fp, dp]
        L.des_oracle_ggx_sample.restype = None
        L.des_oracle_ggx_sample.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.des_oracle_blur_envmap.restype = None
        L.des_oracle_blur_envmap.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int,
                                             ctypes.c_double, fp, dp]
        _lib = L
    return _lib


def _d(x):
    if x is None:
        return None, None
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _typed(x, dtype, ctype):
    if x is None:
        return None, None
    a = np.ascontiguousarray(np.asarray(x).astype(dtype, copy=False))
    return a, a.ctypes.data_as(ctypes.POINTER(ctype))


def max_threads() -> int:
    return int(lib().des_oracle_max_threads())


def texel_table(H: int, W: int, rotmat=None):
    """Rotated texel directions (T x 3) and solid angles (T,) in FP64."""
    M = np.eye(3) if rotmat is None else np.asarray(rotmat, dtype=np.float64)
    M, Mp = _d(M)
    T = H * W
    l = np.zeros((T, 3))
    dw = np.zeros(T)
    lib().des_oracle_texel_table(H, W, Mp, l.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 dw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return l, dw


def _common(inp):
    normal = inp["normal"]
    P = int(np.asarray(normal).shape[0])
    env = np.asarray(inp["envmap"])
    C, H, W = env.shape
    if C != 3:
        raise RuntimeError("Channel number not supported!")
    keep = []
    ptrs = []
    for k in ("normal", "view_dir", "albedo", "specular_albedo", "roughness", "envmap",
              "light2obj_rotmat"):
        a, p = _d(inp.get(k))
        keep.append(a)
        ptrs.append(p)
    fvl, fvlp = _typed(inp.get("face_vertex_list"), np.int32, ctypes.c_int32)
    tri, trip = _typed(inp.get("nearest_triangle_id"), np.int64, ctypes.c_int64)
    bary, baryp = _d(inp.get("barycentric_coord"))
    vis = inp.get("visibility")
    if vis is not None:
        vis = np.ascontiguousarray(np.asarray(vis).view(np.uint32) if np.asarray(vis).dtype
                                   == np.int32 else np.asarray(vis, dtype=np.uint32))
        visp = vis.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    else:
        visp = None
    keep += [fvl, tri, bary, vis]
    N_face = int(fvl.shape[1]) if fvl is not None else 0
    N_vertex = int(vis.shape[1]) if vis is not None else 0
    return P, H, W, N_face, N_vertex, ptrs + [fvlp, trip, baryp, visp], keep


def forward(variant: str, inp: dict, enable_diffuse: bool = True, enable_specular: bool = True,
            nthreads: int = 0):
    """Returns (shading, diffuse_shading or None) as float64 arrays (P x 3)."""
    v = VARIANTS[variant]
    P, H, W, NF, NV, ptrs, keep = _common(inp)
    sh = np.zeros((P, 3))
    ds = np.zeros((P, 3)) if variant in ("v2", "vis") else None
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib().des_oracle_forward(v, int(enable_diffuse), int(enable_specular), P, H, W, NF, NV,
                                  *ptrs, sh.ctypes.data_as(dp),
                                  ds.ctypes.data_as(dp) if ds is not None else None, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle forward failed ({rc})")
    return sh, ds


def backward(variant: str, inp: dict, grad_shading, grad_diffuse_shading=None,
             enable_diffuse: bool = True, enable_specular: bool = True, nthreads: int = 0,
             roughness_abs: bool = False):
    """Returns dict of float64 grads: normal, view_dir, albedo, specular_albedo,
    roughness (P x 1) and envmap (3 x H x W); with roughness_abs also
    "roughness_abs" (P x 1): the magnitudes of the terms summed into the
    alpha^2 and k gradients (two each per texel), through the chain factors
    of the roughness gradient -- the scale of that sum's rounding error."""
    v = VARIANTS[variant]
    P, H, W, NF, NV, ptrs, keep = _common(inp)
    gsh, gshp = _d(grad_shading)
    gds, gdsp = _d(grad_diffuse_shading)
    out = {k: np.zeros((P, 3)) for k in ("normal", "view_dir", "albedo", "specular_albedo")}
    out["roughness"] = np.zeros((P, 1))
    out["envmap"] = np.zeros((3, H, W))
    dp = ctypes.POINTER(ctypes.c_double)
    if roughness_abs:
        out["roughness_abs"] = np.zeros((P, 1))
    rc = lib().des_oracle_backward_ex(
        v, int(enable_diffuse), int(enable_specular), P, H, W, NF, NV, gshp, gdsp, *ptrs,
        *(out[k].ctypes.data_as(dp) for k in ("normal", "view_dir", "albedo", "specular_albedo",
                                              "roughness", "envmap")),
        out["roughness_abs"].ctypes.data_as(dp) if roughness_abs else None, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle backward failed ({rc})")
    return out


# ---- mesh utilities (mesh_oracle.c; etc.cu of the reference) ----------------------

def closest_point(x0, x1, x2, q):
    """(s, t) of the closest point x0 + s (x1 - x0) + t (x2 - x0) to q."""
    L = lib()
    a = [_d(v) for v in (x0, x1, x2, q)]
    s, t = ctypes.c_double(), ctypes.c_double()
    L.des_oracle_closest_point(*(p for _, p in a), ctypes.byref(s), ctypes.byref(t))
    return s.value, t.value


def uv_of_triangle_forward(triangles, query_pos):
    """triangles P x 3 x 3, query_pos P x 3 -> barycentric P x 3 (FP64)."""
    tri, tp = _d(triangles)
    q, qp = _d(query_pos)
    P = q.shape[0]
    out = np.zeros((P, 3))
    rc = lib().des_oracle_uv_of_triangle_forward(
        P, tp, qp, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return out


def uv_of_triangle_backward(triangles, query_pos, grad_barycentric):
    tri, tp = _d(triangles)
    q, qp = _d(query_pos)
    g, gp = _d(grad_barycentric)
    P = q.shape[0]
    out = np.zeros((P, 3))
    rc = lib().des_oracle_uv_of_triangle_backward(
        P, tp, qp, gp, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return out


def nearest_mesh_points_forward(adjacency_head, adjacency_list, face_vertex_list, vertex_pos,
                                query_pos, idxs):
    """-> (nearest_triangle_id P int64, barycentric_coord 3 x P FP64)."""
    ah, ahp = _typed(adjacency_head, np.int32, ctypes.c_int32)
    al, alp = _typed(adjacency_list, np.int32, ctypes.c_int32)
    fv, fvp = _typed(face_vertex_list, np.int32, ctypes.c_int32)
    vp, vpp = _d(vertex_pos)
    q, qp = _d(query_pos)
    ix, ixp = _typed(idxs, np.int64, ctypes.c_int64)
    N_vertex, N_face, P, K = ah.shape[1], fv.shape[1], q.shape[1], ix.shape[1]
    tri = np.zeros(P, dtype=np.int64)
    bary = np.zeros((3, P))
    rc = lib().des_oracle_nearest_forward(
        N_vertex, N_face, P, K, ahp, alp, al.size, fvp, vpp, qp, ixp,
        tri.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        bary.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return tri, bary


def nearest_mesh_points_backward(face_vertex_list, vertex_pos, query_pos, nearest_triangle_id,
                                 grad_barycentric_coord):
    """-> grad_query_pos 3 x P (FP64), upstream gradient from grad_barycentric_coord."""
    fv, fvp = _typed(face_vertex_list, np.int32, ctypes.c_int32)
    vp, vpp = _d(vertex_pos)
    q, qp = _d(query_pos)
    tr, trp = _typed(nearest_triangle_id, np.int64, ctypes.c_int64)
    g, gp = _d(grad_barycentric_coord)
    N_vertex, N_face, P = vp.shape[1], fv.shape[1], q.shape[1]
    out = np.zeros((3, P))
    rc = lib().des_oracle_nearest_backward(
        N_vertex, N_face, P, fvp, vpp, qp, trp, gp,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return out


def extract_bitfield(bitfield, N_bit, i_start, i_end):
    """bitfield N_slot x N int32 -> (i_end - i_start) x N_bit float64 of 0/1."""
    b, bp = _typed(np.asarray(bitfield).view(np.uint32), np.uint32, ctypes.c_uint32)
    N = b.shape[1]
    out = np.zeros((max(i_end - i_start, 0), N_bit))
    lib().des_oracle_extract_bitfield(bp, N, N_bit, i_start, i_end,
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


# ---- env-map utilities (envmap_oracle.c; etc.cu of the reference) ----------------------

def philox(counter, key):
    """Philox4x32-10 of a 4-word counter and a 2-word key -> 4 words."""
    c = np.ascontiguousarray(counter, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib().des_oracle_philox(c.ctypes.data_as(u32p), k.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
    return out


def sample_envmap(query_point, envmap):
    """query_point 3 x P, envmap C x H x W -> C x P (FP64)."""
    q, qp = _d(query_point)
    e = np.ascontiguousarray(envmap, dtype=np.float32)
    C, H, W = e.shape
    P = q.shape[1]
    out = np.zeros((C, P))
    lib().des_oracle_sample_envmap(P, C, H, W, qp, e.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def ggx_sample(r0, r1, alpha):
    """GGX half-vector in the local frame (z up) from two uniforms."""
    out = np.zeros(3)
    lib().des_oracle_ggx_sample(float(r0), float(r1), float(alpha),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def blur_envmap(envmap, tar_H, tar_W, N_samples, alpha):
    """envmap C x H x W -> C x tar_H x tar_W (FP64), Philox4x32-10 samples."""
    e = np.ascontiguousarray(envmap, dtype=np.float32)
    C, H, W = e.shape
    out = np.zeros((C, tar_H, tar_W))
    lib().des_oracle_blur_envmap(C, H, W, tar_H, tar_W, N_samples, float(np.float32(alpha)),
                                 e.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out
