"""Seeded synthetic inputs for the shader workloads (numpy only).

The same generator feeds the parity tests, the oracle and bench.py, so every
consumer sees bit-identical inputs for a given (workload, seed).

Point sets: unit normals uniform on the sphere; view directions
normalize((0,0,3) - pos) with pos ~ U[-0.5,0.5]^3; albedo U[0,1]^3; roughness
U[0.1,1] (P x 1); specular intensity s ~ U[0,0.1] broadcast to 3 channels.
Env maps: lat-long 3 x H x W, lognormal texels.  Visibility: T bits per point,
bit t = Bernoulli(p) AND [l_t . n > 0], packed little-endian into 32-bit words
laid out [N_slot, N] (word for slot s of point p at s*N + p; bit t%32 of slot
t/32), which is the reference's per-vertex layout (bruteforce_shader.cu:2181,
2186) with one vertex per point.
"""
from __future__ import annotations

import math

import numpy as np


def texel_dirs(H: int, W: int, rotmat=None) -> tuple[np.ndarray, np.ndarray]:
    """Rotated texel-centre directions (T x 3) and solid angles (T,), FP64.

    Storage texel t -> row r = t // W, column x = t % W, flipped row
    iy = H-1-r (top row = +y), direction (cos th cos ph, sin th, cos th sin ph)
    with th = ((iy+0.5)/H - 0.5) pi and ph = 2 pi (x+0.5)/W
    (bruteforce_shader.cu:133-147, config.h:11).
    """
    t = np.arange(H * W)
    x = t % W
    iy = H - 1 - t // W
    th = ((iy + 0.5) / H - 0.5) * math.pi
    ph = (x + 0.5) / W * 2.0 * math.pi
    d = np.stack([np.cos(th) * np.cos(ph), np.sin(th), np.cos(th) * np.sin(ph)], axis=1)
    dw = (np.sin(math.pi / H * (iy + 1) - math.pi / 2) - np.sin(math.pi / H * iy - math.pi / 2)) \
        * (2 * math.pi / W)
    M = np.eye(3) if rotmat is None else np.asarray(rotmat, dtype=np.float64)
    return d @ M.T, dw


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q.astype(np.float32)


def make_points(rng: np.random.Generator, n: int) -> dict:
    g = rng.normal(size=(n, 3))
    normal = g / np.linalg.norm(g, axis=1, keepdims=True)
    pos = rng.uniform(-0.5, 0.5, size=(n, 3))
    v = np.array([0.0, 0.0, 3.0]) - pos
    view = v / np.linalg.norm(v, axis=1, keepdims=True)
    albedo = rng.uniform(0.0, 1.0, size=(n, 3))
    rough = rng.uniform(0.1, 1.0, size=(n, 1))
    spec = np.repeat(rng.uniform(0.0, 0.1, size=(n, 1)), 3, axis=1)
    f = np.float32
    return {
        "normal": normal.astype(f), "view_dir": view.astype(f), "albedo": albedo.astype(f),
        "specular_albedo": spec.astype(f), "roughness": rough.astype(f),
    }


def make_envmap(rng: np.random.Generator, H: int, W: int, sigma: float = 1.0,
                n_bright: int = 0, bright_scale: float = 50.0) -> np.ndarray:
    env = np.exp(rng.normal(0.0, sigma, size=(3, H, W)))
    if n_bright:
        idx = rng.choice(H * W, size=n_bright, replace=False)
        flat = env.reshape(3, H * W)
        flat[:, idx] *= bright_scale
    return env.astype(np.float32)


def pack_bits(bits: np.ndarray) -> np.ndarray:
    """bits (N x T) bool -> words (ceil(T/32) x N) int32, bit t%32 of slot t//32."""
    n, T = bits.shape
    n_slot = (T + 31) // 32
    padded = np.zeros((n, n_slot * 32), dtype=bool)
    padded[:, :T] = bits
    by = np.packbits(padded.reshape(n, n_slot * 4, 8), axis=2, bitorder="little").reshape(n, -1)
    words = by.view("<u4").reshape(n, n_slot)
    return np.ascontiguousarray(words.T).view(np.int32)


def unpack_bits(words: np.ndarray, T: int) -> np.ndarray:
    """Inverse of pack_bits: words (N_slot x N) -> bits (N x T) bool."""
    w = np.ascontiguousarray(np.asarray(words).view(np.uint32).T)
    n = w.shape[0]
    by = w.astype("<u4").view(np.uint8).reshape(n, -1)
    return np.unpackbits(by, axis=1, bitorder="little")[:, :T].astype(bool)


def make_visibility(rng: np.random.Generator, normal: np.ndarray, H: int, W: int,
                    rotmat=None, p: float = 0.7, chunk: int = 16384) -> np.ndarray:
    """Visibility words (N_slot x N) int32: Bernoulli(p) AND upper hemisphere."""
    l, _ = texel_dirs(H, W, rotmat)
    n = normal.shape[0]
    T = H * W
    out = np.empty(((T + 31) // 32, n), dtype=np.int32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        up = (normal[s:e].astype(np.float64) @ l.T) > 0.0
        keep = rng.random(size=(e - s, T)) < p
        out[:, s:e] = pack_bits(up & keep)
    return out


def identity_mesh(n: int) -> dict:
    """Reference-API mapping of per-point visibility: one vertex and one
    degenerate face per point, barycentric (1,0,0), so vis == bit exactly."""
    ar = np.arange(n, dtype=np.int32)
    bary = np.zeros((3, n), dtype=np.float32)
    bary[0] = 1.0
    return {
        "face_vertex_list": np.stack([ar, ar, ar]),
        "nearest_triangle_id": np.arange(n, dtype=np.int64),
        "barycentric_coord": bary,
    }


def random_mesh(rng: np.random.Generator, n: int, n_vertex: int, n_face: int) -> dict:
    """General vertex-indexed visibility inputs: random faces over n_vertex
    vertices, random triangle per point, random barycentrics summing to 1."""
    fvl = rng.integers(0, n_vertex, size=(3, n_face)).astype(np.int32)
    tri = rng.integers(0, n_face, size=n).astype(np.int64)
    b = rng.random(size=(3, n)) + 1e-3
    b = (b / b.sum(axis=0, keepdims=True)).astype(np.float32)
    return {"face_vertex_list": fvl, "nearest_triangle_id": tri, "barycentric_coord": b}


def upstream_grads(rng: np.random.Generator, n: int) -> dict:
    return {
        "grad_shading": rng.normal(size=(n, 3)).astype(np.float32),
        "grad_diffuse_shading": rng.normal(size=(n, 3)).astype(np.float32),
    }


def plain_small(seed: int = 0, n: int = 4096, H: int = 16, W: int = 32, rotmat=None) -> dict:
    """Plain method, small: 4096 points, 16x32 lognormal(0,1) env map, no
    visibility, identity rotation unless given."""
    rng = np.random.default_rng(seed)
    d = make_points(rng, n)
    d["envmap"] = make_envmap(rng, H, W, 1.0)
    d["light2obj_rotmat"] = (np.eye(3, dtype=np.float32) if rotmat is None
                             else np.asarray(rotmat, dtype=np.float32))
    d.update(upstream_grads(rng, n))
    return d


def dropin_headline(seed: int = 0, n: int = 70000, H: int = 16, W: int = 32) -> dict:
    """Drop-in headline: 70,000 points, 16x32 lognormal(0,1) env with 3 bright
    texels x50, 512-bit per-point visibility (16 words) Bernoulli(0.7) masked to
    the normal's hemisphere, passed through the reference's vertex API with the
    identity mesh."""
    rng = np.random.default_rng(seed)
    d = make_points(rng, n)
    d["envmap"] = make_envmap(rng, H, W, 1.0, n_bright=3)
    d["light2obj_rotmat"] = np.eye(3, dtype=np.float32)
    d["visibility"] = make_visibility(rng, d["normal"], H, W)
    d.update(identity_mesh(n))
    d.update(upstream_grads(rng, n))
    return d


def make_visibility_device(normal, H: int, W: int, rotmat=None, p: float = 0.7, seed: int = 0,
                           chunk: int = 32768):
    """Device-side twin of make_visibility for large point sets (torch):
    Bernoulli(p) AND upper hemisphere, packed [N_slot, N] int32 words.  Uses
    torch's seeded generator, so the bits differ from make_visibility's."""
    import torch
    dev = normal.device
    l, _ = texel_dirs(H, W, rotmat)
    lt = torch.from_numpy(l).to(dev, torch.float32)
    n = normal.shape[0]
    T = H * W
    ns = (T + 31) // 32
    out = torch.empty((ns, n), dtype=torch.int32, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    weights = (torch.ones(32, dtype=torch.int64, device=dev) << torch.arange(32, device=dev))
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        up = (normal[s:e].to(torch.float32) @ lt.T) > 0
        keep = torch.rand((e - s, T), generator=g, device=dev) < p
        bits = torch.zeros((e - s, ns * 32), dtype=torch.int64, device=dev)
        bits[:, :T] = (up & keep).to(torch.int64)
        words = (bits.view(e - s, ns, 32) * weights).sum(-1)
        words = torch.where(words >= 2 ** 31, words - 2 ** 32, words)
        out[:, s:e] = words.to(torch.int32).T
    return out


def log_uniform_sizes(rng: np.random.Generator, n: int, lo: int, hi: int) -> np.ndarray:
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))).astype(int)


def uneven_batch(seed: int = 0, n_sets: int = 48, n_min: int = 2000, n_max: int = 300000,
                 device="cuda"):
    """Uneven batch: point sets with N log-uniform in [n_min, n_max], each with
    its own 16x32 or 32x64 (p = 0.5) lognormal(0,1) env map, visibility as in
    the drop-in headline (per-point words through the identity mesh), and its
    own materials and upstream gradients.  Returns a list of dicts of device
    tensors."""
    import torch
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(rng, n_sets, n_min, n_max)
    sets = []
    for i, n in enumerate(sizes):
        H, W = (16, 32) if rng.random() < 0.5 else (32, 64)
        d = make_points(rng, int(n))
        d["envmap"] = make_envmap(rng, H, W, 1.0)
        d["light2obj_rotmat"] = np.eye(3, dtype=np.float32)
        d.update(identity_mesh(int(n)))
        d.update(upstream_grads(rng, int(n)))
        t = {k: torch.from_numpy(v).to(device) for k, v in d.items()}
        t["visibility"] = make_visibility_device(t["normal"], H, W, seed=seed * 1000 + i)
        sets.append(t)
    return sets


# ---- meshes and nearest-mesh-point queries ---------------------------------------------

def icosphere(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: vertices (V, 3) float64, faces (F, 3) int32 with
    V = 10 * 4^subdiv + 2 and F = 20 * 4^subdiv, counter-clockwise outward."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.stack(verts), np.asarray(faces, dtype=np.int32)


def vertex_adjacency(faces: np.ndarray, n_vertex: int) -> tuple[np.ndarray, np.ndarray]:
    """Faces around each vertex in the reference's layout (etc.cu:779-806):
    adjacency_head 2 x N_vertex int32 (row 0 the face count of vertex v, row 1
    its inclusive prefix sum) and adjacency_list int32, vertex v's faces at
    [head[1][v] - head[0][v], head[1][v]) in increasing face order."""
    F = faces.shape[0]
    vid = faces.reshape(-1).astype(np.int64)
    fid = np.repeat(np.arange(F, dtype=np.int64), 3)
    order = np.lexsort((fid, vid))
    count = np.bincount(vid, minlength=n_vertex).astype(np.int32)
    head = np.stack([count, np.cumsum(count, dtype=np.int64).astype(np.int32)])
    return head, fid[order].astype(np.int32)


def mesh_queries(seed: int = 0, subdiv: int = 5, n: int = 70000, K: int = 8,
                 jitter: float = 0.02, spread: float = 0.05) -> dict:
    """Nearest-mesh-point workload: an icosphere of the given subdivision with
    its vertices moved radially by U[-jitter, jitter] (a non-convex closed
    surface), n query points at radius 1 + N(0, spread) around it, and the K
    nearest vertices of each query (idxs, ascending distance, from a k-d
    tree).  Arrays in the reference's layouts: face_vertex_list 3 x F,
    vertex_pos 3 x V, query_pos 3 x n, idxs n x K int64."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    verts, faces = icosphere(subdiv)
    verts = verts * (1.0 + rng.uniform(-jitter, jitter, size=(verts.shape[0], 1)))
    verts = verts.astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    q = (d * (1.0 + spread * rng.normal(size=(n, 1)))).astype(np.float32)
    head, adj = vertex_adjacency(faces, verts.shape[0])
    _, idxs = cKDTree(verts.astype(np.float64)).query(q.astype(np.float64), k=K)
    idxs = np.asarray(idxs, dtype=np.int64).reshape(n, K)
    return {
        "adjacency_head": head,
        "adjacency_list": adj,
        "face_vertex_list": np.ascontiguousarray(faces.T),
        "vertex_pos": np.ascontiguousarray(verts.T),
        "query_pos": np.ascontiguousarray(q.T),
        "idxs": idxs,
    }


def mesh_shading(seed: int = 0, subdiv: int = 5, n: int = 70000, K: int = 8, H: int = 16,
                 W: int = 32) -> dict:
    """Vertex-indexed visibility at mesh scale: the points and surface of
    mesh_queries, outward normals (the point's direction from the centre), the
    drop-in headline's materials, 16x32 env map with 3 bright texels and
    upstream gradients, and per-vertex 512-bit visibility (Bernoulli(0.7)
    masked to the vertex's outward hemisphere, N_slot x N_vertex).  The
    nearest face and barycentrics of each point are left to the caller (the
    nearest-mesh-point search); the mesh arrays are under "mesh"."""
    m = mesh_queries(seed=seed, subdiv=subdiv, n=n, K=K)
    rng = np.random.default_rng(seed + 1)
    d = make_points(rng, n)
    q = m["query_pos"].T.astype(np.float64)
    d["normal"] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    d["envmap"] = make_envmap(rng, H, W, 1.0, n_bright=3)
    d["light2obj_rotmat"] = np.eye(3, dtype=np.float32)
    vp = m["vertex_pos"].T.astype(np.float64)
    d["visibility"] = make_visibility(rng, vp / np.linalg.norm(vp, axis=1, keepdims=True), H, W)
    d["face_vertex_list"] = m["face_vertex_list"]
    d.update(upstream_grads(rng, n))
    d["mesh"] = m
    return d
