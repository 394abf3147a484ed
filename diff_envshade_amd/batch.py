"""Batched shading: many point sets, each with its own env map and rotation,
in one persistent launch per direction (include/diff_envshade.h des_batch_*).

    b = ShadeBatch("vis", enable_specular=True)
    outs = b.prepare(sets, with_backward=True)   # allocates outputs, uploads descriptors
    b.forward(); b.backward()                    # or: b.capture(); b.replay()

Each set is a dict of CUDA tensors with the reference binding's argument names
(normal, view_dir, albedo, specular_albedo, roughness, envmap,
light2obj_rotmat, and for the visibility variant face_vertex_list,
nearest_triangle_id, barycentric_coord, visibility; grad_shading /
grad_diffuse_shading when with_backward).  prepare() returns, per set, a dict
of its output tensors (shading, diffuse_shading, grad_*).

A set's results equal the single-set binding's with the same launch shape, so
they do not depend on which other sets share the batch.  The descriptor table
lives in device memory: a captured graph replays for any later batch prepared
into the same workspace (same or smaller sizes).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

VARIANTS = {"diffuse": 0, "specular": 1, "clamp": 2, "v2": 3, "vis": 4}

_vp = ctypes.c_void_p


class DesSet(ctypes.Structure):
    """Mirror of des_set (include/diff_envshade.h)."""
    _fields_ = [
        ("P", ctypes.c_int64), ("H", ctypes.c_int32), ("W", ctypes.c_int32),
        ("N_face", ctypes.c_int32), ("N_vertex", ctypes.c_int32), ("N_slot", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("normal", _vp), ("view_dir", _vp), ("albedo", _vp), ("specular_albedo", _vp),
        ("roughness", _vp), ("envmap", _vp), ("light2obj_rotmat", _vp),
        ("face_vertex_list", _vp), ("nearest_triangle_id", _vp), ("barycentric_coord", _vp),
        ("visibility", _vp), ("shading", _vp), ("diffuse_shading", _vp),
        ("grad_shading", _vp), ("grad_diffuse_shading", _vp), ("grad_normal", _vp),
        ("grad_view_dir", _vp), ("grad_albedo", _vp), ("grad_specular_albedo", _vp),
        ("grad_roughness", _vp), ("grad_envmap", _vp),
    ]


def _bind(L):
    if getattr(L, "_batch_bound", False):
        return L
    sp = ctypes.POINTER(DesSet)
    szp = ctypes.POINTER(ctypes.c_size_t)
    i32, c_int = ctypes.c_int32, ctypes.c_int
    L.des_batch_plan.argtypes = [i32, sp, i32, szp, szp]
    L.des_batch_plan.restype = c_int
    L.des_batch_prepare.argtypes = [i32, c_int, c_int, sp, i32, i32, i32, c_int, _vp,
                                    ctypes.c_size_t, _vp, ctypes.c_size_t, _vp]
    L.des_batch_prepare.restype = c_int
    L.des_batch_forward.argtypes = [i32, c_int, c_int, i32, _vp, _vp]
    L.des_batch_forward.restype = c_int
    L.des_batch_backward.argtypes = [i32, c_int, c_int, i32, _vp, _vp]
    L.des_batch_backward.restype = c_int
    L.des_batch_sort.argtypes = [_vp, _vp]
    L.des_batch_sort.restype = c_int
    L._batch_bound = True
    return L


def _p(t):
    return t.data_ptr() if (t is not None and t.numel() > 0) else None


class ShadeBatch:
    """One persistent forward / backward launch over many point sets."""

    def __init__(self, variant: str = "vis", enable_diffuse: bool = True,
                 enable_specular: bool = True, ntc_forward: int = 8, ntc_backward: int = 8,
                 device=None, workspace=None):
        """workspace: a device uint8 tensor to prepare batches into (fixed size;
        a batch that needs more is refused); by default the batch allocates one
        and grows it as needed."""
        self.variant = variant
        self.vid = VARIANTS[variant]
        self.enable_diffuse = bool(enable_diffuse)
        self.enable_specular = bool(enable_specular)
        self.ntc_forward = int(ntc_forward)
        self.ntc_backward = int(ntc_backward)
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.L = _bind(_lib.load())
        self.fixed_workspace = workspace is not None
        self.workspace = workspace
        self.staging = None
        self._staging_ready = None
        self.graph = None
        self.outputs = []
        self._sets = None

    # -- setup ---------------------------------------------------------------
    def _descriptors(self, sets, with_backward):
        arr = (DesSet * max(1, len(sets)))()
        outs = []
        dout = self.variant in ("v2", "vis")
        for i, s in enumerate(sets):
            n = s["normal"]
            P = n.shape[0]
            C, H, W = s["envmap"].shape
            if C != 3:
                raise RuntimeError("Channel number not supported!")
            dev, dt = n.device, n.dtype
            o = {"shading": torch.empty((P, 3), dtype=dt, device=dev)}
            if dout:
                o["diffuse_shading"] = torch.empty((P, 3), dtype=dt, device=dev)
            if with_backward:
                o["grad_normal"] = torch.empty((P, 3), dtype=dt, device=dev)
                o["grad_albedo"] = torch.empty((P, 3), dtype=dt, device=dev)
                o["grad_envmap"] = torch.empty((3, H, W), dtype=dt, device=dev)
                if self.variant != "diffuse":
                    o["grad_view_dir"] = torch.empty((P, 3), dtype=dt, device=dev)
                    o["grad_specular_albedo"] = torch.empty((P, 3), dtype=dt, device=dev)
                    o["grad_roughness"] = torch.empty((P, 1), dtype=dt, device=dev)
            d = arr[i]
            d.P, d.H, d.W = P, H, W
            if self.variant == "vis":
                d.N_face = s["face_vertex_list"].shape[1]
                d.N_vertex = s["visibility"].shape[1]
                d.N_slot = s["visibility"].shape[0]
            for k in ("normal", "view_dir", "albedo", "specular_albedo", "roughness", "envmap",
                      "light2obj_rotmat", "face_vertex_list", "nearest_triangle_id",
                      "barycentric_coord", "visibility", "grad_shading",
                      "grad_diffuse_shading"):
                t = s.get(k)
                if t is not None:
                    if not (t.is_cuda and t.is_contiguous()):
                        raise RuntimeError(f"set {i}: {k} must be a contiguous CUDA tensor")
                    setattr(d, k, _p(t))
            for k, t in o.items():
                setattr(d, k, _p(t))
            outs.append(o)
        return arr, outs

    def prepare(self, sets, with_backward: bool = True):
        """Validate the sets, allocate their outputs and upload the descriptor
        table on the current stream.  Returns one dict of outputs per set."""
        arr, outs = self._descriptors(sets, with_backward)
        n = len(sets)
        ws = ctypes.c_size_t()
        st = ctypes.c_size_t()
        _lib.check(self.L.des_batch_plan(self.vid, arr, n, ctypes.byref(ws), ctypes.byref(st)))
        if self.workspace is None or self.workspace.numel() < ws.value:
            if self.graph is not None or self.fixed_workspace:
                have = 0 if self.workspace is None else self.workspace.numel()
                raise RuntimeError(f"batch needs a workspace of {ws.value} bytes, the fixed "
                                   f"one holds {have}")
            self.workspace = torch.empty(max(ws.value, 256), dtype=torch.uint8,
                                         device=self.device)
        # the previous upload must have consumed the staging bytes before they are
        # rewritten, and before the old pinned block goes back to torch's cache
        # (which cannot see that HIP still reads it: the pointer went through ctypes)
        if self._staging_ready is not None:
            self._staging_ready.synchronize()
        if self.staging is None or self.staging.numel() < st.value:
            self.staging = torch.empty(max(st.value, 256), dtype=torch.uint8, pin_memory=True)
        stream = torch.cuda.current_stream(self.device)
        _lib.check(self.L.des_batch_prepare(
            self.vid, int(self.enable_diffuse), int(self.enable_specular), arr, n,
            self.ntc_forward, self.ntc_backward, int(bool(with_backward)),
            self.staging.data_ptr(), self.staging.numel(), self.workspace.data_ptr(),
            self.workspace.numel(), stream.cuda_stream))
        self._staging_ready = torch.cuda.Event()
        self._staging_ready.record(stream)
        self.outputs = outs
        self._sets = sets  # keep inputs alive while the device may read them
        return outs

    # -- launches --------------------------------------------------------------
    def forward(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.L.des_batch_forward(self.vid, int(self.enable_diffuse),
                                            int(self.enable_specular), self.ntc_forward,
                                            self.workspace.data_ptr(), s))

    def backward(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.L.des_batch_backward(self.vid, int(self.enable_diffuse),
                                             int(self.enable_specular), self.ntc_backward,
                                             self.workspace.data_ptr(), s))

    def sort(self):
        """Recompute the point order of the prepared sets from the normals'
        current values (prepare() computes it once; des_batch_sort)."""
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.L.des_batch_sort(self.workspace.data_ptr(), s))

    def release(self):
        """Drop the references to the last batch's inputs and outputs (call once
        its launches have completed; the workspace stays)."""
        self.outputs = []
        self._sets = None

    def capture(self, with_backward: bool = True, sort: bool = True):
        """Capture sort + forward (+ backward) into a HIP graph; replay() reruns
        it for whatever batch was last prepared into this workspace.  sort=False
        leaves the point order out of the graph (prepare() computes it)."""
        if self.workspace is None:
            raise RuntimeError("prepare() a batch before capturing")
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            if sort:
                self.sort()
            self.forward()
            if with_backward:
                self.backward()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()
