"""Continuous-batching shading server.

Requests (a point set, an env map from a device-resident pool, forward-only
or forward+backward) arrive over time.  The loop

  * admits waiting requests in arrival order while their device memory
    (outputs and backward workspace) fits the HBM budget; a request that does
    not fit waits, and later ones may pass it only if it has waited less than
    `max_bypass_s` (so large requests are not starved);
  * packs the admitted requests of each kind into one batch (one persistent
    launch per direction, diff_envshade_amd.batch) and launches it on the next
    free stream slot, so batches admitted while others run overlap on the GPU;
  * retires a batch when its completion event fires, releasing its memory
    budget, and records each request's latency (completion minus arrival).

Every request's results are bitwise those of the same request shaded alone:
tiles of different point sets never mix and the launch shape is fixed per
server, so batch composition cannot change any result.

The scheduling core (`Scheduler`) is independent of the GPU: it drives an
executor with launch / poll hooks, so its admission and ordering rules are
tested on the CPU with a simulated executor.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time
from typing import Callable, Optional


@dataclasses.dataclass
class Request:
    rid: int
    arrival_s: float
    points: int
    texels: int
    backward: bool
    payload: object = None           # executor-specific inputs
    bytes: int = 0                   # device memory charged while in flight
    admitted_s: Optional[float] = None
    done_s: Optional[float] = None
    result: object = None

    @property
    def latency_s(self) -> Optional[float]:
        return None if self.done_s is None else self.done_s - self.arrival_s


REDUCE_ROWS = 16  # partial rows per stage-1 group of the env-gradient reduce
ORDER_MIN_PAIRS = 1 << 24  # sets of this many point-texel pairs are sorted (both directions)
ORDER_BLOCK = 1024  # points per block of the counting sort
ORDER_KEYS = 1024  # keys (octahedral cells) of the counting sort


def _align256(x: int) -> int:
    return (x + 255) // 256 * 256


def _order_bytes(points: int) -> int:
    """The point-order workspace of a sorted set (order, key/rank words, per-key
    block counts, key totals; the library's order_layout)."""
    nblk_pad = (-(-points // ORDER_BLOCK) + 3) & ~3
    return 2 * _align256(points * 4) + _align256(ORDER_KEYS * nblk_pad * 4) + \
        _align256(ORDER_KEYS * 4)


def request_bytes(points: int, texels: int, backward: bool, variant: str = "vis") -> int:
    """Device bytes a request holds while in flight: its outputs, its share of
    the batch workspace (env-map-gradient partials, and the point order of a set
    large enough to be sorted), plus the backward gradients."""
    out = points * 3 * 4 * (2 if variant in ("v2", "vis") else 1)
    tiles = (points + 63) // 64
    # per-tile fp32 partials, then fp64 sums of REDUCE_ROWS tiles each, both
    # 256-byte aligned (the library's batch workspace holds them for every set)
    ws = _align256(tiles * 3 * texels * 4) + \
        _align256(((tiles + REDUCE_ROWS - 1) // REDUCE_ROWS) * 3 * texels * 8)
    if points * texels >= ORDER_MIN_PAIRS:
        ws += _order_bytes(points)
    if not backward:
        return out + ws
    grads = points * (3 + 3 + 3 + 3 + 1) * 4 + 3 * texels * 4
    return out + grads + ws


class Scheduler:
    """Admission + continuous batching over an executor.

    executor.launch(kind, requests) -> handle   (kind: "fwd" or "fwdbwd")
    executor.poll(handle) -> bool              (True once the batch is done)
    executor.slots -> int                      (batches allowed in flight)
    """

    def __init__(self, executor, hbm_budget: int, max_batch: int = 64,
                 max_bypass_s: float = 0.05, clock: Callable[[], float] = time.perf_counter):
        self.ex = executor
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        self.max_bypass_s = max_bypass_s
        self.clock = clock
        self.in_use = 0
        self.peak_in_use = 0
        self.batches = 0
        self.batch_sizes = []

    def _admit(self, waiting, now):
        """Pick requests to launch now: the kind (forward-only or forward +
        backward) of the oldest waiting request, then requests of that kind
        in arrival order while the budget allows."""
        chosen = []
        blocked_since = None
        kind = waiting[0].backward
        for r in list(waiting):
            if len(chosen) >= self.max_batch:
                break
            if r.backward != kind:
                continue
            if r.bytes > self.budget:
                raise RuntimeError(f"request {r.rid} needs {r.bytes} bytes, above the whole "
                                   f"HBM budget of {self.budget}")
            if self.in_use + r.bytes <= self.budget:
                if blocked_since is not None and now - blocked_since > self.max_bypass_s:
                    break  # the blocked head has waited long enough: stop passing it
                chosen.append(r)
                self.in_use += r.bytes
                waiting.remove(r)
            elif blocked_since is None:
                blocked_since = r.arrival_s
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        return chosen

    def run(self, requests, realtime: bool = True, on_idle: Optional[Callable] = None):
        """Serve `requests` (sorted by arrival_s, seconds from start).  Returns
        the requests with admitted_s / done_s (seconds from start) filled in."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival_s))
        waiting = collections.deque()
        inflight = []  # (handle, [requests])
        t0 = self.clock()
        while pending or waiting or inflight:
            now = self.clock() - t0
            if not realtime and not waiting and not inflight and pending:
                now = max(now, pending[0].arrival_s)
            while pending and pending[0].arrival_s <= now:
                waiting.append(pending.popleft())
            still = []
            for h, rs in inflight:
                if self.ex.poll(h):
                    done = self.clock() - t0
                    for r in rs:
                        r.done_s = done
                        self.in_use -= r.bytes
                else:
                    still.append((h, rs))
            inflight = still
            launched = False
            while waiting and len(inflight) < self.ex.slots:
                chosen = self._admit(waiting, now)
                if not chosen:
                    break
                for r in chosen:
                    r.admitted_s = now
                kind = "fwdbwd" if chosen[0].backward else "fwd"
                inflight.append((self.ex.launch(kind, chosen), chosen))
                self.batches += 1
                self.batch_sizes.append(len(chosen))
                launched = True
            if not launched and on_idle is not None:
                on_idle()
        return requests


def summarize(requests, pairs_per_request: Callable[[Request], float]):
    lat = sorted(r.latency_s for r in requests)

    def pct(q):
        k = (len(lat) - 1) * q
        f = int(math.floor(k))
        c = min(f + 1, len(lat) - 1)
        return lat[f] + (lat[c] - lat[f]) * (k - f)

    start = min(r.arrival_s for r in requests)
    end = max(r.done_s for r in requests)
    pairs = sum(pairs_per_request(r) for r in requests)
    return {
        "requests": len(requests),
        "p50_ms": pct(0.5) * 1e3,
        "p99_ms": pct(0.99) * 1e3,
        "max_ms": lat[-1] * 1e3,
        "makespan_s": end - start,
        "pairs_per_s": pairs / max(end - start, 1e-12),
    }


# ---- GPU executor ------------------------------------------------------------------

class GpuExecutor:
    """Launches batches through diff_envshade_amd.batch.ShadeBatch on `slots`
    streams; each slot owns its own batch objects and workspaces, reused only
    after the slot's previous batch has completed."""

    def __init__(self, variant: str = "vis", slots: int = 3, ntc_forward: int = 8,
                 ntc_backward: int = 8, device=None):
        import torch

        from .batch import ShadeBatch
        self.torch = torch
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.slots = slots
        self._free = list(range(slots))
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)]
        self._batches = [{k: ShadeBatch(variant, ntc_forward=ntc_forward,
                                        ntc_backward=ntc_backward, device=self.device)
                          for k in ("fwd", "fwdbwd")} for _ in range(slots)]

    def launch(self, kind, requests):
        torch = self.torch
        slot = self._free.pop(0)
        stream = self._streams[slot]
        b = self._batches[slot][kind]
        with torch.cuda.stream(stream):
            sets = [r.payload() if callable(r.payload) else r.payload for r in requests]
            outs = b.prepare(sets, with_backward=(kind == "fwdbwd"))
            b.forward()
            if kind == "fwdbwd":
                b.backward()
            ev = torch.cuda.Event()
            ev.record(stream)
        for r, o in zip(requests, outs):
            r.result = o
        return (slot, ev)

    def poll(self, handle):
        slot, ev = handle
        if ev.query():
            self._free.append(slot)
            return True
        return False


class PoissonWorkload:
    """Requests with Poisson arrivals at `rate` per second; sizes log-uniform
    in [n_min, n_max]; a `fwd_only` fraction forward-only; env maps drawn
    from a pool of `n_env` maps whose sizes alternate between `env_shapes`."""

    def __init__(self, n_requests=2000, rate=1000.0, n_min=2000, n_max=300000, fwd_only=0.3,
                 n_env=32, env_shapes=((16, 32), (32, 64)), seed=0):
        import numpy as np
        rng = np.random.default_rng(seed)
        self.env_shapes = [env_shapes[i % len(env_shapes)] for i in range(n_env)]
        gaps = rng.exponential(1.0 / rate, size=n_requests)
        self.arrivals = np.cumsum(gaps) - gaps[0]
        self.sizes = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max),
                                                size=n_requests))).astype(int)
        self.env_ids = rng.integers(0, n_env, size=n_requests)
        self.backward = rng.random(n_requests) >= fwd_only
        self.n_requests = n_requests

    def requests(self, payload_fn=None, variant="vis"):
        out = []
        for i in range(self.n_requests):
            H, W = self.env_shapes[self.env_ids[i]]
            r = Request(rid=i, arrival_s=float(self.arrivals[i]), points=int(self.sizes[i]),
                        texels=H * W, backward=bool(self.backward[i]))
            r.bytes = request_bytes(r.points, r.texels, r.backward, variant)
            if payload_fn is not None:
                r.payload = payload_fn(r, int(self.env_ids[i]))
            out.append(r)
        return out


class DevicePool:
    """Device-resident point sets and env maps that requests slice into.

    Points: one pool of `max_points` points with visibility words for the
    largest env map (64 slots cover 2048 texels; a 512-texel map reads the
    first 16).  A request of n points is the window [o, o+n) of the pool,
    addressed through the reference's vertex API: its nearest_triangle_id
    is arange(o, o+n) into the pool-wide identity face list, so inputs and
    visibility are views of the pool (no per-request copies).  Env maps: one
    tensor per pool entry."""

    def __init__(self, max_points: int, env_shapes, seed: int = 0, device=None):
        import numpy as np
        import torch

        from . import workloads as WL
        dev = torch.device(device) if device is not None else torch.device("cuda", 0)
        rng = np.random.default_rng(seed)
        d = WL.make_points(rng, max_points)
        Hm, Wm = max(env_shapes, key=lambda s: s[0] * s[1])
        vis = WL.make_visibility(rng, d["normal"], Hm, Wm)
        g = WL.upstream_grads(rng, max_points)
        self.max_points = max_points
        self.t = {k: torch.from_numpy(v).to(dev) for k, v in d.items()}
        self.vis = torch.from_numpy(vis).to(dev)
        self.grads = {k: torch.from_numpy(v).to(dev) for k, v in g.items()}
        ar = torch.arange(max_points, dtype=torch.int32, device=dev)
        self.fvl = torch.stack([ar, ar, ar]).contiguous()
        self.tri = torch.arange(max_points, dtype=torch.int64, device=dev)
        self.envs = [torch.from_numpy(WL.make_envmap(rng, H, W, 1.0)).to(dev)
                     for H, W in env_shapes]
        self.rot = torch.eye(3, device=dev)
        self.device = dev
        self._bary = {}

    def _bary_for(self, n):
        import torch
        b = self._bary.get(n)
        if b is None:
            b = torch.zeros((3, n), device=self.device)
            b[0] = 1.0
            if len(self._bary) < 4096:
                self._bary[n] = b
        return b

    def request_payload(self, r: Request, env_id: int, offset: int = 0) -> dict:
        """Inputs of a request as views of the pool plus its env map."""
        n = r.points
        o = max(0, min(offset, self.max_points - n))
        p = {k: v[o:o + n] for k, v in self.t.items()}
        p["envmap"] = self.envs[env_id]
        p["light2obj_rotmat"] = self.rot
        p["visibility"] = self.vis
        p["face_vertex_list"] = self.fvl
        p["nearest_triangle_id"] = self.tri[o:o + n]
        p["barycentric_coord"] = self._bary_for(n)
        if r.backward:
            p["grad_shading"] = self.grads["grad_shading"][o:o + n]
            p["grad_diffuse_shading"] = self.grads["grad_diffuse_shading"][o:o + n]
        return p
