"""Continuous-batching shading server.

Requests (a point set, an env map from a device-resident pool, forward-only
or forward+backward) arrive over time.  The loop

  * admits waiting requests in arrival order while their device memory fits
    the HBM budget; a request that does not fit waits, and later ones may pass
    it only if it has waited less than `max_bypass_s` (so large requests are
    not starved);
  * packs the admitted requests of each kind into one batch (one persistent
    launch per direction, diff_envshade_amd.batch) and launches it on the next
    free stream slot, so batches admitted while others run overlap on the GPU;
  * retires a batch when its completion event has fired, releasing its memory
    budget, and records each request's latency (the device time of the
    batch's completion event minus the request's arrival);
  * when it can launch nothing it does not spin: it blocks on the completion
    event of the oldest batch in flight (requests wait for a slot or for
    memory), or sleeps until the next arrival.

The GPU executor gives each stream slot one workspace of fixed size, and per
kind one HIP graph of the batched launches captured on it once; a batch is
prepared into the slot's workspace (descriptor upload and point order) and the
graph replayed.  The HBM budget covers every byte the server holds on the
device: the slots' workspaces are charged once, up front, and each request in
flight is charged its outputs and gradients; a batch's share of a workspace
(the env-map-gradient partials and point orders of its sets) is bounded by the
workspace size instead.

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
    bytes: int = 0                   # device bytes of the request (outputs, gradients, workspace share)
    ws_bytes: int = 0                # of which its share of a batch workspace
    admitted_s: Optional[float] = None
    done_s: Optional[float] = None
    result: object = None

    @property
    def latency_s(self) -> Optional[float]:
        return None if self.done_s is None else self.done_s - self.arrival_s


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


def request_parts(points: int, texels: int, backward: bool, variant: str = "vis"):
    """(outputs, gradients, workspace share) in device bytes of one request:
    its shading outputs, its gradient outputs (backward only), and its share
    of the batch workspace (env-map-gradient partials, and the point order of a
    set large enough to be sorted), as the library lays them out."""
    out = points * 3 * 4 * (2 if variant in ("v2", "vis") else 1)
    tiles = (points + 63) // 64
    # per-tile fp32 partials, 256-byte aligned (the library's batch workspace
    # holds them for every set)
    ws = _align256(tiles * 3 * texels * 4)
    if points * texels >= ORDER_MIN_PAIRS:
        ws += _order_bytes(points)
    grads = points * (3 + 3 + 3 + 3 + 1) * 4 + 3 * texels * 4 if backward else 0
    return out, grads, ws


def request_bytes(points: int, texels: int, backward: bool, variant: str = "vis") -> int:
    """All device bytes of a request: outputs, gradients and workspace share."""
    return sum(request_parts(points, texels, backward, variant))


class Scheduler:
    """Admission + continuous batching over an executor.

    executor.launch(kind, requests) -> handle   (kind: "fwd" or "fwdbwd")
    executor.poll(handle) -> bool              (True once the batch is done)
    executor.slots -> int                      (batches allowed in flight)

    Optional executor members:
    executor.fixed_bytes             device bytes the executor holds whatever
                                     runs (charged against the budget up front)
    executor.batch_workspace_bytes   the workspace a batch's sets share; when
                                     set, a request is charged r.bytes - r.ws_bytes
                                     and a batch's ws_bytes must fit this
    executor.start(clock)            called once per run with the run's clock
    executor.wait(handle)            block until the batch is done
    executor.completion_time(handle) the batch's completion, in run-clock seconds
    """

    def __init__(self, executor, hbm_budget: int, max_batch: int = 64,
                 max_bypass_s: float = 0.05, clock: Callable[[], float] = time.perf_counter):
        self.ex = executor
        self.hbm_budget = int(hbm_budget)
        self.fixed = int(getattr(executor, "fixed_bytes", 0))
        self.budget = self.hbm_budget - self.fixed
        if self.budget <= 0:
            raise RuntimeError(f"the executor holds {self.fixed} bytes, more than the HBM "
                               f"budget of {self.hbm_budget}")
        self.ws_limit = getattr(executor, "batch_workspace_bytes", None)
        self.max_batch = max_batch
        self.max_bypass_s = max_bypass_s
        self.clock = clock
        self.in_use = 0
        self.peak_in_use = 0
        self.batches = 0
        self.batch_sizes = []
        self.batch_log = []  # (launch time, kind, request ids)

    def charge(self, r: Request) -> int:
        """Budget bytes of a request while in flight."""
        return r.bytes - r.ws_bytes if self.ws_limit is not None else r.bytes

    def _admit(self, waiting, now):
        """Pick requests to launch now: the kind (forward-only or forward +
        backward) of the oldest waiting request, then requests of that kind
        in arrival order while the budget (and the batch workspace) allow."""
        chosen = []
        blocked_since = None
        kind = waiting[0].backward
        ws = 0
        for r in list(waiting):
            if len(chosen) >= self.max_batch:
                break
            if r.backward != kind:
                continue
            c = self.charge(r)
            if c > self.budget:
                raise RuntimeError(f"request {r.rid} needs {c} bytes, above the whole "
                                   f"HBM budget of {self.budget}")
            if self.ws_limit is not None:
                if r.ws_bytes > self.ws_limit:
                    raise RuntimeError(f"request {r.rid} needs {r.ws_bytes} bytes of batch "
                                       f"workspace, above the {self.ws_limit} of a slot")
                if ws + r.ws_bytes > self.ws_limit:
                    break  # the batch workspace is full
            if self.in_use + c <= self.budget:
                if blocked_since is not None and now - blocked_since > self.max_bypass_s:
                    break  # the blocked head has waited long enough: stop passing it
                chosen.append(r)
                self.in_use += c
                ws += r.ws_bytes
                waiting.remove(r)
            elif blocked_since is None:
                blocked_since = r.arrival_s
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        return chosen

    def run(self, requests, realtime: bool = True, on_idle: Optional[Callable] = None):
        """Serve `requests` (sorted by arrival_s, seconds from start).  Returns
        the requests with admitted_s / done_s (seconds from start) filled in.
        on_idle: called instead of waiting when nothing could be launched (a
        simulated executor advances its clock there)."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival_s))
        waiting = collections.deque()
        inflight = []  # (handle, [requests])
        t0 = self.clock()
        rel = lambda: self.clock() - t0  # noqa: E731
        if hasattr(self.ex, "start"):
            self.ex.start(rel)
        timed = hasattr(self.ex, "completion_time")
        while pending or waiting or inflight:
            now = rel()
            if not realtime and not waiting and not inflight and pending:
                now = max(now, pending[0].arrival_s)
            while pending and pending[0].arrival_s <= now:
                waiting.append(pending.popleft())
            still = []
            for h, rs in inflight:
                if self.ex.poll(h):
                    done = self.ex.completion_time(h) if timed else rel()
                    for r in rs:
                        r.done_s = done
                        self.in_use -= self.charge(r)
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
                self.batch_log.append((now, kind, [r.rid for r in chosen]))
                launched = True
            if launched:
                continue
            if on_idle is not None:
                on_idle()
            elif inflight and (waiting or not pending) and hasattr(self.ex, "wait"):
                # requests wait for a slot or for memory, or only batches are
                # left: block until the oldest batch in flight completes
                self.ex.wait(inflight[0][0])
            elif pending and realtime:
                dt = pending[0].arrival_s - rel()
                if dt > 0:
                    time.sleep(dt)
            elif inflight:
                time.sleep(1e-4)
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


def max_overlap(intervals) -> int:
    """Largest number of [start, end) intervals that overlap at one instant."""
    ev = sorted([(a, 1) for a, _ in intervals] + [(b, -1) for _, b in intervals],
                key=lambda x: (x[0], x[1]))
    cur = best = 0
    for _, d in ev:
        cur += d
        best = max(best, cur)
    return best


# ---- GPU executor ------------------------------------------------------------------

class GpuExecutor:
    """Launches batches through diff_envshade_amd.batch.ShadeBatch on `slots`
    streams.  Each slot owns one device workspace of `workspace_bytes` and,
    per kind, one batch object whose forward (+ backward) launches are captured
    once into a HIP graph on that workspace (`graphs=True`); a batch is then
    prepared into the workspace (descriptor upload and point order, on the
    slot's stream) and the graph replayed.  A slot is reused only after its
    previous batch has completed.

    Completion: each batch records a start and an end event on its stream; the
    end event's device time, against a reference event recorded at start(),
    gives the batch's completion time, and the pairs of events form the device
    timeline (`timeline`: slot, kind, requests, start and end in run-clock
    seconds).

    Results: `on_result(request, outputs)` receives each request's output
    tensors when its batch retires; without it they are left in
    request.result.  The executor drops its own references then, so what the
    caller does not keep is freed.

    Launch shape: 4 waves per 64-point tile in both directions by default
    (four tiles per CU, the batched launches' best); 8 waves per tile finish a
    small batch's few tiles sooner but served 16% fewer requests per second
    on the mixed stream at the same p99."""

    def __init__(self, variant: str = "vis", slots: int = 3, ntc_forward: int = 4,
                 ntc_backward: int = 4, workspace_bytes: int = 2 << 30, graphs: bool = True,
                 on_result: Optional[Callable] = None, device=None):
        import torch

        from .batch import ShadeBatch
        self.torch = torch
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.slots = slots
        self.ntc_forward = ntc_forward  # the launch shape: results are bitwise those of a
        self.ntc_backward = ntc_backward  # lone ShadeBatch with these
        self.graphs = graphs
        self.on_result = on_result
        self._free = list(range(slots))
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)]
        self._ws = [torch.empty(int(workspace_bytes), dtype=torch.uint8, device=self.device)
                    for _ in range(slots)]
        self._batches = [{k: ShadeBatch(variant, ntc_forward=ntc_forward,
                                        ntc_backward=ntc_backward, device=self.device,
                                        workspace=self._ws[i])
                          for k in ("fwd", "fwdbwd")} for i in range(slots)]
        self.fixed_bytes = slots * int(workspace_bytes)
        # the descriptor table sits in front of the sets' shares (a few hundred
        # bytes per set); keep 1 MiB of each workspace for it
        self.batch_workspace_bytes = int(workspace_bytes) - (1 << 20)
        self.timeline = []
        self._ref = None
        self._t_ref = 0.0
        self._captured = False

    def capture(self, sample_set: dict):
        """Capture every slot's graphs on a batch of one sample set (a dict of
        inputs with upstream gradients).  Called by start() if not before."""
        if not self.graphs or self._captured:
            return
        fwd_only = {k: v for k, v in sample_set.items() if not k.startswith("grad_")}
        for i in range(self.slots):
            for kind, s in (("fwd", fwd_only), ("fwdbwd", sample_set)):
                b = self._batches[i][kind]
                b.prepare([s], with_backward=(kind == "fwdbwd"))
                b.capture(with_backward=(kind == "fwdbwd"), sort=False)
                b.release()
        self.torch.cuda.synchronize(self.device)
        self._captured = True

    def start(self, clock):
        """Set the time base: `clock()` is the run's clock (seconds)."""
        torch = self.torch
        if self.graphs and not self._captured:
            raise RuntimeError("capture(sample_set) before serving with graphs")
        self._ref = torch.cuda.Event(enable_timing=True)
        self._ref.record(torch.cuda.current_stream(self.device))
        self._ref.synchronize()
        self._t_ref = clock()
        self.timeline = []

    def launch(self, kind, requests):
        torch = self.torch
        slot = self._free.pop(0)
        stream = self._streams[slot]
        b = self._batches[slot][kind]
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            ev0.record(stream)
            sets = [r.payload() if callable(r.payload) else r.payload for r in requests]
            outs = b.prepare(sets, with_backward=(kind == "fwdbwd"))
            if self.graphs:
                b.replay()
            else:
                b.forward()
                if kind == "fwdbwd":
                    b.backward()
            ev1.record(stream)
        return {"slot": slot, "kind": kind, "ev0": ev0, "ev1": ev1, "requests": requests,
                "outs": outs, "done": False}

    def _retire(self, h):
        if h["done"]:
            return
        h["done"] = True
        self._free.append(h["slot"])
        self._batches[h["slot"]][h["kind"]].release()
        for r, o in zip(h["requests"], h["outs"]):
            if self.on_result is not None:
                self.on_result(r, o)
            else:
                r.result = o
        h["outs"] = None
        if self._ref is not None:
            t0 = self._t_ref + self._ref.elapsed_time(h["ev0"]) * 1e-3
            t1 = self._t_ref + self._ref.elapsed_time(h["ev1"]) * 1e-3
            h["t"] = (t0, t1)
            self.timeline.append((h["slot"], h["kind"], [r.rid for r in h["requests"]], t0, t1))

    def poll(self, h):
        if h["done"]:
            return True
        if h["ev1"].query():
            self._retire(h)
            return True
        return False

    def wait(self, h):
        h["ev1"].synchronize()

    def completion_time(self, h):
        return h["t"][1] if "t" in h else None


class PoissonWorkload:
    """Requests with Poisson arrivals at `rate` per second; sizes log-uniform
    in [n_min, n_max]; a `fwd_only` fraction forward-only; env maps drawn
    from a pool of `n_env` maps whose sizes alternate between `env_shapes`.
    burst=(i0, n): requests i0 .. i0+n-1 arrive together, at request i0's
    arrival time (later arrivals keep their Poisson gaps)."""

    def __init__(self, n_requests=2000, rate=1000.0, n_min=2000, n_max=300000, fwd_only=0.3,
                 n_env=32, env_shapes=((16, 32), (32, 64)), seed=0, burst=None):
        import numpy as np
        rng = np.random.default_rng(seed)
        self.env_shapes = [env_shapes[i % len(env_shapes)] for i in range(n_env)]
        gaps = rng.exponential(1.0 / rate, size=n_requests)
        self.arrivals = np.cumsum(gaps) - gaps[0]
        if burst is not None:
            i0, n = burst
            i1 = min(n_requests, i0 + n)
            span = self.arrivals[i1 - 1] - self.arrivals[i0]
            self.arrivals[i0:i1] = self.arrivals[i0]
            self.arrivals[i1:] -= span
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
            parts = request_parts(r.points, r.texels, r.backward, variant)
            r.bytes = sum(parts)
            r.ws_bytes = parts[2]
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
    is arange(o, o+n) into the pool-wide identity face list, and its
    barycentric_coord the first 3n entries of a constant 1/3 array (the
    three vertices of each face coincide, so the weights only need to sum to
    one), so every input is a view of the pool (no per-request copies or
    allocations).  Env maps: one tensor per pool entry.  The numpy originals
    are kept in `host` for checking results.  Everything is built on the
    current stream and synchronised before the constructor returns, so the
    server's streams may read it at once."""

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
        envs = [WL.make_envmap(rng, H, W, 1.0) for H, W in env_shapes]
        self.max_points = max_points
        self.host = {"points": d, "visibility": vis, "grads": g, "envs": envs}
        self.t = {k: torch.from_numpy(v).to(dev) for k, v in d.items()}
        self.vis = torch.from_numpy(vis).to(dev)
        self.grads = {k: torch.from_numpy(v).to(dev) for k, v in g.items()}
        ar = torch.arange(max_points, dtype=torch.int32, device=dev)
        self.fvl = torch.stack([ar, ar, ar]).contiguous()
        self.tri = torch.arange(max_points, dtype=torch.int64, device=dev)
        self.bary = torch.full((3 * max_points,), 1.0 / 3.0, dtype=torch.float32, device=dev)
        self.envs = [torch.from_numpy(e).to(dev) for e in envs]
        self.rot = torch.eye(3, device=dev)
        self.device = dev
        torch.cuda.synchronize(dev)

    def window(self, r: Request, offset: int) -> int:
        return max(0, min(offset, self.max_points - r.points))

    def request_payload(self, r: Request, env_id: int, offset: int = 0) -> dict:
        """Inputs of a request as views of the pool plus its env map."""
        n = r.points
        o = self.window(r, offset)
        p = {k: v[o:o + n] for k, v in self.t.items()}
        p["envmap"] = self.envs[env_id]
        p["light2obj_rotmat"] = self.rot
        p["visibility"] = self.vis
        p["face_vertex_list"] = self.fvl
        p["nearest_triangle_id"] = self.tri[o:o + n]
        p["barycentric_coord"] = self.bary[:3 * n].view(3, n)
        if r.backward:
            p["grad_shading"] = self.grads["grad_shading"][o:o + n]
            p["grad_diffuse_shading"] = self.grads["grad_diffuse_shading"][o:o + n]
        return p

    def host_problem(self, r: Request, env_id: int, offset: int = 0) -> dict:
        """The same request as numpy arrays in the oracle's form (per-point
        visibility words through the identity mesh)."""
        import numpy as np

        from . import workloads as WL
        n = r.points
        o = self.window(r, offset)
        h = self.host
        d = {k: v[o:o + n] for k, v in h["points"].items()}
        env = h["envs"][env_id]
        d["envmap"] = env
        d["light2obj_rotmat"] = np.eye(3, dtype=np.float32)
        ns = (env.shape[1] * env.shape[2] + 31) // 32
        d["visibility"] = np.ascontiguousarray(h["visibility"][:ns, o:o + n])
        d.update(WL.identity_mesh(n))
        # the weights the device reads (see request_payload)
        d["barycentric_coord"] = np.full((3, n), 1.0 / 3.0, dtype=np.float32)
        for k, v in h["grads"].items():
            d[k] = v[o:o + n]
        return d
