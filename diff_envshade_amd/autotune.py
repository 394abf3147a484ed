"""Launch-shape autotuner for the shading kernels.

The only free launch parameter is NTC, the number of waves that split the
texel range of one 64-point tile (1..8).  Small
point sets want many waves per tile to fill 256 CUs; large ones want few, to
save the per-wave setup and the cross-wave reduction.  The tuner times every
candidate on the device with HIP events and keeps the fastest, per batch
signature, in memory and optionally in a JSON file.

    from diff_envshade_amd.autotune import tune_batch
    ntc_f, ntc_b = tune_batch(batch, sets)      # batch: ShadeBatch
"""
from __future__ import annotations

import json
import math
import os
import threading

import torch

FWD_CANDIDATES = (1, 2, 4, 8)
BWD_CANDIDATES = (1, 2, 4, 8)

_cache: dict = {}
_lock = threading.Lock()


def _cache_path():
    return os.environ.get("DES_AUTOTUNE_CACHE")


def _load_file():
    path = _cache_path()
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                _cache.update(json.load(f))
        except (OSError, ValueError):
            pass


def _save_file():
    path = _cache_path()
    if not path:
        return
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(_cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def signature(variant: str, enable_specular: bool, sets) -> str:
    """Tuning key: device, variant and a coarse histogram of (tiles, texels)."""
    dev = torch.cuda.get_device_name()
    texels = {}
    for s in sets:
        T = int(s["envmap"].shape[1] * s["envmap"].shape[2])
        tiles = (int(s["normal"].shape[0]) + 63) // 64
        texels[T] = texels.get(T, 0) + tiles
    parts = [f"T{T}:tiles2^{round(math.log2(max(1, n)))}" for T, n in sorted(texels.items())]
    return f"{dev}|{variant}|spec{int(bool(enable_specular))}|" + ",".join(parts)


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(reps)]
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    ts = sorted(a.elapsed_time(b) for a, b in evs)
    return ts[len(ts) // 2]


def tune_batch(batch, sets, with_backward: bool = True, reps: int = 5, force: bool = False):
    """Pick (ntc_forward, ntc_backward) for `sets` on `batch`; leaves the
    batch configured with the winners and re-prepared for `sets`."""
    key = signature(batch.variant, batch.enable_specular, sets)
    with _lock:
        if not _cache:
            _load_file()
        hit = None if force else _cache.get(key)
    if hit is None:
        timings = {"forward": {}, "backward": {}}
        best_f, best_b = None, None
        for f in FWD_CANDIDATES:
            batch.ntc_forward = f
            batch.prepare(sets, with_backward=False)
            t = _time(batch.forward, reps)
            timings["forward"][f] = t
            if best_f is None or t < timings["forward"][best_f]:
                best_f = f
        batch.ntc_forward = best_f
        if with_backward:
            for b in BWD_CANDIDATES:
                batch.ntc_backward = b
                batch.prepare(sets, with_backward=True)
                t = _time(batch.backward, reps)
                timings["backward"][b] = t
                if best_b is None or t < timings["backward"][best_b]:
                    best_b = b
        else:
            best_b = batch.ntc_backward
        hit = {"ntc_forward": best_f, "ntc_backward": best_b, "ms": timings}
        with _lock:
            _cache[key] = hit
            _save_file()
    batch.ntc_forward = int(hit["ntc_forward"])
    batch.ntc_backward = int(hit["ntc_backward"])
    batch.prepare(sets, with_backward=with_backward)
    return batch.ntc_forward, batch.ntc_backward
