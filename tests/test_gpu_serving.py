"""Continuous-batching server on the GPU at the serving configuration: 2000
Poisson requests, N log-uniform in [2k, 300k], 30% forward-only, env maps from
a pool of 32 device-resident 16x32 / 32x64 maps, a 64 GB HBM budget, with a
burst of simultaneous arrivals.  Every request completes; each result is
bitwise the same as the same request shaded alone; a sample of results (the
largest, a 300k-point request on a 2048-texel map, and forward-only ones) is
within 1e-5 of the FP64 oracle; batches of several requests form and batches
on different stream slots overlap on the device; the device bytes the server
allocates stay within what it charges against the budget."""
import numpy as np
import pytest
import torch

from _helpers import TOL, compare, oracle_run

pytestmark = pytest.mark.gpu

FWD = ["shading", "diffuse_shading"]
BWD = ["grad_normal", "grad_view_dir", "grad_albedo", "grad_specular_albedo", "grad_roughness",
       "grad_envmap"]


def _alone(b, payload, backward):
    o = b.prepare([payload], with_backward=backward)
    b.forward()
    if backward:
        b.backward()
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in o[0].items()}
    b.release()
    return res


def _serve(pool, wl, offsets, budget, slots=3, workspace_bytes=2 << 30):
    from diff_envshade_amd.serving import GpuExecutor, Request, Scheduler
    reqs = wl.requests(lambda r, e: (lambda: pool.request_payload(r, e, int(offsets[r.rid]))))
    got = {}

    def on_result(r, o):
        got[r.rid] = {k: v.cpu().numpy() for k, v in o.items()}

    def requested():
        torch.cuda.synchronize()
        return torch.cuda.memory_stats()["requested_bytes.all.current"]

    before = requested()
    ex = GpuExecutor("vis", slots=slots, workspace_bytes=workspace_bytes, on_result=on_result)
    ex.capture(pool.request_payload(Request(-1, 0.0, 64, 512, True), 0, 0))
    setup = requested() - before
    # the slots' workspaces, plus the 16 bytes of random-generator state (seed
    # and offset) that the first HIP-graph capture of a process allocates
    assert ex.fixed_bytes <= setup <= ex.fixed_bytes + 1024, (setup, ex.fixed_bytes)
    base = requested()
    torch.cuda.reset_peak_memory_stats()
    sch = Scheduler(ex, budget)
    print(f"serving {len(reqs)} requests on {slots} slots", flush=True)
    sch.run(reqs)
    torch.cuda.synchronize()
    peak = torch.cuda.memory_stats()["requested_bytes.all.peak"] - base
    return reqs, got, ex, sch, peak


def _check_alone(pool, wl, offsets, reqs, got, ex):
    from diff_envshade_amd.batch import ShadeBatch
    b = ShadeBatch("vis", ntc_forward=ex.ntc_forward, ntc_backward=ex.ntc_backward)
    for r in reqs:
        a = _alone(b, pool.request_payload(r, int(wl.env_ids[r.rid]), int(offsets[r.rid])),
                   r.backward)
        assert set(a) == set(got[r.rid]), r.rid
        for k in a:
            assert np.array_equal(a[k], got[r.rid][k]), (r.rid, k)


def _check_oracle(pool, wl, offsets, r, res):
    d = pool.host_problem(r, int(wl.env_ids[r.rid]), int(offsets[r.rid]))
    print(f"oracle check: request {r.rid}, {r.points} points x {r.texels} texels, "
          f"backward {r.backward}", flush=True)
    fwd = [res[k] for k in FWD]
    if r.backward:
        ref = oracle_run("vis", d)
        bwd = [res[k] for k in BWD]
        return compare("vis", (fwd, bwd), ref, tol=TOL)
    from oracle import oracle as O
    sh, ds = O.forward("vis", d)
    return compare("vis", (fwd, []), ([sh, ds], []), tol=TOL)


def test_serving_config_alone_vs_mix_oracle_overlap_and_memory():
    from diff_envshade_amd.serving import DevicePool, PoissonWorkload, max_overlap, summarize
    shapes = [(16, 32), (32, 64)] * 16
    pool = DevicePool(max_points=300000, env_shapes=shapes, seed=0)
    # requests 600..899 arrive at once: a burst on top of the Poisson stream
    wl = PoissonWorkload(n_requests=2000, rate=2000.0, n_min=2000, n_max=300000, fwd_only=0.3,
                         n_env=32, env_shapes=((16, 32), (32, 64)), seed=21, burst=(600, 300))
    # the largest request: 300 000 points on a 32 x 64 map, forward + backward
    i_big = int(np.argmax(wl.sizes))
    wl.sizes[i_big] = 300000
    wl.env_ids[i_big] = 1
    wl.backward[i_big] = True
    rng = np.random.default_rng(22)
    offsets = rng.integers(0, 300000, size=wl.n_requests)
    budget = 64 << 30
    reqs, got, ex, sch, peak = _serve(pool, wl, offsets, budget)

    assert all(r.done_s is not None and r.done_s >= r.arrival_s for r in reqs)
    assert len(got) == len(reqs)
    assert sch.in_use == 0
    # memory: what the server allocated while serving is at most what it
    # charged, and with its workspaces within the budget
    assert sch.peak_in_use <= sch.budget
    assert peak <= sch.peak_in_use, (peak, sch.peak_in_use)
    assert ex.fixed_bytes + sch.peak_in_use <= budget
    # batching: the burst forms batches of several requests
    burst = set(range(600, 900))
    sizes = [len(ids) for _, _, ids in sch.batch_log if burst.intersection(ids)]
    assert np.mean(sizes) > 1.0 and max(sizes) > 1
    assert max(sch.batch_sizes) > 1
    # overlap on the device: batches of different slots run at the same time
    iv = [(t0, t1) for _, _, _, t0, t1 in ex.timeline]
    assert len(iv) == sch.batches
    assert max_overlap(iv) >= 2
    slots_busy = {s for s, _, _, _, _ in ex.timeline}
    assert len(slots_busy) >= 2
    s = summarize(reqs, lambda r: r.points * r.texels)
    print(f"served: p50 {s['p50_ms']:.2f} ms, p99 {s['p99_ms']:.2f} ms, {sch.batches} batches, "
          f"mean batch {np.mean(sch.batch_sizes):.2f}, max overlap {max_overlap(iv)}", flush=True)

    # FP64 oracle on a sample: the largest, two forward-only, two others
    fwd_only = [r for r in reqs if not r.backward]
    others = [r for r in reqs if r.backward and r.rid != i_big]
    sample = [reqs[i_big]] + fwd_only[:2] + others[:2]
    assert reqs[i_big].points == 300000 and reqs[i_big].texels == 2048
    for r in sample:
        _check_oracle(pool, wl, offsets, r, got[r.rid])

    # every request alone gives bitwise the result it got in the mix
    _check_alone(pool, wl, offsets, reqs, got, ex)


def test_tight_budget_bounds_what_the_server_allocates():
    # a budget a few batches wide: admission holds requests back, and the
    # device bytes allocated while serving stay within the charges
    from diff_envshade_amd.serving import DevicePool, PoissonWorkload
    shapes = [(16, 32), (32, 64)] * 2
    pool = DevicePool(max_points=60000, env_shapes=shapes, seed=5)
    wl = PoissonWorkload(n_requests=150, rate=20000.0, n_min=2000, n_max=60000, fwd_only=0.3,
                         n_env=4, seed=6)
    rng = np.random.default_rng(7)
    offsets = rng.integers(0, 60000, size=wl.n_requests)
    ws = 256 << 20
    budget = 2 * ws + (48 << 20)
    reqs, got, ex, sch, peak = _serve(pool, wl, offsets, budget, slots=2, workspace_bytes=ws)
    assert all(r.done_s is not None for r in reqs)
    assert sch.peak_in_use <= budget - ex.fixed_bytes
    assert sch.peak_in_use > (budget - ex.fixed_bytes) // 2  # the budget was what bound
    assert peak <= sch.peak_in_use, (peak, sch.peak_in_use)
    assert ex.fixed_bytes + sch.peak_in_use <= budget
    _check_alone(pool, wl, offsets, reqs, got, ex)
