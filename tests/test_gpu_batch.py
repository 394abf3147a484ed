"""Batched persistent launches (GPU): every set of a batch matches the oracle,
matches the single-set binding with the same launch shape bit for bit, and a
captured graph replays correctly for later batches."""
import numpy as np
import pytest
import torch

from _helpers import TOL, compare, gpu_inputs, gpu_run, oracle_run, problem

pytestmark = pytest.mark.gpu


def _sets(specs, seed=0):
    """specs: list of (n, H, W) -> (numpy problems, device dicts)."""
    nps, devs = [], []
    for i, (n, H, W) in enumerate(specs):
        d = problem(seed=seed + i, n=n, H=H, W=W, vis="identity")
        nps.append(d)
        devs.append(gpu_inputs(d))
    return nps, devs


def _np(outs, variant="vis"):
    fwd = [outs["shading"], outs["diffuse_shading"]]
    bwd = [outs[k] for k in ("grad_normal", "grad_view_dir", "grad_albedo",
                             "grad_specular_albedo", "grad_roughness", "grad_envmap")]
    return [x.cpu().numpy() for x in fwd], [x.cpu().numpy() for x in bwd]


SPECS = [(300, 16, 32), (5000, 32, 64), (1, 4, 8), (0, 16, 32), (2049, 16, 32), (777, 8, 16)]


@pytest.mark.parametrize("ntc_f,ntc_b", [(8, 8), (4, 2), (1, 1)])
def test_batch_matches_oracle_and_single_set(ntc_f, ntc_b):
    from diff_envshade_amd import _lib
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = _sets(SPECS, seed=20)
    b = ShadeBatch("vis", enable_specular=True, ntc_forward=ntc_f, ntc_backward=ntc_b)
    outs = b.prepare(devs, with_backward=True)
    b.forward()
    b.backward()
    torch.cuda.synchronize()
    try:
        _lib.set_tiling_override(ntc_f, ntc_b)
        for d, o in zip(nps, outs):
            got = _np(o)
            if d["normal"].shape[0] > 0:
                compare("vis", got, oracle_run("vis", d))
            single = gpu_run("vis", d)
            for a, s in zip(got[0] + got[1], single[0] + single[1]):
                assert np.array_equal(a.reshape(s.shape), s)
    finally:
        _lib.set_tiling_override(0, 0)


def test_batch_results_independent_of_batch_mates():
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = _sets(SPECS, seed=30)
    b = ShadeBatch("vis")
    b.prepare(devs)
    b.forward()
    b.backward()
    torch.cuda.synchronize()
    together = [_np(o) for o in b.outputs]
    for i, dev in enumerate(devs):
        alone = ShadeBatch("vis")
        o = alone.prepare([dev])
        alone.forward()
        alone.backward()
        torch.cuda.synchronize()
        a = _np(o[0])
        for x, y in zip(a[0] + a[1], together[i][0] + together[i][1]):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("order", [0, 1], ids=["order-auto", "sorted"])
def test_graph_capture_and_replay_new_batches(order):
    # the graph holds sort + forward + backward; every size is read from the
    # workspace, so it replays for new batches (sorted ones too, order=1)
    from diff_envshade_amd import _lib
    from diff_envshade_amd.batch import ShadeBatch
    _lib.set_point_order(order)
    try:
        _replay_new_batches()
    finally:
        _lib.set_point_order(0)


def _replay_new_batches():
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = _sets([(4000, 32, 64), (3000, 16, 32), (100, 16, 32)], seed=40)
    b = ShadeBatch("vis")
    b.prepare(devs)
    b.capture()
    for rep, specs in enumerate([[(2500, 16, 32), (700, 32, 64)], [(4000, 32, 64)],
                                 [(10, 4, 8), (3000, 16, 32), (2000, 16, 32)]]):
        nps2, devs2 = _sets(specs, seed=50 + 10 * rep)
        outs = b.prepare(devs2)
        b.replay()
        torch.cuda.synchronize()
        graph_res = [_np(o) for o in outs]
        # eager work between replays, then an eager run of the same batch
        e = ShadeBatch("vis")
        eo = e.prepare(devs2)
        e.forward()
        e.backward()
        torch.cuda.synchronize()
        for g, x, d in zip(graph_res, eo, nps2):
            ex = _np(x)
            for a, c in zip(g[0] + g[1], ex[0] + ex[1]):
                assert np.array_equal(a, c)
            compare("vis", g, oracle_run("vis", d))


def test_forward_only_batch():
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = _sets([(1000, 16, 32), (64, 8, 8)], seed=60)
    for d in devs:
        d.pop("grad_shading")
        d.pop("grad_diffuse_shading")
    b = ShadeBatch("vis")
    outs = b.prepare(devs, with_backward=False)
    b.forward()
    with pytest.raises(RuntimeError, match="without backward"):
        b.backward()  # rejected before any launch
    torch.cuda.synchronize()
    for d, o in zip(nps, outs):
        sh, ds = oracle_run("vis", d)[0]
        from _helpers import relerr
        assert relerr(o["shading"].cpu().numpy(), sh) <= TOL
        assert relerr(o["diffuse_shading"].cpu().numpy(), ds) <= TOL


def test_autotuner_picks_valid_shapes():
    from diff_envshade_amd.autotune import tune_batch
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = _sets([(20000, 16, 32), (9000, 32, 64)], seed=70)
    b = ShadeBatch("vis")
    f, bw = tune_batch(b, devs, reps=2, force=True)
    assert f in (1, 2, 4, 8) and bw in (1, 2, 4, 8)
    b.forward()
    b.backward()
    torch.cuda.synchronize()
    for d, o in zip(nps, b.outputs):
        compare("vis", _np(o), oracle_run("vis", d))


@pytest.mark.parametrize("variant", ["diffuse", "specular", "clamp", "v2"])
def test_batch_other_variants(variant):
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = [], []
    for i, (n, H, W) in enumerate([(500, 8, 16), (130, 16, 32)]):
        d = problem(seed=80 + i, n=n, H=H, W=W)
        if variant not in ("v2",):
            d.pop("grad_diffuse_shading")
        nps.append(d)
        devs.append(gpu_inputs(d))
    b = ShadeBatch(variant, enable_diffuse=True, enable_specular=True)
    outs = b.prepare(devs)
    b.forward()
    b.backward()
    torch.cuda.synchronize()
    for d, o in zip(nps, outs):
        fwd = [o["shading"].cpu().numpy()]
        if variant == "v2":
            fwd.append(o["diffuse_shading"].cpu().numpy())
        if variant == "diffuse":
            names = ("grad_normal", "grad_albedo", "grad_envmap")
        else:
            names = ("grad_normal", "grad_view_dir", "grad_albedo", "grad_specular_albedo",
                     "grad_roughness", "grad_envmap")
        bwd = [o[k].cpu().numpy() for k in names]
        compare(variant, (fwd, bwd), oracle_run(variant, d))


@pytest.mark.parametrize("ntc_f,ntc_b", [(4, 4), (8, 2)])
def test_sorted_batch_matches_oracle_and_single_set(ntc_f, ntc_b):
    # with the point order forced on, prepare sorts every set once and both
    # directions shade in that order, as the single-set entry points do
    from diff_envshade_amd import _lib
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = _sets(SPECS, seed=40)
    try:
        _lib.set_point_order(1)
        b = ShadeBatch("vis", enable_specular=True, ntc_forward=ntc_f, ntc_backward=ntc_b)
        outs = b.prepare(devs, with_backward=True)
        b.forward()
        b.backward()
        torch.cuda.synchronize()
        _lib.set_tiling_override(ntc_f, ntc_b)
        for d, o in zip(nps, outs):
            got = _np(o)
            if d["normal"].shape[0] > 0:
                compare("vis", got, oracle_run("vis", d))
            single = gpu_run("vis", d)
            for a, s in zip(got[0] + got[1], single[0] + single[1]):
                assert np.array_equal(a.reshape(s.shape), s)
    finally:
        _lib.set_tiling_override(0, 0)
        _lib.set_point_order(0)


def test_automatic_order_policy_in_a_batch():
    # one set above the backward's sort threshold (70k x 512 pairs) beside small
    # ones: the batch sorts it for the backward only, like the single-set calls
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = _sets([(70000, 16, 32), (3000, 16, 32)], seed=50)
    b = ShadeBatch("vis", enable_specular=True, ntc_forward=4, ntc_backward=4)
    outs = b.prepare(devs, with_backward=True)
    b.forward()
    b.backward()
    torch.cuda.synchronize()
    from diff_envshade_amd import _lib
    try:
        _lib.set_tiling_override(4, 4)
        for d, o in zip(nps, outs):
            got = _np(o)
            single = gpu_run("vis", d)
            for a, s in zip(got[0] + got[1], single[0] + single[1]):
                assert np.array_equal(a.reshape(s.shape), s)
    finally:
        _lib.set_tiling_override(0, 0)


@pytest.mark.parametrize("order", [-1, 1])
def test_envgrad_reduce_of_single_set_matches_batch_at_many_groups(order):
    # the single-set backward and the batch reduce the env-map partials with
    # the same two stages in different launches (one set vs a descriptor table);
    # both add in the same order, so every output agrees bit for bit.  137 and
    # 42 groups of 16 tiles: a second round of eight loads per lane in the final
    # sum (lanes q < 9 of 16), partial rounds, and partial column blocks (96
    # and 768 columns)
    from diff_envshade_amd import _lib
    from diff_envshade_amd.batch import ShadeBatch
    nps, devs = _sets([(140000, 4, 8), (43000, 8, 32)], seed=50)
    try:
        _lib.set_point_order(order)
        b = ShadeBatch("vis", enable_specular=True, ntc_forward=4, ntc_backward=4)
        outs = b.prepare(devs, with_backward=True)
        b.forward()
        b.backward()
        torch.cuda.synchronize()
        _lib.set_tiling_override(4, 4)
        for d, o in zip(nps, outs):
            got = _np(o)
            single = gpu_run("vis", d)
            for a, s in zip(got[0] + got[1], single[0] + single[1]):
                assert np.array_equal(a.reshape(s.shape), s)
    finally:
        _lib.set_tiling_override(0, 0)
        _lib.set_point_order(0)


def test_launch_must_match_what_was_prepared():
    # a launch with another variant, other flags or another ntc than the
    # preparation recorded is rejected before it runs (it would read pointers
    # the preparation did not check, or deal groups to missing waves)
    from diff_envshade_amd import _lib
    from diff_envshade_amd.batch import ShadeBatch, _bind
    L = _bind(_lib.load())
    nps, devs = _sets([(3000, 16, 32)], seed=61)
    b = ShadeBatch("vis", enable_specular=False, ntc_forward=4, ntc_backward=2)
    b.prepare(devs, with_backward=True)
    ws = b.workspace.data_ptr()
    s = torch.cuda.current_stream().cuda_stream
    assert L.des_batch_forward(4, 1, 1, 4, ws, s) == -1      # specular on, prepared off
    assert b"specular" in L.des_last_error()
    assert L.des_batch_forward(4, 1, 0, 2, ws, s) == -1      # another ntc_forward
    assert b"ntc_forward" in L.des_last_error()
    assert L.des_batch_backward(4, 1, 0, 4, ws, s) == -1     # another ntc_backward
    assert b"ntc_backward" in L.des_last_error()
    assert L.des_batch_backward(3, 1, 0, 2, ws, s) == -1     # another variant
    assert L.des_batch_forward(4, 1, 0, 4, ws, s) == 0
    assert L.des_batch_backward(4, 1, 0, 2, ws, s) == 0
    torch.cuda.synchronize()
    d = nps[0]
    ref = oracle_run("vis", d, spec=False)
    got = ([b.outputs[0]["shading"].cpu().numpy(), b.outputs[0]["diffuse_shading"].cpu().numpy()],
           [b.outputs[0][k].cpu().numpy() for k in ("grad_normal", "grad_view_dir", "grad_albedo",
                                                    "grad_specular_albedo", "grad_roughness",
                                                    "grad_envmap")])
    compare("vis", got, ref)
