"""Point order by normal direction (GPU): sorting the points into
normal-coherent tiles, skipping texels that light no lane of a tile and
dealing the lit texel groups evenly to the tile's waves may change the
outputs only at fp32 rounding level (every sum is regrouped), and the
results stay bitwise reproducible.  Oracle parity of the sorted path is
covered by the parity suite, which runs every case with the order forced
on."""
import numpy as np
import pytest

from _helpers import TOL, compare, gpu_run, oracle_run, problem, relerr

pytestmark = pytest.mark.gpu


def _run(mode, variant, d, diff=True, spec=True):
    from diff_envshade_amd import _lib
    try:
        _lib.set_point_order(mode)
        return gpu_run(variant, d, diff, spec)
    finally:
        _lib.set_point_order(0)


@pytest.mark.parametrize("variant,diff,spec", [("vis", True, True), ("v2", True, True),
                                               ("clamp", True, True), ("specular", False, True),
                                               ("diffuse", True, False)])
def test_point_order_changes_outputs_only_by_rounding(variant, diff, spec):
    # every sum is regrouped; sums that cancel (grad_roughness) move by up to
    # ~2e-6 normwise, so the two orders are held to the 1e-5 accuracy bar they
    # each meet against FP64 in the parity suite
    d = problem(seed=21, n=20000, H=16, W=32, vis="identity" if variant == "vis" else "none")
    (f0, b0), (f1, b1) = _run(-1, variant, d, diff, spec), _run(1, variant, d, diff, spec)
    for a, b in zip(f0 + b0, f1 + b1):
        assert relerr(b, a) <= TOL


def test_sorted_outputs_are_bitwise_reproducible():
    d = problem(seed=22, n=30000, H=16, W=32, vis="identity")
    f0, b0 = _run(1, "vis", d)
    f1, b1 = _run(1, "vis", d)
    for a, b in zip(f0 + b0, f1 + b1):
        assert np.array_equal(a, b)


def test_coherent_normals_skip_half_the_map():
    # every normal within ~3 degrees of +y: the lower hemisphere of the map is
    # dark for whole tiles, so whole groups of texels take the skip path
    rng = np.random.default_rng(23)
    d = problem(seed=23, n=5000, H=16, W=32, rotation=False, vis="identity")
    n = np.array([0.0, 1.0, 0.0]) + 0.05 * rng.normal(size=(5000, 3))
    d["normal"] = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    from diff_envshade_amd.workloads import make_visibility
    d["visibility"] = make_visibility(rng, d["normal"], 16, 32)
    ref = oracle_run("vis", d)
    compare("vis", _run(1, "vis", d), ref)
    compare("vis", _run(-1, "vis", d), ref)


def test_degenerate_normals_are_ordered():
    # zero and non-finite-free degenerate normals get key 0; every point is
    # still shaded exactly once
    d = problem(seed=24, n=3000, H=8, W=16, vis="none")
    d["normal"][::7] = 0.0
    d["normal"][1::11] = np.array([0.0, 0.0, -1.0], dtype=np.float32)
    compare("v2", _run(1, "v2", d), oracle_run("v2", d))


@pytest.mark.parametrize("n", [4096 * 3, 4096 * 64 + 4000])
def test_order_block_boundaries(n):
    # 12288 points: three full 4096-point sort blocks (fused scan); 266144
    # points: 65 blocks, past the fused limit (separate per-key scan)
    d = problem(seed=25, n=n, H=8, W=16, vis="identity")
    (f0, b0), (f1, b1) = _run(-1, "vis", d), _run(1, "vis", d)
    for a, b in zip(f0 + b0, f1 + b1):
        assert relerr(b, a) <= TOL
