"""Env-map sampling and blur on the GPU against the FP64 oracle
(oracle/envmap_oracle.c).

Both sides evaluate the same FP64 formulas with the same random numbers
(Philox4x32-10, counter = sample, target texel); they differ only where their
math libraries' sin, cos, atan2 and asin round differently, which moves a
lookup by ~1e-16 of a texel.  Outputs must agree to 1e-7 normwise relative
error, float32 rounding included.
"""
import numpy as np
import pytest
import torch

from diff_envshade_amd import _diff_envshade as de
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = 1e-7


def rel(x, ref):
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.abs(np.asarray(x, dtype=np.float64) - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("C,H,W,P", [(3, 16, 32, 10007), (4, 7, 13, 999), (1, 1, 1, 64),
                                     (5, 32, 64, 4096), (3, 128, 256, 100000)])
def test_sample_matches_oracle(C, H, W, P):
    rng = np.random.default_rng(C * 1000 + H)
    env = np.exp(rng.normal(size=(C, H, W))).astype(np.float32)
    q = rng.normal(size=(3, P)).astype(np.float32)
    out = de.sample_envmap_forward(torch.from_numpy(q).to(DEV), torch.from_numpy(env).to(DEV))
    assert out.shape == (C, P)
    assert rel(out.cpu().numpy(), O.sample_envmap(q, env)) < TOL


def test_sample_edge_cases():
    rng = np.random.default_rng(1)
    env = rng.random((3, 8, 16)).astype(np.float32)
    # a zero-length query samples NaN; the poles and the seam are finite
    q = np.array([[0.0, 0.0, 0.0, 1.0, 1.0], [0.0, 1.0, -1.0, 0.0, 0.0],
                  [0.0, 0.0, 0.0, 0.0, -1e-9]], dtype=np.float32)
    out = de.sample_envmap_forward(torch.from_numpy(q).to(DEV), torch.from_numpy(env).to(DEV))
    out = out.cpu().numpy()
    assert np.all(np.isnan(out[:, 0])) and np.all(np.isfinite(out[:, 1:]))
    assert rel(out[:, 1:], O.sample_envmap(q[:, 1:], env)) < TOL
    empty = de.sample_envmap_forward(torch.zeros(3, 0, device=DEV), torch.from_numpy(env).to(DEV))
    assert empty.shape == (3, 0)


@pytest.mark.parametrize("C,H,W,tH,tW,N,alpha", [
    (3, 16, 32, 8, 16, 1000, 0.3),
    (4, 9, 17, 5, 7, 63, 0.8),     # fewer samples than lanes
    (3, 8, 16, 4, 8, 64, 0.05),
    (3, 8, 16, 3, 5, 65, 1.0),     # one sample more than lanes
    (3, 32, 64, 32, 64, 1, 0.5),
    (4, 16, 32, 16, 32, 16, 0.0),  # zero roughness at the source size: the map itself
    (4, 16, 32, 6, 10, 300, 0.6),  # four waves per target texel
    (3, 8, 16, 2, 3, 256, 0.2),
])
def test_blur_matches_oracle(C, H, W, tH, tW, N, alpha):
    rng = np.random.default_rng(N)
    env = np.exp(rng.normal(size=(C, H, W))).astype(np.float32)
    out = de.blur_envmap_forward(torch.from_numpy(env).to(DEV), tH, tW, N, alpha)
    assert out.shape == (C, tH, tW)
    ref = O.blur_envmap(env, tH, tW, N, alpha)
    assert rel(out.cpu().numpy(), ref) < TOL
    if alpha == 0.0 and (tH, tW) == (H, W):
        assert np.allclose(out.cpu().numpy(), env, rtol=1e-6, atol=0)


def test_blur_is_reproducible_and_checks_arguments():
    rng = np.random.default_rng(2)
    env = torch.from_numpy(rng.random((3, 16, 32)).astype(np.float32)).to(DEV)
    a = de.blur_envmap_forward(env, 16, 32, 333, 0.4)
    b = de.blur_envmap_forward(env, 16, 32, 333, 0.4)
    assert torch.equal(a, b)
    with pytest.raises(RuntimeError, match="Channel number not supported!"):
        de.blur_envmap_forward(torch.zeros(5, 4, 8, device=DEV), 4, 8, 8, 0.1)
    with pytest.raises(RuntimeError, match="N_samples"):
        de.blur_envmap_forward(env, 4, 8, 0, 0.1)
    assert de.blur_envmap_forward(env, 0, 8, 8, 0.1).shape == (3, 0, 8)
