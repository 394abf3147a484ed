"""Pins of the FP64 oracle's env-map utilities (oracle/envmap_oracle.c) against
what the mathematics fixes, on the CPU:

  * Philox4x32-10 against the known-answer vectors published with the
    algorithm (Salmon et al., SC 2011, Random123 distribution);
  * the GGX sampler (etc.cu:588-605): tan^2(theta) of its directions has the
    GGX CDF t / (alpha^2 + t) (Walter et al. 2007) and its azimuth is uniform;
  * the bilinear lookup (etc.cu:527-586): a texel centre's direction returns
    the texel, the seam interpolates between the last and first column, the
    poles clamp to the first / last row, the result is linear in the map and
    independent of the query's length;
  * the blur (etc.cu:603-685): alpha = 0 at the source size returns the map, a
    constant map stays constant, and a large-sample blur converges to the
    GGX-weighted mean of the lookup computed by quadrature over the sphere.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_philox_known_answers():
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert [int(x) for x in O.philox(ctr, key)] == want


@pytest.mark.parametrize("alpha", [0.05, 0.3, 1.0])
def test_ggx_sampler_has_the_ggx_distribution(alpha):
    rng = np.random.default_rng(1)
    n = 20000
    r = rng.random((n, 2))
    l = np.array([O.ggx_sample(a, b, alpha) for a, b in r])
    assert np.allclose(np.linalg.norm(l, axis=1), 1.0, atol=1e-12)
    t = np.sort((1.0 - l[:, 2] ** 2) / l[:, 2] ** 2)
    cdf = t / (alpha * alpha + t)
    ks = np.max(np.abs(cdf - (np.arange(1, n + 1) / n)))
    assert ks < 1.63 / math.sqrt(n)  # 99% Kolmogorov-Smirnov bound
    phi = np.sort(np.mod(np.arctan2(l[:, 1], l[:, 0]), 2 * math.pi) / (2 * math.pi))
    assert np.max(np.abs(phi - np.arange(1, n + 1) / n)) < 1.63 / math.sqrt(n)


def texel_dir(r, x, H, W):
    """Direction of the centre of texel (row r, column x) in the lookup's convention."""
    u = (x + 0.5) / W
    v = 1.0 - (r + 0.5) / H
    theta = (v - 0.5) * math.pi
    phi = 2 * math.pi * u
    return np.array([math.cos(theta) * math.cos(phi), math.sin(theta),
                     math.cos(theta) * math.sin(phi)])


def test_lookup_at_texel_centres_returns_the_texel():
    rng = np.random.default_rng(2)
    C, H, W = 4, 7, 13
    env = rng.random((C, H, W)).astype(np.float32)
    q = np.stack([texel_dir(r, x, H, W) for r in range(H) for x in range(W)], axis=1)
    out = O.sample_envmap(q, env)
    assert np.allclose(out, env.reshape(C, -1), rtol=0, atol=1e-12)


def test_lookup_wraps_the_seam_and_clamps_the_poles():
    rng = np.random.default_rng(3)
    C, H, W = 3, 6, 10
    env = rng.random((C, H, W)).astype(np.float32).astype(np.float64)
    # on row 2's centre line, a quarter texel past the seam: between columns W-1 and 0
    r = 2
    theta = (1.0 - (r + 0.5) / H - 0.5) * math.pi
    pu = 0.25  # u * W
    phi = 2 * math.pi * pu / W
    q = np.array([[math.cos(theta) * math.cos(phi)], [math.sin(theta)],
                  [math.cos(theta) * math.sin(phi)]])
    kx = pu + 0.5
    want = (1 - kx) * env[:, r, W - 1] + kx * env[:, r, 0]
    assert np.allclose(O.sample_envmap(q, env.astype(np.float32))[:, 0], want, atol=1e-12)
    # straight up: row 0 only, whatever the azimuth blend
    top = O.sample_envmap(np.array([[0.0], [1.0], [0.0]]), env.astype(np.float32))[:, 0]
    assert np.all(top >= env[:, 0].min(axis=1) - 1e-12) and np.all(top <= env[:, 0].max(axis=1) + 1e-12)
    bottom = O.sample_envmap(np.array([[0.3], [-5.0], [0.1]]), env.astype(np.float32))[:, 0]
    assert np.all(bottom >= env[:, -1].min(axis=1) - 1e-12)
    assert np.all(bottom <= env[:, -1].max(axis=1) + 1e-12)


def test_lookup_is_linear_and_scale_free():
    rng = np.random.default_rng(4)
    a = rng.random((3, 8, 16)).astype(np.float32)
    b = rng.random((3, 8, 16)).astype(np.float32)
    q = rng.normal(size=(3, 500))
    sa, sb, sab = O.sample_envmap(q, a), O.sample_envmap(q, b), O.sample_envmap(q, a + b)
    assert np.allclose(sa + sb, sab, atol=1e-6)  # a + b rounded to float32
    assert np.allclose(O.sample_envmap(q * 7.5, a), sa, atol=1e-12)
    assert np.all(np.isnan(O.sample_envmap(np.zeros((3, 1)), a)))


def test_blur_with_zero_roughness_at_source_size_is_the_map():
    rng = np.random.default_rng(5)
    for C, H, W in ((3, 8, 16), (4, 5, 9)):
        env = rng.random((C, H, W)).astype(np.float32)
        out = O.blur_envmap(env, H, W, 7, 0.0)
        assert np.allclose(out, env, atol=1e-12)


def test_blur_keeps_a_constant_map():
    env = np.full((3, 8, 16), 2.5, dtype=np.float32)
    out = O.blur_envmap(env, 5, 11, 33, 0.7)
    assert np.allclose(out, 2.5, atol=1e-12)


def test_blur_converges_to_the_ggx_weighted_mean():
    # E[lookup(h)] over h ~ D(h) (h.n), by quadrature in (theta, phi) around n,
    # against a 200,000-sample blur of the same texel
    rng = np.random.default_rng(6)
    H, W = 8, 16
    env = rng.random((3, H, W)).astype(np.float32)
    alpha, tar_H, tar_W = 0.4, 2, 3
    N = 200000
    blur = O.blur_envmap(env, tar_H, tar_W, N, alpha)
    nt, nph = 400, 256
    for ih, iw in ((0, 1), (1, 2)):
        theta = (ih + 0.5) / tar_H * math.pi
        phi = (iw + 0.5) / tar_W * 2 * math.pi
        n = np.array([math.sin(theta) * math.cos(phi), math.cos(theta),
                      math.sin(theta) * math.sin(phi)])
        a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e1 = np.cross(n, a)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        # t = tan^2(theta_h) has CDF t / (alpha^2 + t): midpoint quadrature in that CDF
        cdf = (np.arange(nt) + 0.5) / nt
        t = alpha * alpha * cdf / (1 - cdf)
        ct = 1 / np.sqrt(1 + t)
        st = np.sqrt(1 - ct * ct)
        ph = (np.arange(nph) + 0.5) / nph * 2 * math.pi
        dirs = (st[:, None, None] * np.cos(ph)[None, :, None] * e1 +
                st[:, None, None] * np.sin(ph)[None, :, None] * e2 +
                ct[:, None, None] * n).reshape(-1, 3)
        mean = O.sample_envmap(dirs.T.copy(), env).mean(axis=1)
        samples_sd = O.sample_envmap(dirs.T.copy(), env).std(axis=1)
        err = np.abs(blur[:, ih, iw] - mean)
        assert np.all(err < 4 * samples_sd / math.sqrt(N) + 2e-3 * np.abs(mean)), (err, mean)
