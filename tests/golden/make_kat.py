"""Generate tests/golden/kat_single_texel.json: known-answer vectors for the
shader, computed from textbook formulas in angle form, independently of the
oracle's vector code.

The reference ships no tests or golden vectors, so these known answers are
built from the mathematics the reference implements
(zjumsj/diff_envshade bruteforce_shader.cu):

  * lat-long geometry: storage row r is latitude band r counted from the TOP
    (config.h:11 LEFT_TOP_AS_ORIGIN), band edges at latitudes
    pi/2 - r pi/H ... ; column x is longitude phi = 2 pi (x + 1/2) / W
    measured from +x towards +z; texel solid angle = (2 pi / W) * (sin of
    upper edge latitude - sin of lower edge latitude) (bs:124-147);
  * one texel lit (one-hot env map), the rotation M maps that texel's
    direction d onto a chosen light direction l (bs:305-307);
  * n = +y, l at angle a from n and v (unit) at angle -b from n in the x-y
    plane, so the half vector is at angle (a - b)/2 from n and l.h =
    cos((a + b)/2);
  * Lambert: E * albedo * cos(a) / pi * dOmega (bs:2204-2217);
  * Cook-Torrance with Schlick F (F0 = 0.04), GGX D with alpha = r^2 and
    Smith-Schlick G1(x) = x / (x (1 - k) + k), k = (r + 1)^2 / 8:
    f = D F G1(n.l) G1(n.v) / (4 n.l n.v), radiance = E * s * f * n.l * dOmega
    (bs:2218-2234, written there as F G D / 4 with G = G1 G1 / n.v);
  * a view at or below the horizon (n.v <= 0): the reference clamps n.v to 0
    inside G (bs:161-171), so G = G1(n.l) / k -- the n.v -> 0+ limit of
    G1(n.v) / n.v -- and the non-clamp variants still shade the specular
    lobe, with F saturating to 1 where l.h < 0; the clamp variant skips the
    whole point (bs:1111, 1125).

Run:  python tests/golden/make_kat.py   (writes the JSON next to this file)
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "kat_single_texel.json")

F0 = 0.04


def texel_geometry(H, W, r, x):
    lat_top = math.pi / 2 - r * math.pi / H
    lat_bot = math.pi / 2 - (r + 1) * math.pi / H
    lat = 0.5 * (lat_top + lat_bot)
    phi = 2 * math.pi * (x + 0.5) / W
    d = np.array([math.cos(lat) * math.cos(phi), math.sin(lat), math.cos(lat) * math.sin(phi)])
    domega = 2 * math.pi / W * (math.sin(lat_top) - math.sin(lat_bot))
    return d, domega


def rotation_taking(a_vec, b_vec):
    """Proper rotation R with R a = b (unit vectors), Rodrigues' formula."""
    a = a_vec / np.linalg.norm(a_vec)
    b = b_vec / np.linalg.norm(b_vec)
    v = np.cross(a, b)
    c = float(np.dot(a, b))
    if np.linalg.norm(v) < 1e-12:
        if c > 0:
            return np.eye(3)
        axis = np.cross(a, [1.0, 0.0, 0.0])
        if np.linalg.norm(axis) < 1e-6:
            axis = np.cross(a, [0.0, 1.0, 0.0])
        axis /= np.linalg.norm(axis)
        return 2 * np.outer(axis, axis) - np.eye(3)
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx * (1.0 / (1.0 + c))


def ggx_D(cos_h, alpha):
    # D = alpha^2 / (pi cos^4 th_h (alpha^2 + tan^2 th_h)^2)
    tan2 = (1.0 - cos_h * cos_h) / (cos_h * cos_h)
    return alpha ** 2 / (math.pi * cos_h ** 4 * (alpha ** 2 + tan2) ** 2)


def schlick_F(cos_d):
    return F0 + (1.0 - F0) * (1.0 - cos_d) ** 5


def smith_G1(x, k):
    return x / (x * (1.0 - k) + k)


def case(name, H, W, r, x, a_deg, b_deg, rough, E, albedo, spec, vscale=1.0, view=None):
    a = math.radians(a_deg)
    b = math.radians(b_deg)
    d, dw = texel_geometry(H, W, r, x)
    l = np.array([math.sin(a), math.cos(a), 0.0])
    M = rotation_taking(d, l)
    n = np.array([0.0, 1.0, 0.0])
    v = vscale * np.array([-math.sin(b), math.cos(b), 0.0]) if view is None else np.array(view)
    alpha = rough * rough
    k = (rough + 1.0) ** 2 / 8.0
    cos_l, cos_v = math.cos(a), math.cos(b)
    cos_h = math.cos((a - b) / 2.0)
    cos_d = math.cos((a + b) / 2.0)
    diffuse = [E[c] * max(cos_l, 0.0) / math.pi * dw for c in range(3)]
    nv = float(np.dot(n, v))
    if cos_l > 0:
        # the reference's G uses the unnormalised view vector's n.v; v is unit
        # here unless vscale != 1, in which case n.v and l.h change with |v|
        h = (l + v) / np.linalg.norm(l + v)
        cos_h = float(np.dot(n, h))
        cos_d = max(float(np.dot(l, h)), 0.0)  # Schlick takes sat(l.h)
        # G1(n.v) / n.v, or its n.v -> 0+ limit 1/k where the reference clamps n.v to 0
        g1v_over_nv = smith_G1(nv, k) / nv if nv > 0 else 1.0 / k
        f = ggx_D(cos_h, alpha) * schlick_F(cos_d) * smith_G1(cos_l, k) * g1v_over_nv \
            / (4.0 * cos_l)
        specular = [E[c] * spec[c] * f * cos_l * dw for c in range(3)]
    else:
        specular = [0.0, 0.0, 0.0]
    shading = [albedo[c] * diffuse[c] + specular[c] for c in range(3)]
    env = np.zeros((3, H, W))
    env[:, r, x] = E
    t = r * W + x
    return {
        "name": name, "H": H, "W": W, "texel": t, "E": list(map(float, E)),
        "normal": n.tolist(), "view_dir": v.tolist(), "albedo": list(map(float, albedo)),
        "specular_albedo": list(map(float, spec)), "roughness": rough,
        "light2obj_rotmat": M.tolist(),
        "expected": {
            "diffuse_shading": diffuse,
            "shading_v2": shading,
            "shading_specular_only": specular,
            # clamp variant, both branches on: the point is skipped unless n.v > 0
            "shading_clamp": shading if nv > 0 else [0.0, 0.0, 0.0],
        },
    }


def main():
    cases = [
        case("normal_incidence_mirror", 8, 16, 3, 5, 0.0, 0.0, 0.5, [1.0, 2.0, 3.0],
             [0.5, 0.25, 1.0], [0.1, 0.2, 0.3]),
        case("oblique_mirror", 8, 16, 1, 2, 40.0, 40.0, 0.3, [2.0, 1.0, 0.5],
             [0.2, 0.4, 0.6], [0.05, 0.05, 0.05]),
        case("off_specular", 16, 32, 6, 30, 25.0, 55.0, 0.6, [1.5, 1.5, 1.5],
             [1.0, 1.0, 1.0], [0.08, 0.04, 0.02]),
        case("rough_one", 4, 8, 0, 7, 60.0, 10.0, 1.0, [3.0, 0.5, 1.0],
             [0.3, 0.3, 0.3], [0.1, 0.1, 0.1]),
        case("grazing", 16, 32, 15, 0, 80.0, 70.0, 0.2, [1.0, 1.0, 1.0],
             [0.7, 0.1, 0.2], [0.09, 0.09, 0.09]),
        case("unnormalised_view", 8, 16, 4, 9, 30.0, 20.0, 0.45, [1.0, 2.0, 0.5],
             [0.6, 0.6, 0.6], [0.07, 0.03, 0.01], vscale=1.7),
        case("below_horizon", 8, 16, 2, 11, 120.0, 30.0, 0.5, [5.0, 5.0, 5.0],
             [1.0, 1.0, 1.0], [0.1, 0.1, 0.1]),
        # views at and below the horizon (n.v <= 0)
        case("back_facing_view", 8, 16, 2, 4, 30.0, 120.0, 0.4, [2.0, 1.0, 3.0],
             [0.5, 0.5, 0.5], [0.1, 0.2, 0.05]),
        case("view_on_horizon", 16, 32, 5, 20, 40.0, 90.0, 0.35, [1.0, 3.0, 2.0],
             [0.2, 0.8, 0.4], [0.06, 0.06, 0.06], view=[-1.0, 0.0, 0.0]),
        case("back_facing_l_dot_h_negative", 8, 16, 3, 13, 70.0, 150.0, 0.7, [4.0, 2.0, 1.0],
             [0.9, 0.3, 0.6], [0.1, 0.1, 0.1]),
    ]
    with open(OUT, "w") as f:
        json.dump({"source": "tests/golden/make_kat.py", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
