// shade_math.h -- device math shared by the shading kernels: constants,
// fast reciprocal / rsqrt, the FP64 texel geometry and the per-pair
// Lambert + GGX lobe.
//
// Semantics follow zjumsj/diff_envshade bruteforce_shader.cu ("bs"):
// texel geometry bs:124-147 with the LEFT_TOP_AS_ORIGIN row flip (config.h:11),
// Fresnel bs:149-159, G1 bs:161-171, NDF bs:203-210, constants bs:30-31.
#pragma once
#include <hip/hip_runtime.h>

namespace des {
namespace math {

constexpr float kF0 = 0.04f;      // Schlick R0 (bs:30)
constexpr float kThres = 1e-9f;   // NDF denominator floor (bs:31)
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kPi = 3.14159265358979323846f;
constexpr double kPiD = 3.14159265358979323846;

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sat(float x) { return __builtin_amdgcn_fmed3f(x, 0.f, 1.f); }

// Solid angle of lat-long row band times cos(theta): dOmega = kdw * cos(theta) with
// kdw = 2 sin(pi/(2H)) 2pi/W, the cancellation-free form of
// (sin(pi(iy+1)/H - pi/2) - sin(pi iy/H - pi/2)) 2pi/W.
__device__ __forceinline__ double solid_angle_scale(int H, int W) {
  return 4.0 * kPiD * sinpi(0.5 / (double)H) / (double)W;
}

// Storage texel t -> rotated direction l = M d (fp32, rounded once from FP64),
// |l|^2 and dOmega (FP64).  Row r = t / W, column x = t % W, flipped row
// iy = H-1-r; theta = ((iy+.5)/H - .5) pi, phi = 2 pi (x+.5)/W,
// d = (cos th cos ph, sin th, cos th sin ph).
struct TexelGeo {
  float lx, ly, lz, ll;
  double dw;
};

__device__ __forceinline__ TexelGeo texel_geometry(int t, int H, int W, const double* M,
                                                   double kdw) {
  const int r = t / W;
  const int x = t - r * W;
  const int iy = H - 1 - r;
  double sth, cth, sph, cph;
  sincospi(((double)iy + 0.5) / (double)H - 0.5, &sth, &cth);
  sincospi((2.0 * (double)x + 1.0) / (double)W, &sph, &cph);
  const double dx = cth * cph, dy = sth, dz = cth * sph;
  TexelGeo g;
  g.lx = (float)(M[0] * dx + M[1] * dy + M[2] * dz);
  g.ly = (float)(M[3] * dx + M[4] * dy + M[5] * dz);
  g.lz = (float)(M[6] * dx + M[7] * dy + M[8] * dz);
  g.ll = (float)((double)g.lx * g.lx + (double)g.ly * g.ly + (double)g.lz * g.lz);
  g.dw = cth * kdw;
  return g;
}

// ---- the per-pair lobe -------------------------------------------------------

// Quantities of one (point, texel) pair shared by forward and backward.
struct Lobe {
  float cs;    // l.n
  float nl;    // max(cs, 0)
  float hx, hy, hz;  // l + v, formed componentwise (exact for near-opposite l, v)
  float rs;    // 1/|l+v|
  float ch;    // l.h
  float nh;    // n.h
  float om4;   // (1 - sat(ch))^4
  float F, G, D;
  float ia;    // 1/(nl(1-k)+k)
  float it2;   // 1/max(pi tD^2, thres)
  float tD;    // (n.h)^2 (alpha^2-1) + 1
  float tmp;   // F G D / 4, or 0 where the reference skips the lobe (tmp <= 0 or NaN)
  bool spec_on;
};

// S provides nx..nz, vx..vz, k, omk, ib, a2, am1 (per-point constants).
template <bool SPEC, class S>
__device__ __forceinline__ Lobe eval_lobe(const S& s, const float4& g) {
  Lobe o;
  o.cs = fmaf(g.z, s.nz, fmaf(g.y, s.ny, g.x * s.nx));
  o.nl = fmaxf(o.cs, 0.f);
  if (SPEC) {
    // h = (l+v)/|l+v| (bs:2218-2219); l+v componentwise keeps |l+v|, l.h and
    // n.h accurate to fp32 rounding when l and v are nearly opposite
    o.hx = g.x + s.vx;
    o.hy = g.y + s.vy;
    o.hz = g.z + s.vz;
    const float hb2 = fmaf(o.hz, o.hz, fmaf(o.hy, o.hy, o.hx * o.hx));
    // floor keeps (n.h)^2 finite when v = -l exactly; such pairs are masked below
    o.rs = rsq(fmaxf(hb2, 1e-20f));
    o.ch = fmaf(g.z, o.hz, fmaf(g.y, o.hy, g.x * o.hx)) * o.rs;
    o.nh = fmaf(s.nz, o.hz, fmaf(s.ny, o.hy, s.nx * o.hx)) * o.rs;
    const float om = 1.f - sat(o.ch);
    const float om2 = om * om;
    o.om4 = om2 * om2;
    o.F = fmaf(o.om4 * om, 1.f - kF0, kF0);
    o.ia = rcp(fmaf(o.nl, s.omk, s.k));
    o.G = o.nl * o.ia * s.ib;
    o.tD = fmaf(o.nh * o.nh, s.am1, 1.f);
    o.it2 = rcp(fmaxf(kPi * o.tD * o.tD, kThres));
    o.D = s.a2 * o.it2;
    const float tmp = o.F * o.G * o.D * 0.25f;
    // the reference skips the lobe unless tmp > 0; |l+v| = 0 gives NaN there
    o.spec_on = (tmp > 0.f) && (hb2 > 0.f);
    o.tmp = o.spec_on ? tmp : 0.f;
  } else {
    o.hx = o.hy = o.hz = 0.f;
    o.rs = o.ch = o.nh = o.om4 = o.F = o.G = o.D = o.ia = o.it2 = o.tD = 0.f;
    o.spec_on = false;
    o.tmp = 0.f;
  }
  return o;
}

}  // namespace math
}  // namespace des
