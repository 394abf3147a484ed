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
// the floor on pi tD^2 restated as a floor on tD^2
constexpr float kThresOverPi = 3.18309886183790671538e-10f;
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

// One texel as the shading loops read it (48 bytes, 16-byte aligned).
//   d*: rotated light direction l = M d in FP64;  l*: the same rounded to fp32;
//   e*: E_c * dOmega (fp32, rounded once).
struct Texel {
  double dx, dy, dz;
  float lx, ly, lz;
  float e0, e1, e2;
};
static_assert(sizeof(Texel) == 48, "texel entry layout");

// Storage texel t -> rotated direction and solid angle (FP64).  Row
// r = t / W, column x = t % W, flipped row iy = H-1-r; theta = ((iy+.5)/H - .5) pi,
// phi = 2 pi (x+.5)/W, d = (cos th cos ph, sin th, cos th sin ph), l = M d.
struct TexelGeo {
  double lx, ly, lz;
  double dw;
};

// sin and cos of theta for flipped row iy, and of phi for column x
__device__ __forceinline__ void row_angle(int iy, int H, double& s, double& c) {
  sincospi(((double)iy + 0.5) / (double)H - 0.5, &s, &c);
}
__device__ __forceinline__ void col_angle(int x, int W, double& s, double& c) {
  sincospi((2.0 * (double)x + 1.0) / (double)W, &s, &c);
}

__device__ __forceinline__ TexelGeo texel_from_angles(double sth, double cth, double sph,
                                                      double cph, const double* M, double kdw) {
  const double dx = cth * cph, dy = sth, dz = cth * sph;
  TexelGeo g;
  g.lx = M[0] * dx + M[1] * dy + M[2] * dz;
  g.ly = M[3] * dx + M[4] * dy + M[5] * dz;
  g.lz = M[6] * dx + M[7] * dy + M[8] * dz;
  g.dw = cth * kdw;
  return g;
}

__device__ __forceinline__ TexelGeo texel_geometry(int t, int H, int W, const double* M,
                                                   double kdw) {
  const int r = t / W;
  const int x = t - r * W;
  double sth, cth, sph, cph;
  row_angle(H - 1 - r, H, sth, cth);
  col_angle(x, W, sph, cph);
  return texel_from_angles(sth, cth, sph, cph, M, kdw);
}

// The same from a table of the map's row angles (ang[iy], iy < H) and column
// angles (ang[H + x]) as {sin, cos}: the same values, computed once per map
// instead of twice per texel.
__device__ __forceinline__ TexelGeo texel_geometry(int t, int H, int W, const double2* ang,
                                                   const double* M, double kdw) {
  const int r = t / W;
  const int x = t - r * W;
  const double2 a = ang[H - 1 - r], b = ang[H + x];
  return texel_from_angles(a.x, a.y, b.x, b.y, M, kdw);
}

__device__ __forceinline__ Texel make_texel(const TexelGeo& g, float E0, float E1, float E2) {
  Texel o;
  o.dx = g.lx; o.dy = g.ly; o.dz = g.lz;
  o.lx = (float)g.lx; o.ly = (float)g.ly; o.lz = (float)g.lz;
  o.e0 = (float)((double)E0 * g.dw);
  o.e1 = (float)((double)E1 * g.dw);
  o.e2 = (float)((double)E2 * g.dw);
  return o;
}

__device__ __forceinline__ Texel zero_texel() {
  Texel o;
  o.dx = o.dy = o.dz = 0.0;
  o.lx = o.ly = o.lz = 0.f;
  o.e0 = o.e1 = o.e2 = 0.f;
  return o;
}

// One texel as the forward reads it (48 bytes): l = M d and |l|^2 in FP64,
// E_c * dOmega in fp32.  The forward forms every dot product of its lobe in
// FP64 from these, so it needs no fp32 copy of l.
struct FwdTexel {
  double dx, dy, dz;
  double l2;
  float e0, e1, e2;
  float pad;
};
static_assert(sizeof(FwdTexel) == sizeof(Texel), "forward texel entry layout");

__device__ __forceinline__ FwdTexel make_fwd_texel(const TexelGeo& g, float E0, float E1,
                                                   float E2) {
  FwdTexel o;
  o.dx = g.lx; o.dy = g.ly; o.dz = g.lz;
  o.l2 = fma(g.lz, g.lz, fma(g.ly, g.ly, g.lx * g.lx));
  o.e0 = (float)((double)E0 * g.dw);
  o.e1 = (float)((double)E1 * g.dw);
  o.e2 = (float)((double)E2 * g.dw);
  o.pad = 0.f;
  return o;
}

__device__ __forceinline__ FwdTexel zero_fwd_texel() {
  FwdTexel o;
  o.dx = o.dy = o.dz = o.l2 = 0.0;
  o.e0 = o.e1 = o.e2 = o.pad = 0.f;
  return o;
}

// The light direction of a table entry rounded to fp32 (the cone test's input).
__device__ __forceinline__ void texel_dir(const Texel& g, float& x, float& y, float& z) {
  x = g.lx; y = g.ly; z = g.lz;
}
__device__ __forceinline__ void texel_dir(const FwdTexel& g, float& x, float& y, float& z) {
  x = (float)g.dx; y = (float)g.dy; z = (float)g.dz;
}

// ---- the per-pair lobe -------------------------------------------------------

// Quantities of one (point, texel) pair shared by forward and backward.
struct Lobe {
  float cs;    // l.n
  float nl;    // max(cs, 0)
  float hx, hy, hz;  // hb = l + v (formed in FP64)
  float ux, uy, uz;  // hb x (n x hb) = n |hb|^2 - hb (n.hb) (formed in FP64)
  float rs;    // 1/|l+v|
  float rs2;   // rs^2
  float ch;    // l.h
  float sch;   // sat(l.h)
  float nh;    // n.h
  float om4;   // (1 - sat(ch))^4
  float F;
  float Gq;    // G / 4, G = nl ia ib (without IB: nl ia, the per-point ib / 4 left out)
  float D;     // alpha^2 / max(pi tD^2, thres)
  float ia;    // 1/(nl(1-k)+k)
  float tD;    // (n.h)^2 (alpha^2-1) + 1
  float tD2;   // tD^2
  float itD2;  // 1/max(tD^2, thres/pi) = pi / max(pi tD^2, thres)
  float tmp;   // F Gq D, or 0 where the reference skips the lobe (tmp <= 0 or NaN)
  bool spec_on;
};

// S provides nx..nz, vx..vz (fp32), nxd..nzd, vxd..vzd (the same in FP64),
// k, omk, ib4, a2, a2pi (per-point constants).
//
// Geometry of the half vector in FP64 (v_fma_f64 runs at the fp32 vector rate
// on gfx950): hb = l + v, |hb|^2 and n.hb are formed exactly enough that the
// NDF denominator
//   tD = (n.h)^2 (alpha^2 - 1) + 1 = alpha^2 (n.h)^2 + (1 - (n.h)^2),
//   1 - (n.h)^2 = (|hb|^2 - (n.hb)^2) / |hb|^2   (numerator in FP64)
// keeps full fp32 relative accuracy at the peak of sharp lobes, where the
// reference's fp32 form cancels (1 - (n.h)^2 ~ alpha^2 ~ 1e-4).  The identity
// holds for any |n|, as the reference's form does.  The backward's
// n - (n.h) h = (hb x (n x hb)) / |hb|^2 is formed the same way, as
// n |hb|^2 - hb (n.hb) in FP64 (size ~ sin(theta_h), no 1/sin cancellation).
// The constant factors 1/4 (of F G D / 4) and 1/pi (of D) ride on the
// per-point constants ib4 = ib/4 and a2pi = alpha^2/pi.  Without IB, Gq and
// tmp leave out ib4, which every term of a pair carries, so that a caller can
// apply it to its sums once.  `mask` (0 or ~0, a 0/1 visibility bit) is
// applied to max(l.n, 0) as a bit mask: G and with it tmp are 0 where it is 0.
template <bool SPEC, bool IB = true, class S>
__device__ __forceinline__ Lobe eval_lobe(const S& s, const Texel& g, uint32_t mask = ~0u) {
  Lobe o;
  o.cs = fmaf(g.lz, s.nz, fmaf(g.ly, s.ny, g.lx * s.nx));
  o.nl = __uint_as_float(__float_as_uint(fmaxf(o.cs, 0.f)) & mask);
  if (SPEC) {
    // h = (l+v)/|l+v| (bs:2218-2219)
    const double hx = g.dx + s.vxd, hy = g.dy + s.vyd, hz = g.dz + s.vzd;
    const double hb2 = fma(hz, hz, fma(hy, hy, hx * hx));
    const double nhb = fma(s.nzd, hz, fma(s.nyd, hy, s.nxd * hx));
    const float sn2 = (float)fma(-nhb, nhb, hb2);  // |hb|^2 (1 - (n.h)^2)
    o.ux = (float)fma(s.nxd, hb2, -hx * nhb);
    o.uy = (float)fma(s.nyd, hb2, -hy * nhb);
    o.uz = (float)fma(s.nzd, hb2, -hz * nhb);
    o.hx = (float)hx;
    o.hy = (float)hy;
    o.hz = (float)hz;
    const float hb2f = (float)hb2;
    // floor keeps (n.h)^2 finite when v = -l exactly; such pairs are masked below
    o.rs = rsq(fmaxf(hb2f, 1e-20f));
    o.rs2 = o.rs * o.rs;
    o.ch = fmaf(g.lz, o.hz, fmaf(g.ly, o.hy, g.lx * o.hx)) * o.rs;
    o.sch = sat(o.ch);
    o.nh = (float)nhb * o.rs;
    const float om = 1.f - o.sch;
    const float om2 = om * om;
    o.om4 = om2 * om2;
    o.F = fmaf(o.om4 * om, 1.f - kF0, kF0);
    o.ia = rcp(fmaf(o.nl, s.omk, s.k));
    o.Gq = IB ? o.nl * o.ia * s.ib4 : o.nl * o.ia;
    o.tD = fmaf(o.nh * o.nh, s.a2, sn2 * o.rs2);
    o.tD2 = o.tD * o.tD;
    o.itD2 = rcp(fmaxf(o.tD2, kThresOverPi));
    o.D = s.a2pi * o.itD2;
    const float tmp = o.F * o.Gq * o.D;
    // the reference skips the lobe unless tmp > 0; |l+v| = 0 gives NaN there
    o.spec_on = (tmp > 0.f) && (hb2f > 0.f);
    o.tmp = o.spec_on ? tmp : 0.f;
  } else {
    o.hx = o.hy = o.hz = 0.f;
    o.ux = o.uy = o.uz = 0.f;
    o.rs = o.rs2 = o.ch = o.sch = o.nh = o.om4 = o.F = o.Gq = o.D = o.ia = 0.f;
    o.tD = o.tD2 = o.itD2 = 0.f;
    o.spec_on = false;
    o.tmp = 0.f;
  }
  return o;
}

// Per-point constants in the form eval_lobe reads them.
struct PointGeom {
  float nx, ny, nz, vx, vy, vz;
  double nxd, nyd, nzd, vxd, vyd, vzd;
  float nv, k, omk, ib4, a2, a2pi, c4;
  float n2c;  // 1 - |n|^2 (formed in FP64), for eval_weights_f32
  double nvd, v2d;  // n.v and |v|^2 in FP64, for fwd_pair
  double a2m1d;     // alpha^2 - 1 in FP64, for fwd_pair
  float cD;         // ib4 * a2pi: the per-point factor of F G D / 4, for fwd_pair
};

__device__ __forceinline__ PointGeom point_geom(float nx, float ny, float nz, float vx, float vy,
                                                float vz, float r) {
  PointGeom c;
  c.nx = nx; c.ny = ny; c.nz = nz;
  c.vx = vx; c.vy = vy; c.vz = vz;
  c.nxd = nx; c.nyd = ny; c.nzd = nz;
  c.vxd = vx; c.vyd = vy; c.vzd = vz;
  c.nv = nx * vx + ny * vy + nz * vz;
  // bs:2126-2129: alpha2 = r^4, k = (r+1)^2 / 8
  const float r2 = r * r;
  c.a2 = r2 * r2;
  c.a2pi = c.a2 * kInvPi;
  c.c4 = -4.f * (c.a2 - 1.f);  // d tD / d(n.h) = 2 (alpha^2 - 1) n.h, times the -2 of d D / d tD
  c.k = (r + 1.f) * (r + 1.f) * 0.125f;
  c.omk = 1.f - c.k;
  c.ib4 = 0.25f / (fmaxf(c.nv, 0.f) * c.omk + c.k);
  c.n2c = (float)(1.0 - (c.nxd * c.nxd + c.nyd * c.nyd + c.nzd * c.nzd));
  c.nvd = fma(c.nzd, c.vzd, fma(c.nyd, c.vyd, c.nxd * c.vxd));
  c.v2d = fma(c.vzd, c.vzd, fma(c.vyd, c.vyd, c.vxd * c.vxd));
  c.a2m1d = (double)c.a2 - 1.0;
  c.cD = c.ib4 * c.a2pi;
  return c;
}

// The forward's two weights of one pair, each times the pair's visibility:
//   x = vis max(l.n, 0),   y = vis F G D / 4  (0 where the reference skips the lobe).
// MASK: vis is 0 or 1, given as mask = 0 or ~0 and applied to max(l.n, 0) as a
// bit mask; otherwise vis is the float weight vis_f.  G = max(l.n,0) ia ib is
// linear in max(l.n,0) (ia from the unmasked value), so y carries the weight
// through x.
//
// Every dot product is formed in FP64 from the table's FP64 direction l:
//   l.n, l.v;  l.hb = |l|^2 + l.v;  |hb|^2 = l.hb + (l.v + |v|^2);  n.hb = l.n + n.v,
// so the NDF denominator
//   tD = (n.h)^2 alpha^2 + (1 - (n.h)^2) = (|hb|^2 + (alpha^2 - 1) (n.hb)^2) / |hb|^2
// takes its numerator from FP64, rounded once (no cancellation at the peak of
// sharp lobes), and l.h = l.hb / |hb| gives the Fresnel term.  The two
// reciprocals of G1(l) and D share one: F G D / 4 = F x (ib/4) (alpha^2/pi) /
// (max(tD^2, thres/pi) (max(l.n,0)(1-k) + k)).  Same terms as eval_lobe.
template <bool SPEC, bool MASK>
__device__ __forceinline__ void fwd_pair(const PointGeom& s, const FwdTexel& g, uint32_t mask,
                                         float vis_f, float& x, float& y) {
  const double nl = fma(g.dz, s.nzd, fma(g.dy, s.nyd, g.dx * s.nxd));
  const float nlf = fmaxf((float)nl, 0.f);
  x = MASK ? __uint_as_float(__float_as_uint(nlf) & mask) : nlf * vis_f;
  y = 0.f;
  if (!SPEC) return;
  const double lv = fma(g.dz, s.vzd, fma(g.dy, s.vyd, g.dx * s.vxd));
  const double lhb = lv + g.l2;
  const double hb2 = lhb + (lv + s.v2d);
  const double nhb = nl + s.nvd;
  const float tdn = (float)fma(s.a2m1d, nhb * nhb, hb2);  // |hb|^2 tD
  const float hb2f = (float)hb2;
  // floor keeps the terms finite when v = -l exactly; such pairs are masked below
  const float rs = rsq(fmaxf(hb2f, 1e-20f));
  const float om = 1.f - sat((float)lhb * rs);
  const float om2 = om * om;
  const float F = fmaf(om2 * om2 * om, 1.f - kF0, kF0);
  const float tD = tdn * (rs * rs);
  const float den = fmaxf(tD * tD, kThresOverPi) * fmaf(nlf, s.omk, s.k);
  const float tmp = ((F * x) * s.cD) * rcp(den);
  // the reference skips the lobe unless tmp > 0; |l+v| = 0 gives NaN there
  y = (tmp > 0.f && hb2 > 0.0) ? tmp : 0.f;
}

// The two weights of a pair in fp32 only, for consumers that round them to
// 16 bits (the bf16 / fp16 relighting): w1 = max(l.n, 0) and w2 = F G D / 4
// (0 where the reference skips the lobe).  1 - (n.h)^2 comes from the
// Lagrange identity |hb|^2 - (n.hb)^2 = |n x hb|^2 + (1 - |n|^2) |hb|^2, the
// cross product formed with fused multiply-adds, so the NDF denominator keeps
// a few ulps of relative accuracy at the peak of sharp lobes instead of the
// cancellation of the reference's 1 - (n.h)^2.
template <bool SPEC>
__device__ __forceinline__ void eval_weights_f32(const PointGeom& s, const Texel& g, float& w1,
                                                 float& w2) {
  const float cs = fmaf(g.lz, s.nz, fmaf(g.ly, s.ny, g.lx * s.nx));
  w1 = fmaxf(cs, 0.f);
  w2 = 0.f;
  if (!SPEC) return;
  const float hx = g.lx + s.vx, hy = g.ly + s.vy, hz = g.lz + s.vz;
  const float hb2 = fmaf(hz, hz, fmaf(hy, hy, hx * hx));
  const float nhb = fmaf(s.nz, hz, fmaf(s.ny, hy, s.nx * hx));
  const float cx = fmaf(s.ny, hz, -s.nz * hy);
  const float cy = fmaf(s.nz, hx, -s.nx * hz);
  const float cz = fmaf(s.nx, hy, -s.ny * hx);
  const float sn2 = fmaf(s.n2c, hb2, fmaf(cz, cz, fmaf(cy, cy, cx * cx)));
  const float rs = rsq(fmaxf(hb2, 1e-20f));
  const float rs2 = rs * rs;
  const float ch = fmaf(g.lz, hz, fmaf(g.ly, hy, g.lx * hx)) * rs;
  const float nh = nhb * rs;
  const float om = 1.f - sat(ch);
  const float om2 = om * om;
  const float F = fmaf(om2 * om2 * om, 1.f - kF0, kF0);
  const float Gq = w1 * rcp(fmaf(w1, s.omk, s.k)) * s.ib4;
  const float tD = fmaf(nh * nh, s.a2, sn2 * rs2);
  const float D = s.a2pi * rcp(fmaxf(tD * tD, kThresOverPi));
  const float tmp = F * Gq * D;
  w2 = (tmp > 0.f && hb2 > 0.f) ? tmp : 0.f;
}

}  // namespace math
}  // namespace des
