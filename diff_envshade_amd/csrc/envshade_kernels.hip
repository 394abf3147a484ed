// envshade_kernels.hip -- gfx950 kernels for the brute-force environment-light
// shader (forward, fused backward, deterministic env-map-gradient reduce).
//
// Work decomposition:
//   * a workgroup owns one tile of 64 points (lane = point) and NTC waves;
//     per staged super-chunk of texels the waves share the tile's texel
//     table and visibility words in LDS, with one dark bit per texel (the
//     texels that light no lane of the tile); the lit texels (forward) or
//     lit groups of 4 texels (backward) are dealt to the waves in turn;
//   * the per-texel table (rotated light direction in FP64, E_c * dOmega) is
//     computed from closed forms in FP64 and rounded once where it is fp32;
//     every lane reads the same entry (LDS broadcast);
//   * per-point sums are reduced over the NTC waves through LDS in fixed
//     order; per-texel env-map-gradient sums are reduced over the 64 lanes
//     with permlane32/permlane16 swaps + row DPP (4 texels per reduction) and
//     stored as per-tile partials, which one more launch sums in FP64 in a
//     fixed order (bitwise reproducible, no atomics);
//   * the tiles beyond those resident at once run as texel slices, their
//     per-point sums combined after the launch.
//
// Semantics follow zjumsj/diff_envshade bruteforce_shader.cu ("bs"):
//   texel geometry bs:124-147 with the LEFT_TOP_AS_ORIGIN row flip (config.h:11);
//   Lambert + GGX (F bs:149-159, G bs:161-201, D bs:203-227);
//   variants: diffuse bs:231-530, specular bs:532-1038, clamp bs:1040-1552,
//   v2 bs:1554-2066, visibility v2 bs:2068-2683.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "envshade_internal.h"
#include "shade_math.h"

namespace des {

namespace {

using namespace math;

// ---- cross-lane helpers (wave64) -------------------------------------------------

template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}

// Sum of the 16 lanes of each row, returned in every lane of that row.
__device__ __forceinline__ float row_allreduce(float s) {
  s += dpp<0x128>(s);  // row_ror:8
  s += dpp<0x124>(s);  // row_ror:4
  s += dpp<0x122>(s);  // row_ror:2
  s += dpp<0x121>(s);  // row_ror:1
  return s;
}

// Reduce four per-lane values q0..q3 (one per texel) over all 64 lanes.
// Result: every lane of row k (lanes 16k..16k+15) holds the total of texel
// kRowTexel[k] = {0, 2, 1, 3}[k].
__device__ __forceinline__ float reduce4(float q0, float q1, float q2, float q3) {
  // upper 32 lanes of the first operand <-> lower 32 lanes of the second
  auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(q0), __float_as_uint(q1), false, false);
  auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(q2), __float_as_uint(q3), false, false);
  float s01 = __uint_as_float(a[0]) + __uint_as_float(a[1]);  // rows 0,1: q0   rows 2,3: q1
  float s23 = __uint_as_float(b[0]) + __uint_as_float(b[1]);  // rows 0,1: q2   rows 2,3: q3
  // odd rows of the first operand <-> even rows of the second
  auto c = __builtin_amdgcn_permlane16_swap(__float_as_uint(s01), __float_as_uint(s23), false, false);
  float s = __uint_as_float(c[0]) + __uint_as_float(c[1]);  // row0 q0, row1 q2, row2 q1, row3 q3
  return row_allreduce(s);
}

// ---- per-point state ---------------------------------------------------------

template <int V, bool SPEC>
struct PointState : PointGeom {
  float live;  // 0 for padding lanes and culled points (clamp variant)
  // visibility
  int vert0, vert1, vert2;
  float c0, c1, c2;
};

template <int V, bool SPEC>
__device__ __forceinline__ void load_point(const Problem& pr, int64_t p, bool valid,
                                           PointState<V, SPEC>& s) {
  float nx = 0.f, ny = 0.f, nz = 0.f, vx = 0.f, vy = 0.f, vz = 0.f, r = 0.f;
  if (valid) {
    nx = pr.normal[3 * p + 0];
    ny = pr.normal[3 * p + 1];
    nz = pr.normal[3 * p + 2];
    if (SPEC || V == kClamp) {
      vx = pr.view_dir[3 * p + 0];
      vy = pr.view_dir[3 * p + 1];
      vz = pr.view_dir[3 * p + 2];
    }
    if (SPEC) r = pr.rough[p];
  }
  static_cast<PointGeom&>(s) = point_geom(nx, ny, nz, vx, vy, vz, r);
  s.live = valid ? 1.f : 0.f;
  if (V == kClamp && !(s.nv > 0.f)) s.live = 0.f;  // per-point back-face cull (bs:1111)
  s.vert0 = s.vert1 = s.vert2 = 0;
  s.c0 = s.c1 = s.c2 = 0.f;
  if (V == kVis && valid) {
    // bs:2145-2151: nearest triangle -> 3 vertices, barycentric weights (3 x P)
    const int64_t f = pr.tri[p];
    if (f >= 0 && f < (int64_t)pr.N_face) {
      int a = pr.fvl[f], b = pr.fvl[(int64_t)pr.N_face + f], c = pr.fvl[2 * (int64_t)pr.N_face + f];
      // out-of-range ids read as occluded instead of faulting
      const bool ok = a >= 0 && a < pr.N_vertex && b >= 0 && b < pr.N_vertex && c >= 0 &&
                      c < pr.N_vertex;
      if (ok) {
        s.vert0 = a;
        s.vert1 = b;
        s.vert2 = c;
        s.c0 = pr.bary[p];
        s.c1 = pr.bary[pr.P + p];
        s.c2 = pr.bary[2 * pr.P + p];
      }
    }
  }
}

// ---- the tile's point data, gathered once per tile ------------------------------
//
// Every wave of a tile shades the same 64 points, and in sorted order each
// point's fields are scattered over the arrays: a load instruction touches 64
// cache lines, and the tile's start is bound by how many of them the CU's
// vector memory path can take.  So each field group is gathered by one wave
// (group g by wave g mod ntc), written to LDS (slot-major, [slot][lane]), and
// read by every wave after a barrier.  Slots:
//   0-2 normal, 3-5 view_dir, 6 roughness, 7-9 the face's vertex ids (as int
//   bits; 0 with zero weights where the face or a vertex id is out of range),
//   10-12 barycentric weights; backward: 13-15 grad_shading, 16-18 albedo,
//   19-21 grad_diffuse_shading, 22-24 specular albedo.
// kPointXchSlots x 64 floats at LDS offset 0, before the texel table is built
// there.
template <int V, bool DIFF, bool SPEC, bool BWD>
__device__ __forceinline__ void gather_point(const Problem& pr, int64_t p, bool valid, int wv,
                                             int ntc, int lane, float* xch) {
  constexpr bool DOUT = (V == kV2 || V == kVis);
  auto put = [&](int slot, float x) { xch[slot * 64 + lane] = x; };
  auto put3 = [&](int slot, const float* a) {
    put(slot, valid ? a[3 * p] : 0.f);
    put(slot + 1, valid ? a[3 * p + 1] : 0.f);
    put(slot + 2, valid ? a[3 * p + 2] : 0.f);
  };
  constexpr int NG = BWD ? 5 : 3;
  for (int g = wv; g < NG; g += ntc) {
    if (g == 0) {
      put3(0, pr.normal);
      put(6, (SPEC && valid) ? pr.rough[p] : 0.f);
    } else if (g == 1) {
      if (SPEC || V == kClamp) put3(3, pr.view_dir);
    } else if (g == 2) {
      if (V == kVis) {
        // bs:2145-2151: nearest triangle -> 3 vertices, barycentric weights (3 x P)
        int a = 0, b = 0, c = 0;
        float c0 = 0.f, c1 = 0.f, c2 = 0.f;
        if (valid) {
          const float b0 = pr.bary[p], b1 = pr.bary[pr.P + p], b2 = pr.bary[2 * pr.P + p];
          const int64_t f = pr.tri[p];
          if (f >= 0 && f < (int64_t)pr.N_face) {
            const int fa = pr.fvl[f], fb = pr.fvl[(int64_t)pr.N_face + f];
            const int fc = pr.fvl[2 * (int64_t)pr.N_face + f];
            // out-of-range ids read as occluded instead of faulting
            if (fa >= 0 && fa < pr.N_vertex && fb >= 0 && fb < pr.N_vertex && fc >= 0 &&
                fc < pr.N_vertex) {
              a = fa; b = fb; c = fc;
              c0 = b0; c1 = b1; c2 = b2;
            }
          }
        }
        put(7, __int_as_float(a));
        put(8, __int_as_float(b));
        put(9, __int_as_float(c));
        put(10, c0);
        put(11, c1);
        put(12, c2);
      }
    } else if (g == 3) {
      put3(13, pr.g_sh);
      if (DIFF) put3(16, pr.albedo);
    } else {
      if (DIFF && DOUT) put3(19, pr.g_ds);
      if (SPEC) put3(22, pr.spec);
    }
  }
}

// The point state from the exchanged fields (as load_point builds it).
template <int V, bool SPEC>
__device__ __forceinline__ void read_point(const float* xch, bool valid, int lane,
                                           PointState<V, SPEC>& s) {
  auto get = [&](int slot) { return xch[slot * 64 + lane]; };
  const bool hv = SPEC || V == kClamp;
  static_cast<PointGeom&>(s) = point_geom(get(0), get(1), get(2), hv ? get(3) : 0.f,
                                          hv ? get(4) : 0.f, hv ? get(5) : 0.f,
                                          SPEC ? get(6) : 0.f);
  s.live = valid ? 1.f : 0.f;
  if (V == kClamp && !(s.nv > 0.f)) s.live = 0.f;  // per-point back-face cull (bs:1111)
  s.vert0 = s.vert1 = s.vert2 = 0;
  s.c0 = s.c1 = s.c2 = 0.f;
  if (V == kVis) {
    s.vert0 = __float_as_int(get(7));
    s.vert1 = __float_as_int(get(8));
    s.vert2 = __float_as_int(get(9));
    s.c0 = get(10);
    s.c1 = get(11);
    s.c2 = get(12);
  }
}

// Visibility words of one 32-texel slot (bs:2180-2185): word for vertex v of
// slot s at visibility[s * N_vertex + v].
struct VisWords {
  uint32_t w0, w1, w2;
};

// vis weight for bit b: sum_j c_j * bit_j (bs:2186-2190); the live flag of the
// other variants.  The wave-uniform case where all three vertices of every
// lane coincide (one word per slot, weight bit * (c0+c1+c2)) is taken by the
// shading loops as a bit mask instead.
template <int V, bool SPEC>
__device__ __forceinline__ float vis_weight(const PointState<V, SPEC>& s, const VisWords& w, int b) {
  if (V != kVis) return s.live;
  const float b0 = (float)__builtin_amdgcn_ubfe(w.w0, b, 1);
  const float b1 = (float)__builtin_amdgcn_ubfe(w.w1, b, 1);
  const float b2 = (float)__builtin_amdgcn_ubfe(w.w2, b, 1);
  return fmaf(b2, s.c2, fmaf(b1, s.c1, b0 * s.c0));
}

// The tile's visibility words for the slots of the super-chunk at s0 (sn
// texels) are staged in LDS once per super-chunk -- word of vertex j of lane l
// for local slot sl at vis[(sl * 3 + j) * 64 + l] -- so that the texel loop
// reads them from LDS instead of waiting on global memory per group.
template <int V, bool SPEC>
__device__ __forceinline__ void stage_vis(const Problem& pr, const PointState<V, SPEC>& st, int s0,
                                          int sn, bool single, uint32_t* vis, int wv, int ntc,
                                          int lane) {
  if (V != kVis) return;
  const int nsl = (sn + 31) >> 5;
  for (int sl = wv; sl < nsl; sl += ntc) {
    const uint32_t* base = pr.vis + (int64_t)((s0 >> 5) + sl) * pr.N_vertex;
    vis[(sl * 3 + 0) * 64 + lane] = base[st.vert0];
    if (!single) {
      vis[(sl * 3 + 1) * 64 + lane] = base[st.vert1];
      vis[(sl * 3 + 2) * 64 + lane] = base[st.vert2];
    }
  }
}

template <int V>
__device__ __forceinline__ VisWords staged_words(const uint32_t* vis, int sl, bool single,
                                                 int lane) {
  VisWords w{0u, 0u, 0u};
  if (V == kVis) {
    w.w0 = vis[(sl * 3 + 0) * 64 + lane];
    if (!single) {
      w.w1 = vis[(sl * 3 + 1) * 64 + lane];
      w.w2 = vis[(sl * 3 + 2) * 64 + lane];
    }
  }
  return w;
}

// ---- dark texels ----------------------------------------------------------------
//
// A pair adds exactly nothing -- to any output or gradient -- unless l.n > 0:
// the Lambert term is gated on l.n > 0, and the GGX term vanishes with G = 0
// at max(l.n, 0) = 0.  If the normals of a tile's contributing lanes lie in a
// cone (unit axis a, half-angle beta), texel l lights none of them when the
// angle between l and a is at least 90 deg + beta, i.e. when
//   l.a <= -sin(beta) |l|.
// The test widens the cone (cos(beta) lowered by 4e-6, sin(beta) raised by
// 1e-4), far more than the fp32 rounding of either side: a texel it calls dark
// has l.n < -1e-4 |l| |n| on every lane, which no lane's own fp32 cosine can
// round up to a positive value, so skipping it changes no result.
// The texels the cone test leaves lit with l.a <= 0 (a band about a great
// circle, about 15% of a 16x32 map per sorted tile) are then tested
// exactly: dark when the pair's own gate -- the same expression, bit for bit,
// as the shading loop's -- is closed on every contributing lane.  Texels with
// l.a > 0 light some lane (sum over lanes of l.n/|n| = |sum n/|n|| l.a > 0).
struct Cone {
  float ax, ay, az;  // axis
  float sb;          // widened sin(beta); -4: every texel dark, 4: none
};

__device__ __forceinline__ float uniform(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// Sum / minimum over the wave, the same in every lane: the 16 lanes of each
// row combined with row-rotate DPP, then the four rows' values (lanes 0, 16,
// 32, 48) in order -- no LDS round trips.
__device__ __forceinline__ float lane_value(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

__device__ __forceinline__ float wave_sum(float x) {
  x = row_allreduce(x);
  return ((lane_value(x, 0) + lane_value(x, 16)) + lane_value(x, 32)) + lane_value(x, 48);
}

__device__ __forceinline__ float wave_min(float x) {
  x = fminf(x, dpp<0x128>(x));  // row_ror:8
  x = fminf(x, dpp<0x124>(x));  // row_ror:4
  x = fminf(x, dpp<0x122>(x));  // row_ror:2
  x = fminf(x, dpp<0x121>(x));  // row_ror:1
  return fminf(fminf(lane_value(x, 0), lane_value(x, 16)),
               fminf(lane_value(x, 32), lane_value(x, 48)));
}

// Lanes that can contribute: live, with a finite non-zero normal.
template <int V, bool SPEC>
__device__ __forceinline__ bool cone_lane(const PointState<V, SPEC>& s) {
  const float n2 = fmaf(s.nz, s.nz, fmaf(s.ny, s.ny, s.nx * s.nx));
  return s.live != 0.f && n2 > 0.f && n2 < INFINITY;
}

// Cone of the normals of the lanes that can contribute.
template <int V, bool SPEC>
__device__ __forceinline__ Cone tile_cone(const PointState<V, SPEC>& s) {
  Cone c{0.f, 0.f, 0.f, 4.f};
  const float n2 = fmaf(s.nz, s.nz, fmaf(s.ny, s.ny, s.nx * s.nx));
  const bool inc = cone_lane(s);
  if (!__any(inc)) {
    c.sb = -4.f;
    return c;
  }
  const float r = inc ? rsq(n2) : 0.f;
  const float ux = inc ? s.nx * r : 0.f, uy = inc ? s.ny * r : 0.f, uz = inc ? s.nz * r : 0.f;
  const float sx = uniform(wave_sum(ux)), sy = uniform(wave_sum(uy)), sz = uniform(wave_sum(uz));
  const float s2 = fmaf(sz, sz, fmaf(sy, sy, sx * sx));
  if (!(s2 > 1e-6f)) return c;  // normals cancel: no useful cone
  const float ra = rsq(s2);
  const float ax = sx * ra, ay = sy * ra, az = sz * ra;
  const float cmin = uniform(wave_min(inc ? fmaf(uz, az, fmaf(uy, ay, ux * ax)) : 1.f)) - 4e-6f;
  if (!(cmin > 0.f)) return c;  // wider than a hemisphere
  c.ax = ax;
  c.ay = ay;
  c.az = az;
  c.sb = __builtin_amdgcn_sqrtf(fmaxf(0.f, fmaf(-cmin, cmin, 1.f))) + 1e-4f;
  return c;
}

// Dark bits of the 64 texels [base, base + 64) of a staged super-chunk, one
// texel per lane (have: the lane's texel is one of the sn staged, with fp32
// direction l): bit i set when texel base + i lights no lane of the tile, by
// the cone test or, for the texels it leaves in doubt, by lit_any(i) (the
// exact test, wave-uniform) returning false.  Texels past sn are dark.
template <class LitAny>
__device__ __forceinline__ uint64_t dark_bits(bool have, float lx, float ly, float lz,
                                              const Cone& c, LitAny&& lit_any) {
  bool dark = true, doubt = false;
  if (have) {
    const float len = __builtin_amdgcn_sqrtf(fmaf(lz, lz, fmaf(ly, ly, lx * lx)));
    const float la = fmaf(lz, c.az, fmaf(ly, c.ay, lx * c.ax));
    dark = la <= -c.sb * len;
    doubt = !dark && la <= 0.f;
  }
  uint64_t dk = __ballot(dark);
  uint64_t todo = __ballot(doubt);
#pragma unroll 1
  while (todo != 0ull) {
    const int i = __builtin_ctzll(todo);
    todo &= todo - 1ull;
    if (!lit_any(i)) dk |= 1ull << i;
  }
  return dk;
}

// The dark words of a staged super-chunk: word k (texels 64k .. 64k+63) at
// dmask[k], written by the wave that builds those table entries.  A cone
// wider than a hemisphere darkens no texel: all words 0.
__device__ __forceinline__ uint64_t read_dark_word(const uint64_t* dmask, int k) {
  const uint64_t w = dmask[k];
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)w);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(w >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// ---- per-texel table ------------------------------------------------------------

// The map's row and column angles ({sin, cos} of theta for flipped row iy at
// ang[iy], of phi for column x at ang[H + x]) computed into LDS once per tile,
// for the texel table; null when the map has more than kAngleSlots rows +
// columns (the table build then computes them per texel).  A barrier must
// follow before the table is built.  The last wave takes the first entries:
// wave 0 has the most point fields to gather.
__device__ __forceinline__ const double2* stage_angles(const Problem& pr, const Tiling& tl) {
  extern __shared__ float4 smem[];
  const int n = pr.H + pr.W;
  if (n > kAngleSlots) return nullptr;
  double2* ang = reinterpret_cast<double2*>(reinterpret_cast<char*>(smem) + tl.ang_off);
  for (int i = (int)(blockDim.x - 1 - threadIdx.x); i < n; i += (int)blockDim.x) {
    double sn, cs;
    if (i < pr.H)
      row_angle(i, pr.H, sn, cs);
    else
      col_angle(i - pr.H, pr.W, sn, cs);
    ang[i] = make_double2(sn, cs);
  }
  return ang;
}

// Texels [t0, t0+n) of the lat-long map into LDS entries [0, n); entries
// [n, n_alloc) are zero so padded texels contribute exactly nothing.
//   storage texel t: row r = t / W, column x = t % W, flipped row iy = H-1-r;
//   theta = ((iy+.5)/H - .5) pi, phi = 2 pi (x+.5)/W, d = (cos th cos ph, sin th, cos th sin ph)
//   dOmega = (sin(pi(iy+1)/H - pi/2) - sin(pi iy/H - pi/2)) 2pi/W
//          = 2 sin(pi/(2H)) cos(theta) 2pi/W   (exact identity, no cancellation)
//   l = M d with M row-major.
// Wave w builds the 64-entry blocks w, w + ntc, ... and with each block writes
// its dark word dmask[block] (dark_bits against the tile's lanes); every word
// up to entry n_alloc is written.
template <int V, bool SPEC>
__device__ void build_table(const Problem& pr, const PointState<V, SPEC>& st, const Cone& cone,
                            int t0, int n, int n_alloc, Texel* tab, uint64_t* dmask,
                            const double2* ang) {
  double M[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) M[i] = (double)pr.rotmat[i];
  const double kdw = solid_angle_scale(pr.H, pr.W);
  const int T = pr.T;
  const int lane = threadIdx.x & 63;
  const bool wide = cone.sb >= 1.f;
  const bool inc = cone_lane(st);
  for (int base = (int)threadIdx.x - lane; base < n_alloc; base += (int)blockDim.x) {
    const int i = base + lane;
    Texel e = zero_texel();
    if (i < n) {
      const int t = t0 + i;
      const TexelGeo geo = ang ? texel_geometry(t, pr.H, pr.W, ang, M, kdw)
                               : texel_geometry(t, pr.H, pr.W, M, kdw);
      e = make_texel(geo, pr.envmap[t], pr.envmap[T + t], pr.envmap[2 * T + t]);
    }
    if (i < n_alloc) tab[i] = e;
    uint64_t dk = 0ull;
    if (!wide) {
      dk = dark_bits(i < n, e.lx, e.ly, e.lz, cone, [&](int j) {
        const float lx = lane_value(e.lx, j), ly = lane_value(e.ly, j), lz = lane_value(e.lz, j);
        // the backward's gate: eval_lobe's l.n > 0, the same expression
        return __ballot(inc && fmaf(lz, st.nz, fmaf(ly, st.ny, lx * st.nx)) > 0.f) != 0ull;
      });
    }
    if (lane == 0) dmask[base >> 6] = dk;
  }
}

__device__ __forceinline__ double lane_value(double x, int l) {
  const int2 v = *reinterpret_cast<const int2*>(&x);
  int2 r;
  r.x = __builtin_amdgcn_readlane(v.x, l);
  r.y = __builtin_amdgcn_readlane(v.y, l);
  return *reinterpret_cast<const double*>(&r);
}

// The forward's table (FwdTexel entries), same texels, zero padding and dark
// words; the exact test is the forward's gate, fwd_pair's FP64 l.n > 0.
template <int V, bool SPEC>
__device__ void build_fwd_table(const Problem& pr, const PointState<V, SPEC>& st,
                                const Cone& cone, int t0, int n, int n_alloc, FwdTexel* tab,
                                uint64_t* dmask, const double2* ang) {
  double M[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) M[i] = (double)pr.rotmat[i];
  const double kdw = solid_angle_scale(pr.H, pr.W);
  const int T = pr.T;
  const int lane = threadIdx.x & 63;
  const bool wide = cone.sb >= 1.f;
  const bool inc = cone_lane(st);
  for (int base = (int)threadIdx.x - lane; base < n_alloc; base += (int)blockDim.x) {
    const int i = base + lane;
    FwdTexel e = zero_fwd_texel();
    if (i < n) {
      const int t = t0 + i;
      const TexelGeo geo = ang ? texel_geometry(t, pr.H, pr.W, ang, M, kdw)
                               : texel_geometry(t, pr.H, pr.W, M, kdw);
      e = make_fwd_texel(geo, pr.envmap[t], pr.envmap[T + t], pr.envmap[2 * T + t]);
    }
    if (i < n_alloc) tab[i] = e;
    uint64_t dk = 0ull;
    if (!wide) {
      float lx, ly, lz;
      texel_dir(e, lx, ly, lz);
      dk = dark_bits(i < n, lx, ly, lz, cone, [&](int j) {
        const double dx = lane_value(e.dx, j), dy = lane_value(e.dy, j), dz = lane_value(e.dz, j);
        // max((float)nl, 0) > 0 needs nl > 0
        return __ballot(inc && fma(dz, st.nzd, fma(dy, st.nyd, dx * st.nxd)) > 0.0) != 0ull;
      });
    }
    if (lane == 0) dmask[base >> 6] = dk;
  }
}

// Texel groups.  A staged super-chunk of sn texels is cut into groups of 4
// (a group never straddles a 32-texel visibility slot).  Every wave of the
// tile walks the groups in order: the lit groups -- those that light some
// lane -- are dealt to the ntc waves in turn, so each wave shades the same
// number of them (+-1) whatever way the tile faces, and a dark group k belongs
// to wave k mod ntc, which only writes its zero env-map partials.  The deal
// depends only on the tile's dark bits, so the sums are reproducible.
// lit(tg, d4) / dark(tg) get the group's first texel and, for lit groups, its
// 4 dark bits.
template <class Lit, class Dark>
__device__ __forceinline__ void for_groups(const uint64_t* dmask, int sn, int wv, int ntc,
                                           int lane, Lit&& lit, Dark&& dark) {
  const int ngt = (sn + 3) >> 2;
  int c = 0;  // lit groups before group k0 (wave-uniform)
#pragma unroll 1
  for (int k0 = 0; k0 < ngt; k0 += 16) {
    const int kn = min(16, ngt - k0);
    // Which of these 16 groups this wave owns, one group per lane (lane i
    // looks at group k0 + i), before any shading: the shading loop below then
    // has one exit, and its accumulators need no copies at control-flow joins.
    const uint64_t dk = read_dark_word(dmask, k0 >> 4);
    const uint64_t all4 = dk & (dk >> 1) & (dk >> 2) & (dk >> 3);  // bit 4 i: group i dark
    const bool in = lane < kn;
    const bool gdark = in && ((all4 >> (4 * (lane & 15))) & 1ull) != 0ull;
    const bool glit = in && !gdark;
    const uint64_t litm = __ballot(glit);
    const int rank = c + (int)__builtin_popcountll(litm & ((1ull << lane) - 1ull));
    uint32_t own = (uint32_t)__ballot(glit && (rank & (ntc - 1)) == wv);
    uint32_t own_dark = (uint32_t)__ballot(gdark && ((k0 + lane) & (ntc - 1)) == wv);
    c += (int)__builtin_popcountll(litm);
#pragma unroll 1
    while (own_dark != 0u) {
      const int i = __builtin_ctz(own_dark);
      own_dark &= own_dark - 1u;
      dark(4 * (k0 + i));
    }
#pragma unroll 1
    while (own != 0u) {
      const int i = __builtin_ctz(own);
      own &= own - 1u;
      lit(4 * (k0 + i), (uint32_t)(dk >> (4 * i)) & 0xfu);
    }
  }
}

// Lit texels of a staged super-chunk, for the forward (which writes nothing
// per texel): the lit texels -- those that light some lane -- are dealt to
// the ntc waves in turn (the j-th lit texel to wave j mod ntc).  Each wave
// first lists its own texels, in order, in LDS (list: room for ceil(sn/ntc)
// entries, 8-byte aligned, written and read by this wave only), then shades
// them four at a time (lit4: the four lobes form one block the compiler
// interleaves) and the last 0-3 one at a time (lit1), in loops without other
// branches, so the accumulators need no copies at control-flow joins.  The
// deal depends only on the tile's dark bits, so the sums are reproducible.
template <class Lit4, class Lit1>
__device__ __forceinline__ void for_lit_texels(const uint64_t* dmask, int sn, int wv, int ntc,
                                               int lane, uint16_t* list, Lit4&& lit4,
                                               Lit1&& lit1) {
  const uint64_t below = (1ull << lane) - 1ull;
  int c = 0;      // lit texels dealt before word k
  int nown = 0;   // this wave's texels listed so far
#pragma unroll 1
  for (int k = 0; 64 * k < sn; ++k) {
    const int left = sn - 64 * k;
    uint64_t litm = ~read_dark_word(dmask, k);
    if (left < 64) litm &= (1ull << left) - 1ull;  // only the staged texels
    const bool lit = ((litm >> lane) & 1ull) != 0ull;
    const int rank = c + (int)__builtin_popcountll(litm & below);
    const uint64_t own = __ballot(lit && (rank & (ntc - 1)) == wv);
    if ((own >> lane) & 1ull) list[nown + (int)__builtin_popcountll(own & below)] = (uint16_t)(64 * k + lane);
    nown += (int)__builtin_popcountll(own);
    c += (int)__builtin_popcountll(litm);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the list, written above, read below
  const uint2* l4 = reinterpret_cast<const uint2*>(list);
  int b = 0;
#pragma unroll 1
  for (; b + 4 <= nown; b += 4) {
    const uint2 e = l4[b >> 2];
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)e.x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)e.y);
    lit4((int)(lo & 0xffffu), (int)(lo >> 16), (int)(hi & 0xffffu), (int)(hi >> 16));
  }
#pragma unroll 1
  for (; b < nown; ++b) lit1(__builtin_amdgcn_readfirstlane((int)list[b]));
}

// ---- forward -----------------------------------------------------------------

// One wave's share of the lit texels of one staged super-chunk (sn texels),
// four pairs per block; when the tile's cone is wider than a hemisphere (no
// texel can be dark: unsorted points) the groups of 4 texels are dealt in a
// fixed stride instead, without the lists.  SINGLE: the wave-uniform case of one visibility word
// per slot, chosen once per super-chunk; the word's bit is applied as a mask
// and the lane's weight (the sum of its barycentrics) by the caller, once per
// sum.  Other variants' per-lane weight (live) is also applied by the caller.
// acc = {sum E vis max(l.n,0) per channel, sum E vis F G D / 4 per channel}.
template <int V, bool DIFF, bool SPEC, bool SINGLE>
__device__ __forceinline__ void forward_chunk(const PointState<V, SPEC>& st, const FwdTexel* tab,
                                              const uint32_t* svis, const uint64_t* dmask,
                                              uint16_t* list, bool wide, int sn, int wv, int ntc,
                                              int lane, float (&acc)[6]) {
  constexpr bool MASK = (V != kVis) || SINGLE;
  auto shade = [&](int t) {
    const FwdTexel& g = tab[t];
    uint32_t mask = ~0u;
    float vis = 1.f;
    if (V == kVis) {
      const VisWords w = staged_words<V>(svis, t >> 5, SINGLE, lane);
      if (SINGLE)
        mask = (uint32_t)__builtin_amdgcn_sbfe((int)w.w0, t & 31, 1);  // 0 or ~0
      else
        vis = vis_weight<V, SPEC>(st, w, t & 31);
    }
    float x, y;
    fwd_pair<SPEC, MASK>(st, g, mask, vis, x, y);
    if (DIFF) {
      acc[0] = fmaf(g.e0, x, acc[0]);
      acc[1] = fmaf(g.e1, x, acc[1]);
      acc[2] = fmaf(g.e2, x, acc[2]);
    }
    if (SPEC) {
      acc[3] = fmaf(g.e0, y, acc[3]);
      acc[4] = fmaf(g.e1, y, acc[4]);
      acc[5] = fmaf(g.e2, y, acc[5]);
    }
  };
  if (wide) {
    // no texel is dark (unsorted points): wave w takes groups of 4 texels
    // w, w + ntc, ...; texels past sn are zero entries of the table
#pragma unroll 1
    for (int tg = 4 * wv; tg < sn; tg += 4 * ntc) {
      shade(tg);
      shade(tg + 1);
      shade(tg + 2);
      shade(tg + 3);
    }
    return;
  }
  for_lit_texels(
      dmask, sn, wv, ntc, lane, list,
      [&](int t0, int t1, int t2, int t3) {
        shade(t0);
        shade(t1);
        shade(t2);
        shade(t3);
      },
      [&](int t) { shade(t); });
}

// One 64-point tile of one problem; called by every wave of the workgroup.
// Waves other than wave 0 return after the cross-chunk reduction; wave 0
// writes the outputs.  No barrier follows the last one executed here, so a
// caller reusing LDS must synchronise first.
// order: point of each tile position (tile * 64 + lane), or null for input order.
// Outputs of point p from its sums acc (kForwardAcc entries: sum E vis
// max(l.n,0) and sum E vis F G D / 4 per channel).
template <int V, bool DIFF, bool SPEC>
__device__ __forceinline__ void write_forward_outputs(const Problem& pr, int64_t p,
                                                      const float* acc) {
  constexpr bool DOUT = (V == kV2 || V == kVis);
  float dsh[3], out[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    dsh[c] = acc[c] * kInvPi;
    float o = 0.f;
    if (DIFF) o = pr.albedo[3 * p + c] * dsh[c];
    if (SPEC) o = fmaf(pr.spec[3 * p + c], acc[3 + c], o);
    out[c] = o;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) pr.sh[3 * p + c] = out[c];
  if (DOUT) {
#pragma unroll
    for (int c = 0; c < 3; ++c) pr.ds[3 * p + c] = dsh[c];
  }
}

// Texels [t_begin, t_end) of one 64-point tile (t_begin a multiple of 32);
// slice_item >= 0 marks a tail slice, whose per-point sums go to
// pr.spart[slice_item] instead of the outputs.
template <int V, bool DIFF, bool SPEC>
__device__ __forceinline__ void forward_tile(const Problem& pr, const Tiling& tl, int64_t tile,
                                             const int32_t* order, int t_begin, int t_end,
                                             int64_t slice_item) {
  extern __shared__ float4 smem[];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int64_t q = tile * 64 + lane;
  const bool valid = q < pr.P;
  const int64_t p = (order != nullptr && valid) ? (int64_t)order[q] : q;

  PointState<V, SPEC> st;
  const double2* ang;
  {
    float* xch = reinterpret_cast<float*>(smem);
    gather_point<V, DIFF, SPEC, false>(pr, p, valid, wv, tl.ntc, lane, xch);
    ang = stage_angles(pr, tl);
    __syncthreads();
    read_point<V, SPEC>(xch, valid, lane, st);
  }
  bool single = true;
  float csum = 0.f;
  if (V == kVis) {
    single = __all(st.vert0 == st.vert1 && st.vert1 == st.vert2);
    csum = st.c0 + st.c1 + st.c2;
  }

  const Cone cone = tile_cone(st);

  constexpr int NACC = kForwardAcc;
  float acc[NACC] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};

  FwdTexel* tab = reinterpret_cast<FwdTexel*>(smem);
  uint32_t* svis = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(smem) + tl.vis_off);
  uint64_t* dmask = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(smem) + tl.dark_off);
  uint16_t* list = reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(smem) + tl.list_off) +
                   wv * tl.list_cap;
  for (int s0 = t_begin; s0 < t_end; s0 += tl.st) {
    const int sn = min(tl.st, t_end - s0);
    __syncthreads();
    build_fwd_table(pr, st, cone, s0, sn, min(tl.st, (sn + 31) & ~31), tab, dmask, ang);
    stage_vis(pr, st, s0, sn, single, svis, wv, tl.ntc, lane);
    __syncthreads();
    const bool wide = cone.sb >= 1.f;  // a cone wider than a hemisphere: no dark texel
    if (single)
      forward_chunk<V, DIFF, SPEC, true>(st, tab, svis, dmask, list, wide, sn, wv, tl.ntc, lane,
                                         acc);
    else
      forward_chunk<V, DIFF, SPEC, false>(st, tab, svis, dmask, list, wide, sn, wv, tl.ntc, lane,
                                          acc);
  }
  // the per-lane weight the masked sums leave out: the barycentric sum of a
  // single-word visibility tile, or the live flag of the other variants
  const float wl = (V == kVis) ? (single ? csum : 1.f) : st.live;
#pragma unroll
  for (int j = 0; j < NACC; ++j) acc[j] *= wl;

  // fixed-order reduction over the NTC texel chunks
  if (tl.ntc > 1) {
    __syncthreads();
    float* red = reinterpret_cast<float*>(smem);
#pragma unroll
    for (int j = 0; j < NACC; ++j) red[(wv * NACC + j) * 64 + lane] = acc[j];
    __syncthreads();
    if (wv != 0) return;
#pragma unroll
    for (int j = 0; j < NACC; ++j) {
      float sum = red[j * 64 + lane];
      for (int w = 1; w < tl.ntc; ++w) sum += red[(w * NACC + j) * 64 + lane];
      acc[j] = sum;
    }
  }
  if (slice_item >= 0) {
    // a tail slice: its per-point sums, combined after the launch
    float* o = pr.spart + (int64_t)slice_item * NACC * 64;
#pragma unroll
    for (int j = 0; j < NACC; ++j) o[j * 64 + lane] = acc[j];
    return;
  }
  if (!valid) return;
  write_forward_outputs<V, DIFF, SPEC>(pr, p, acc);
}


// Per-point texel sums of the backward (one wave's share of the texels).
struct BwdSums {
  float ds0, ds1, ds2;  // sum E vis max(l.n,0)
  float sp0, sp1, sp2;  // sum E vis F G D / 4
  float gnx, gny, gnz;  // sum coef_n l + rs gdnh hb
  float gvx, gvy, gvz;  // sum rs gch l - rs^2 gch (l.h) hb + rs^3 gdnh (hb x (n x hb))
  float Sq;             // sum grad_a G / 4  (the G1(v) term: Sb = -4 Sq)
  float ga2a, ga2b;     // alpha^2 gradient: sum pi gD it2 and sum (n.h)^2 t
  float gk;             // sum grad_a nl (nl - 1) ia^2
};

// One wave's share of the texel groups of one staged super-chunk at s0 (sn
// texels): per-point sums into a, and per group the tile's env-map partials
// of its 4 texels into part.  The backward deals groups, not texels as the
// forward does, because its env-map partials are reduced per group of 4
// (dealing lit texels in batches of 4 was measured slower: each texel then
// computes its own table and visibility addresses).  Dark texels of a lit
// group are skipped one by one: four backward lobes in one block need more
// than the 128 registers that keep four waves per SIMD.
//
// The per-point factor ib/4 of F G D / 4 is left out of the pair's lobe (Gq =
// G1(l) = nl ia, y = vis F G1(l) D) and rides on B (B_c ib/4); the sums of y
// and of grad_a G1(l) are scaled by it once per tile (backward_points).  SINGLE
// is the visibility variant's one-word tile: the pair's bit masks max(l.n, 0)
// and the per-lane weight c0+c1+c2 rides on A and B (and on the tile's sums of
// x and y), so the pair needs no float weight.
template <int V, bool DIFF, bool SPEC, bool SINGLE>
__device__ __forceinline__ void backward_chunk(const PointState<V, SPEC>& st, const Texel* tab,
                                               const uint32_t* svis, const uint64_t* dmask, int s0,
                                               int sn, int wv, int ntc, int lane, const float* A,
                                               const float* B, int T, float* part, int row_texel,
                                               int ch_out, BwdSums& a) {
  static_assert(!SINGLE || V == kVis, "one-word tiles are the visibility variant's");
  for_groups(dmask, sn, wv, ntc, lane, [&](int tg, uint32_t d4) {
    const int t_out = s0 + tg + row_texel;
    const VisWords w = staged_words<V>(svis, tg >> 5, SINGLE, lane);
    const int b4 = tg & 31;
    {
      float q[4][3];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        q[i][0] = q[i][1] = q[i][2] = 0.f;
        if ((d4 >> i) & 1u) continue;
        const Texel& g = tab[tg + i];
        // the pair's weight: a bit mask (SINGLE) or a float (vis)
        const uint32_t m = SINGLE ? (uint32_t)__builtin_amdgcn_sbfe((int)w.w0, b4 + i, 1) : ~0u;
        const float vis = SINGLE ? 1.f : vis_weight<V, SPEC>(st, w, b4 + i);
        const Lobe lb = eval_lobe<SPEC, false>(st, g, m);
        float coef_n = 0.f;
        float x = 0.f, y = 0.f;
        if (DIFF) {
          x = SINGLE ? lb.nl : lb.nl * vis;
          a.ds0 = fmaf(g.e0, x, a.ds0);
          a.ds1 = fmaf(g.e1, x, a.ds1);
          a.ds2 = fmaf(g.e2, x, a.ds2);
          // gdot = vis [l.n > 0] sum_c A_c E'_c   (bs:2426-2438)
          const float AE = fmaf(A[2], g.e2, fmaf(A[1], g.e1, A[0] * g.e0));
          if (SINGLE)
            coef_n = (lb.nl > 0.f) ? AE : 0.f;
          else
            coef_n = (lb.cs > 0.f) ? AE * vis : 0.f;
        }
        if (SPEC) {
          y = SINGLE ? lb.tmp : lb.tmp * vis;
          a.sp0 = fmaf(g.e0, y, a.sp0);
          a.sp1 = fmaf(g.e1, y, a.sp1);
          a.sp2 = fmaf(g.e2, y, a.sp2);
          const float BE = fmaf(B[2], g.e2, fmaf(B[1], g.e1, B[0] * g.e0));
          // grad of tmp = F G D / 4, times ib/4 (carried by B)
          const float gt = lb.spec_on ? (SINGLE ? BE : BE * vis) : 0.f;
          const float PF = gt * lb.F;
          const float gF = gt * (lb.Gq * lb.D);
          const float gD = PF * lb.Gq;
          // G1 VJP (bs:173-201), with grad_a = gG ib = (gt F D / 4) ib and
          // grad_b = -gG G ib = -4 grad_a G/4; l.n > 0 holds wherever gt != 0
          const float grad_a = PF * lb.D;
          const float w = grad_a * (lb.ia * lb.ia);
          const float gdl = w * st.k;
          a.gk = fmaf(w, fmaf(lb.nl, lb.nl, -lb.nl), a.gk);
          a.Sq = fmaf(grad_a, lb.Gq, a.Sq);
          // Fresnel VJP (bs:155-159): 5 (F0-1) gF (1-sat c)^4 where 0 <= c <= 1
          const float gch = (lb.sch == lb.ch) ? (5.f * (kF0 - 1.f)) * gF * lb.om4 : 0.f;
          // NDF VJP (bs:212-227): D = a2 / (pi tD^2) gives dD/da2 = it2 and
          // dD/dtD = -2 D / tD = -2 D tD itD2 (1/tD = tD itD2 where the floor is
          // not hit; there the reference's gradient through tD is 0)
          const float gDi = gD * lb.itD2;  // pi gD it2
          a.ga2a += gDi;
          const float t = (lb.tD2 > kThresOverPi) ? gDi * (lb.D * lb.tD) : 0.f;  // g_t = -2 t
          a.ga2b = fmaf(lb.nh * lb.nh, t, a.ga2b);
          const float gdnh = t * lb.nh * st.c4;  // 2 g_t (alpha^2 - 1) n.h
          // h = hb rs with hb = l + v; grad_h = gch l + gdnh n.
          //   NDF term of grad_normal:   gdnh h = (rs gdnh) hb
          //   normalize VJP (bs:62-71):  rs (grad_h - h (h.grad_h))
          //     = rs gch (l - (l.h) h) + rs gdnh (n - (n.h) h)
          //     = rs gch l - rs^2 gch (l.h) hb + rs^3 gdnh (hb x (n x hb))
          // The last form of n - (n.h) h (size sin(theta_h)) avoids the
          // 1/sin(theta_h) cancellation of sharp lobes; hb is kept as a vector
          // because splitting it into l and v parts would cancel by ~1/|l+v|
          // when l and v are nearly opposite.
          const float rg = lb.rs * gdnh;
          const float q2 = lb.rs2 * gch * lb.ch;
          const float r3 = lb.rs2 * rg;
          const float coef_v = lb.rs * gch;
          coef_n += gdl;
          a.gnx = fmaf(rg, lb.hx, a.gnx);
          a.gny = fmaf(rg, lb.hy, a.gny);
          a.gnz = fmaf(rg, lb.hz, a.gnz);
          a.gvx = fmaf(r3, lb.ux, fmaf(-q2, lb.hx, fmaf(coef_v, g.lx, a.gvx)));
          a.gvy = fmaf(r3, lb.uy, fmaf(-q2, lb.hy, fmaf(coef_v, g.ly, a.gvy)));
          a.gvz = fmaf(r3, lb.uz, fmaf(-q2, lb.hz, fmaf(coef_v, g.lz, a.gvz)));
        }
        a.gnx = fmaf(coef_n, g.lx, a.gnx);
        a.gny = fmaf(coef_n, g.ly, a.gny);
        a.gnz = fmaf(coef_n, g.lz, a.gnz);
#pragma unroll
        for (int c = 0; c < 3; ++c) q[i][c] = fmaf(A[c], x, B[c] * y);
      }
      // env-map gradient: sum over the 64 points of this tile, 4 texels at a time
      const float r0 = reduce4(q[0][0], q[1][0], q[2][0], q[3][0]);
      const float r1 = reduce4(q[0][1], q[1][1], q[2][1], q[3][1]);
      const float r2 = reduce4(q[0][2], q[1][2], q[2][2], q[3][2]);
      if (ch_out < 3 && t_out < T) {
        const float val = (ch_out == 0) ? r0 : ((ch_out == 1) ? r1 : r2);
        part[(int64_t)ch_out * T + t_out] = val;
      }
    }
  }, [&](int tg) {
    // the 4 texels light no lane of the tile: their partials are exactly 0
    const int t_out = s0 + tg + row_texel;
    if (ch_out < 3 && t_out < T) part[(int64_t)ch_out * T + t_out] = 0.f;
  });
}

// Per-point gradients of point p from its reduced sums acc (kBackwardAcc
// entries, see BwdSums) and its upstream shading gradient gs.
template <int V, bool DIFF, bool SPEC>
__device__ __forceinline__ void write_backward_outputs(const Problem& pr,
                                                       const PointGeom& st, int64_t p,
                                                       const float* gs, const float* acc) {
  // gn = sum (coef_n l + rs gdnh hb) + v [n.v>0](1-k) Sb
  // gv = sum (rs gch l - rs^2 gch (l.h) hb + rs^3 gdnh (hb x (n x hb)))
  //      + n [n.v>0](1-k) Sb
  const float Sb = -4.f * acc[12];
  const float nvpos = (st.nv > 0.f) ? 1.f : 0.f;
  const float g1v = SPEC ? nvpos * st.omk * Sb : 0.f;
  pr.gn[3 * p + 0] = fmaf(st.vx, g1v, acc[6]);
  pr.gn[3 * p + 1] = fmaf(st.vy, g1v, acc[7]);
  pr.gn[3 * p + 2] = fmaf(st.vz, g1v, acc[8]);
  if (V != kDiffuse) {
    // disabled branches are written as zeros (bs:2889-2905, bs:3205-3214)
    pr.gv[3 * p + 0] = SPEC ? fmaf(st.nx, g1v, acc[9]) : 0.f;
    pr.gv[3 * p + 1] = SPEC ? fmaf(st.ny, g1v, acc[10]) : 0.f;
    pr.gv[3 * p + 2] = SPEC ? fmaf(st.nz, g1v, acc[11]) : 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) pr.gs[3 * p + c] = SPEC ? gs[c] * acc[3 + c] : 0.f;
    float gr = 0.f;
    if (SPEC) {
      const float r = pr.rough[p];
      const float gkt = fmaf(1.f - fmaxf(st.nv, 0.f), Sb, acc[15]);
      // grad of alpha^2: sum gD it2 + sum (n.h)^2 g_t = ga2a / pi - 2 ga2b
      const float ga2 = fmaf(acc[13], kInvPi, -2.f * acc[14]);
      // bs:2501-2502
      gr = fmaf(ga2 * 4.f * r * r, r, gkt * (1.f + r) * 0.25f);
    }
    pr.gr[p] = gr;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) pr.ga[3 * p + c] = DIFF ? gs[c] * acc[c] * kInvPi : 0.f;
}

// ---- backward ----------------------------------------------------------------
//
// Per pair (with x = vis max(l.n,0), y = vis F G D / 4, E'_c = E_c dOmega,
// A_c = (gs_c a_c [+ gd_c]) / pi, B_c = gs_c s_c):
//   grad_albedo_c   = gs_c / pi * sum E'_c x
//   grad_spec_c     = gs_c * sum E'_c y
//   gE_c(t)         = dOmega_t * sum_p (A_c x + B_c y)
//   l-coefficients of grad_normal / grad_view_dir come from the diffuse term,
//   the G1 VJP (bs:173-201) and the normalize VJP (bs:62-71); the half-vector
//   terms of the NDF and normalize VJPs are accumulated along l + v and along
//   hb x (n x hb); the G1 parts along n and v are per-point constants times a
//   scalar sum (Sb), applied once at the end.
// The rest of a tile once its points are read (backward_tile): the texel
// loop, the reduction over the tile's waves and the outputs.  SINGLE is the
// visibility variant's one-word tile.
template <int V, bool DIFF, bool SPEC, bool SINGLE>
__device__ __forceinline__ void backward_points(const Problem& pr, const Tiling& tl, int64_t tile,
                                                const PointState<V, SPEC>& st, int64_t p,
                                                bool valid, const float* gs, float* A, float* B,
                                                const double2* ang, int t_begin, int t_end,
                                                int64_t slice_item, int lane, int wv) {
  extern __shared__ float4 smem[];
  // the pair weight's per-lane factor: c0+c1+c2 on a one-word tile
  const float cw = SINGLE ? st.c0 + st.c1 + st.c2 : 1.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    A[c] *= cw;
    B[c] *= cw * st.ib4;
  }

  const Cone cone = tile_cone(st);

  BwdSums a{};

  Texel* tab = reinterpret_cast<Texel*>(smem);
  uint32_t* svis = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(smem) + tl.vis_off);
  uint64_t* dmask = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(smem) + tl.dark_off);
  float* part = pr.part + tile * 3 * pr.T;
  // lanes (lane & 15) < 3 of row k store channel (lane & 15) of texel kRowTexel[k]
  const int row = lane >> 4;
  const int row_texel = (row == 1) ? 2 : ((row == 2) ? 1 : row);
  const int ch_out = lane & 15;

  for (int s0 = t_begin; s0 < t_end; s0 += tl.st) {
    const int sn = min(tl.st, t_end - s0);
    __syncthreads();
    build_table(pr, st, cone, s0, sn, min(tl.st, (sn + 31) & ~31), tab, dmask, ang);
    stage_vis(pr, st, s0, sn, SINGLE, svis, wv, tl.ntc, lane);
    __syncthreads();
    backward_chunk<V, DIFF, SPEC, SINGLE>(st, tab, svis, dmask, s0, sn, wv, tl.ntc, lane, A, B,
                                          pr.T, part, row_texel, ch_out, a);
  }
  // the per-point factors the pair sums leave out: c0+c1+c2 of a one-word
  // tile on the sums of x and y, ib/4 on those of y and grad_a G1(l)
  a.ds0 *= cw;
  a.ds1 *= cw;
  a.ds2 *= cw;
  a.sp0 *= cw * st.ib4;
  a.sp1 *= cw * st.ib4;
  a.sp2 *= cw * st.ib4;
  a.Sq *= st.ib4;

  // fixed-order reduction of the per-point sums over the NTC texel chunks
  constexpr int NACC = kBackwardAcc;
  float acc[NACC] = {a.ds0, a.ds1, a.ds2, a.sp0, a.sp1, a.sp2, a.gnx, a.gny,
                     a.gnz, a.gvx, a.gvy, a.gvz, a.Sq,  a.ga2a, a.ga2b, a.gk};
  if (tl.ntc > 1) {
    __syncthreads();
    float* red = reinterpret_cast<float*>(smem);
#pragma unroll
    for (int j = 0; j < NACC; ++j) red[(wv * NACC + j) * 64 + lane] = acc[j];
    __syncthreads();
    if (wv != 0) return;
#pragma unroll
    for (int j = 0; j < NACC; ++j) {
      float sum = red[j * 64 + lane];
      for (int w = 1; w < tl.ntc; ++w) sum += red[(w * NACC + j) * 64 + lane];
      acc[j] = sum;
    }
  }
  if (slice_item >= 0) {
    // a tail slice: its per-point sums, combined after the launch
    float* out = pr.spart + (int64_t)slice_item * NACC * 64;
#pragma unroll
    for (int j = 0; j < NACC; ++j) out[j * 64 + lane] = acc[j];
    return;
  }
  if (!valid) return;
  write_backward_outputs<V, DIFF, SPEC>(pr, st, p, gs, acc);
}

// Texels [t_begin, t_end) of one 64-point tile (t_begin a multiple of 32);
// slice_item >= 0 marks a tail slice, whose per-point sums go to
// pr.spart[slice_item] instead of the outputs.
// ANGLES: the texel table reads the map's angles from LDS (stage_angles).  The
// single-problem launch computes them per texel instead: with the LDS table
// the one-word code path of its tail slices spilled and the headline backward
// ran 95.3 us against 92.7 (interleaved runs); the persistent batch launch
// gains 1.5% per step with it.
template <int V, bool DIFF, bool SPEC, bool ANGLES>
__device__ __forceinline__ void backward_tile(const Problem& pr, const Tiling& tl, int64_t tile,
                                              const int32_t* order, int t_begin, int t_end,
                                              int64_t slice_item) {
  extern __shared__ float4 smem[];
  constexpr bool DOUT = (V == kV2 || V == kVis);
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int64_t q0 = tile * 64 + lane;
  const bool valid = q0 < pr.P;
  const int64_t p = (order != nullptr && valid) ? (int64_t)order[q0] : q0;

  PointState<V, SPEC> st;
  float gs[3] = {0.f, 0.f, 0.f}, A[3] = {0.f, 0.f, 0.f}, B[3] = {0.f, 0.f, 0.f};
  const double2* ang;
  {
    float* xch = reinterpret_cast<float*>(smem);
    gather_point<V, DIFF, SPEC, true>(pr, p, valid, wv, tl.ntc, lane, xch);
    ang = ANGLES ? stage_angles(pr, tl) : nullptr;
    __syncthreads();
    read_point<V, SPEC>(xch, valid, lane, st);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gs[c] = xch[(13 + c) * 64 + lane];
      if (DIFF) {
        float up = gs[c] * xch[(16 + c) * 64 + lane];
        if (DOUT) up += xch[(19 + c) * 64 + lane];
        A[c] = up * kInvPi;
      }
      if (SPEC) B[c] = gs[c] * xch[(22 + c) * 64 + lane];
    }
  }
  // the visibility variant's one-word tiles take their own code path, so that
  // its registers are allocated apart from the general path's
  const bool single = (V == kVis) && __all(st.vert0 == st.vert1 && st.vert1 == st.vert2);
  if (single)
    backward_points<V, DIFF, SPEC, V == kVis>(pr, tl, tile, st, p, valid, gs, A, B, ang, t_begin,
                                              t_end, slice_item, lane, wv);
  else
    backward_points<V, DIFF, SPEC, false>(pr, tl, tile, st, p, valid, gs, A, B, ang, t_begin,
                                          t_end, slice_item, lane, wv);
}

// ---- single-problem launches: one workgroup per tile ---------------------------

// Tail item j of a single-problem backward launch -> tail tile k (tile split0 +
// k) and slice sl.  The slices of one tile gather the same points; workgroup ids go to
// the 8 XCDs in turn, so each group of 8 tail tiles is dealt slice-major (item
// o of the group: tile o mod m, slice o / m, m = the group's tiles) and the
// slices of a full group's tile share an XCD and its L2.
__device__ __forceinline__ void tail_item(int64_t j, int64_t ntail, int slices, int64_t& k,
                                          int& sl) {
  const int64_t g = j / (8 * slices);
  const int64_t o = j - g * 8 * slices;
  const int64_t m = min((int64_t)8, ntail - 8 * g);
  k = 8 * g + o % m;
  sl = (int)(o / m);
}

template <int V, bool DIFF, bool SPEC>
__global__ __launch_bounds__(1024) void shade_forward_kernel(Problem pr, Tiling tl) {
  const int64_t b = blockIdx.x;
  if (b < tl.split0) {
    forward_tile<V, DIFF, SPEC>(pr, tl, b, pr.order, 0, pr.T, -1);
  } else {
    // tail slices, dispatched after the whole tiles, tile-major (the backward's
    // grouping by XCD, tail_item, made the forward 0.65 us slower)
    const int64_t j = b - tl.split0;
    const int64_t tile = tl.split0 + j / tl.slices;
    const int t0 = (int)(j % tl.slices) * tl.slice_texels;
    forward_tile<V, DIFF, SPEC>(pr, tl, tile, pr.order, t0, min(pr.T, t0 + tl.slice_texels), j);
  }
}

// The forward's tail slices: per point the slices' sums in slice order, then
// the point's outputs.  One tail tile per 64-thread block, so the scattered
// loads and stores spread over as many CUs as there are tail tiles.
template <int V, bool DIFF, bool SPEC>
__global__ __launch_bounds__(64) void forward_combine_kernel(Problem pr, Tiling tl) {
  const int64_t k = blockIdx.x;  // tail tile k: tile split0 + k (< tiles, checked at launch)
  const int lane = (int)threadIdx.x;
  const int64_t q = (tl.split0 + k) * 64 + lane;
  if (q >= pr.P) return;
  const int64_t p = pr.order != nullptr ? (int64_t)pr.order[q] : q;
  constexpr int NACC = kForwardAcc;
  float acc[NACC];
#pragma unroll
  for (int j = 0; j < NACC; ++j) acc[j] = 0.f;
  for (int sl = 0; sl < tl.slices; ++sl) {
    const float* in = pr.spart + (k * tl.slices + sl) * NACC * 64;
#pragma unroll
    for (int j = 0; j < NACC; ++j) acc[j] += in[j * 64 + lane];
  }
  write_forward_outputs<V, DIFF, SPEC>(pr, p, acc);
}

// Four waves per SIMD (<= 128 VGPRs): the measured best for this kernel.
template <int V, bool DIFF, bool SPEC>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void shade_backward_kernel(
    Problem pr, Tiling tl) {
  const int64_t b = blockIdx.x;
  if (b < tl.split0) {
    backward_tile<V, DIFF, SPEC, false>(pr, tl, b, pr.order, 0, pr.T, -1);
  } else {
    // tail slices, dispatched after the whole tiles
    int64_t k;
    int sl;
    tail_item(b - tl.split0, (pr.P + 63) / 64 - tl.split0, tl.slices, k, sl);
    const int t0 = sl * tl.slice_texels;
    backward_tile<V, DIFF, SPEC, false>(pr, tl, tl.split0 + k, pr.order, t0,
                                        min(pr.T, t0 + tl.slice_texels), k * tl.slices + sl);
  }
}

// ---- batched launches: persistent workgroups pull tiles of many problems ------
//
// The descriptor table (BatchHeader + BatchSet[n]) lives in device memory, so a
// captured graph of these launches can be replayed for any batch that fits the
// workspace: sizes are read on the device, not baked into the launch.

// Highest set index whose first item is <= g (sets with no items share the
// next set's start, so the search lands on the set that owns g).
template <typename F>
__device__ __forceinline__ int find_set(int n, int64_t g, F first_item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first_item(mid) <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int64_t next_item(unsigned long long* counter) {
  __shared__ int64_t s_item;
  __syncthreads();
  if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(counter, 1ull);
  __syncthreads();
  return s_item;
}

template <int V, bool DIFF, bool SPEC>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void batch_forward_kernel(const BatchHeader* __restrict__ hdr,
                                                             const BatchSet* __restrict__ sets,
                                                             unsigned long long* counter,
                                                             int32_t sig) {
  if (hdr->sig_fwd != sig) return;  // the table was prepared for another launch
  const int64_t total = hdr->total_tiles;
  const int n = hdr->n_sets;
  for (;;) {
    const int64_t g = next_item(counter);
    if (g >= total) break;
    const int s = __builtin_amdgcn_readfirstlane(
        find_set(n, g, [&](int i) { return sets[i].tile0; }));
    forward_tile<V, DIFF, SPEC>(sets[s].pr, sets[s].tl_fwd, g - sets[s].tile0, sets[s].order_fwd,
                                0, sets[s].pr.T, -1);
  }
}

template <int V, bool DIFF, bool SPEC>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void batch_backward_kernel(
    const BatchHeader* __restrict__ hdr, const BatchSet* __restrict__ sets,
    unsigned long long* counter, int32_t sig) {
  if (!hdr->has_backward || hdr->sig_bwd != sig) return;
  const int64_t total = hdr->total_tiles;
  const int n = hdr->n_sets;
  for (;;) {
    const int64_t g = next_item(counter);
    if (g >= total) break;
    const int s = __builtin_amdgcn_readfirstlane(
        find_set(n, g, [&](int i) { return sets[i].tile0; }));
    backward_tile<V, DIFF, SPEC, true>(sets[s].pr, sets[s].tl_bwd, g - sets[s].tile0, sets[s].order_bwd,
                                 0, sets[s].pr.T, -1);
  }
}

// ---- env-map gradient: fixed-order FP64 reduction of the per-tile partials -----

// gE[c][t] = dOmega_t * sum over the tile rows of part[row][c*T + t], in FP64,
// rounded once to fp32, in one fixed order: wave w of a 1024-thread block
// covers rows [w*nrows/16, (w+1)*nrows/16) of a block of columns; within it
// the rows r with (r - r0) mod 4 = k are summed in order into sum k, the four
// sums are added as ((s0 + s1) + s2) + s3, and wave 0 adds the 16 wave sums in
// order.  The order depends only on nrows, so grad_envmap is bitwise
// reproducible, and the two thread mappings below give the same bits:
//   COLS = 16: lane l reads column l & 15 of the rows r0 + (l >> 4) + 4 i --
//     four 64-byte row segments per load, 96 blocks at the 16x32 headline
//     (latency-bound: many blocks, eight loads in flight per lane);
//   COLS = 64: lane l reads column l of every row, keeping the four sums --
//     one 256-byte row segment per load, for the batched launches whose
//     partials run to hundreds of MB (bandwidth-bound).
constexpr int kReduceWaves = 16;
constexpr int kReduceCols = 16;       // single-problem launches
constexpr int kReduceColsBatch = 64;  // batched launches

template <int COLS>
__device__ __forceinline__ void reduce_cols_item(const float* __restrict__ part,
                                                 float* __restrict__ gE, int64_t nrows, int H,
                                                 int W, int cb) {
  static_assert(COLS == 16 || COLS == 64, "column block");
  __shared__ double s_sum[kReduceWaves][64];
  const int T = H * W;
  const int ncol = 3 * T;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int j = (COLS == 16) ? (lane & 15) : lane;
  const int col = cb * COLS + j;
  const int64_t r0 = nrows * w / kReduceWaves;
  const int64_t r1 = nrows * (w + 1) / kReduceWaves;
  double tot = 0.0;
  if (COLS == 16) {
    const int rs = lane >> 4;
    double s = 0.0;
    if (col < ncol) {
      const float* src = part + col;
      int64_t r = r0 + rs;
      for (; r + 28 < r1; r += 32) {
        float v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = src[(r + 4 * i) * ncol];
#pragma unroll
        for (int i = 0; i < 8; ++i) s += (double)v[i];
      }
      for (; r < r1; r += 4) s += (double)src[r * ncol];
    }
    // sub-rows 0..3 of each column, in order: lanes j, j+16, j+32, j+48
    const double s1 = __shfl(s, j + 16, 64), s2 = __shfl(s, j + 32, 64);
    const double s3 = __shfl(s, j + 48, 64);
    tot = ((s + s1) + s2) + s3;
  } else {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (col < ncol) {
      const float* src = part + col;
      int64_t r = r0;
      for (; r + 8 <= r1; r += 8) {
        float v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = src[(r + i) * ncol];
#pragma unroll
        for (int i = 0; i < 8; ++i) s[i & 3] += (double)v[i];  // r - r0 = 8 k + i
      }
      for (; r < r1; ++r) s[(r - r0) & 3] += (double)src[r * ncol];
    }
    tot = ((s[0] + s[1]) + s[2]) + s[3];
  }
  if (COLS == 64 || lane < 16) s_sum[w][j] = tot;
  __syncthreads();
  if (w == 0 && (COLS == 64 || lane < 16) && col < ncol) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kReduceWaves; ++k) t += s_sum[k][j];
    const int tt = col % T;
    const int iy = H - 1 - tt / W;
    const double dw = solid_angle_scale(H, W) * sinpi(((double)iy + 0.5) / (double)H);
    gE[col] = (float)(t * dw);
  }
  __syncthreads();  // s_sum is reused by the next item of a persistent block
}

// One launch: the first ncomb blocks combine the backward's tail slices, one
// tile per block in its first wave (spread over as many CUs as there are tail
// tiles, since the combine is a chain of scattered loads and stores): per
// point the slices' sums in slice order, then the point's gradients.  They
// come first so that their loads overlap the column sums of the other blocks,
// which sum one block of env-map-gradient columns each.
// Column blocks 2m and 2m + 1 (16 columns, 64 bytes of each partials row) share
// the rows' 128-byte lines; workgroup ids are dealt to the 8 XCDs in turn, so
// item i of each 16 and item i + 8, which run on one XCD, take the two halves
// of one line and its L2 fetches the line once.  A last partial group of 16
// keeps the identity.
__device__ __forceinline__ int line_paired_block(int i, int n) {
  const int g = i & ~15;
  if (g + 16 > n) return i;
  return g + 2 * (i & 7) + ((i >> 3) & 1);
}

template <int V, bool DIFF, bool SPEC>
__global__ __launch_bounds__(1024) void envgrad_reduce_kernel(Problem pr, Tiling tl, int ncomb) {
  const int64_t nrows = (pr.P + 63) / 64;
  if ((int)blockIdx.x >= ncomb) {
    const int ncb = (int)gridDim.x - ncomb;
    reduce_cols_item<kReduceCols>(pr.part, pr.gE, nrows, pr.H, pr.W,
                                  line_paired_block((int)blockIdx.x - ncomb, ncb));
    return;
  }
  if (threadIdx.x >= 64) return;
  const int64_t k = blockIdx.x;  // tail tile k: tile split0 + k (< nrows, checked at launch)
  const int lane = (int)threadIdx.x;
  const int64_t q = (tl.split0 + k) * 64 + lane;
  if (q >= pr.P) return;
  const int64_t p = pr.order != nullptr ? (int64_t)pr.order[q] : q;
  constexpr int NACC = kBackwardAcc;
  float acc[NACC];
#pragma unroll
  for (int j = 0; j < NACC; ++j) acc[j] = 0.f;
  for (int sl = 0; sl < tl.slices; ++sl) {
    const float* in = pr.spart + (k * tl.slices + sl) * NACC * 64;
#pragma unroll
    for (int j = 0; j < NACC; ++j) acc[j] += in[j * 64 + lane];
  }
  PointState<V, SPEC> st;
  load_point<V, SPEC>(pr, p, true, st);
  float gs[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) gs[c] = pr.g_sh[3 * p + c];
  write_backward_outputs<V, DIFF, SPEC>(pr, st, p, gs, acc);
}

// The same column sums for every set of a batch, persistent blocks over the
// (set, column block) items.
__global__ __launch_bounds__(1024) void batch_envgrad_reduce_kernel(
    const BatchHeader* __restrict__ hdr, const BatchSet* __restrict__ sets, int32_t sig) {
  if (!hdr->has_backward || hdr->sig_bwd != sig) return;
  const int64_t total = hdr->total_red;
  const int n = hdr->n_sets;
  for (int64_t item = blockIdx.x; item < total; item += gridDim.x) {
    const int s = __builtin_amdgcn_readfirstlane(
        find_set(n, item, [&](int i) { return sets[i].red_item0; }));
    const int cb = (int)(item - sets[s].red_item0);
    const Problem& pr = sets[s].pr;
    reduce_cols_item<kReduceColsBatch>(pr.part, pr.gE, (pr.P + 63) / 64, pr.H, pr.W, cb);
  }
}

__global__ __launch_bounds__(64) void zero_counters_kernel(unsigned long long* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(256) void fill_zero_kernel(float* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = 0.f;
}

template <int V, bool DIFF, bool SPEC>
hipError_t launch_fwd(const Problem& pr, const Tiling& tl, hipStream_t stream, hipEvent_t ev0,
                      hipEvent_t ev1) {
  const int64_t tiles = (pr.P + 63) / 64;
  if (tl.split0 > tiles || tl.slices < 1) return hipErrorInvalidValue;
  const unsigned nblk = (unsigned)(tl.split0 + (tiles - tl.split0) * tl.slices);
  hipError_t e;
  if (ev0 != nullptr && (e = hipEventRecord(ev0, stream)) != hipSuccess) return e;
  hipLaunchKernelGGL((shade_forward_kernel<V, DIFF, SPEC>), dim3(nblk), dim3(64 * tl.ntc),
                     tl.lds_fwd, stream, pr, tl);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (ev1 != nullptr && (e = hipEventRecord(ev1, stream)) != hipSuccess) return e;
  if (tl.split0 == tiles) return hipSuccess;
  hipLaunchKernelGGL((forward_combine_kernel<V, DIFF, SPEC>), dim3((unsigned)(tiles - tl.split0)),
                     dim3(64), 0, stream, pr, tl);
  return hipGetLastError();
}

template <int V, bool DIFF, bool SPEC>
hipError_t launch_bwd(const Problem& pr, const Tiling& tl, hipStream_t stream) {
  const int64_t tiles = (pr.P + 63) / 64;
  const unsigned nblk = (unsigned)(tl.split0 + (tiles - tl.split0) * tl.slices);
  hipLaunchKernelGGL((shade_backward_kernel<V, DIFF, SPEC>), dim3(nblk), dim3(64 * tl.ntc),
                     tl.lds_bwd, stream, pr, tl);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_forward(int variant, bool diff, bool spec, const Problem& pr, const Tiling& tl,
                          hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1) {
  switch (variant) {
    case kDiffuse: return launch_fwd<kDiffuse, true, false>(pr, tl, stream, ev0, ev1);
    case kSpecular:
      if (diff && spec) return launch_fwd<kSpecular, true, true>(pr, tl, stream, ev0, ev1);
      if (diff) return launch_fwd<kSpecular, true, false>(pr, tl, stream, ev0, ev1);
      if (spec) return launch_fwd<kSpecular, false, true>(pr, tl, stream, ev0, ev1);
      return launch_fwd<kSpecular, false, false>(pr, tl, stream, ev0, ev1);
    case kClamp:
      if (diff && spec) return launch_fwd<kClamp, true, true>(pr, tl, stream, ev0, ev1);
      if (diff) return launch_fwd<kClamp, true, false>(pr, tl, stream, ev0, ev1);
      if (spec) return launch_fwd<kClamp, false, true>(pr, tl, stream, ev0, ev1);
      return launch_fwd<kClamp, false, false>(pr, tl, stream, ev0, ev1);
    case kV2:
      if (spec) return launch_fwd<kV2, true, true>(pr, tl, stream, ev0, ev1);
      return launch_fwd<kV2, true, false>(pr, tl, stream, ev0, ev1);
    case kVis:
      if (spec) return launch_fwd<kVis, true, true>(pr, tl, stream, ev0, ev1);
      return launch_fwd<kVis, true, false>(pr, tl, stream, ev0, ev1);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_backward(int variant, bool diff, bool spec, const Problem& pr, const Tiling& tl,
                           hipStream_t stream) {
  hipError_t e;
  switch (variant) {
    case kDiffuse: e = launch_bwd<kDiffuse, true, false>(pr, tl, stream); break;
    case kSpecular:
      if (diff && spec) e = launch_bwd<kSpecular, true, true>(pr, tl, stream);
      else if (diff) e = launch_bwd<kSpecular, true, false>(pr, tl, stream);
      else if (spec) e = launch_bwd<kSpecular, false, true>(pr, tl, stream);
      else e = launch_bwd<kSpecular, false, false>(pr, tl, stream);
      break;
    case kClamp:
      if (diff && spec) e = launch_bwd<kClamp, true, true>(pr, tl, stream);
      else if (diff) e = launch_bwd<kClamp, true, false>(pr, tl, stream);
      else if (spec) e = launch_bwd<kClamp, false, true>(pr, tl, stream);
      else e = launch_bwd<kClamp, false, false>(pr, tl, stream);
      break;
    case kV2:
      e = spec ? launch_bwd<kV2, true, true>(pr, tl, stream) : launch_bwd<kV2, true, false>(pr, tl, stream);
      break;
    case kVis:
      e = spec ? launch_bwd<kVis, true, true>(pr, tl, stream) : launch_bwd<kVis, true, false>(pr, tl, stream);
      break;
    default: return hipErrorInvalidValue;
  }
  return e;
}

namespace {

template <int V, bool DIFF, bool SPEC>
hipError_t launch_reduce(const Problem& pr, const Tiling& tl, hipStream_t stream) {
  const int ncol = 3 * pr.T;
  const int64_t nrows = (pr.P + 63) / 64;
  const int ncomb = (int)(nrows - tl.split0);
  const unsigned cb = (unsigned)((ncol + kReduceCols - 1) / kReduceCols);
  hipLaunchKernelGGL((envgrad_reduce_kernel<V, DIFF, SPEC>), dim3(cb + ncomb),
                     dim3(64 * kReduceWaves), 0, stream, pr, tl, ncomb);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_envgrad_reduce(int variant, bool diff, bool spec, const Problem& pr,
                                 const Tiling& tl, hipStream_t stream) {
  const int64_t nrows = (pr.P + 63) / 64;
  if (tl.split0 > nrows || tl.slices < 1) return hipErrorInvalidValue;
  switch (variant) {
    case kDiffuse: return launch_reduce<kDiffuse, true, false>(pr, tl, stream);
    case kSpecular:
      if (diff && spec) return launch_reduce<kSpecular, true, true>(pr, tl, stream);
      if (diff) return launch_reduce<kSpecular, true, false>(pr, tl, stream);
      if (spec) return launch_reduce<kSpecular, false, true>(pr, tl, stream);
      return launch_reduce<kSpecular, false, false>(pr, tl, stream);
    case kClamp:
      if (diff && spec) return launch_reduce<kClamp, true, true>(pr, tl, stream);
      if (diff) return launch_reduce<kClamp, true, false>(pr, tl, stream);
      if (spec) return launch_reduce<kClamp, false, true>(pr, tl, stream);
      return launch_reduce<kClamp, false, false>(pr, tl, stream);
    case kV2:
      return spec ? launch_reduce<kV2, true, true>(pr, tl, stream) : launch_reduce<kV2, true, false>(pr, tl, stream);
    case kVis:
      return spec ? launch_reduce<kVis, true, true>(pr, tl, stream) : launch_reduce<kVis, true, false>(pr, tl, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fill_zero(float* x, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, x,
                     n);
  return hipGetLastError();
}


namespace {

template <int V, bool DIFF, bool SPEC>
hipError_t launch_bfwd(const BatchLaunch& bl, hipStream_t stream) {
  hipLaunchKernelGGL((batch_forward_kernel<V, DIFF, SPEC>), dim3(bl.grid), dim3(64 * bl.ntc),
                     bl.lds, stream, bl.hdr, bl.sets, &bl.hdr_mut->fwd_counter, bl.sig);
  return hipGetLastError();
}

template <int V, bool DIFF, bool SPEC>
hipError_t launch_bbwd(const BatchLaunch& bl, hipStream_t stream) {
  hipLaunchKernelGGL((batch_backward_kernel<V, DIFF, SPEC>), dim3(bl.grid), dim3(64 * bl.ntc),
                     bl.lds, stream, bl.hdr, bl.sets, &bl.hdr_mut->bwd_counter, bl.sig);
  return hipGetLastError();
}

template <bool FWD, int V, bool DIFF, bool SPEC>
hipError_t launch_b(const BatchLaunch& bl, hipStream_t stream) {
  return FWD ? launch_bfwd<V, DIFF, SPEC>(bl, stream) : launch_bbwd<V, DIFF, SPEC>(bl, stream);
}

template <bool FWD>
hipError_t dispatch_batch(int variant, bool diff, bool spec, const BatchLaunch& bl,
                          hipStream_t stream) {
  switch (variant) {
    case kDiffuse: return launch_b<FWD, kDiffuse, true, false>(bl, stream);
    case kSpecular:
      if (diff && spec) return launch_b<FWD, kSpecular, true, true>(bl, stream);
      if (diff) return launch_b<FWD, kSpecular, true, false>(bl, stream);
      if (spec) return launch_b<FWD, kSpecular, false, true>(bl, stream);
      return launch_b<FWD, kSpecular, false, false>(bl, stream);
    case kClamp:
      if (diff && spec) return launch_b<FWD, kClamp, true, true>(bl, stream);
      if (diff) return launch_b<FWD, kClamp, true, false>(bl, stream);
      if (spec) return launch_b<FWD, kClamp, false, true>(bl, stream);
      return launch_b<FWD, kClamp, false, false>(bl, stream);
    case kV2:
      return spec ? launch_b<FWD, kV2, true, true>(bl, stream) : launch_b<FWD, kV2, true, false>(bl, stream);
    case kVis:
      return spec ? launch_b<FWD, kVis, true, true>(bl, stream) : launch_b<FWD, kVis, true, false>(bl, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_batch_forward(int variant, bool diff, bool spec, const BatchLaunch& bl,
                                hipStream_t stream) {
  hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, stream, &bl.hdr_mut->fwd_counter, 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return dispatch_batch<true>(variant, diff, spec, bl, stream);
}

hipError_t launch_batch_backward(int variant, bool diff, bool spec, const BatchLaunch& bl,
                                 hipStream_t stream) {
  hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, stream, &bl.hdr_mut->bwd_counter, 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = dispatch_batch<false>(variant, diff, spec, bl, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(batch_envgrad_reduce_kernel, dim3(bl.reduce_grid), dim3(64 * kReduceWaves), 0,
                     stream, bl.hdr, bl.sets, bl.sig);
  return hipGetLastError();
}

int batch_blocks_per_cu(int variant, bool diff, bool spec, bool fwd, int ntc, size_t lds) {
  // all instantiations of one direction share launch bounds; use the heaviest one
  int n = 0;
  hipError_t e;
  if (fwd)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_forward_kernel<kVis, true, true>,
                                                     64 * ntc, lds);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_backward_kernel<kVis, true, true>,
                                                     64 * ntc, lds);
  (void)variant; (void)diff; (void)spec;
  return (e == hipSuccess && n > 0) ? n : 1;
}

}  // namespace des
