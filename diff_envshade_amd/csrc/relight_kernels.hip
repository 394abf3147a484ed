// relight_kernels.hip -- many env maps, one point set: the per-pair weights
// W1 = vis max(l.n, 0) and W2 = vis F G D / 4 are generated tile by tile in
// registers, directly in the MFMA operand layout, and multiplied on the matrix
// cores against the env maps (forward) or the upstream gradients (env-map
// gradient).  See relight_internal.h for the algebra.
//
// MFMA: v_mfma_f32_16x16x32_{bf16,f16}.  Operand layouts (wave64):
//   A (16 x 32): lane l holds A[l % 16][8 (l / 16) + j], j = 0..7
//   B (32 x 16): lane l holds B[8 (l / 16) + j][l % 16]
//   D (16 x 16): lane l holds D[4 (l / 16) + i][l % 16], i = 0..3
// Forward:  D[p, col] += W[p, t] L[t, col]   (A = W tile: 16 points x 32 texels)
// Env grad: D[t, col] += W^T[t, p] AB[p, col] (A = W^T tile: 16 texels x 32 points)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "relight_internal.h"
#include "shade_math.h"

namespace des {

namespace {

using namespace math;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int PREC>
__device__ __forceinline__ f32x4 mma(const uint4& a, const uint4& b, const f32x4& c) {
  if constexpr (PREC == kPrecFP16) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
}

// Eight fp32 values -> one 16-byte fragment (round to nearest even); for the
// split precision also the fragment of the rounding residual.
template <int PREC>
__device__ __forceinline__ void to_frag(const float* w, uint4& hi, uint4& lo) {
  if constexpr (PREC == kPrecFP16) {
    f16x8 h;
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = (_Float16)w[j];
    hi = __builtin_bit_cast(uint4, h);
    lo = make_uint4(0, 0, 0, 0);
  } else {
    bf16x8 h;
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = (__bf16)w[j];
    hi = __builtin_bit_cast(uint4, h);
    if constexpr (PREC == kPrecBF16x3) {
      bf16x8 l;
#pragma unroll
      for (int j = 0; j < 8; ++j) l[j] = (__bf16)(w[j] - (float)h[j]);
      lo = __builtin_bit_cast(uint4, l);
    } else {
      lo = make_uint4(0, 0, 0, 0);
    }
  }
}

// Padding points get benign constants (finite lobe) and zero visibility.
__device__ __forceinline__ PointGeom load_point_consts(const RelightArgs& a, int64_t p, bool spec) {
  if (p >= a.P) return point_geom(0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.5f);
  const float nx = a.normal[3 * p], ny = a.normal[3 * p + 1], nz = a.normal[3 * p + 2];
  float vx = 0.f, vy = 0.f, vz = 1.f, r = 0.5f;
  if (spec) {
    vx = a.view_dir[3 * p];
    vy = a.view_dir[3 * p + 1];
    vz = a.view_dir[3 * p + 2];
    r = a.rough[p];
  }
  return point_geom(nx, ny, nz, vx, vy, vz, r);
}

// The weights of one pair.  The split precision carries ~16 bits of each
// weight and takes them from the FP64-geometry lobe the drop-in uses; bf16 and
// fp16 keep 8 / 11 bits, for which the fp32 lobe is accurate to a few ulps.
template <int PREC, bool SPEC>
__device__ __forceinline__ void pair_weights(const PointGeom& pc, const Texel& g, float vis,
                                             float& w1, float& w2) {
  if constexpr (PREC == kPrecBF16x3) {
    const Lobe lb = eval_lobe<SPEC>(pc, g);
    w1 = lb.nl * vis;
    w2 = SPEC ? lb.tmp * vis : 0.f;
  } else {
    eval_weights_f32<SPEC>(pc, g, w1, w2);
    w1 *= vis;
    w2 *= vis;
  }
}

// ---- preparation kernels ----------------------------------------------------------

// texel table in chunk order (entry i = chunk i / 32, entry i % 32), zero
// beyond T (tp entries); the radiance fields are unused here
__global__ __launch_bounds__(256) void relight_texels_kernel(RelightArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.tp) return;
  Texel g = zero_texel();
  const int t = i < 32 * a.kc ? chunk_texel(a, i >> 5, i & 31) : a.T;
  if (t < a.T) {
    double M[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k] = (double)a.rotmat[k];
    g = make_texel(texel_geometry(t, a.H, a.W, M, solid_angle_scale(a.H, a.W)), 0.f, 0.f, 0.f);
  }
  a.geo[i] = g;
}

// vsort[slot][q] = vis[slot][order[q]], thread per (slot, q).  Workgroups are
// dealt to the 8 XCDs round-robin by id; the id -> (slot, q block) map keeps
// every q block of one slot on one XCD, so each slot row (P words) is fetched
// into one L2 and the random gathers of its words hit there.
__global__ __launch_bounds__(256) void relight_permute_vis_kernel(const uint32_t* __restrict__ vis,
                                                                  const int32_t* __restrict__ order,
                                                                  int64_t P, int nslot, int nqb,
                                                                  uint32_t* __restrict__ vsort) {
  const int L = blockIdx.x;
  const int k = L >> 3;
  const int sl = (k / nqb) * 8 + (L & 7);
  const int64_t q = (int64_t)(k % nqb) * 256 + threadIdx.x;
  if (sl >= nslot || q >= P) return;
  vsort[(int64_t)sl * P + q] = vis[(int64_t)sl * P + order[q]];
}

// L fragments: block (chunk c, n-tile n), lane l, j -> t = 32 c + 8 (l/16) + j,
// column = col0 + 16 n + l % 16 = 3 e + channel; L = E_e[channel, t] dOmega_t.
template <int PREC>
__global__ __launch_bounds__(64) void relight_L_kernel(RelightArgs a) {
  const int c = blockIdx.x, n = blockIdx.y, l = threadIdx.x;
  const int col = a.col0 + 16 * n + (l & 15);
  const double kdw = solid_angle_scale(a.H, a.W);
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int t = chunk_texel(a, c, 8 * (l >> 4) + j);
    float x = 0.f;
    if (t < a.T && col < a.ncol) {
      const int iy = a.H - 1 - t / a.W;
      const double dw = kdw * sinpi(((double)iy + 0.5) / (double)a.H);
      x = (float)((double)a.envmaps[(int64_t)col * a.T + t] * dw);
    }
    v[j] = x;
  }
  uint4 hi, lo;
  to_frag<PREC>(v, hi, lo);
  uint4* dst = a.lfrag + ((int64_t)(c * a.ntg + n) * a.nf) * 64;
  dst[l] = hi;
  if (a.nf == 2) dst[64 + l] = lo;
}

// per-point constants of shading position q (point order[q], or q)
__global__ __launch_bounds__(256) void relight_points_kernel(RelightArgs a, bool spec) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= (int64_t)a.kp * 32) return;
  const int64_t p = (a.order != nullptr && q < a.P) ? (int64_t)a.order[q] : q;
  a.pconst[q] = load_point_consts(a, p, spec);
}

// A / B fragments of the env-gradient GEMM: block (k-step ks, n-tile n), lane l,
// j -> point p = 32 ks + 8 (l/16) + j, column = col0 + 16 n + l % 16.
template <int PREC>
__global__ __launch_bounds__(64) void relight_AB_kernel(RelightArgs a, bool spec) {
  const int64_t ks = blockIdx.x;
  const int n = blockIdx.y, l = threadIdx.x;
  const int col = a.col0 + 16 * n + (l & 15);
  const int ch = col % 3;
  float va[8], vb[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t q = 32 * ks + 8 * (l >> 4) + j;
    float A = 0.f, B = 0.f;
    if (q < a.P && col < a.ncol) {
      const int64_t p = a.order != nullptr ? (int64_t)a.order[q] : q;
      const float gs = a.g_sh[p * a.ncol + col];
      const float gd = a.g_ds[p * a.ncol + col];
      A = fmaf(gs, a.albedo[3 * p + ch], gd) * kInvPi;
      if (spec) B = gs * a.spec[3 * p + ch];
    }
    va[j] = A;
    vb[j] = B;
  }
  uint4 hi, lo;
  uint4* dst = a.abfrag + ((ks * a.ntg + n) * 2) * a.nf * 64;
  to_frag<PREC>(va, hi, lo);
  dst[l] = hi;
  if (a.nf == 2) dst[64 + l] = lo;
  to_frag<PREC>(vb, hi, lo);
  dst[a.nf * 64 + l] = hi;
  if (a.nf == 2) dst[a.nf * 64 + 64 + l] = lo;
}

// ---- forward: 4 waves x 16 points per workgroup, all texels, all columns ------------

template <int PREC, bool SPEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void relight_forward_kernel(
    RelightArgs a) {
  constexpr int NF = (PREC == kPrecBF16x3) ? 2 : 1;
  __shared__ uint4 s_frag[kRelightMaxNT * NF * 64];
  __shared__ Texel s_geo[32];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  // shading position q; its point p (the rows of 16 positions are normal-sorted
  // when a.order is set)
  const int64_t q = (int64_t)blockIdx.x * 64 + wv * 16 + (lane & 15);
  const bool valid = q < a.P;
  const int64_t p = (a.order != nullptr && valid) ? (int64_t)a.order[q] : q;
  const PointGeom pc = load_point_consts(a, p, SPEC);
  const int tb = 8 * (lane >> 4);

  f32x4 acc1[kRelightMaxNT], acc2[kRelightMaxNT];
#pragma unroll
  for (int n = 0; n < kRelightMaxNT; ++n) {
    acc1[n] = f32x4{0.f, 0.f, 0.f, 0.f};
    acc2[n] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  const int nfrag = a.ntg * NF * 64;
  for (int c = 0; c < a.kc; ++c) {
    __syncthreads();
    const uint4* src = a.lfrag + (int64_t)c * nfrag;
    for (int i = threadIdx.x; i < nfrag; i += 256) s_frag[i] = src[i];
    if (threadIdx.x < 32) s_geo[threadIdx.x] = a.geo[32 * c + threadIdx.x];
    __syncthreads();
    // the visibility of the lane's 8 texels (consecutive bits of one word)
    uint32_t bits = valid ? 0xffu : 0u;
    if (a.vis != nullptr && valid) {
      const int t0 = chunk_texel(a, c, tb);
      bits = (a.vsort[(int64_t)(t0 >> 5) * a.P + q] >> (t0 & 31)) & 0xffu;
    }
    // a chunk that the 16 points of this wave all see none of adds exactly 0
    if (!__any(bits != 0u)) continue;
    float w1[8], w2[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float vis = (float)__builtin_amdgcn_ubfe(bits, j, 1);
      pair_weights<PREC, SPEC>(pc, s_geo[tb + j], vis, w1[j], w2[j]);
      // one pair at a time: interleaving the eight lobes would hold all their
      // temporaries at once and cost occupancy
      __builtin_amdgcn_sched_barrier(0);
    }
    uint4 a1h, a1l, a2h, a2l;
    to_frag<PREC>(w1, a1h, a1l);
    if (SPEC) to_frag<PREC>(w2, a2h, a2l);
#pragma unroll
    for (int n = 0; n < kRelightMaxNT; ++n) {
      if (n < a.ntg) {
        const uint4 bh = s_frag[(n * NF) * 64 + lane];
        acc1[n] = mma<PREC>(a1h, bh, acc1[n]);
        if (SPEC) acc2[n] = mma<PREC>(a2h, bh, acc2[n]);
        if constexpr (NF == 2) {
          const uint4 bl = s_frag[(n * NF + 1) * 64 + lane];
          acc1[n] = mma<PREC>(a1h, bl, acc1[n]);
          acc1[n] = mma<PREC>(a1l, bh, acc1[n]);
          if (SPEC) {
            acc2[n] = mma<PREC>(a2h, bl, acc2[n]);
            acc2[n] = mma<PREC>(a2l, bh, acc2[n]);
          }
        }
      }
    }
  }
  // epilogue: D[position row][column]; outputs are P x E x 3 (column = 3 e + channel)
  const int64_t qrow0 = (int64_t)blockIdx.x * 64 + wv * 16 + 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t qq = qrow0 + i;
    if (qq >= a.P) continue;
    const int64_t pp = a.order != nullptr ? (int64_t)a.order[qq] : qq;
    float al[3] = {0.f, 0.f, 0.f}, sp[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      al[ch] = a.albedo[3 * pp + ch];
      if (SPEC) sp[ch] = a.spec[3 * pp + ch];
    }
#pragma unroll
    for (int n = 0; n < kRelightMaxNT; ++n) {
      const int col = a.col0 + 16 * n + (lane & 15);
      if (n < a.ntg && col < a.ncol) {
        const int ch = col % 3;
        const float alb = ch == 0 ? al[0] : (ch == 1 ? al[1] : al[2]);
        const float spc = ch == 0 ? sp[0] : (ch == 1 ? sp[1] : sp[2]);
        const float d1 = acc1[n][i] * kInvPi;
        a.ds[pp * a.ncol + col] = d1;
        a.sh[pp * a.ncol + col] = SPEC ? fmaf(spc, acc2[n][i], alb * d1) : alb * d1;
      }
    }
  }
}

// ---- env-map gradient: 4 waves x 16 texels per workgroup, one point chunk ----------

template <int PREC, bool SPEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void relight_envgrad_kernel(
    RelightArgs a) {
  constexpr int NF = (PREC == kPrecBF16x3) ? 2 : 1;
  __shared__ uint4 s_ab[kRelightMaxNT * 2 * NF * 64];
  __shared__ PointGeom s_pc[32];
  __shared__ uint32_t s_vis[8][32];  // [8-texel group of the 2 chunks][position]
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  // Workgroup id -> (texel block bx, point chunk by): the workgroups of one
  // point chunk read the same fragments of A and B, so they are kept on one
  // XCD (ids are dealt to the 8 XCDs round-robin) and share its L2.
  const int tblocks = a.tp / kRelightTexBlock;
  const int k8 = (int)blockIdx.x >> 3;
  const int by = (k8 / tblocks) * 8 + ((int)blockIdx.x & 7);
  const int bx = k8 % tblocks;
  if (by >= a.npc) return;
  const int i = (bx * 4 + wv) * 16 + (lane & 15);  // chunk-order texel, < tp
  const Texel g = a.geo[i];
  const int cg = (((i >> 5) - 2 * bx) << 2) | ((i & 31) >> 3);  // 0..7
  const int tbit = (chunk_texel(a, i >> 5, i & 24) & 31) + (i & 7);
  const int pb = 8 * (lane >> 4);

  f32x4 acc[kRelightMaxNT];
#pragma unroll
  for (int n = 0; n < kRelightMaxNT; ++n) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int ks0 = by * a.ks_per_chunk;
  const int ks1 = min(a.kp, ks0 + a.ks_per_chunk);
  const int nab = a.ntg * 2 * NF * 64;
  for (int ks = ks0; ks < ks1; ++ks) {
    __syncthreads();
    {
      // word of 8-texel group (chunk 2 bx + cgi / 4, group cgi % 4) for position pt
      const int cgi = threadIdx.x >> 5, pt = threadIdx.x & 31;
      const int64_t q = (int64_t)ks * 32 + pt;
      const int c = 2 * bx + (cgi >> 2);
      uint32_t w = 0u;
      if (q < a.P && c < a.kc) {
        const int t0 = chunk_texel(a, c, (cgi & 3) * 8);
        if (t0 < a.T) w = (a.vis != nullptr) ? a.vsort[(int64_t)(t0 >> 5) * a.P + q] : 0xffffffffu;
      }
      s_vis[cgi][pt] = w;
    }
    __syncthreads();
    // the visibility of this texel for the lane's 8 positions
    uint32_t bits = 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) bits |= ((s_vis[cg][pb + j] >> tbit) & 1u) << j;
    // 64 texels that none of the 32 positions sees add exactly 0: skip the
    // k-step before staging its fragments (the barrier also orders the s_vis
    // reads above before the next k-step overwrites it)
    if (!__syncthreads_or(bits != 0u)) continue;
    const uint4* src = a.abfrag + (int64_t)ks * nab;
    for (int i2 = threadIdx.x; i2 < nab; i2 += 256) s_ab[i2] = src[i2];
    if (threadIdx.x < 32) s_pc[threadIdx.x] = a.pconst[(int64_t)ks * 32 + threadIdx.x];
    __syncthreads();
    // 16 texels that none of the 32 positions sees add exactly 0
    if (!__any(bits != 0u)) continue;
    float w1[8], w2[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int pt = pb + j;
      const PointGeom& pc = s_pc[pt];
      const float vis = (float)__builtin_amdgcn_ubfe(bits, j, 1);
      pair_weights<PREC, SPEC>(pc, g, vis, w1[j], w2[j]);
      __builtin_amdgcn_sched_barrier(0);  // one pair at a time, as in the forward
    }
    uint4 a1h, a1l, a2h, a2l;
    to_frag<PREC>(w1, a1h, a1l);
    if (SPEC) to_frag<PREC>(w2, a2h, a2l);
#pragma unroll
    for (int n = 0; n < kRelightMaxNT; ++n) {
      if (n < a.ntg) {
        const uint4* fr = s_ab + (n * 2) * NF * 64;
        const uint4 Ah = fr[lane];
        acc[n] = mma<PREC>(a1h, Ah, acc[n]);
        if constexpr (NF == 2) {
          const uint4 Al = fr[64 + lane];
          acc[n] = mma<PREC>(a1h, Al, acc[n]);
          acc[n] = mma<PREC>(a1l, Ah, acc[n]);
        }
        if (SPEC) {
          const uint4 Bh = fr[NF * 64 + lane];
          acc[n] = mma<PREC>(a2h, Bh, acc[n]);
          if constexpr (NF == 2) {
            const uint4 Bl = fr[NF * 64 + 64 + lane];
            acc[n] = mma<PREC>(a2h, Bl, acc[n]);
            acc[n] = mma<PREC>(a2l, Bh, acc[n]);
          }
        }
      }
    }
  }
  // epilogue: D[texel row][column] -> part[chunk][t][column]
  const int nc = a.ntg * 16;
  float* dst = a.part + (int64_t)by * a.tp * nc;
  const int trow0 = (bx * 4 + wv) * 16 + 4 * (lane >> 4);
#pragma unroll
  for (int n = 0; n < kRelightMaxNT; ++n) {
    if (n < a.ntg) {
#pragma unroll
      for (int i = 0; i < 4; ++i) dst[(int64_t)(trow0 + i) * nc + 16 * n + (lane & 15)] = acc[n][i];
    }
  }
}

// gE_e[c, t] = dOmega_t * sum over point chunks (FP64, fixed order), rounded once
__global__ __launch_bounds__(256) void relight_reduce_kernel(RelightArgs a) {
  const int nc = a.ntg * 16;
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)(32 * a.kc) * nc) return;
  const int i = (int)(idx / nc);  // chunk-order row
  const int cl = (int)(idx % nc);
  const int col = a.col0 + cl;
  if (col >= a.ncol) return;
  const int t = chunk_texel(a, i >> 5, i & 31);
  if (t >= a.T) return;
  double s = 0.0;
  for (int k = 0; k < a.npc; ++k) s += (double)a.part[((int64_t)k * a.tp + i) * nc + cl];
  const int iy = a.H - 1 - t / a.W;
  const double dw = solid_angle_scale(a.H, a.W) * sinpi(((double)iy + 0.5) / (double)a.H);
  a.gE[(int64_t)col * a.T + t] = (float)(s * dw);
}

template <int PREC>
hipError_t launch_fwd_p(bool spec, const RelightArgs& a, hipStream_t s) {
  const unsigned grid = (unsigned)((a.P + 63) / 64);
  if (spec)
    hipLaunchKernelGGL((relight_forward_kernel<PREC, true>), dim3(grid), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((relight_forward_kernel<PREC, false>), dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_eg_p(bool spec, const RelightArgs& a, hipStream_t s) {
  // 1-D: (texel block, point chunk) from the id, the point chunks rounded up to
  // a multiple of the 8 XCDs (see relight_envgrad_kernel)
  const dim3 grid((unsigned)(a.tp / kRelightTexBlock * ((a.npc + 7) / 8) * 8));
  if (spec)
    hipLaunchKernelGGL((relight_envgrad_kernel<PREC, true>), grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((relight_envgrad_kernel<PREC, false>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t relight_prepare_texels(const RelightArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(relight_texels_kernel, dim3((unsigned)((a.tp + 255) / 256)), dim3(256), 0, s,
                     a);
  return hipGetLastError();
}

hipError_t relight_prepare_L(int prec, const RelightArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.kc, (unsigned)a.ntg);
  switch (prec) {
    case kPrecBF16: hipLaunchKernelGGL(relight_L_kernel<kPrecBF16>, grid, dim3(64), 0, s, a); break;
    case kPrecFP16: hipLaunchKernelGGL(relight_L_kernel<kPrecFP16>, grid, dim3(64), 0, s, a); break;
    case kPrecBF16x3: hipLaunchKernelGGL(relight_L_kernel<kPrecBF16x3>, grid, dim3(64), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t relight_launch_forward(int prec, bool spec, const RelightArgs& a, hipStream_t s) {
  switch (prec) {
    case kPrecBF16: return launch_fwd_p<kPrecBF16>(spec, a, s);
    case kPrecFP16: return launch_fwd_p<kPrecFP16>(spec, a, s);
    case kPrecBF16x3: return launch_fwd_p<kPrecBF16x3>(spec, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t relight_prepare_points(bool spec, const RelightArgs& a, hipStream_t s) {
  const unsigned grid = (unsigned)(((int64_t)a.kp * 32 + 255) / 256);
  hipLaunchKernelGGL(relight_points_kernel, dim3(grid), dim3(256), 0, s, a, spec);
  return hipGetLastError();
}

hipError_t relight_prepare_AB(int prec, bool spec, const RelightArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.kp, (unsigned)a.ntg);
  switch (prec) {
    case kPrecBF16: hipLaunchKernelGGL(relight_AB_kernel<kPrecBF16>, grid, dim3(64), 0, s, a, spec); break;
    case kPrecFP16: hipLaunchKernelGGL(relight_AB_kernel<kPrecFP16>, grid, dim3(64), 0, s, a, spec); break;
    case kPrecBF16x3: hipLaunchKernelGGL(relight_AB_kernel<kPrecBF16x3>, grid, dim3(64), 0, s, a, spec); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t relight_launch_envgrad(int prec, bool spec, const RelightArgs& a, hipStream_t s) {
  switch (prec) {
    case kPrecBF16: return launch_eg_p<kPrecBF16>(spec, a, s);
    case kPrecFP16: return launch_eg_p<kPrecFP16>(spec, a, s);
    case kPrecBF16x3: return launch_eg_p<kPrecBF16x3>(spec, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t relight_reduce_envgrad(const RelightArgs& a, hipStream_t s) {
  const int64_t n = (int64_t)(32 * a.kc) * a.ntg * 16;
  hipLaunchKernelGGL(relight_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace des

namespace des {

hipError_t relight_permute_vis(const RelightArgs& a, int nslot, uint32_t* vsort, hipStream_t s) {
  const int nqb = (int)((a.P + 255) / 256);
  const unsigned grid = (unsigned)(nqb * ((nslot + 7) / 8) * 8);
  hipLaunchKernelGGL(relight_permute_vis_kernel, dim3(grid), dim3(256), 0, s, a.vis, a.order, a.P,
                     nslot, nqb, vsort);
  return hipGetLastError();
}

}  // namespace des
