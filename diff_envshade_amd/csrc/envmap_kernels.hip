// envmap_kernels.hip -- env-map preprocessing beside the shader on gfx950:
// bilinear sampling of a lat-long map along query directions and its GGX blur
// (zjumsj/diff_envshade etc.cu:527-705).
//
// Both run in FP64 (directions, angles, bilinear weights, sums) and round once
// to float at the end.  Conventions as the reference's config.h: row 0 is the
// top (+y, LEFT_TOP_AS_ORIGIN), the seam at phi = 0 wraps (FIX_ENVMAP_SEAM).
//
// The blur's random numbers: Philox4x32-10 (Salmon et al., "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011), key (0, 0), counter (sample, target
// texel, 0, 0), uniforms (x + 0.5) / 2^32 of the first two words -- counter
// based, so any lane can draw any sample and the result does not depend on
// the launch shape.  The reference draws from curand's XORWOW instead; the
// estimator is the same, the samples are not.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "envmap_internal.h"

#pragma clang fp contract(off)

namespace des {
namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ void philox4x32_10(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

__device__ __forceinline__ double u01(uint32_t x) {
  return ((double)x + 0.5) * (1.0 / 4294967296.0);
}

// Bilinear taps of unit direction (x, y, z) (SampleEnvmap, etc.cu:527-574):
// plane offsets of the four texels and the weights kx, ky; nan = the
// direction is not finite.
struct Taps {
  int o00, o01, o10, o11;
  double kx, ky;
  bool nan;
};

__device__ __forceinline__ Taps taps(double x, double y, double z, int H, int W) {
  Taps t;
  t.nan = !(isfinite(x) && isfinite(y) && isfinite(z));
  if (t.nan) {
    t.o00 = t.o01 = t.o10 = t.o11 = 0;
    t.kx = t.ky = 0.0;
    return t;
  }
  double phi = atan2(z, x);
  if (phi < 0.0) phi += 2.0 * kPi;
  y = y < -1.0 ? -1.0 : (y > 1.0 ? 1.0 : y);
  const double theta = asin(y);
  const double u = phi / (2.0 * kPi);
  const double v = 1.0 - (theta / kPi + 0.5);
  const double pu = u * W, pv = v * H;
  int ux0 = (int)floor(pu - 0.5), uy0 = (int)floor(pv - 0.5);
  int ux1 = ux0 + 1, uy1 = uy0 + 1;
  t.kx = pu - (double)ux0 - 0.5;
  t.ky = pv - (double)uy0 - 0.5;
  if (ux0 < 0) ux0 += W;
  if (ux0 >= W) ux0 -= W;
  if (ux1 < 0) ux1 += W;
  if (ux1 >= W) ux1 -= W;
  uy0 = min(max(uy0, 0), H - 1);
  uy1 = min(max(uy1, 0), H - 1);
  t.o00 = uy0 * W + ux0;
  t.o01 = uy0 * W + ux1;
  t.o10 = uy1 * W + ux0;
  t.o11 = uy1 * W + ux1;
  return t;
}

// etc.cu:575-585
__device__ __forceinline__ double blend(const Taps& t, const float* __restrict__ plane) {
  if (t.nan) return __builtin_nan("");
  const double v00 = plane[t.o00], v01 = plane[t.o01], v10 = plane[t.o10], v11 = plane[t.o11];
  return (1.0 - t.ky) * ((1.0 - t.kx) * v00 + t.kx * v01) +
         t.ky * ((1.0 - t.kx) * v10 + t.kx * v11);
}

// SampleEnvmap_Forward (etc.cu:687-705): one thread per query point
// (query_point 3 x P, normalised here), output C x P.
__global__ __launch_bounds__(256) void sample_envmap_kernel(int64_t P, int C, int H, int W,
                                                            const float* __restrict__ q,
                                                            const float* __restrict__ env,
                                                            float* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  const double x = q[p], y = q[P + p], z = q[2 * P + p];
  const double L = sqrt(x * x + y * y + z * z);
  const Taps t = taps(x / L, y / L, z / L, H, W);
  const int64_t plane = (int64_t)H * W;
  for (int c = 0; c < C; ++c) out[c * P + p] = (float)blend(t, env + c * plane);
}

// BlurEnvmap_Forward (etc.cu:603-685): WPT waves per target texel (1 or all 4
// of the block).  The texel's lanes take samples l, l + 64 WPT, ...; each adds
// its GGX-distributed lookups in FP64, each wave sums its lanes in a fixed tree
// order and the block sums its waves in order, so the result is reproducible
// bit for bit.
template <int C, int WPT>
__global__ __launch_bounds__(256) void blur_envmap_kernel(int H, int W, int tar_H, int tar_W,
                                                          int N, float alpha,
                                                          const float* __restrict__ env,
                                                          float* __restrict__ out) {
  __shared__ double part[4][C];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int sub = wv % WPT;  // wave within the texel
  const int64_t T = (int64_t)tar_H * tar_W;
  const int64_t idx = (int64_t)blockIdx.x * (4 / WPT) + wv / WPT;
  if (idx >= T) return;  // whole waves; with WPT = 4 the grid holds exactly T blocks
  const int ih = (int)(idx / tar_W), iw = (int)(idx % tar_W);
  // the texel's centre direction and a frame around it (etc.cu:616-656)
  const double theta = (ih + 0.5) / tar_H * kPi;
  const double phi = (iw + 0.5) / tar_W * 2.0 * kPi;
  double st, ct, sp, cp;
  sincos(theta, &st, &ct);
  sincos(phi, &sp, &cp);
  double n[3] = {st * cp, ct, st * sp};
  double L = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= L;
  n[1] /= L;
  n[2] /= L;
  double vx[3];
  if (fabs(n[0]) > 0.9) {
    vx[0] = n[2];
    vx[1] = 0.0;
    vx[2] = -n[0];
  } else {
    vx[0] = 0.0;
    vx[1] = -n[2];
    vx[2] = n[1];
  }
  L = sqrt(vx[0] * vx[0] + vx[1] * vx[1] + vx[2] * vx[2]);
  vx[0] /= L;
  vx[1] /= L;
  vx[2] /= L;
  double vy[3] = {n[1] * vx[2] - n[2] * vx[1], n[2] * vx[0] - n[0] * vx[2],
                  n[0] * vx[1] - n[1] * vx[0]};
  L = sqrt(vy[0] * vy[0] + vy[1] * vy[1] + vy[2] * vy[2]);
  vy[0] /= L;
  vy[1] /= L;
  vy[2] /= L;
  const double a = alpha;
  const int64_t plane = (int64_t)H * W;
  double acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = 0.0;
  for (int i = sub * 64 + lane; i < N; i += 64 * WPT) {
    uint32_t r[4] = {(uint32_t)i, (uint32_t)idx, 0u, 0u};
    philox4x32_10(r, 0u, 0u);
    // TrowbridgeReitzDistributionSample (etc.cu:588-605)
    const double r0 = u01(r[0]), r1 = u01(r[1]);
    const double ps = 2.0 * kPi * r1;
    const double tan2 = a * a * r0 / (1.0 - r0);
    const double cos_t = 1.0 / sqrt(1.0 + tan2);
    double tmp = 1.0 - cos_t * cos_t;
    if (tmp < 0.0) tmp = 0.0;
    const double sin_t = sqrt(tmp);
    double sps, cps;
    sincos(ps, &sps, &cps);
    const double l0 = sin_t * cps, l1 = sin_t * sps, l2 = cos_t;
    double d[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) d[j] = vx[j] * l0 + vy[j] * l1 + n[j] * l2;
    L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const Taps t = taps(d[0] / L, d[1] / L, d[2] / L, H, W);
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] += blend(t, env + c * plane);
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) acc[c] += __shfl_xor(acc[c], d, 64);
  }
  if (WPT > 1) {
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < C; ++c) part[wv][c] = acc[c];
    }
    __syncthreads();
    if (sub != 0) return;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      double s = part[wv][c];
      for (int w = 1; w < WPT; ++w) s += part[wv + w][c];
      acc[c] = s;
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < C; ++c) out[c * T + idx] = (float)(acc[c] / N);
  }
}

}  // namespace

hipError_t launch_sample_envmap(int64_t P, int C, int H, int W, const float* q, const float* env,
                                float* out, hipStream_t stream) {
  if (P <= 0) return hipSuccess;
  hipLaunchKernelGGL(sample_envmap_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream,
                     P, C, H, W, q, env, out);
  return hipGetLastError();
}

// Waves per target texel: all four of a block when one wave per texel would
// leave most SIMDs without a wave and each texel has samples enough for four
// waves (at 2048 texels one wave per texel is as fast).
hipError_t launch_blur_envmap(int C, int H, int W, int tar_H, int tar_W, int N, float alpha,
                              const float* env, float* out, hipStream_t stream) {
  const int64_t T = (int64_t)tar_H * tar_W;
  if (T <= 0) return hipSuccess;
  const bool wide = T < 1024 && N >= 256;
  const dim3 grid((unsigned)(wide ? T : (T + 3) / 4));
  if (C == 3) {
    if (wide)
      hipLaunchKernelGGL((blur_envmap_kernel<3, 4>), grid, dim3(256), 0, stream, H, W, tar_H,
                         tar_W, N, alpha, env, out);
    else
      hipLaunchKernelGGL((blur_envmap_kernel<3, 1>), grid, dim3(256), 0, stream, H, W, tar_H,
                         tar_W, N, alpha, env, out);
  } else {
    if (wide)
      hipLaunchKernelGGL((blur_envmap_kernel<4, 4>), grid, dim3(256), 0, stream, H, W, tar_H,
                         tar_W, N, alpha, env, out);
    else
      hipLaunchKernelGGL((blur_envmap_kernel<4, 1>), grid, dim3(256), 0, stream, H, W, tar_H,
                         tar_W, N, alpha, env, out);
  }
  return hipGetLastError();
}

}  // namespace des
