/* envmap_oracle.c -- FP64 restatement of the reference's env-map utilities
 * (zjumsj/diff_envshade etc.cu), TEST INFRASTRUCTURE ONLY: loaded by tests/ and
 * bench.py's CPU baseline through oracle/oracle.py, never by the product.
 *
 *   bilinear lookup of a direction, seam wrapped      SampleEnvmap etc.cu:527-586
 *   sample_envmap_forward                             etc.cu:687-705, 1145-1178
 *   GGX (Trowbridge-Reitz) direction sampling         etc.cu:588-605
 *   blur_envmap_forward                               etc.cu:603-685, 1103-1143
 *
 * Conventions of config.h (LEFT_TOP_AS_ORIGIN, FIX_ENVMAP_SEAM): row 0 is the
 * top (+y), the seam at phi = 0 wraps.
 *
 * Random numbers: the reference draws from curand's XORWOW (seed 0, one
 * subsequence per target texel).  This restatement and the HIP kernel draw
 * from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC 2011) with key (0, 0) and counter (sample, target
 * texel, 0, 0); a uniform is (x + 0.5) / 2^32 of the first two output words.
 * The estimator is the reference's; its samples are not.
 *
 * Deliberate differences kept identical in the HIP path: a direction that is
 * not finite (a zero-length query) samples NaN; the output of
 * sample_envmap_forward has C rows (the reference allocates 3, etc.cu:1164).
 */
#include <math.h>
#include <stdint.h>

#define DES_PI 3.14159265358979323846

/* Philox4x32-10: ten rounds of two 32x32->64 multiplies, the key bumped by the
 * Weyl constants between rounds. */
static void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
  }
}

void des_oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; i++) out[i] = c[i];
}

static double u01(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

/* SampleEnvmap (etc.cu:527-586): bilinear lookup of unit direction (x, y, z)
 * in a C x H x W lat-long map, out[c * stride] for c < C. */
static void lookup(double x, double y, double z, int H, int W, int C, const float* env,
                   double* out, int64_t stride) {
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) {
    for (int c = 0; c < C; c++) out[c * stride] = NAN;
    return;
  }
  double phi = atan2(z, x);
  if (phi < 0.0) phi += 2.0 * DES_PI;
  if (y < -1.0) y = -1.0;
  else if (y > 1.0) y = 1.0;
  double theta = asin(y);
  double u = phi / (2.0 * DES_PI);
  double v = 1.0 - (theta / DES_PI + 0.5);
  double pu = u * W, pv = v * H;
  int ux0 = (int)floor(pu - 0.5), uy0 = (int)floor(pv - 0.5);
  int ux1 = ux0 + 1, uy1 = uy0 + 1;
  double kx = pu - (double)ux0 - 0.5, ky = pv - (double)uy0 - 0.5;
  if (ux0 < 0) ux0 += W;
  if (ux0 >= W) ux0 -= W;
  if (ux1 < 0) ux1 += W;
  if (ux1 >= W) ux1 -= W;
  if (uy0 < 0) uy0 = 0;
  if (uy0 >= H) uy0 = H - 1;
  if (uy1 < 0) uy1 = 0;
  if (uy1 >= H) uy1 = H - 1;
  for (int c = 0; c < C; c++) {
    const float* e = env + (int64_t)c * H * W;
    double v00 = e[uy0 * W + ux0], v01 = e[uy0 * W + ux1];
    double v10 = e[uy1 * W + ux0], v11 = e[uy1 * W + ux1];
    out[c * stride] = (1.0 - ky) * ((1.0 - kx) * v00 + kx * v01) + ky * ((1.0 - kx) * v10 + kx * v11);
  }
}

/* SampleEnvmap_Forward (etc.cu:687-705): query_point 3 x P (normalised here),
 * output C x P. */
void des_oracle_sample_envmap(int64_t P, int C, int H, int W, const double* q, const float* env,
                              double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < P; p++) {
    double x = q[p], y = q[P + p], z = q[2 * P + p];
    double L = sqrt(x * x + y * y + z * z);
    lookup(x / L, y / L, z / L, H, W, C, env, out + p, P);
  }
}

/* TrowbridgeReitzDistributionSample (etc.cu:588-605): a GGX half-vector in the
 * local frame (z = the mean direction) from two uniforms. */
void des_oracle_ggx_sample(double r0, double r1, double alpha, double* l) {
  double phi = 2.0 * DES_PI * r1;
  double tan2 = alpha * alpha * r0 / (1.0 - r0);
  double cos_t = 1.0 / sqrt(1.0 + tan2);
  double tmp = 1.0 - cos_t * cos_t;
  if (tmp < 0.0) tmp = 0.0;
  double sin_t = sqrt(tmp);
  l[0] = sin_t * cos(phi);
  l[1] = sin_t * sin(phi);
  l[2] = cos_t;
}

/* BlurEnvmap_Forward (etc.cu:603-685): for each texel of a tar_H x tar_W map,
 * the mean of N bilinear lookups along GGX-distributed directions around the
 * texel's centre direction.  Output C x tar_H x tar_W. */
void des_oracle_blur_envmap(int C, int H, int W, int tar_H, int tar_W, int N, double alpha,
                            const float* env, double* out) {
  const int64_t T = (int64_t)tar_H * tar_W;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t idx = 0; idx < T; idx++) {
    int ih = (int)(idx / tar_W), iw = (int)(idx % tar_W);
    double theta = (ih + 0.5) / tar_H * DES_PI;
    double phi = (iw + 0.5) / tar_W * 2.0 * DES_PI;
    double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
    double n[3] = {st * cp, ct, st * sp};
    double L = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int j = 0; j < 3; j++) n[j] /= L;
    double vx[3], vy[3];
    if (fabs(n[0]) > 0.9) {
      vx[0] = n[2]; vx[1] = 0.0; vx[2] = -n[0];
    } else {
      vx[0] = 0.0; vx[1] = -n[2]; vx[2] = n[1];
    }
    L = sqrt(vx[0] * vx[0] + vx[1] * vx[1] + vx[2] * vx[2]);
    for (int j = 0; j < 3; j++) vx[j] /= L;
    vy[0] = n[1] * vx[2] - n[2] * vx[1];
    vy[1] = n[2] * vx[0] - n[0] * vx[2];
    vy[2] = n[0] * vx[1] - n[1] * vx[0];
    L = sqrt(vy[0] * vy[0] + vy[1] * vy[1] + vy[2] * vy[2]);
    for (int j = 0; j < 3; j++) vy[j] /= L;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < N; i++) {
      uint32_t c4[4] = {(uint32_t)i, (uint32_t)idx, 0u, 0u};
      philox4x32_10(c4, 0u, 0u);
      double l[3], d[3], o[4];
      des_oracle_ggx_sample(u01(c4[0]), u01(c4[1]), alpha, l);
      for (int j = 0; j < 3; j++) d[j] = vx[j] * l[0] + vy[j] * l[1] + n[j] * l[2];
      L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      lookup(d[0] / L, d[1] / L, d[2] / L, H, W, C, env, o, 1);
      for (int c = 0; c < C; c++) acc[c] += o[c];
    }
    for (int c = 0; c < C; c++) out[(int64_t)c * T + idx] = acc[c] / N;
  }
}
