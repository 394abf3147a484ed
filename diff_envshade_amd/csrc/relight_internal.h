// relight_internal.h -- many env maps lighting one point set on MFMA.
//
// shading_e[p, c]   = a_c / pi * (W1 L_e)[p, c] + s_c * (W2 L_e)[p, c]
// diffuse_e[p, c]   = (W1 L_e)[p, c] / pi
// gE_e[c, t]        = dOmega_t * sum_p (W1[p, t] A_e[p, c] + W2[p, t] B_e[p, c])
// with W1[p, t] = vis max(l_t . n_p, 0), W2[p, t] = vis F G D / 4 (the per-pair
// Lambert and GGX weights of the reference's visibility shader,
// bruteforce_shader.cu:2204-2234), L_e[t, c] = E_e[c, t] dOmega_t,
// A_e = (gs_e a + gd_e) / pi and B_e = gs_e s (bruteforce_shader.cu:2661-2675).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "shade_math.h"

namespace des {

enum RelightPrecision : int {
  kPrecBF16 = 0,    // W, L (and A, B) rounded to bf16, one MFMA per product
  kPrecFP16 = 1,    // the same in fp16
  kPrecBF16x3 = 2,  // hi + lo bf16 split of both operands, three MFMAs per product
};

constexpr int kRelightMaxNT = 12;   // n-tiles of 16 columns per launch (64 maps x 3)
constexpr int kRelightTexBlock = 64;  // texels per envgrad workgroup (4 waves x 16)

struct RelightArgs {
  int64_t P;
  int H, W, T;
  int E;           // env maps
  int col0;        // first (map, channel) column of this launch's column group
  int ncol;        // 3 E
  int ntg;         // n-tiles in this column group (<= kRelightMaxNT)
  int nf;          // fragments per operand: 1, or 2 for the hi/lo split
  int kc;          // 32-texel chunks = ceil(T / 32)
  int tp;          // padded texel count (multiple of 64)
  int kp;          // 32-point k-steps = ceil(P / 32)
  int npc;         // point chunks of the env-gradient pass
  int ks_per_chunk;
  const float* normal;
  const float* view_dir;
  const float* albedo;
  const float* spec;
  const float* rough;
  const float* envmaps;   // E x 3 x H x W
  const float* rotmat;
  const uint32_t* vis;    // N_slot x P (per-point words) or NULL = all visible
  const int32_t* order;   // shading order: point of sorted position q, or NULL (input order)
  const uint32_t* vsort;  // visibility words of sorted position q: [slot][q] (vis if unsorted)
  int blocked;            // K chunks are 4 x 8 texel blocks of the map (else 32-texel runs)
  const float* g_sh;      // P x E x 3
  const float* g_ds;      // P x E x 3
  float* sh;              // P x E x 3
  float* ds;              // P x E x 3
  float* gE;              // E x 3 x H x W
  // workspace
  math::Texel* geo;       // tp texel entries
  uint4* lfrag;           // [kc][ntg][nf][64] B-operand fragments of L
  math::PointGeom* pconst;  // kp*32 per-point constants
  uint4* abfrag;          // [kp][ntg][2][nf][64] B-operand fragments of A and B
  float* part;            // [npc][tp][ntg*16]
};

// Chunk layout.  The MFMA's K dimension runs over 32-texel chunks in any
// order of texels; with H % 4 == 0 and W % 8 == 0 chunk c is a 4 x 8 block of
// the map (rows 4 br .. 4 br + 3, columns 8 bc .. 8 bc + 7), entry j at row
// 4 br + j / 8, column 8 bc + j % 8, so that the chunk's directions are
// compact and more often dark for a row of normal-sorted points; otherwise
// chunk c is texels 32 c .. 32 c + 31.  Either way entries 8 g .. 8 g + 7 are
// consecutive storage texels inside one 32-texel visibility slot.
__host__ __device__ inline int chunk_texel(const RelightArgs& a, int c, int j) {
  if (a.blocked) {
    const int nbc = a.W >> 3;
    const int br = c / nbc, bc = c - br * nbc;
    return (4 * br + (j >> 3)) * a.W + 8 * bc + (j & 7);
  }
  return 32 * c + j;
}

hipError_t relight_prepare_texels(const RelightArgs& a, hipStream_t s);
// visibility words in shading order: vsort[slot][q] = vis[slot][order[q]]
hipError_t relight_permute_vis(const RelightArgs& a, int nslot, uint32_t* vsort, hipStream_t s);
hipError_t relight_prepare_L(int prec, const RelightArgs& a, hipStream_t s);
hipError_t relight_launch_forward(int prec, bool spec, const RelightArgs& a, hipStream_t s);
hipError_t relight_prepare_points(bool spec, const RelightArgs& a, hipStream_t s);
hipError_t relight_prepare_AB(int prec, bool spec, const RelightArgs& a, hipStream_t s);
hipError_t relight_launch_envgrad(int prec, bool spec, const RelightArgs& a, hipStream_t s);
hipError_t relight_reduce_envgrad(const RelightArgs& a, hipStream_t s);

}  // namespace des
