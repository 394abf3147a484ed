// relight_capi.cpp -- C ABI of the many-env-map relighting path on MFMA
// (include/diff_envshade.h, des_relight_*): validation, workspace layout,
// column groups of up to 64 maps per launch, and the launch sequence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/diff_envshade.h"
#include "envshade_internal.h"
#include "relight_internal.h"

namespace des_capi {
int fail(int code, const char* fmt, ...);
int ok();
}  // namespace des_capi

namespace {

using des::RelightArgs;
using des_capi::fail;
using des_capi::ok;

size_t a256(size_t x) { return (x + 255) & ~size_t(255); }

struct RelightLayout {
  size_t geo, lfrag, pconst, abfrag, part, order, vsort, total;
  int kc, tp, kp, ntg_max, nf, npc, ks_per_chunk;
};

// With visibility the points are shaded in normal-sorted order (rows of 16 /
// 32 points facing one way see whole chunks dark, which the kernels skip), so
// the workspace also holds the order and the visibility words in that order.
RelightLayout relight_layout(int64_t P, int E, int H, int W, int prec, bool envgrad,
                             bool with_vis) {
  RelightLayout L{};
  const int T = H * W;
  L.kc = (T + 31) / 32;
  L.tp = (T + des::kRelightTexBlock - 1) / des::kRelightTexBlock * des::kRelightTexBlock;
  L.kp = (int)((P + 31) / 32);
  L.ntg_max = std::min(des::kRelightMaxNT, (3 * E + 15) / 16);
  L.nf = prec == des::kPrecBF16x3 ? 2 : 1;
  const int tblocks = L.tp / des::kRelightTexBlock;
  int npc = std::max(1, (1024 + tblocks - 1) / tblocks);
  npc = std::min(npc, std::max(1, L.kp));
  L.ks_per_chunk = std::max(1, (L.kp + npc - 1) / npc);
  L.npc = std::max(1, (L.kp + L.ks_per_chunk - 1) / L.ks_per_chunk);
  size_t off = 0;
  L.geo = off;
  off += a256(size_t(L.tp) * sizeof(des::math::Texel));
  L.lfrag = off;
  off += a256(size_t(L.kc) * L.ntg_max * L.nf * 64 * 16);
  L.pconst = L.abfrag = L.part = off;
  if (envgrad) {
    L.pconst = off;
    off += a256(size_t(L.kp) * 32 * sizeof(des::math::PointGeom));
    L.abfrag = off;
    off += a256(size_t(L.kp) * L.ntg_max * 2 * L.nf * 64 * 16);
    L.part = off;
    off += a256(size_t(L.npc) * L.tp * L.ntg_max * 16 * 4);
  }
  L.order = L.vsort = off;
  if (with_vis && P > 0) {
    L.order = off;
    off += a256(des::order_layout(P).total);
    L.vsort = off;
    off += a256(size_t(L.kc) * size_t(P) * sizeof(uint32_t));
  }
  L.total = off;
  return L;
}

int check_relight(int64_t P, int32_t E, int32_t H, int32_t W, int32_t N_slot, int32_t prec) {
  if (P < 0) return fail(DES_ERR_INVALID_ARGUMENT, "P must be >= 0");
  if (E < 1) return fail(DES_ERR_INVALID_ARGUMENT, "need at least one env map (E = %d)", E);
  if (H <= 0 || W <= 0 || int64_t(H) * W > (int64_t(1) << 24))
    return fail(DES_ERR_INVALID_ARGUMENT, "env maps must be E x 3 x H x W with 1 <= H*W <= 2^24");
  if (prec < des::kPrecBF16 || prec > des::kPrecBF16x3)
    return fail(DES_ERR_INVALID_ARGUMENT, "unknown precision %d", prec);
  if (int64_t(E) * 3 > (int64_t(1) << 20)) return fail(DES_ERR_INVALID_ARGUMENT, "too many env maps");
  if ((P + 31) / 32 > int64_t(0x7fffffff)) return fail(DES_ERR_INVALID_ARGUMENT, "P too large");
  if (N_slot >= 0 && N_slot < (H * W + 31) / 32)
    return fail(DES_ERR_INVALID_ARGUMENT, "visibility has %d slots, the %d-texel maps need %d", N_slot,
                H * W, (H * W + 31) / 32);
  return DES_OK;
}

RelightArgs make_args(const RelightLayout& L, int64_t P, int E, int H, int W, void* ws) {
  RelightArgs a{};
  a.blocked = (H % 4 == 0 && W % 8 == 0) ? 1 : 0;
  a.P = P;
  a.H = H;
  a.W = W;
  a.T = H * W;
  a.E = E;
  a.ncol = 3 * E;
  a.nf = L.nf;
  a.kc = L.kc;
  a.tp = L.tp;
  a.kp = L.kp;
  a.npc = L.npc;
  a.ks_per_chunk = L.ks_per_chunk;
  char* w = static_cast<char*>(ws);
  a.geo = reinterpret_cast<des::math::Texel*>(w + L.geo);
  a.lfrag = reinterpret_cast<uint4*>(w + L.lfrag);
  a.pconst = reinterpret_cast<des::math::PointGeom*>(w + L.pconst);
  a.abfrag = reinterpret_cast<uint4*>(w + L.abfrag);
  a.part = reinterpret_cast<float*>(w + L.part);
  return a;
}

int launch_err(hipError_t e, const char* what) {
  return fail(DES_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
}

// With visibility: sort the points by normal direction and gather their
// visibility words into that order; without, shade in input order.
hipError_t prepare_order(const RelightLayout& L, RelightArgs& a, void* ws, hipStream_t s) {
  a.order = nullptr;
  a.vsort = a.vis;
  if (a.vis == nullptr) return hipSuccess;
  char* w = static_cast<char*>(ws);
  hipError_t e = des::launch_point_order(a.normal, a.P, w + L.order, s);
  if (e != hipSuccess) return e;
  a.order = reinterpret_cast<const int32_t*>(w + L.order + des::order_layout(a.P).order_off);
  uint32_t* vs = reinterpret_cast<uint32_t*>(w + L.vsort);
  e = des::relight_permute_vis(a, a.kc, vs, s);
  a.vsort = vs;
  return e;
}

}  // namespace

extern "C" {

DES_API size_t des_relight_workspace_size(int64_t P, int32_t E, int32_t H, int32_t W,
                                          int32_t precision, int with_envgrad) {
  if (P < 0 || E < 1 || H <= 0 || W <= 0 || precision < 0 || precision > 2) return 0;
  return relight_layout(P, E, H, W, precision, with_envgrad != 0, true).total;
}

DES_API int des_relight_forward(int64_t P, int32_t E, int32_t H, int32_t W, int32_t N_slot,
                                const float* normal, const float* view_dir, const float* albedo,
                                const float* specular_albedo, const float* roughness,
                                const float* envmaps, const float* light2obj_rotmat,
                                const int32_t* visibility, int enable_specular, int32_t precision,
                                float* shading, float* diffuse_shading, void* workspace,
                                size_t workspace_bytes, void* stream) {
  if (int rc = check_relight(P, E, H, W, visibility ? N_slot : -1, precision)) return rc;
  if (P == 0) return ok();
  if (!normal || !albedo || !envmaps || !light2obj_rotmat || !shading || !diffuse_shading)
    return fail(DES_ERR_INVALID_ARGUMENT, "normal, albedo, envmaps, light2obj_rotmat and outputs must not be NULL");
  if (enable_specular && (!view_dir || !specular_albedo || !roughness))
    return fail(DES_ERR_INVALID_ARGUMENT, "view_dir, specular_albedo and roughness must not be NULL");
  const RelightLayout L = relight_layout(P, E, H, W, precision, false, visibility != nullptr);
  if (!workspace || workspace_bytes < L.total)
    return fail(DES_ERR_WORKSPACE_TOO_SMALL, "workspace of %zu bytes is smaller than the %zu required",
                workspace_bytes, L.total);
  if ((reinterpret_cast<uintptr_t>(workspace) & 255) != 0)
    return fail(DES_ERR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  RelightArgs a = make_args(L, P, E, H, W, workspace);
  a.normal = normal; a.view_dir = view_dir; a.albedo = albedo; a.spec = specular_albedo;
  a.rough = roughness; a.envmaps = envmaps; a.rotmat = light2obj_rotmat;
  a.vis = reinterpret_cast<const uint32_t*>(visibility);
  a.sh = shading; a.ds = diffuse_shading;
  const bool spec = enable_specular != 0;
  hipError_t e = prepare_order(L, a, workspace, s);
  if (e != hipSuccess) return launch_err(e, "point order");
  e = des::relight_prepare_texels(a, s);
  if (e != hipSuccess) return launch_err(e, "texel table");
  for (int col0 = 0; col0 < a.ncol; col0 += des::kRelightMaxNT * 16) {
    a.col0 = col0;
    a.ntg = std::min(des::kRelightMaxNT, (a.ncol - col0 + 15) / 16);
    if ((e = des::relight_prepare_L(precision, a, s)) != hipSuccess) return launch_err(e, "env-map fragments");
    if ((e = des::relight_launch_forward(precision, spec, a, s)) != hipSuccess) return launch_err(e, "relight forward");
  }
  return ok();
}

DES_API int des_relight_backward_envmap(int64_t P, int32_t E, int32_t H, int32_t W, int32_t N_slot,
                                        const float* normal, const float* view_dir,
                                        const float* albedo, const float* specular_albedo,
                                        const float* roughness, const float* light2obj_rotmat,
                                        const int32_t* visibility, const float* grad_shading,
                                        const float* grad_diffuse_shading, int enable_specular,
                                        int32_t precision, float* grad_envmaps, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  if (int rc = check_relight(P, E, H, W, visibility ? N_slot : -1, precision)) return rc;
  if (!grad_envmaps) return fail(DES_ERR_INVALID_ARGUMENT, "grad_envmaps must not be NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (P == 0) {
    hipError_t e = des::launch_fill_zero(grad_envmaps, int64_t(3) * E * H * W, s);
    return e == hipSuccess ? ok() : launch_err(e, "zero fill");
  }
  if (!normal || !albedo || !light2obj_rotmat || !grad_shading || !grad_diffuse_shading)
    return fail(DES_ERR_INVALID_ARGUMENT, "normal, albedo, light2obj_rotmat and upstream gradients must not be NULL");
  if (enable_specular && (!view_dir || !specular_albedo || !roughness))
    return fail(DES_ERR_INVALID_ARGUMENT, "view_dir, specular_albedo and roughness must not be NULL");
  const RelightLayout L = relight_layout(P, E, H, W, precision, true, visibility != nullptr);
  if (!workspace || workspace_bytes < L.total)
    return fail(DES_ERR_WORKSPACE_TOO_SMALL, "workspace of %zu bytes is smaller than the %zu required",
                workspace_bytes, L.total);
  if ((reinterpret_cast<uintptr_t>(workspace) & 255) != 0)
    return fail(DES_ERR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
  RelightArgs a = make_args(L, P, E, H, W, workspace);
  a.normal = normal; a.view_dir = view_dir; a.albedo = albedo; a.spec = specular_albedo;
  a.rough = roughness; a.rotmat = light2obj_rotmat;
  a.vis = reinterpret_cast<const uint32_t*>(visibility);
  a.g_sh = grad_shading; a.g_ds = grad_diffuse_shading; a.gE = grad_envmaps;
  const bool spec = enable_specular != 0;
  hipError_t e = prepare_order(L, a, workspace, s);
  if (e != hipSuccess) return launch_err(e, "point order");
  e = des::relight_prepare_texels(a, s);
  if (e != hipSuccess) return launch_err(e, "texel table");
  if ((e = des::relight_prepare_points(spec, a, s)) != hipSuccess) return launch_err(e, "point constants");
  for (int col0 = 0; col0 < a.ncol; col0 += des::kRelightMaxNT * 16) {
    a.col0 = col0;
    a.ntg = std::min(des::kRelightMaxNT, (a.ncol - col0 + 15) / 16);
    if ((e = des::relight_prepare_AB(precision, spec, a, s)) != hipSuccess) return launch_err(e, "gradient fragments");
    if ((e = des::relight_launch_envgrad(precision, spec, a, s)) != hipSuccess) return launch_err(e, "relight env grad");
    if ((e = des::relight_reduce_envgrad(a, s)) != hipSuccess) return launch_err(e, "env grad reduce");
  }
  return ok();
}

}  // extern "C"
