// envmap_capi.cpp -- C ABI of the env-map utilities (include/diff_envshade.h,
// des_sample_envmap_forward, des_blur_envmap_forward): argument validation,
// then one launch each on the caller's stream.  The reference's host wrappers
// are etc.cu:1103-1178.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/diff_envshade.h"
#include "envmap_internal.h"

namespace des_capi {
int fail(int code, const char* fmt, ...);
int ok();
}  // namespace des_capi

namespace {

using des_capi::fail;
using des_capi::ok;

int launch_err(hipError_t e, const char* what) {
  return fail(DES_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
}

int check_map(int32_t C, int32_t H, int32_t W) {
  if (C < 1 || H < 1 || W < 1) return fail(DES_ERR_INVALID_ARGUMENT, "envmap must be C x H x W with C, H, W >= 1");
  if (int64_t(H) * W >= (int64_t(1) << 31)) return fail(DES_ERR_INVALID_ARGUMENT, "envmap has 2^31 or more texels");
  return 0;
}

}  // namespace

DES_API int des_sample_envmap_forward(int64_t P, int32_t C, int32_t H, int32_t W,
                                      const float* query_point, const float* envmap, float* output,
                                      void* stream) {
  if (int rc = check_map(C, H, W)) return rc;
  if (P < 0 || (P + 255) / 256 >= (int64_t(1) << 32)) return fail(DES_ERR_INVALID_ARGUMENT, "P out of range");
  if (P == 0) return ok();
  if (!query_point || !envmap || !output)
    return fail(DES_ERR_INVALID_ARGUMENT, "query_point, envmap and output must not be NULL");
  const hipError_t e = des::launch_sample_envmap(P, C, H, W, query_point, envmap, output,
                                                 static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_err(e, "sample envmap");
  return ok();
}

DES_API int des_blur_envmap_forward(int32_t C, int32_t H, int32_t W, int32_t tar_H, int32_t tar_W,
                                    int32_t N_samples, float alpha, const float* envmap,
                                    float* output, void* stream) {
  if (C != 3 && C != 4) return fail(DES_ERR_UNSUPPORTED_CHANNELS, "Channel number not supported!");
  if (int rc = check_map(C, H, W)) return rc;
  if (tar_H < 0 || tar_W < 0) return fail(DES_ERR_INVALID_ARGUMENT, "tar_H and tar_W must be >= 0");
  if ((int64_t(tar_H) * tar_W + 3) / 4 >= (int64_t(1) << 32))
    return fail(DES_ERR_INVALID_ARGUMENT, "target map too large");
  if (N_samples < 1) return fail(DES_ERR_INVALID_ARGUMENT, "N_samples must be >= 1");
  if (!std::isfinite(alpha)) return fail(DES_ERR_INVALID_ARGUMENT, "alpha must be finite");
  if (int64_t(tar_H) * tar_W == 0) return ok();
  if (!envmap || !output) return fail(DES_ERR_INVALID_ARGUMENT, "envmap and output must not be NULL");
  const hipError_t e = des::launch_blur_envmap(C, H, W, tar_H, tar_W, N_samples, alpha, envmap,
                                               output, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_err(e, "blur envmap");
  return ok();
}
