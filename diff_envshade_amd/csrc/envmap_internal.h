// envmap_internal.h -- launch interface of the env-map kernels
// (envmap_kernels.hip) used by the C ABI (envmap_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace des {

// query_point 3 x P, env C x H x W, out C x P
hipError_t launch_sample_envmap(int64_t P, int C, int H, int W, const float* q, const float* env,
                                float* out, hipStream_t stream);
// env C x H x W (C = 3 or 4), out C x tar_H x tar_W
hipError_t launch_blur_envmap(int C, int H, int W, int tar_H, int tar_W, int N, float alpha,
                              const float* env, float* out, hipStream_t stream);

}  // namespace des
