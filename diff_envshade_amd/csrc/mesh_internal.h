// mesh_internal.h -- launch interface of the mesh kernels (mesh_kernels.hip)
// used by the C ABI (mesh_capi.cpp).  Shapes as zjumsj/diff_envshade etc.cu
// documents them: adjacency_head 2 x N_vertex (faces around a vertex, then
// their inclusive prefix sum), adjacency_list N_adj, face_vertex_list
// 3 x N_face, vertex_pos 3 x N_vertex, query_pos 3 x P, idxs P x K.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace des {

struct MeshArgs {
  int32_t N_vertex;
  int32_t K;
  int64_t N_face, P, n_adj;
  const int32_t* adj_head;
  const int32_t* adj_list;
  const int32_t* fvl;
  const float* vpos;
  const float* query;
  const int64_t* idxs;
  int64_t* tri_out;      // forward: nearest_triangle_id (P)
  float* bary_out;       // forward: barycentric_coord (3 x P)
  const int64_t* tri_in; // backward: nearest_triangle_id (P)
  const float* gbary;    // backward: grad_barycentric_coord (3 x P)
  float* gquery;         // backward: grad_query_pos (3 x P)
};

hipError_t launch_nearest_forward(const MeshArgs& m, hipStream_t stream);
hipError_t launch_nearest_backward(const MeshArgs& m, hipStream_t stream);
hipError_t launch_uv_forward(int64_t P, const float* tri, const float* query, float* bary,
                             hipStream_t stream);
hipError_t launch_uv_backward(int64_t P, const float* tri, const float* query, const float* gbary,
                              float* gquery, hipStream_t stream);
hipError_t launch_bitfield(const uint32_t* bits, int64_t N, int N_bit, int64_t i_start, int64_t nb,
                           float* out, hipStream_t stream);

}  // namespace des
