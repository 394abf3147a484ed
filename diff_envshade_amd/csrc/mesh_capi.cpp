// mesh_capi.cpp -- C ABI of the mesh utilities (include/diff_envshade.h,
// des_get_nearest_mesh_points_*, des_get_uv_of_triangle_*,
// des_extract_bitfield_forward): argument validation, then one launch each on
// the caller's stream.  The reference's host wrappers are etc.cu:1044-1071
// and etc.cu:1218-1365.
#include <hip/hip_runtime.h>

#include "../../include/diff_envshade.h"
#include "mesh_internal.h"

namespace des_capi {
int fail(int code, const char* fmt, ...);
int ok();
}  // namespace des_capi

namespace {

using des_capi::fail;
using des_capi::ok;

constexpr int64_t kMaxPoints = int64_t(1) << 31;  // the reference indexes points with int
// adjacency entries: a group of 16 lanes sums 16 face counts in 32 bits
constexpr int64_t kMaxAdjacency = int64_t(1) << 26;

int launch_err(hipError_t e, const char* what) {
  return fail(DES_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
}

int check_mesh(int32_t N_vertex, int64_t N_face, int64_t P) {
  if (N_vertex < 0 || N_face < 0) return fail(DES_ERR_INVALID_ARGUMENT, "N_vertex and N_face must be >= 0");
  if (N_face >= kMaxPoints) return fail(DES_ERR_INVALID_ARGUMENT, "N_face must be < 2^31");
  if (P < 0 || P >= kMaxPoints) return fail(DES_ERR_INVALID_ARGUMENT, "P must be in [0, 2^31)");
  return 0;
}

}  // namespace

DES_API int des_get_nearest_mesh_points_forward(
    int32_t N_vertex, int64_t N_face, int64_t P, int32_t K, const int32_t* adjacency_head,
    const int32_t* adjacency_list, int64_t N_adjacency, const int32_t* face_vertex_list,
    const float* vertex_pos, const float* query_pos, const int64_t* idxs,
    int64_t* nearest_triangle_id, float* barycentric_coord, void* stream) {
  if (int rc = check_mesh(N_vertex, N_face, P)) return rc;
  if (K < 0) return fail(DES_ERR_INVALID_ARGUMENT, "K must be >= 0");
  if (N_adjacency < 0 || N_adjacency >= kMaxAdjacency)
    return fail(DES_ERR_INVALID_ARGUMENT, "N_adjacency must be in [0, 2^26)");
  if (P == 0) return ok();
  if (!query_pos || !nearest_triangle_id || !barycentric_coord || (K > 0 && !idxs))
    return fail(DES_ERR_INVALID_ARGUMENT, "query_pos, idxs and outputs must not be NULL");
  if (K > 0 && N_vertex > 0 && (!adjacency_head || (N_adjacency > 0 && !adjacency_list)))
    return fail(DES_ERR_INVALID_ARGUMENT, "adjacency_head and adjacency_list must not be NULL");
  if (N_face > 0 && N_vertex > 0 && (!face_vertex_list || !vertex_pos))
    return fail(DES_ERR_INVALID_ARGUMENT, "face_vertex_list and vertex_pos must not be NULL");
  des::MeshArgs m{};
  m.N_vertex = N_vertex;
  m.K = K;
  m.N_face = N_face;
  m.P = P;
  m.n_adj = N_adjacency;
  m.adj_head = adjacency_head;
  m.adj_list = adjacency_list;
  m.fvl = face_vertex_list;
  m.vpos = vertex_pos;
  m.query = query_pos;
  m.idxs = idxs;
  m.tri_out = nearest_triangle_id;
  m.bary_out = barycentric_coord;
  const hipError_t e = des::launch_nearest_forward(m, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_err(e, "nearest mesh points forward");
  return ok();
}

DES_API int des_get_nearest_mesh_points_backward(
    int32_t N_vertex, int64_t N_face, int64_t P, const int32_t* face_vertex_list,
    const float* vertex_pos, const float* query_pos, const int64_t* nearest_triangle_id,
    const float* grad_barycentric_coord, float* grad_query_pos, void* stream) {
  if (int rc = check_mesh(N_vertex, N_face, P)) return rc;
  if (P == 0) return ok();
  if (!query_pos || !nearest_triangle_id || !grad_barycentric_coord || !grad_query_pos)
    return fail(DES_ERR_INVALID_ARGUMENT, "query_pos, nearest_triangle_id, gradients must not be NULL");
  if (N_face > 0 && N_vertex > 0 && (!face_vertex_list || !vertex_pos))
    return fail(DES_ERR_INVALID_ARGUMENT, "face_vertex_list and vertex_pos must not be NULL");
  des::MeshArgs m{};
  m.N_vertex = N_vertex;
  m.N_face = N_face;
  m.P = P;
  m.fvl = face_vertex_list;
  m.vpos = vertex_pos;
  m.query = query_pos;
  m.tri_in = nearest_triangle_id;
  m.gbary = grad_barycentric_coord;
  m.gquery = grad_query_pos;
  const hipError_t e = des::launch_nearest_backward(m, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_err(e, "nearest mesh points backward");
  return ok();
}

DES_API int des_get_uv_of_triangle_forward(int64_t P, const float* triangles,
                                           const float* query_pos, float* barycentric_coord,
                                           void* stream) {
  if (P < 0 || P >= kMaxPoints) return fail(DES_ERR_INVALID_ARGUMENT, "P must be in [0, 2^31)");
  if (P == 0) return ok();
  if (!triangles || !query_pos || !barycentric_coord)
    return fail(DES_ERR_INVALID_ARGUMENT, "triangles, query_pos and barycentric_coord must not be NULL");
  const hipError_t e = des::launch_uv_forward(P, triangles, query_pos, barycentric_coord,
                                              static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_err(e, "uv of triangle forward");
  return ok();
}

DES_API int des_get_uv_of_triangle_backward(int64_t P, const float* triangles,
                                            const float* query_pos,
                                            const float* grad_barycentric_coord,
                                            float* grad_query_pos, void* stream) {
  if (P < 0 || P >= kMaxPoints) return fail(DES_ERR_INVALID_ARGUMENT, "P must be in [0, 2^31)");
  if (P == 0) return ok();
  if (!triangles || !query_pos || !grad_barycentric_coord || !grad_query_pos)
    return fail(DES_ERR_INVALID_ARGUMENT, "triangles, query_pos and gradients must not be NULL");
  const hipError_t e = des::launch_uv_backward(P, triangles, query_pos, grad_barycentric_coord,
                                               grad_query_pos, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_err(e, "uv of triangle backward");
  return ok();
}

DES_API int des_extract_bitfield_forward(const int32_t* bitfield, int32_t N_slot, int64_t N,
                                         int32_t N_bit, int64_t i_start, int64_t i_end,
                                         float* output, void* stream) {
  if (N_slot < 0 || N < 0 || N_bit < 0) return fail(DES_ERR_INVALID_ARGUMENT, "N_slot, N and N_bit must be >= 0");
  if (int64_t(N_bit) > int64_t(N_slot) * 32)
    return fail(DES_ERR_INVALID_ARGUMENT, "N_bit = %d needs %d slots, the bitfield has %d", N_bit,
                (N_bit + 31) / 32, N_slot);
  if (i_start < 0 || i_start > i_end || i_end > N)
    return fail(DES_ERR_INVALID_ARGUMENT, "need 0 <= i_start <= i_end <= N (got %lld, %lld, N = %lld)",
                (long long)i_start, (long long)i_end, (long long)N);
  const int64_t nb = i_end - i_start;
  if (nb == 0 || N_bit == 0) return ok();
  if (!bitfield || !output) return fail(DES_ERR_INVALID_ARGUMENT, "bitfield and output must not be NULL");
  if (nb >= (int64_t(1) << 37))  // 64 points per block of the launch grid
    return fail(DES_ERR_INVALID_ARGUMENT, "bitfield range too large");
  const hipError_t e = des::launch_bitfield(reinterpret_cast<const uint32_t*>(bitfield), N, N_bit,
                                            i_start, nb, output, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_err(e, "extract bitfield");
  return ok();
}
