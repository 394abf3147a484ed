/*
 * diff_envshade.h -- C ABI of the MI355X (gfx950) brute-force environment-light
 * shader.  Drop-in boundary for the ten shader bindings of zjumsj/diff_envshade
 * (bindings.cpp:6-19, prototypes gsface_shader.h:3-118, host wrappers
 * bruteforce_shader.cu:2687-3453), for the five mesh utilities that produce
 * its mesh inputs (bindings.cpp:27-30, 33; etc.cu) and for the two env-map
 * utilities beside it (bindings.cpp:21, 23; etc.cu), declared at the end.
 *
 * Conventions (all entry points):
 *   - every pointer is a device pointer to contiguous row-major data on the
 *     current HIP device; float = IEEE binary32;
 *   - shapes: normal, view_dir, albedo, specular_albedo, every P x 3 output
 *     and upstream gradient: P x 3; roughness and grad_roughness: P x 1;
 *     envmap and grad_envmap: C x H x W (channel-major); light2obj_rotmat:
 *     3 x 3 row-major; face_vertex_list: 3 x N_face int32;
 *     nearest_triangle_id: P int64; barycentric_coord: 3 x P;
 *     visibility: N_slot x N_vertex 32-bit words (bit t%32 of word
 *     [t/32][vertex] is texel t, storage order);
 *   - C must be 3 (DES_ERR_UNSUPPORTED_CHANNELS otherwise, as the reference's
 *     "Channel number not supported!");
 *   - `stream` is a hipStream_t (NULL = default stream); work is enqueued
 *     asynchronously and nothing is allocated: outputs are caller-owned and
 *     every output element is written (disabled branches as zeros), so the
 *     caller may pass uninitialised buffers;
 *   - the backward entry points need a caller-provided workspace of at least
 *     des_backward_workspace_size(P, H, W) bytes (256-byte aligned) for the
 *     per-tile env-map-gradient partials (and the point order, below);
 *     grad_envmap is summed from them in a fixed order in FP64, so it is
 *     bitwise reproducible run to run;
 *   - point order: for large problems the entry points first sort the points
 *     by normal direction (a stable counting sort, deterministic), so that a
 *     64-point tile shares a hemisphere and texels lit by none of its points
 *     are skipped, and the texel groups a tile lights are dealt evenly to its
 *     waves.  Outputs change with the order at fp32 rounding level only (each
 *     sum is regrouped) and stay bitwise reproducible run to run: the order and
 *     the deal depend only on the inputs.  The forward entry points take an
 *     optional workspace of des_forward_workspace_size(P, H, W) bytes for the
 *     order (NULL = shade in input order); the backward workspace includes it;
 *   - return 0 (DES_OK) or a negative DES_ERR_* code; des_last_error() gives
 *     the message of the calling thread's last failure.
 *   - calls on distinct streams with distinct buffers are independent; calls
 *     sharing a workspace must be ordered on one stream.
 *
 * Deliberate differences from the reference: launches go to the given stream
 * (not the legacy default stream), shapes are validated (N_slot >= ceil(H*W/32)),
 * grad_envmap is deterministic, and out-of-range face / vertex ids read as
 * occluded instead of faulting.
 */
#ifndef DIFF_ENVSHADE_H
#define DIFF_ENVSHADE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DES_API __attribute__((visibility("default")))

enum {
  DES_OK = 0,
  DES_ERR_INVALID_ARGUMENT = -1,
  DES_ERR_UNSUPPORTED_CHANNELS = -2,
  DES_ERR_WORKSPACE_TOO_SMALL = -3,
  DES_ERR_LAUNCH = -4
};

/* Version of this ABI (bumped on any signature change). */
DES_API int des_abi_version(void);

/* Message of the calling thread's most recent failed call ("" if none). */
DES_API const char* des_last_error(void);

/* Workspace bytes needed by every *_backward* entry point for P points and an
 * H x W env map (0 when P == 0). */
DES_API size_t des_backward_workspace_size(int64_t P, int32_t H, int32_t W);

/* Workspace bytes the *_forward* entry points use for the point order of P
 * points and an H x W env map (0 when they would not reorder). */
DES_API size_t des_forward_workspace_size(int64_t P, int32_t H, int32_t W);

/* Point-order policy (no counterpart in the reference), process-wide:
 * 0 = automatic (both directions reorder from 2^24 point-texel pairs
 * P x H x W), 1 = always, -1 = never.
 * The workspace sizes above follow the policy in force when they are asked.
 * Process-wide on purpose: PyTorch's autograd runs a graph's backward on its
 * own thread, so a per-thread policy would shade the forward and the backward
 * of one graph under different settings.  Results move with it at fp32
 * rounding level only; set it before any thread shades, not while they do. */
DES_API int des_set_point_order(int32_t mode);

/* Point order shared between calls (no counterpart in the reference).  The
 * calling thread's next call into the library, if it is a shading entry point
 * for P points that would reorder them, takes the order from
 * `order_workspace` instead of sorting: the workspace of an earlier
 * *_forward* call for P points that sorted (des_forward_workspace_size bytes,
 * 256-byte aligned), ordered before this call on the stream.  A call for
 * another number of points ignores it.  With the same normals the outputs are
 * bitwise those of a call that sorts by itself; every order of P points is a
 * permutation, so with other normals they stay correct and move at rounding
 * level only.  The next call consumes the hook whatever it is; NULL clears
 * it. */
DES_API int des_reuse_point_order_once(const void* order_workspace, int64_t P);

/* bindings.cpp:6 -> bruteforce_diffuse_shader_forward (bruteforce_shader.cu:2687-2731).
 * Lambert only: shading = sum_t [l.n>0] E a (l.n/pi) dOmega. */
DES_API int des_bruteforce_diffuse_shader_forward(
    int64_t P, int32_t C, int32_t H, int32_t W,
    const float* normal, const float* albedo, const float* envmap,
    const float* light2obj_rotmat,
    float* shading, void* workspace, size_t workspace_bytes, void* stream);

/* bindings.cpp:7 -> bruteforce_diffuse_shader_backward (bruteforce_shader.cu:2733-2792).
 * Outputs [grad_normal, grad_albedo, grad_envmap]. */
DES_API int des_bruteforce_diffuse_shader_backward(
    int64_t P, int32_t C, int32_t H, int32_t W,
    const float* grad_shading, const float* normal, const float* albedo,
    const float* envmap, const float* light2obj_rotmat,
    float* grad_normal, float* grad_albedo, float* grad_envmap,
    void* workspace, size_t workspace_bytes, void* stream);

/* bindings.cpp:9 -> bruteforce_specular_shader_forward (bruteforce_shader.cu:2796-2851).
 * Lambert (if enable_diffuse) + GGX (if enable_specular). */
DES_API int des_bruteforce_specular_shader_forward(
    int64_t P, int32_t C, int32_t H, int32_t W,
    const float* normal, const float* view_dir, const float* albedo,
    const float* specular_albedo, const float* roughness, const float* envmap,
    const float* light2obj_rotmat, int enable_diffuse, int enable_specular,
    float* shading, void* workspace, size_t workspace_bytes, void* stream);

/* bindings.cpp:10 -> bruteforce_specular_shader_backward (bruteforce_shader.cu:2853-2949).
 * Outputs [grad_normal, grad_view_dir, grad_albedo, grad_specular_albedo,
 * grad_roughness, grad_envmap]; grad_albedo is zero if !enable_diffuse,
 * grad_view_dir / grad_specular_albedo / grad_roughness zero if !enable_specular. */
DES_API int des_bruteforce_specular_shader_backward(
    int64_t P, int32_t C, int32_t H, int32_t W,
    const float* grad_shading, const float* normal, const float* view_dir,
    const float* albedo, const float* specular_albedo, const float* roughness,
    const float* envmap, const float* light2obj_rotmat,
    int enable_diffuse, int enable_specular,
    float* grad_normal, float* grad_view_dir, float* grad_albedo,
    float* grad_specular_albedo, float* grad_roughness, float* grad_envmap,
    void* workspace, size_t workspace_bytes, void* stream);

/* bindings.cpp:12 -> bruteforce_specular_shader_clamp_forward (bruteforce_shader.cu:2953-3008).
 * As the specular forward, but points with n.v <= 0 shade to 0. */
DES_API int des_bruteforce_specular_shader_clamp_forward(
    int64_t P, int32_t C, int32_t H, int32_t W,
    const float* normal, const float* view_dir, const float* albedo,
    const float* specular_albedo, const float* roughness, const float* envmap,
    const float* light2obj_rotmat, int enable_diffuse, int enable_specular,
    float* shading, void* workspace, size_t workspace_bytes, void* stream);

/* bindings.cpp:13 -> bruteforce_specular_shader_clamp_backward (bruteforce_shader.cu:3010-3106). */
DES_API int des_bruteforce_specular_shader_clamp_backward(
    int64_t P, int32_t C, int32_t H, int32_t W,
    const float* grad_shading, const float* normal, const float* view_dir,
    const float* albedo, const float* specular_albedo, const float* roughness,
    const float* envmap, const float* light2obj_rotmat,
    int enable_diffuse, int enable_specular,
    float* grad_normal, float* grad_view_dir, float* grad_albedo,
    float* grad_specular_albedo, float* grad_roughness, float* grad_envmap,
    void* workspace, size_t workspace_bytes, void* stream);

/* bindings.cpp:15 -> bruteforce_specular_shader_forward2 (bruteforce_shader.cu:3108-3165).
 * Diffuse always on; outputs [shading, diffuse_shading]. */
DES_API int des_bruteforce_specular_shader_forward2(
    int64_t P, int32_t C, int32_t H, int32_t W,
    const float* normal, const float* view_dir, const float* albedo,
    const float* specular_albedo, const float* roughness, const float* envmap,
    const float* light2obj_rotmat, int enable_specular,
    float* shading, float* diffuse_shading, void* workspace, size_t workspace_bytes,
    void* stream);

/* bindings.cpp:16 -> bruteforce_specular_shader_backward2 (bruteforce_shader.cu:3167-3259). */
DES_API int des_bruteforce_specular_shader_backward2(
    int64_t P, int32_t C, int32_t H, int32_t W,
    const float* grad_shading, const float* grad_diffuse_shading,
    const float* normal, const float* view_dir, const float* albedo,
    const float* specular_albedo, const float* roughness, const float* envmap,
    const float* light2obj_rotmat, int enable_specular,
    float* grad_normal, float* grad_view_dir, float* grad_albedo,
    float* grad_specular_albedo, float* grad_roughness, float* grad_envmap,
    void* workspace, size_t workspace_bytes, void* stream);

/* bindings.cpp:18 -> bruteforce_specularvisibility_shader_forward2
 * (bruteforce_shader.cu:3261-3336).  v2 times barycentric-interpolated
 * per-vertex visibility bits. */
DES_API int des_bruteforce_specularvisibility_shader_forward2(
    int64_t P, int32_t C, int32_t H, int32_t W,
    int32_t N_face, int32_t N_vertex, int32_t N_slot,
    const float* normal, const float* view_dir, const float* albedo,
    const float* specular_albedo, const float* roughness, const float* envmap,
    const float* light2obj_rotmat,
    const int32_t* face_vertex_list, const int64_t* nearest_triangle_id,
    const float* barycentric_coord, const int32_t* visibility,
    int enable_specular,
    float* shading, float* diffuse_shading, void* workspace, size_t workspace_bytes,
    void* stream);

/* bindings.cpp:19 -> bruteforce_specularvisibility_shader_backward2
 * (bruteforce_shader.cu:3338-3453). */
DES_API int des_bruteforce_specularvisibility_shader_backward2(
    int64_t P, int32_t C, int32_t H, int32_t W,
    int32_t N_face, int32_t N_vertex, int32_t N_slot,
    const float* grad_shading, const float* grad_diffuse_shading,
    const float* normal, const float* view_dir, const float* albedo,
    const float* specular_albedo, const float* roughness, const float* envmap,
    const float* light2obj_rotmat,
    const int32_t* face_vertex_list, const int64_t* nearest_triangle_id,
    const float* barycentric_coord, const int32_t* visibility,
    int enable_specular,
    float* grad_normal, float* grad_view_dir, float* grad_albedo,
    float* grad_specular_albedo, float* grad_roughness, float* grad_envmap,
    void* workspace, size_t workspace_bytes, void* stream);

/* Launch-shape override for tuning (no counterpart in the reference): number
 * of waves splitting the texel range per 64-point tile, forward and backward
 * (each a power of two, forward <= 16, backward <= 8; 0 restores the built-in
 * choice).  Process-wide, like des_set_point_order and for the same reason;
 * results do not depend on it beyond fp32 rounding. */
DES_API int des_set_tiling_override(int32_t ntc_forward, int32_t ntc_backward);

/* ---- backward without allocation (no counterpart in the reference) ---------
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the
 * visibility backward (des_bruteforce_specularvisibility_shader_backward2,
 * same arguments, same results bit for bit) with its workspace in a static
 * per-device scratch array of des_static_scratch_bytes() bytes (32 MiB) that
 * the library holds from its load to the end of the process.  Problems whose
 * des_backward_workspace_size exceeds it are rejected with
 * DES_ERR_INVALID_ARGUMENT: every problem of up to 276 480 points with a
 * 16 x 32 map, or up to 77 184 points with a 32 x 64 map, fits (70 000 points
 * at 16 x 32 use 9.9 MB).
 * Nothing is allocated: the
 * first call on a device only looks up that device's scratch address.  All
 * calls on one GPU share its scratch, so calls that may overlap on one GPU
 * (other streams, other threads) race and must be serialised by the caller. */
DES_API size_t des_static_scratch_bytes(void);

DES_API int des_bruteforce_specularvisibility_shader_backward2_noalloc(
    int64_t P, int32_t C, int32_t H, int32_t W, int32_t N_face, int32_t N_vertex, int32_t N_slot,
    const float* grad_shading, const float* grad_diffuse_shading, const float* normal,
    const float* view_dir, const float* albedo, const float* specular_albedo,
    const float* roughness, const float* envmap, const float* light2obj_rotmat,
    const int32_t* face_vertex_list, const int64_t* nearest_triangle_id,
    const float* barycentric_coord, const int32_t* visibility, int enable_specular,
    float* grad_normal, float* grad_view_dir, float* grad_albedo, float* grad_specular_albedo,
    float* grad_roughness, float* grad_envmap, void* stream);

/* ---- batched launches (no counterpart in the reference) --------------------
 * Many point sets, each with its own env map and rotation, shaded by one
 * persistent launch per direction.  All sets of a batch use one variant and
 * one pair of flags.  Per-set results equal those of the single-set entry
 * points with the same launch shape (des_set_tiling_override), so a set's
 * results do not depend on which other sets share its batch.
 *
 *   1. des_batch_plan     -> workspace (device) and staging (host) sizes
 *   2. des_batch_prepare  -> validates the sets, writes the descriptor table
 *                            into `staging` and enqueues its upload into the
 *                            workspace on `stream` (keep `staging` alive, and
 *                            preferably pinned, until the copy has run), then
 *                            the point order of every set large enough to be
 *                            sorted (the single-set policy, des_set_point_order)
 *   3. des_batch_forward / des_batch_backward -> persistent launches that read
 *      every size from the workspace; they can be captured once in a HIP graph
 *      and replayed for any later batch prepared into the same workspace.
 * A set's forward and backward shade in the order computed at prepare time;
 * des_batch_sort recomputes it from the normals' current values (any order
 * gives correct results, a stale one only skips fewer texels), and is
 * capturable like the launches.
 */
typedef struct des_set {
  int64_t P;
  int32_t H, W;
  int32_t N_face, N_vertex, N_slot; /* visibility variant only */
  int32_t reserved;
  const float* normal;
  const float* view_dir;
  const float* albedo;
  const float* specular_albedo;
  const float* roughness;
  const float* envmap;
  const float* light2obj_rotmat;
  const int32_t* face_vertex_list;
  const int64_t* nearest_triangle_id;
  const float* barycentric_coord;
  const int32_t* visibility;
  float* shading;
  float* diffuse_shading;           /* v2 / visibility variants */
  const float* grad_shading;        /* backward only (NULL for forward-only batches) */
  const float* grad_diffuse_shading;
  float* grad_normal;
  float* grad_view_dir;
  float* grad_albedo;
  float* grad_specular_albedo;
  float* grad_roughness;
  float* grad_envmap;
} des_set;

/* Variant ids of the batched entry points. */
enum {
  DES_VARIANT_DIFFUSE = 0,
  DES_VARIANT_SPECULAR = 1,
  DES_VARIANT_CLAMP = 2,
  DES_VARIANT_V2 = 3,
  DES_VARIANT_VISIBILITY = 4
};

DES_API int des_batch_plan(int32_t variant, const des_set* sets, int32_t n_sets,
                           size_t* workspace_bytes, size_t* staging_bytes);

/* ntc_forward / ntc_backward: waves per 64-point tile (powers of two, <= 8),
 * usually from the autotuner; they must match the later launches.  with_backward:
 * the sets carry upstream gradients and gradient outputs.  The library records
 * (variant, flags, ntc_forward, ntc_backward, with_backward) per workspace and
 * the table carries them too: des_batch_forward / des_batch_backward /
 * des_batch_sort on a workspace that was not prepared, or with another variant,
 * other flags or another ntc, return DES_ERR_INVALID_ARGUMENT before any
 * launch, and a replayed graph whose table was re-prepared for another launch
 * does nothing on the device. */
DES_API int des_batch_prepare(int32_t variant, int enable_diffuse, int enable_specular,
                              const des_set* sets, int32_t n_sets, int32_t ntc_forward,
                              int32_t ntc_backward, int with_backward, void* staging,
                              size_t staging_bytes, void* workspace, size_t workspace_bytes,
                              void* stream);

DES_API int des_batch_forward(int32_t variant, int enable_diffuse, int enable_specular,
                              int32_t ntc_forward, void* workspace, void* stream);

/* DES_ERR_INVALID_ARGUMENT when the batch was prepared without with_backward. */
DES_API int des_batch_backward(int32_t variant, int enable_diffuse, int enable_specular,
                               int32_t ntc_backward, void* workspace, void* stream);

/* Recomputes the point order of the prepared batch's sorted sets (three
 * persistent launches that read the job table from the workspace; a batch with
 * no sorted set costs three near-empty launches). */
DES_API int des_batch_sort(void* workspace, void* stream);

/* ---- many env maps lighting one point set, on MFMA (no counterpart in the reference) ----
 * shading[p, e, c] and diffuse_shading[p, e, c] (P x E x 3) for E env maps
 * (E x 3 x H x W) sharing one rotation, with per-point visibility words
 * (N_slot x P, bit t%32 of word [t/32][p]; NULL = all visible).  The per-pair
 * Lambert and GGX weights are generated in registers in the MFMA operand
 * layout and multiplied against the maps on the matrix cores.
 * des_relight_backward_envmap gives grad_envmaps (E x 3 x H x W) from the
 * upstream gradients (P x E x 3 each); the cross-block sum is FP64 in fixed
 * order.  precision: DES_PREC_BF16, DES_PREC_FP16, or DES_PREC_BF16X3 (hi+lo
 * split of both operands, three MFMAs per product).  fp16 holds weights up to
 * 65504 and a GGX weight reaches 2 / (pi r^4), so with specular on fp16 needs
 * roughness >= 0.0558; these entry points do not check it (a sharper lobe
 * gives inf), the Python calls in diff_envshade_amd/relight.py refuse it. */
enum { DES_PREC_BF16 = 0, DES_PREC_FP16 = 1, DES_PREC_BF16X3 = 2 };

DES_API size_t des_relight_workspace_size(int64_t P, int32_t E, int32_t H, int32_t W,
                                          int32_t precision, int with_envgrad);

DES_API int des_relight_forward(int64_t P, int32_t E, int32_t H, int32_t W, int32_t N_slot,
                                const float* normal, const float* view_dir, const float* albedo,
                                const float* specular_albedo, const float* roughness,
                                const float* envmaps, const float* light2obj_rotmat,
                                const int32_t* visibility, int enable_specular, int32_t precision,
                                float* shading, float* diffuse_shading, void* workspace,
                                size_t workspace_bytes, void* stream);

DES_API int des_relight_backward_envmap(int64_t P, int32_t E, int32_t H, int32_t W, int32_t N_slot,
                                        const float* normal, const float* view_dir,
                                        const float* albedo, const float* specular_albedo,
                                        const float* roughness, const float* light2obj_rotmat,
                                        const int32_t* visibility, const float* grad_shading,
                                        const float* grad_diffuse_shading, int enable_specular,
                                        int32_t precision, float* grad_envmaps, void* workspace,
                                        size_t workspace_bytes, void* stream);

/* ---- mesh utilities: the producers of the shader's mesh inputs ----
 * Same shapes and meaning as the reference's bindings (bindings.cpp:27-33,
 * prototypes gsface_shader.h:122-172); geometry in FP64 with the reference's
 * closest-point region rule (etc.cu:205-393), each operation correctly
 * rounded, so results are reproducible and exactly checkable.  Deliberate
 * differences: indices out of range are skipped instead of read, a point
 * with no valid candidate face gets triangle id -1 and zero barycentrics (the
 * reference traps), and the candidate set is not capped at 128 faces. */

/* get_nearest_mesh_points_forward (etc.cu:1277-1326, kernel etc.cu:779-862):
 * for each query point (query_pos 3 x P), the nearest point on the faces
 * around its K nearest vertices (idxs P x K int64, negative = none).
 * adjacency_head 2 x N_vertex int32: [0][v] = number of faces around v,
 * [1][v] = inclusive prefix sum of [0], so v's faces are adjacency_list
 * [head[1][v] - head[0][v], head[1][v]) (adjacency_list has N_adjacency
 * int32 entries).  Outputs nearest_triangle_id (P int64) and
 * barycentric_coord (3 x P: 1 - s - t, s, t).  Among faces at equal distance
 * the first in candidate order (vertex by vertex in idxs order, each vertex's
 * faces in list order) wins, as in the reference. */
DES_API int des_get_nearest_mesh_points_forward(
    int32_t N_vertex, int64_t N_face, int64_t P, int32_t K, const int32_t* adjacency_head,
    const int32_t* adjacency_list, int64_t N_adjacency, const int32_t* face_vertex_list,
    const float* vertex_pos, const float* query_pos, const int64_t* idxs,
    int64_t* nearest_triangle_id, float* barycentric_coord, void* stream);

/* get_nearest_mesh_points_backward (etc.cu:1328-1365, kernel etc.cu:864-920):
 * grad_query_pos (3 x P) from grad_barycentric_coord (3 x P).  The reference
 * kernel reads the upstream gradient from its own output buffer
 * (etc.cu:884-888); this one reads grad_barycentric_coord. */
DES_API int des_get_nearest_mesh_points_backward(
    int32_t N_vertex, int64_t N_face, int64_t P, const int32_t* face_vertex_list,
    const float* vertex_pos, const float* query_pos, const int64_t* nearest_triangle_id,
    const float* grad_barycentric_coord, float* grad_query_pos, void* stream);

/* get_uv_of_triangle_forward / _backward (etc.cu:1218-1275, kernels
 * etc.cu:707-777): triangles P x 3 x 3 (vertex, xyz), query_pos P x 3,
 * barycentric_coord and gradients P x 3. */
DES_API int des_get_uv_of_triangle_forward(int64_t P, const float* triangles,
                                           const float* query_pos, float* barycentric_coord,
                                           void* stream);
DES_API int des_get_uv_of_triangle_backward(int64_t P, const float* triangles,
                                            const float* query_pos,
                                            const float* grad_barycentric_coord,
                                            float* grad_query_pos, void* stream);

/* extract_bitfield_forward (etc.cu:1044-1071, kernel etc.cu:1021-1040):
 * output[b][k] = bit k % 32 of word [k / 32][i_start + b] of bitfield
 * (N_slot x N 32-bit words), as 0.0f / 1.0f, for b < i_end - i_start and
 * k < N_bit; output is (i_end - i_start) x N_bit.  Needs
 * 0 <= i_start <= i_end <= N and N_bit <= 32 * N_slot (unchecked in the
 * reference). */
DES_API int des_extract_bitfield_forward(const int32_t* bitfield, int32_t N_slot, int64_t N,
                                         int32_t N_bit, int64_t i_start, int64_t i_end,
                                         float* output, void* stream);

/* ---- env-map preprocessing beside the shader ----
 * FP64 directions, angles, bilinear weights and sums; row 0 is the top (+y),
 * the seam at phi = 0 wraps (the reference's config.h). */

/* sample_envmap_forward (etc.cu:1145-1178, kernel etc.cu:687-705): bilinear
 * lookup of envmap (C x H x W) along each query direction (query_point 3 x P,
 * normalised here).  output is C x P (the reference allocates 3 x P whatever
 * C is, etc.cu:1164); a zero-length query samples NaN. */
DES_API int des_sample_envmap_forward(int64_t P, int32_t C, int32_t H, int32_t W,
                                      const float* query_point, const float* envmap, float* output,
                                      void* stream);

/* blur_envmap_forward (etc.cu:1103-1143, kernel etc.cu:603-685): each texel of
 * a tar_H x tar_W map is the mean of N_samples bilinear lookups along
 * GGX-distributed directions (roughness parameter alpha) around its centre
 * direction; C must be 3 or 4 (DES_ERR_UNSUPPORTED_CHANNELS, "Channel number
 * not supported!").  output is C x tar_H x tar_W.  The random numbers come
 * from Philox4x32-10 (key 0, counter = sample, target texel), where the
 * reference uses curand's XORWOW: the same estimator with other samples.
 * The result is reproducible bit for bit.  N_samples must be >= 1. */
DES_API int des_blur_envmap_forward(int32_t C, int32_t H, int32_t W, int32_t tar_H, int32_t tar_W,
                                    int32_t N_samples, float alpha, const float* envmap,
                                    float* output, void* stream);

/* Timing hook for benchmarks (no counterpart in the reference): when set,
 * the calling thread's next *_backward* call records hipEvent_t `start` right
 * before and `stop` right after its main shading kernel, on the call's stream,
 * then clears the hook.  Pass NULLs to clear it explicitly. */
DES_API int des_set_backward_timing_events(void* start, void* stop);

/* The same for the calling thread's next *_forward* call: `start` right
 * before and `stop` right after its main shading kernel (the tail-slice
 * combine, when there is one, runs after `stop`). */
DES_API int des_set_forward_timing_events(void* start, void* stop);

#ifdef __cplusplus
}
#endif

#endif /* DIFF_ENVSHADE_H */
