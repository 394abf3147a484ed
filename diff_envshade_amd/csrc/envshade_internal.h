// envshade_internal.h -- shared between the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace des {

// Shader variants (one per pair of reference bindings, bindings.cpp:6-19).
enum Variant : int {
  kDiffuse = 0,   // bruteforce_diffuse_shader_*            (bs:231-530)
  kSpecular = 1,  // bruteforce_specular_shader_*           (bs:532-1038)
  kClamp = 2,     // bruteforce_specular_shader_clamp_*     (bs:1040-1552)
  kV2 = 3,        // bruteforce_specular_shader_*2          (bs:1554-2066)
  kVis = 4,       // bruteforce_specularvisibility_shader_*2 (bs:2068-2683)
};

// Device pointers and sizes of one shading problem (plain data, passed by value).
struct Problem {
  int64_t P;
  int H, W, T;
  int N_face, N_vertex;
  const float* normal;    // P x 3
  const float* view_dir;  // P x 3
  const float* albedo;    // P x 3
  const float* spec;      // P x 3 (specular_albedo)
  const float* rough;     // P x 1
  const float* envmap;    // 3 x H x W
  const float* rotmat;    // 3 x 3 row-major (light2obj_rotmat)
  const int32_t* fvl;     // 3 x N_face
  const int64_t* tri;     // P
  const float* bary;      // 3 x P
  const uint32_t* vis;    // N_slot x N_vertex
  const float* g_sh;      // P x 3
  const float* g_ds;      // P x 3
  float* sh;              // P x 3
  float* ds;              // P x 3
  float* gn;
  float* gv;
  float* ga;
  float* gs;
  float* gr;
  float* gE;              // 3 x H x W
  const int32_t* order;   // P: point of each tile lane (tile*64 + lane), or null = identity
  float* part;            // [ceil(P/64)][3][T] per-tile env-map partials
  float* spart;           // backward tail slices: [slice item][kBackwardAcc][64] per-point sums
};

// Launch shape: one 64-point tile per workgroup, NTC waves over texel chunks.
struct Tiling {
  int ntc;          // waves (texel chunks) per workgroup
  int cw;           // super-chunk texels per wave, st / ntc (multiple of 32)
  int st;           // texels per super-chunk = ntc * cw
  int vis_off;      // byte offset of the staged visibility words in dynamic LDS
  int dark_off;     // byte offset of the super-chunk's dark words (one uint64 per 64 texels)
  int list_off;     // forward: byte offset of the waves' lit-texel lists (uint16 entries)
  int list_cap;     // forward: entries per wave, ceil(st / ntc) rounded up to 4
  int ang_off;      // byte offset of the map's row and column angles (kAngleSlots double2)
  size_t lds_fwd;   // dynamic LDS bytes, forward
  size_t lds_bwd;   // dynamic LDS bytes, backward
  // Backward tail slices: tiles [split0, tiles) are shaded as `slices`
  // workgroups each, slice s covering texels [s * slice_texels, ...), after the
  // whole tiles; their per-point sums are combined by the env-gradient reduce launch.
  // split0 = tiles and slices = 1 when nothing is sliced.
  int64_t split0;
  int slices;
  int slice_texels;
};

struct SortJobs;  // the point-order jobs of a batch (below)

// Batched launches: a device-resident descriptor table.
struct BatchSet {
  Problem pr;         // pointers + sizes of one point set (part inside the workspace)
  const int32_t* order_fwd;  // point order the forward shades in, or null (input order)
  const int32_t* order_bwd;  // the same for the backward
  Tiling tl_fwd;      // ntc equals the forward launch's ntc; cw/st follow this set's T
  Tiling tl_bwd;      // same for the backward launch
  int64_t tile0;      // first global tile of this set
  int64_t red_item0;  // first env-gradient reduce item (one per 64 columns of 3 T)
};

// Launch signature of a prepared batch: variant, branch flags and the waves per
// tile of one direction.  The batched kernels compare the signature they were
// launched with against the one recorded at prepare time and do nothing on a
// mismatch (the host entry points reject the launch before that).
__host__ __device__ constexpr int32_t batch_signature(int variant, bool diff, bool spec, int ntc) {
  return int32_t(variant) | (diff ? 0x10 : 0) | (spec ? 0x20 : 0) | (int32_t(ntc) << 8);
}

struct BatchHeader {
  int32_t n_sets;
  int32_t variant;
  int32_t has_backward;  // every set carries upstream gradients and gradient outputs
  int32_t reserved;
  int32_t sig_fwd;       // batch_signature(variant, diff, spec, ntc_forward) at prepare
  int32_t sig_bwd;       // the same with ntc_backward
  int64_t total_tiles;
  int64_t total_red;  // env-gradient reduce items of all sets
  unsigned long long fwd_counter;  // work counters, zeroed by a kernel at each launch
  unsigned long long bwd_counter;
  const SortJobs* sort_jobs;       // the sets' point-order jobs (inside the table)
};

struct BatchLaunch {
  const BatchHeader* hdr;
  BatchHeader* hdr_mut;
  const BatchSet* sets;
  int32_t sig;      // batch_signature of this launch
  int ntc;
  size_t lds;
  unsigned grid;
  unsigned reduce_grid;
};

hipError_t launch_batch_forward(int variant, bool diff, bool spec, const BatchLaunch& bl,
                                hipStream_t stream);
hipError_t launch_batch_backward(int variant, bool diff, bool spec, const BatchLaunch& bl,
                                 hipStream_t stream);
int batch_blocks_per_cu(int variant, bool diff, bool spec, bool fwd, int ntc, size_t lds);

constexpr int kForwardAcc = 6;
constexpr int kBackwardAcc = 16;
constexpr int kMaxSuperChunk = 512;  // texels staged in LDS at once (48 B each, 24 KB)
constexpr int kTexelBytes = 48;
// Row and column angles ({sin, cos} in FP64) a tile keeps in LDS for the texel
// table when the map has at most this many rows + columns; larger maps compute
// them per texel.
constexpr int kAngleSlots = 128;
constexpr int kPointXchSlots = 25;  // per-point fields a tile exchanges through LDS

// Main shading kernel, then (tail slices) the combine; ev_start / ev_stop, when
// given, are recorded right before and after the main kernel.
hipError_t launch_forward(int variant, bool diff, bool spec, const Problem& pr, const Tiling& tl,
                          hipStream_t stream, hipEvent_t ev_start = nullptr,
                          hipEvent_t ev_stop = nullptr);
hipError_t launch_backward(int variant, bool diff, bool spec, const Problem& pr, const Tiling& tl,
                           hipStream_t stream);
// The env-map-gradient reduction and the combine of the backward's tail
// slices, one launch.
hipError_t launch_envgrad_reduce(int variant, bool diff, bool spec, const Problem& pr,
                                 const Tiling& tl, hipStream_t stream);
hipError_t launch_fill_zero(float* x, int64_t n, hipStream_t stream);

// ---- normal-coherent point order (point_order.hip) ----
constexpr int kOrderBlock = 1024;  // points per block of the counting sort
constexpr int kOrderKeys = 1024;   // 32 x 32 octahedral cells of the normal

// Workspace of the point order, offsets from a 256-byte aligned base; the
// order itself (int32 x P) is at order_off and stays valid for the shading
// launches that follow on the same stream.
struct OrderLayout {
  size_t order_off, keyrank_off, count_off, total_off, total;
};
OrderLayout order_layout(int64_t P);
hipError_t launch_point_order(const float* normal, int64_t P, void* ws, hipStream_t stream);

// The sort of one point set as a plain descriptor, so that one launch can sort
// many sets (a batch): its rank and scatter workgroups are blk0 .. blk0+nblk-1
// of the launch, its scan workgroups scan0 .. scan0+sort_scan_groups-1.
struct SortJob {
  const float* normal;
  int64_t P;
  uint32_t* keyrank;
  uint32_t* count;
  uint32_t* total;
  int32_t* order;
  int32_t nblk, nblk_pad, fused;
  int32_t blk0, scan0;
};
SortJob sort_job(const float* normal, int64_t P, void* ws);  // ws: an order_layout(P) workspace
int sort_scan_groups(const SortJob& job);

// A device-resident table of sort jobs: njobs entries, nblocks rank/scatter
// items and nscan scan items in total (jobs[] follows the header).
struct SortJobs {
  int32_t njobs, nblocks, nscan, reserved;
  SortJob jobs[1];
};
// Sorts every job of the table *tabp (a device pointer to a device pointer,
// read when the kernels run) with fixed grids of persistent workgroups
// (max_groups for rank and scatter, 4x that for the scan); an empty table
// costs three near-empty launches.
hipError_t launch_point_order_jobs(const SortJobs* const* tabp, int max_groups,
                                   hipStream_t stream);

}  // namespace des
