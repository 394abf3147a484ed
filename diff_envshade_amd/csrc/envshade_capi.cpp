// envshade_capi.cpp -- C ABI of the shader (include/diff_envshade.h): argument
// validation, launch-shape choice, workspace layout and kernel dispatch.
// Mirrors the checks and allocation semantics of the reference host wrappers
// (bruteforce_shader.cu:2687-3453), without allocating.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/diff_envshade.h"
#include "envshade_internal.h"

namespace des_capi {

thread_local std::string g_last_error;
// Point order handed to the calling thread's next call (des_reuse_point_order_once);
// every call returning through fail() or ok() drops it, the shading calls take it.
thread_local const void* g_reuse_order = nullptr;
thread_local int64_t g_reuse_order_P = -1;

int fail(int code, const char* fmt, ...) {
  g_reuse_order = nullptr;
  g_reuse_order_P = -1;
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int ok() {
  g_reuse_order = nullptr;
  g_reuse_order_P = -1;
  g_last_error.clear();
  return DES_OK;
}

// The handed-over order if it was computed for P points, else null.
const void* take_reuse_order(int64_t P) {
  const void* r = g_reuse_order_P == P ? g_reuse_order : nullptr;
  g_reuse_order = nullptr;
  g_reuse_order_P = -1;
  return r;
}

}  // namespace des_capi

namespace {

using des_capi::fail;
using des_capi::g_last_error;
using des_capi::ok;

thread_local hipEvent_t g_bwd_ev_start = nullptr;
thread_local hipEvent_t g_bwd_ev_stop = nullptr;
thread_local hipEvent_t g_fwd_ev_start = nullptr;
thread_local hipEvent_t g_fwd_ev_stop = nullptr;
std::atomic<int> g_ntc_fwd_override{0};
std::atomic<int> g_ntc_bwd_override{0};
std::atomic<int> g_order_mode{0};  // 0 automatic, 1 always, -1 never

constexpr int kTileP = 64;
constexpr int64_t kMaxTexels = int64_t(1) << 24;
constexpr int kMinWaves = 1024;  // one wave per SIMD on 256 CUs x 4 SIMDs
// Automatic point order from this many point-texel pairs up: below it the
// three sort launches (~15 us) cost about what the skipped texels save.  In
// both directions the same threshold, so that a forward's order can serve the
// backward that follows it (des_reuse_point_order_once); alone, the forward
// about breaks even there (measured on MI355X at 70k x 512 pairs: 81.0 us
// unsorted, 79.5 us sorted including the sort).
constexpr int64_t kOrderMinPairsBackward = int64_t(1) << 24;
constexpr int64_t kOrderMinPairsForward = int64_t(1) << 24;

int64_t n_tiles(int64_t P) { return (P + kTileP - 1) / kTileP; }

bool is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

des::Tiling tiling_for(int ntc, int T, bool backward) {
  const int nslot = (T + 31) / 32;
  const int max_slots = des::kMaxSuperChunk / 32;
  const int slots_per_wave = std::max(1, std::min((nslot + ntc - 1) / ntc, max_slots / ntc));
  des::Tiling tl;
  tl.ntc = ntc;
  tl.cw = slots_per_wave * 32;
  tl.st = ntc * tl.cw;
  // texel table, then the tile's visibility words (3 per point per slot),
  // then the dark words (one bit per texel); the cross-wave reduction at the
  // end reuses the space from offset 0
  const size_t table = size_t(des::kTexelBytes) * tl.st;
  const size_t vis = size_t(tl.st / 32) * 3 * 64 * sizeof(uint32_t);
  const size_t dark = size_t((tl.st + 63) / 64) * sizeof(uint64_t);
  const size_t staged = table + vis + dark;
  tl.vis_off = int(table);
  tl.dark_off = int(table + vis);
  // the forward's per-wave lists of lit texels follow
  tl.list_off = int(staged);
  tl.list_cap = ((tl.st + ntc - 1) / ntc + 3) & ~3;
  const size_t staged_fwd = staged + size_t(ntc) * tl.list_cap * sizeof(uint16_t);
  // (the tile's exchanged point fields also use the space from offset 0, first)
  const size_t xch = size_t(des::kPointXchSlots) * 64 * sizeof(float);
  // the map's row and column angles last, apart from all of the above: they
  // are written with the exchanged point fields and read by the table build
  const size_t ang = size_t(des::kAngleSlots) * 2 * sizeof(double);
  const size_t fwd_end =
      std::max({staged_fwd, xch, size_t(ntc) * des::kForwardAcc * 64 * sizeof(float)});
  const size_t bwd_end = std::max({staged, xch, size_t(ntc) * des::kBackwardAcc * 64 * sizeof(float)});
  const size_t fwd_ang = (fwd_end + 15) & ~size_t(15), bwd_ang = (bwd_end + 15) & ~size_t(15);
  tl.ang_off = int(backward ? bwd_ang : fwd_ang);  // the direction this tiling is for
  tl.lds_fwd = fwd_ang + ang;
  tl.lds_bwd = bwd_ang + ang;
  tl.split0 = INT64_MAX;  // set per problem by choose_tiling (single-set launches only)
  tl.slices = 1;
  tl.slice_texels = T;
  return tl;
}

// Backward tail slices.  A backward launch keeps about 1024 tiles resident on
// MI355X (256 CUs x four 4-wave tiles); the tiles beyond a multiple of that
// would otherwise run as a fifth tile on some CUs while the others idle
// (measured at 70 000 points: 70 CUs did 5 tiles, the rest 4).  When that
// remainder is at most half the resident count, its tiles are shaded as texel
// slices -- eight each in the backward when there are at most 128 of them,
// else four, and four in the forward (measured best at 70 000 points: 8 in the
// forward cost more set-up than it saved) -- after the
// whole tiles, so every CU gets about the same work.  The rule depends on
// (P, T) only, so results do not depend on the device.
constexpr int64_t kResidentTiles = 1024;
constexpr int kTailSlicesForward = 4;
constexpr int kTailSlicesBackward = 8;

struct TailSplit {
  int64_t split0;
  int slices, slice_texels;
};

TailSplit tail_split(int64_t P, int T, bool backward) {
  const int64_t tiles = n_tiles(P);
  const int nslot = (T + 31) / 32;
  const int64_t rem = tiles % kResidentTiles;
  // eight backward slices only for a short tail (a long one balances with four,
  // and its per-point sums would grow the workspace eightfold)
  const int k = (backward && rem <= kResidentTiles / 8) ? kTailSlicesBackward : kTailSlicesForward;
  if (tiles > kResidentTiles && rem > 0 && 2 * rem <= kResidentTiles && nslot >= k)
    return TailSplit{tiles - rem, k, ((nslot + k - 1) / k) * 32};
  return TailSplit{tiles, 1, T};
}

// Launch shape as a function of (P, T) only, so results never depend on the
// device it runs on.
des::Tiling choose_tiling(int64_t P, int T, bool backward) {
  const int64_t tiles = n_tiles(P);
  const int nslot = (T + 31) / 32;
  const int cap = backward ? 8 : 16;
  int ntc = 1;
  const int ov = backward ? g_ntc_bwd_override.load() : g_ntc_fwd_override.load();
  if (ov > 0) {
    ntc = std::min(ov, cap);
  } else {
    // Measured on MI355X from 20k to 262k points with 512-texel super-chunks
    // (37 KB of LDS per tile): a tile runs best over 4 waves in both
    // directions, four 4-wave tiles per CU (the backward is held to 128
    // VGPRs, 4 waves per SIMD).  Problems of fewer than one wave per SIMD at
    // that split use more waves per tile.
    const int base = 4;
    while (ntc * 2 <= base && ntc * 2 <= nslot) ntc *= 2;
    while (ntc * 2 <= cap && ntc * 2 <= nslot && tiles * ntc < kMinWaves) ntc *= 2;
  }
  des::Tiling tl = tiling_for(ntc, T, backward);
  tl.split0 = tiles;
  if (backward && ov <= 0 && ntc == 4) {
    const TailSplit ts = tail_split(P, T, true);
    tl.split0 = ts.split0;
    tl.slices = ts.slices;
    tl.slice_texels = ts.slice_texels;
  }
  return tl;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Whether the single-problem entry points sort the points by normal direction.
bool use_order(int64_t P, int T, bool backward) {
  if (P <= 0 || P > int64_t(0x7fffffff)) return false;  // the order is int32
  const int mode = g_order_mode.load();
  if (mode != 0) return mode > 0;
  return P * int64_t(T) >= (backward ? kOrderMinPairsBackward : kOrderMinPairsForward);
}

struct WsLayout {
  size_t part_off, spart_off, order_off, total;
  bool ordered;
};

// Backward workspace: per-tile fp32 partials, the per-point sums of the tail
// slices (single-set launches), and (when the points are reordered) the
// point-order workspace.
WsLayout ws_layout(int64_t P, int T, bool allow_order) {
  WsLayout w{0, 0, 0, 0, false};
  if (P <= 0) return w;
  const int64_t tiles = n_tiles(P);
  w.part_off = 0;
  w.total = align256(size_t(tiles) * 3 * T * sizeof(float));
  if (allow_order) {
    const TailSplit ts = tail_split(P, T, true);
    w.spart_off = w.total;
    w.total += align256(size_t(tiles - ts.split0) * ts.slices * des::kBackwardAcc * 64 *
                        sizeof(float));
  }
  if (allow_order && use_order(P, T, true)) {
    w.ordered = true;
    w.order_off = w.total;
    w.total += des::order_layout(P).total;
  }
  return w;
}

int check_common(int64_t P, int32_t C, int32_t H, int32_t W) {
  if (P < 0) return fail(DES_ERR_INVALID_ARGUMENT, "P must be >= 0 (got %lld)", (long long)P);
  if (C != 3) return fail(DES_ERR_UNSUPPORTED_CHANNELS, "Channel number not supported!");
  if (H <= 0 || W <= 0)
    return fail(DES_ERR_INVALID_ARGUMENT, "envmap must be C x H x W with H, W >= 1 (got %d x %d)", H, W);
  if (int64_t(H) * W > kMaxTexels)
    return fail(DES_ERR_INVALID_ARGUMENT, "envmap has %lld texels, limit is %lld",
                (long long)(int64_t(H) * W), (long long)kMaxTexels);
  if (n_tiles(P) > int64_t(0x7fffffff))
    return fail(DES_ERR_INVALID_ARGUMENT, "P = %lld exceeds the grid limit", (long long)P);
  return DES_OK;
}

#define REQUIRE(ptr)                                                                     \
  do {                                                                                   \
    if ((ptr) == nullptr) return fail(DES_ERR_INVALID_ARGUMENT, #ptr " must not be NULL"); \
  } while (0)

int check_vis(int64_t P, int T, int32_t N_face, int32_t N_vertex, int32_t N_slot) {
  if (P == 0) return DES_OK;
  if (N_face <= 0) return fail(DES_ERR_INVALID_ARGUMENT, "face_vertex_list must be 3 x N_face with N_face >= 1");
  if (N_vertex <= 0) return fail(DES_ERR_INVALID_ARGUMENT, "visibility must be N_slot x N_vertex with N_vertex >= 1");
  const int need = (T + 31) / 32;
  if (N_slot < need)
    return fail(DES_ERR_INVALID_ARGUMENT,
                "visibility has %d slots of 32 texels, the %d-texel envmap needs %d", N_slot, T, need);
  return DES_OK;
}

int launch_error(hipError_t e, const char* what) {
  return fail(DES_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
}

// Forward workspace (when the points are reordered): the point order, then
// the per-point sums of the tail slices.
struct FwdLayout {
  size_t spart_off, total;
};

FwdLayout fwd_layout(int64_t P, int T) {
  FwdLayout f;
  f.spart_off = align256(des::order_layout(P).total);
  const TailSplit ts = tail_split(P, T, false);
  f.total = f.spart_off +
            align256(size_t(n_tiles(P) - ts.split0) * ts.slices * des::kForwardAcc * 64 *
                     sizeof(float));
  return f;
}

int run_forward(int variant, bool diff, bool spec, des::Problem& pr, void* ws, size_t ws_bytes,
                hipStream_t stream) {
  const void* reuse = des_capi::take_reuse_order(pr.P);
  hipEvent_t ev0 = g_fwd_ev_start, ev1 = g_fwd_ev_stop;
  g_fwd_ev_start = g_fwd_ev_stop = nullptr;
  if (pr.P == 0) return ok();
  const bool ordered = use_order(pr.P, pr.T, false);
  const bool use_ws = ws != nullptr && ordered;
  const FwdLayout fl = fwd_layout(pr.P, pr.T);
  if (use_ws) {
    if (ws_bytes < fl.total)
      return fail(DES_ERR_WORKSPACE_TOO_SMALL, "workspace of %zu bytes is smaller than the %zu required",
                  ws_bytes, fl.total);
    if ((reinterpret_cast<uintptr_t>(ws) & 255) != 0)
      return fail(DES_ERR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
  }
  if (reuse != nullptr && ordered) {
    pr.order = reinterpret_cast<const int32_t*>(static_cast<const char*>(reuse) +
                                                des::order_layout(pr.P).order_off);
  } else if (use_ws) {
    hipError_t e = des::launch_point_order(pr.normal, pr.P, ws, stream);
    if (e != hipSuccess) return launch_error(e, "point order failed");
    pr.order = reinterpret_cast<const int32_t*>(static_cast<char*>(ws) +
                                                des::order_layout(pr.P).order_off);
  }
  des::Tiling tl = choose_tiling(pr.P, pr.T, false);
  // tail slices need the workspace for their per-point sums
  if (use_ws && tl.ntc == 4 && g_ntc_fwd_override.load() <= 0) {
    const TailSplit ts = tail_split(pr.P, pr.T, false);
    tl.split0 = ts.split0;
    tl.slices = ts.slices;
    tl.slice_texels = ts.slice_texels;
    pr.spart = reinterpret_cast<float*>(static_cast<char*>(ws) + fl.spart_off);
  }
  hipError_t e = des::launch_forward(variant, diff, spec, pr, tl, stream, ev0, ev1);
  if (e != hipSuccess) return launch_error(e, "forward launch failed");
  return ok();
}

int run_backward(int variant, bool diff, bool spec, des::Problem& pr, void* ws, size_t ws_bytes,
                 hipStream_t stream) {
  const void* reuse = des_capi::take_reuse_order(pr.P);
  REQUIRE(pr.gE);
  if (pr.P == 0) {
    hipError_t e = des::launch_fill_zero(pr.gE, int64_t(3) * pr.T, stream);
    if (e != hipSuccess) return launch_error(e, "zero fill failed");
    return ok();
  }
  const WsLayout lay = ws_layout(pr.P, pr.T, true);
  if (ws == nullptr || ws_bytes < lay.total)
    return fail(DES_ERR_WORKSPACE_TOO_SMALL, "workspace of %zu bytes is smaller than the %zu required",
                ws_bytes, lay.total);
  if ((reinterpret_cast<uintptr_t>(ws) & 255) != 0)
    return fail(DES_ERR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
  pr.part = reinterpret_cast<float*>(static_cast<char*>(ws) + lay.part_off);
  pr.spart = reinterpret_cast<float*>(static_cast<char*>(ws) + lay.spart_off);
  if (lay.ordered && reuse != nullptr) {
    pr.order = reinterpret_cast<const int32_t*>(static_cast<const char*>(reuse) +
                                                des::order_layout(pr.P).order_off);
  } else if (lay.ordered) {
    char* ows = static_cast<char*>(ws) + lay.order_off;
    hipError_t e = des::launch_point_order(pr.normal, pr.P, ows, stream);
    if (e != hipSuccess) return launch_error(e, "point order failed");
    pr.order = reinterpret_cast<const int32_t*>(ows + des::order_layout(pr.P).order_off);
  }
  const des::Tiling tl = choose_tiling(pr.P, pr.T, true);
  hipEvent_t ev0 = g_bwd_ev_start, ev1 = g_bwd_ev_stop;
  g_bwd_ev_start = g_bwd_ev_stop = nullptr;
  if (ev0 && ev1) {
    hipError_t er = hipEventRecord(ev0, stream);
    if (er != hipSuccess) return launch_error(er, "timing event record failed");
  }
  hipError_t e = des::launch_backward(variant, diff, spec, pr, tl, stream);
  if (e != hipSuccess) return launch_error(e, "backward launch failed");
  if (ev0 && ev1) {
    hipError_t er = hipEventRecord(ev1, stream);
    if (er != hipSuccess) return launch_error(er, "timing event record failed");
  }
  e = des::launch_envgrad_reduce(variant, diff, spec, pr, tl, stream);
  if (e != hipSuccess) return launch_error(e, "env-map gradient reduce failed");
  return ok();
}

des::Problem make_problem(int64_t P, int32_t H, int32_t W) {
  des::Problem pr{};
  pr.P = P;
  pr.H = H;
  pr.W = W;
  pr.T = H * W;
  return pr;
}


// ---- batches ------------------------------------------------------------------

constexpr size_t kHeaderBytes = 256;
static_assert(sizeof(des::BatchHeader) <= kHeaderBytes, "header too large");

// Batch workspace: the descriptor table (header, one BatchSet per set, one
// SortJob per sorted set), then per set its env-gradient partials and, when
// the set is sorted, its point-order workspace.  A set is sorted when either
// direction's policy asks for it (the backward's threshold is the lower one);
// the order is computed once, at prepare time, for both directions.
struct BatchLayout {
  size_t sets_off, jobs_off, table_bytes, total;
  int njobs;
  std::vector<size_t> part_off, order_off;  // order_off: 0 = not sorted
};

BatchLayout batch_layout(const des_set* sets, int n) {
  BatchLayout L;
  L.sets_off = kHeaderBytes;
  L.jobs_off = (kHeaderBytes + sizeof(des::BatchSet) * size_t(n) + 15) & ~size_t(15);
  L.njobs = 0;
  for (int i = 0; i < n; ++i)
    if (use_order(sets[i].P, sets[i].H * sets[i].W, true)) ++L.njobs;
  L.table_bytes = align256(L.jobs_off + offsetof(des::SortJobs, jobs) +
                           sizeof(des::SortJob) * size_t(L.njobs));
  size_t off = L.table_bytes;
  L.part_off.resize(n);
  L.order_off.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    const WsLayout w = ws_layout(sets[i].P, sets[i].H * sets[i].W, false);
    L.part_off[i] = off + w.part_off;
    off += w.total;
    if (use_order(sets[i].P, sets[i].H * sets[i].W, true)) {
      L.order_off[i] = off;
      off += align256(des::order_layout(sets[i].P).total);
    }
  }
  L.total = off;
  return L;
}

int check_set(int variant, bool diff, bool spec, bool with_bwd, const des_set& s, int i) {
  if (int rc = check_common(s.P, 3, s.H, s.W))
    return fail(rc, "set %d: %s", i, g_last_error.c_str());
  if (s.P == 0) {
    if (with_bwd && s.grad_envmap == nullptr) return fail(DES_ERR_INVALID_ARGUMENT, "set %d: grad_envmap must not be NULL", i);
    return DES_OK;
  }
  const bool dout = variant == des::kV2 || variant == des::kVis;
  if (variant == des::kVis) {
    if (int rc = check_vis(s.P, s.H * s.W, s.N_face, s.N_vertex, s.N_slot))
      return fail(rc, "set %d: %s", i, g_last_error.c_str());
    if (!s.face_vertex_list || !s.nearest_triangle_id || !s.barycentric_coord || !s.visibility)
      return fail(DES_ERR_INVALID_ARGUMENT, "set %d: visibility inputs must not be NULL", i);
  }
  if (!s.normal || !s.envmap || !s.light2obj_rotmat || !s.shading)
    return fail(DES_ERR_INVALID_ARGUMENT, "set %d: normal, envmap, light2obj_rotmat and shading must not be NULL", i);
  if (diff && !s.albedo) return fail(DES_ERR_INVALID_ARGUMENT, "set %d: albedo must not be NULL", i);
  if ((spec || variant == des::kClamp) && !s.view_dir)
    return fail(DES_ERR_INVALID_ARGUMENT, "set %d: view_dir must not be NULL", i);
  if (spec && (!s.specular_albedo || !s.roughness))
    return fail(DES_ERR_INVALID_ARGUMENT, "set %d: specular_albedo and roughness must not be NULL", i);
  if (dout && !s.diffuse_shading)
    return fail(DES_ERR_INVALID_ARGUMENT, "set %d: diffuse_shading must not be NULL", i);
  if (with_bwd) {
    if (!s.grad_shading || !s.grad_normal || !s.grad_albedo || !s.grad_envmap)
      return fail(DES_ERR_INVALID_ARGUMENT, "set %d: backward buffers must not be NULL", i);
    if (variant != des::kDiffuse && (!s.grad_view_dir || !s.grad_specular_albedo || !s.grad_roughness))
      return fail(DES_ERR_INVALID_ARGUMENT, "set %d: backward buffers must not be NULL", i);
    if (dout && !s.grad_diffuse_shading)
      return fail(DES_ERR_INVALID_ARGUMENT, "set %d: grad_diffuse_shading must not be NULL", i);
  }
  return DES_OK;
}

int check_ntc(int ntc, int cap, const char* what) {
  if (ntc < 1 || ntc > cap || !is_pow2(ntc))
    return fail(DES_ERR_INVALID_ARGUMENT, "%s must be a power of two in [1, %d] (got %d)", what, cap, ntc);
  return DES_OK;
}

int num_cus() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cached[dev] = n;
  }
  return cached[dev];
}

des::BatchLaunch batch_launch(void* ws, int variant, bool diff, bool spec, bool fwd, int ntc) {
  des::BatchLaunch bl;
  bl.hdr = static_cast<const des::BatchHeader*>(ws);
  bl.hdr_mut = static_cast<des::BatchHeader*>(ws);
  bl.sets = reinterpret_cast<const des::BatchSet*>(static_cast<char*>(ws) + kHeaderBytes);
  bl.sig = des::batch_signature(variant, diff, spec, ntc);
  bl.ntc = ntc;
  // table of one super-chunk, or the cross-chunk reduction scratch
  const des::Tiling tl = tiling_for(ntc, des::kMaxSuperChunk, !fwd);
  bl.lds = fwd ? tl.lds_fwd : tl.lds_bwd;
  const int per_cu = des::batch_blocks_per_cu(variant, diff, spec, fwd, ntc, bl.lds);
  bl.grid = unsigned(num_cus() * per_cu);
  bl.reduce_grid = unsigned(num_cus());
  return bl;
}

// ---- scratch of the entry point that allocates nothing -------------------------
//
// des_bruteforce_specularvisibility_shader_backward2_noalloc keeps its
// workspace (per-tile env-map-gradient partials, tail-slice sums, point order)
// in this array instead of a caller's buffer.  Each GPU holds its own copy
// from the library's load to the end of the process; in host code the name
// is only a stand-in, so the device address of each GPU's copy comes from
// hipGetSymbolAddress, looked up once per device with that device current and
// used wherever the host hands the scratch to the device.  Every call on one
// GPU uses the same array: calls that overlap on one GPU (other streams,
// other threads) race.  Every byte a call reads is written earlier in the same
// call, so nothing needs clearing between calls.
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch into *out (DES_OK), or an error.
int device_scratch(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return launch_error(e, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(DES_ERR_INVALID_ARGUMENT, "device %d is beyond the %d the scratch table covers",
                dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return launch_error(e, "hipGetSymbolAddress of the static scratch failed");
    if (ptr == nullptr)
      return fail(DES_ERR_LAUNCH, "hipGetSymbolAddress gave no address for the static scratch");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return DES_OK;
}

// What des_batch_prepare recorded for a workspace, so that a launch with
// another variant, other branch flags or another launch shape is rejected
// before it runs (its kernels would read pointers the preparation did not
// check, or deal the texel groups to waves that do not exist).
struct PreparedBatch {
  int variant;
  bool diff, spec, with_backward;
  int ntc_fwd, ntc_bwd;
};
std::mutex g_batch_mu;
std::unordered_map<const void*, PreparedBatch> g_batches;

void record_batch(const void* ws, const PreparedBatch& b) {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  g_batches[ws] = b;
}

int check_prepared(const void* ws, int variant, bool diff, bool spec, bool fwd, int ntc) {
  PreparedBatch b;
  {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto it = g_batches.find(ws);
    if (it == g_batches.end())
      return fail(DES_ERR_INVALID_ARGUMENT, "workspace was not prepared by des_batch_prepare");
    b = it->second;
  }
  if (b.variant != variant || b.diff != diff || b.spec != spec)
    return fail(DES_ERR_INVALID_ARGUMENT,
                "batch was prepared for variant %d (diffuse %d, specular %d), launched as "
                "variant %d (diffuse %d, specular %d)",
                b.variant, int(b.diff), int(b.spec), variant, int(diff), int(spec));
  const int want = fwd ? b.ntc_fwd : b.ntc_bwd;
  if (want != ntc)
    return fail(DES_ERR_INVALID_ARGUMENT, "batch was prepared with %s = %d, launched with %d",
                fwd ? "ntc_forward" : "ntc_backward", want, ntc);
  if (!fwd && !b.with_backward)
    return fail(DES_ERR_INVALID_ARGUMENT, "batch was prepared without backward buffers");
  return DES_OK;
}

}  // namespace

extern "C" {

DES_API int des_abi_version(void) { return 2; }

DES_API const char* des_last_error(void) { return g_last_error.c_str(); }

DES_API size_t des_backward_workspace_size(int64_t P, int32_t H, int32_t W) {
  if (P <= 0 || H <= 0 || W <= 0) return 0;
  return ws_layout(P, H * W, true).total;
}

DES_API size_t des_forward_workspace_size(int64_t P, int32_t H, int32_t W) {
  if (P <= 0 || H <= 0 || W <= 0 || !use_order(P, H * W, false)) return 0;
  return fwd_layout(P, H * W).total;
}

DES_API int des_reuse_point_order_once(const void* order_workspace, int64_t P) {
  if ((reinterpret_cast<uintptr_t>(order_workspace) & 255) != 0)
    return fail(DES_ERR_INVALID_ARGUMENT, "order workspace must be 256-byte aligned");
  if (order_workspace != nullptr && P <= 0)
    return fail(DES_ERR_INVALID_ARGUMENT, "P must be >= 1 (got %lld)", (long long)P);
  const int rc = ok();
  des_capi::g_reuse_order = order_workspace;
  des_capi::g_reuse_order_P = order_workspace != nullptr ? P : -1;
  return rc;
}

DES_API int des_set_point_order(int32_t mode) {
  if (mode < -1 || mode > 1) return fail(DES_ERR_INVALID_ARGUMENT, "mode must be -1, 0 or 1");
  g_order_mode.store(mode);
  return ok();
}

DES_API int des_set_tiling_override(int32_t ntc_forward, int32_t ntc_backward) {
  if (ntc_forward < 0 || ntc_forward > 16 || (ntc_forward && !is_pow2(ntc_forward)))
    return fail(DES_ERR_INVALID_ARGUMENT, "ntc_forward must be 0 or a power of two <= 16");
  if (ntc_backward < 0 || ntc_backward > 8 || (ntc_backward && !is_pow2(ntc_backward)))
    return fail(DES_ERR_INVALID_ARGUMENT, "ntc_backward must be 0 or a power of two <= 8");
  g_ntc_fwd_override.store(ntc_forward);
  g_ntc_bwd_override.store(ntc_backward);
  return ok();
}

DES_API int des_set_backward_timing_events(void* start, void* stop) {
  if ((start == nullptr) != (stop == nullptr))
    return fail(DES_ERR_INVALID_ARGUMENT, "pass both timing events or neither");
  g_bwd_ev_start = static_cast<hipEvent_t>(start);
  g_bwd_ev_stop = static_cast<hipEvent_t>(stop);
  return ok();
}

DES_API int des_set_forward_timing_events(void* start, void* stop) {
  if ((start == nullptr) != (stop == nullptr))
    return fail(DES_ERR_INVALID_ARGUMENT, "pass both timing events or neither");
  g_fwd_ev_start = static_cast<hipEvent_t>(start);
  g_fwd_ev_stop = static_cast<hipEvent_t>(stop);
  return ok();
}

DES_API int des_bruteforce_diffuse_shader_forward(int64_t P, int32_t C, int32_t H, int32_t W,
                                                  const float* normal, const float* albedo,
                                                  const float* envmap,
                                                  const float* light2obj_rotmat, float* shading,
                                                  void* workspace, size_t workspace_bytes,
                                                  void* stream) {
  if (int rc = check_common(P, C, H, W)) return rc;
  if (P > 0) {
    REQUIRE(normal); REQUIRE(albedo); REQUIRE(envmap); REQUIRE(light2obj_rotmat); REQUIRE(shading);
  }
  des::Problem pr = make_problem(P, H, W);
  pr.normal = normal; pr.albedo = albedo; pr.envmap = envmap; pr.rotmat = light2obj_rotmat;
  pr.sh = shading;
  return run_forward(des::kDiffuse, true, false, pr, workspace, workspace_bytes,
                     static_cast<hipStream_t>(stream));
}

DES_API int des_bruteforce_diffuse_shader_backward(
    int64_t P, int32_t C, int32_t H, int32_t W, const float* grad_shading, const float* normal,
    const float* albedo, const float* envmap, const float* light2obj_rotmat, float* grad_normal,
    float* grad_albedo, float* grad_envmap, void* workspace, size_t workspace_bytes,
    void* stream) {
  if (int rc = check_common(P, C, H, W)) return rc;
  if (P > 0) {
    REQUIRE(grad_shading); REQUIRE(normal); REQUIRE(albedo); REQUIRE(envmap);
    REQUIRE(light2obj_rotmat); REQUIRE(grad_normal); REQUIRE(grad_albedo);
  }
  des::Problem pr = make_problem(P, H, W);
  pr.g_sh = grad_shading; pr.normal = normal; pr.albedo = albedo; pr.envmap = envmap;
  pr.rotmat = light2obj_rotmat;
  pr.gn = grad_normal; pr.ga = grad_albedo; pr.gE = grad_envmap;
  return run_backward(des::kDiffuse, true, false, pr, workspace, workspace_bytes,
                      static_cast<hipStream_t>(stream));
}

static int specular_forward(int variant, int64_t P, int32_t C, int32_t H, int32_t W,
                            const float* normal, const float* view_dir, const float* albedo,
                            const float* specular_albedo, const float* roughness,
                            const float* envmap, const float* light2obj_rotmat, bool diff,
                            bool spec, float* shading, float* diffuse_shading, void* workspace,
                            size_t workspace_bytes, void* stream) {
  if (int rc = check_common(P, C, H, W)) return rc;
  if (P > 0) {
    REQUIRE(normal); REQUIRE(envmap); REQUIRE(light2obj_rotmat); REQUIRE(shading);
    if (diff) REQUIRE(albedo);
    if (spec) { REQUIRE(view_dir); REQUIRE(specular_albedo); REQUIRE(roughness); }
    if (variant == des::kClamp) REQUIRE(view_dir);
    if (variant == des::kV2 || variant == des::kVis) REQUIRE(diffuse_shading);
  }
  des::Problem pr = make_problem(P, H, W);
  pr.normal = normal; pr.view_dir = view_dir; pr.albedo = albedo; pr.spec = specular_albedo;
  pr.rough = roughness; pr.envmap = envmap; pr.rotmat = light2obj_rotmat;
  pr.sh = shading; pr.ds = diffuse_shading;
  return run_forward(variant, diff, spec, pr, workspace, workspace_bytes,
                     static_cast<hipStream_t>(stream));
}

static int specular_backward(int variant, int64_t P, int32_t C, int32_t H, int32_t W,
                             const float* grad_shading, const float* grad_diffuse_shading,
                             const float* normal, const float* view_dir, const float* albedo,
                             const float* specular_albedo, const float* roughness,
                             const float* envmap, const float* light2obj_rotmat, bool diff,
                             bool spec, float* grad_normal, float* grad_view_dir,
                             float* grad_albedo, float* grad_specular_albedo,
                             float* grad_roughness, float* grad_envmap, void* workspace,
                             size_t workspace_bytes, void* stream, des::Problem* vis_in) {
  if (int rc = check_common(P, C, H, W)) return rc;
  if (P > 0) {
    REQUIRE(grad_shading); REQUIRE(normal); REQUIRE(envmap); REQUIRE(light2obj_rotmat);
    REQUIRE(grad_normal); REQUIRE(grad_view_dir); REQUIRE(grad_albedo);
    REQUIRE(grad_specular_albedo); REQUIRE(grad_roughness);
    if (diff) REQUIRE(albedo);
    if (spec) { REQUIRE(view_dir); REQUIRE(specular_albedo); REQUIRE(roughness); }
    if (variant == des::kClamp) REQUIRE(view_dir);
    if (variant == des::kV2 || variant == des::kVis) REQUIRE(grad_diffuse_shading);
  }
  des::Problem pr = vis_in ? *vis_in : make_problem(P, H, W);
  pr.g_sh = grad_shading; pr.g_ds = grad_diffuse_shading;
  pr.normal = normal; pr.view_dir = view_dir; pr.albedo = albedo; pr.spec = specular_albedo;
  pr.rough = roughness; pr.envmap = envmap; pr.rotmat = light2obj_rotmat;
  pr.gn = grad_normal; pr.gv = grad_view_dir; pr.ga = grad_albedo; pr.gs = grad_specular_albedo;
  pr.gr = grad_roughness; pr.gE = grad_envmap;
  return run_backward(variant, diff, spec, pr, workspace, workspace_bytes,
                      static_cast<hipStream_t>(stream));
}

DES_API int des_bruteforce_specular_shader_forward(
    int64_t P, int32_t C, int32_t H, int32_t W, const float* normal, const float* view_dir,
    const float* albedo, const float* specular_albedo, const float* roughness,
    const float* envmap, const float* light2obj_rotmat, int enable_diffuse, int enable_specular,
    float* shading, void* workspace, size_t workspace_bytes, void* stream) {
  return specular_forward(des::kSpecular, P, C, H, W, normal, view_dir, albedo, specular_albedo,
                          roughness, envmap, light2obj_rotmat, enable_diffuse != 0,
                          enable_specular != 0, shading, nullptr, workspace, workspace_bytes,
                          stream);
}

DES_API int des_bruteforce_specular_shader_backward(
    int64_t P, int32_t C, int32_t H, int32_t W, const float* grad_shading, const float* normal,
    const float* view_dir, const float* albedo, const float* specular_albedo,
    const float* roughness, const float* envmap, const float* light2obj_rotmat,
    int enable_diffuse, int enable_specular, float* grad_normal, float* grad_view_dir,
    float* grad_albedo, float* grad_specular_albedo, float* grad_roughness, float* grad_envmap,
    void* workspace, size_t workspace_bytes, void* stream) {
  return specular_backward(des::kSpecular, P, C, H, W, grad_shading, nullptr, normal, view_dir,
                           albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
                           enable_diffuse != 0, enable_specular != 0, grad_normal, grad_view_dir,
                           grad_albedo, grad_specular_albedo, grad_roughness, grad_envmap,
                           workspace, workspace_bytes, stream, nullptr);
}

DES_API int des_bruteforce_specular_shader_clamp_forward(
    int64_t P, int32_t C, int32_t H, int32_t W, const float* normal, const float* view_dir,
    const float* albedo, const float* specular_albedo, const float* roughness,
    const float* envmap, const float* light2obj_rotmat, int enable_diffuse, int enable_specular,
    float* shading, void* workspace, size_t workspace_bytes, void* stream) {
  return specular_forward(des::kClamp, P, C, H, W, normal, view_dir, albedo, specular_albedo,
                          roughness, envmap, light2obj_rotmat, enable_diffuse != 0,
                          enable_specular != 0, shading, nullptr, workspace, workspace_bytes,
                          stream);
}

DES_API int des_bruteforce_specular_shader_clamp_backward(
    int64_t P, int32_t C, int32_t H, int32_t W, const float* grad_shading, const float* normal,
    const float* view_dir, const float* albedo, const float* specular_albedo,
    const float* roughness, const float* envmap, const float* light2obj_rotmat,
    int enable_diffuse, int enable_specular, float* grad_normal, float* grad_view_dir,
    float* grad_albedo, float* grad_specular_albedo, float* grad_roughness, float* grad_envmap,
    void* workspace, size_t workspace_bytes, void* stream) {
  return specular_backward(des::kClamp, P, C, H, W, grad_shading, nullptr, normal, view_dir,
                           albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
                           enable_diffuse != 0, enable_specular != 0, grad_normal, grad_view_dir,
                           grad_albedo, grad_specular_albedo, grad_roughness, grad_envmap,
                           workspace, workspace_bytes, stream, nullptr);
}

DES_API int des_bruteforce_specular_shader_forward2(
    int64_t P, int32_t C, int32_t H, int32_t W, const float* normal, const float* view_dir,
    const float* albedo, const float* specular_albedo, const float* roughness,
    const float* envmap, const float* light2obj_rotmat, int enable_specular, float* shading,
    float* diffuse_shading, void* workspace, size_t workspace_bytes, void* stream) {
  return specular_forward(des::kV2, P, C, H, W, normal, view_dir, albedo, specular_albedo,
                          roughness, envmap, light2obj_rotmat, true, enable_specular != 0,
                          shading, diffuse_shading, workspace, workspace_bytes, stream);
}

DES_API int des_bruteforce_specular_shader_backward2(
    int64_t P, int32_t C, int32_t H, int32_t W, const float* grad_shading,
    const float* grad_diffuse_shading, const float* normal, const float* view_dir,
    const float* albedo, const float* specular_albedo, const float* roughness,
    const float* envmap, const float* light2obj_rotmat, int enable_specular, float* grad_normal,
    float* grad_view_dir, float* grad_albedo, float* grad_specular_albedo, float* grad_roughness,
    float* grad_envmap, void* workspace, size_t workspace_bytes, void* stream) {
  return specular_backward(des::kV2, P, C, H, W, grad_shading, grad_diffuse_shading, normal,
                           view_dir, albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
                           true, enable_specular != 0, grad_normal, grad_view_dir, grad_albedo,
                           grad_specular_albedo, grad_roughness, grad_envmap, workspace,
                           workspace_bytes, stream, nullptr);
}

DES_API int des_bruteforce_specularvisibility_shader_forward2(
    int64_t P, int32_t C, int32_t H, int32_t W, int32_t N_face, int32_t N_vertex, int32_t N_slot,
    const float* normal, const float* view_dir, const float* albedo,
    const float* specular_albedo, const float* roughness, const float* envmap,
    const float* light2obj_rotmat, const int32_t* face_vertex_list,
    const int64_t* nearest_triangle_id, const float* barycentric_coord,
    const int32_t* visibility, int enable_specular, float* shading, float* diffuse_shading,
    void* workspace, size_t workspace_bytes, void* stream) {
  if (int rc = check_common(P, C, H, W)) return rc;
  if (int rc = check_vis(P, H * W, N_face, N_vertex, N_slot)) return rc;
  if (P > 0) {
    REQUIRE(normal); REQUIRE(albedo); REQUIRE(envmap); REQUIRE(light2obj_rotmat);
    REQUIRE(face_vertex_list); REQUIRE(nearest_triangle_id); REQUIRE(barycentric_coord);
    REQUIRE(visibility); REQUIRE(shading); REQUIRE(diffuse_shading);
    if (enable_specular) { REQUIRE(view_dir); REQUIRE(specular_albedo); REQUIRE(roughness); }
  }
  des::Problem pr = make_problem(P, H, W);
  pr.N_face = N_face; pr.N_vertex = N_vertex;
  pr.normal = normal; pr.view_dir = view_dir; pr.albedo = albedo; pr.spec = specular_albedo;
  pr.rough = roughness; pr.envmap = envmap; pr.rotmat = light2obj_rotmat;
  pr.fvl = face_vertex_list; pr.tri = nearest_triangle_id; pr.bary = barycentric_coord;
  pr.vis = reinterpret_cast<const uint32_t*>(visibility);
  pr.sh = shading; pr.ds = diffuse_shading;
  return run_forward(des::kVis, true, enable_specular != 0, pr, workspace, workspace_bytes,
                     static_cast<hipStream_t>(stream));
}

DES_API int des_bruteforce_specularvisibility_shader_backward2(
    int64_t P, int32_t C, int32_t H, int32_t W, int32_t N_face, int32_t N_vertex, int32_t N_slot,
    const float* grad_shading, const float* grad_diffuse_shading, const float* normal,
    const float* view_dir, const float* albedo, const float* specular_albedo,
    const float* roughness, const float* envmap, const float* light2obj_rotmat,
    const int32_t* face_vertex_list, const int64_t* nearest_triangle_id,
    const float* barycentric_coord, const int32_t* visibility, int enable_specular,
    float* grad_normal, float* grad_view_dir, float* grad_albedo, float* grad_specular_albedo,
    float* grad_roughness, float* grad_envmap, void* workspace, size_t workspace_bytes,
    void* stream) {
  if (int rc = check_common(P, C, H, W)) return rc;
  if (int rc = check_vis(P, H * W, N_face, N_vertex, N_slot)) return rc;
  if (P > 0) {
    REQUIRE(face_vertex_list); REQUIRE(nearest_triangle_id); REQUIRE(barycentric_coord);
    REQUIRE(visibility);
  }
  des::Problem pr = make_problem(P, H, W);
  pr.N_face = N_face; pr.N_vertex = N_vertex;
  pr.fvl = face_vertex_list; pr.tri = nearest_triangle_id; pr.bary = barycentric_coord;
  pr.vis = reinterpret_cast<const uint32_t*>(visibility);
  return specular_backward(des::kVis, P, C, H, W, grad_shading, grad_diffuse_shading, normal,
                           view_dir, albedo, specular_albedo, roughness, envmap, light2obj_rotmat,
                           true, enable_specular != 0, grad_normal, grad_view_dir, grad_albedo,
                           grad_specular_albedo, grad_roughness, grad_envmap, workspace,
                           workspace_bytes, stream, &pr);
}

DES_API size_t des_static_scratch_bytes(void) { return kStaticScratchBytes; }

DES_API int des_bruteforce_specularvisibility_shader_backward2_noalloc(
    int64_t P, int32_t C, int32_t H, int32_t W, int32_t N_face, int32_t N_vertex, int32_t N_slot,
    const float* grad_shading, const float* grad_diffuse_shading, const float* normal,
    const float* view_dir, const float* albedo, const float* specular_albedo,
    const float* roughness, const float* envmap, const float* light2obj_rotmat,
    const int32_t* face_vertex_list, const int64_t* nearest_triangle_id,
    const float* barycentric_coord, const int32_t* visibility, int enable_specular,
    float* grad_normal, float* grad_view_dir, float* grad_albedo, float* grad_specular_albedo,
    float* grad_roughness, float* grad_envmap, void* stream) {
  if (int rc = check_common(P, C, H, W)) return rc;
  const size_t need = des_backward_workspace_size(P, H, W);
  if (need > kStaticScratchBytes)
    return fail(DES_ERR_INVALID_ARGUMENT,
                "P = %lld points with a %d x %d env map need %zu bytes of scratch; the static "
                "scratch of this entry point holds %zu (use the workspace entry point)",
                (long long)P, H, W, need, kStaticScratchBytes);
  void* ws = nullptr;
  if (P > 0)
    if (int rc = device_scratch(&ws)) return rc;
  return des_bruteforce_specularvisibility_shader_backward2(
      P, C, H, W, N_face, N_vertex, N_slot, grad_shading, grad_diffuse_shading, normal, view_dir,
      albedo, specular_albedo, roughness, envmap, light2obj_rotmat, face_vertex_list,
      nearest_triangle_id, barycentric_coord, visibility, enable_specular, grad_normal,
      grad_view_dir, grad_albedo, grad_specular_albedo, grad_roughness, grad_envmap, ws,
      ws ? kStaticScratchBytes : 0, stream);
}

DES_API int des_batch_plan(int32_t variant, const des_set* sets, int32_t n_sets,
                           size_t* workspace_bytes, size_t* staging_bytes) {
  if (variant < des::kDiffuse || variant > des::kVis)
    return fail(DES_ERR_INVALID_ARGUMENT, "unknown variant %d", variant);
  if (n_sets < 0 || (n_sets > 0 && sets == nullptr))
    return fail(DES_ERR_INVALID_ARGUMENT, "sets must hold n_sets >= 0 descriptors");
  for (int i = 0; i < n_sets; ++i)
    if (int rc = check_common(sets[i].P, 3, sets[i].H, sets[i].W))
      return fail(rc, "set %d: %s", i, g_last_error.c_str());
  const BatchLayout L = batch_layout(sets, n_sets);
  if (workspace_bytes) *workspace_bytes = L.total;
  if (staging_bytes) *staging_bytes = L.table_bytes;
  return ok();
}

DES_API int des_batch_prepare(int32_t variant, int enable_diffuse, int enable_specular,
                              const des_set* sets, int32_t n_sets, int32_t ntc_forward,
                              int32_t ntc_backward, int with_backward, void* staging,
                              size_t staging_bytes, void* workspace, size_t workspace_bytes,
                              void* stream) {
  if (variant < des::kDiffuse || variant > des::kVis)
    return fail(DES_ERR_INVALID_ARGUMENT, "unknown variant %d", variant);
  if (n_sets < 0 || (n_sets > 0 && sets == nullptr))
    return fail(DES_ERR_INVALID_ARGUMENT, "sets must hold n_sets >= 0 descriptors");
  if (int rc = check_ntc(ntc_forward, 8, "ntc_forward")) return rc;
  if (int rc = check_ntc(ntc_backward, 8, "ntc_backward")) return rc;
  const bool diff = variant == des::kDiffuse || variant == des::kV2 || variant == des::kVis || enable_diffuse;
  const bool spec = variant != des::kDiffuse && enable_specular;
  for (int i = 0; i < n_sets; ++i)
    if (int rc = check_set(variant, diff, spec, with_backward != 0, sets[i], i)) return rc;
  const BatchLayout L = batch_layout(sets, n_sets);
  if (staging == nullptr || staging_bytes < L.table_bytes)
    return fail(DES_ERR_WORKSPACE_TOO_SMALL, "staging buffer of %zu bytes is smaller than the %zu required",
                staging_bytes, L.table_bytes);
  if (workspace == nullptr || workspace_bytes < L.total)
    return fail(DES_ERR_WORKSPACE_TOO_SMALL, "workspace of %zu bytes is smaller than the %zu required",
                workspace_bytes, L.total);
  if ((reinterpret_cast<uintptr_t>(workspace) & 255) != 0)
    return fail(DES_ERR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
  char* st = static_cast<char*>(staging);
  std::memset(st, 0, L.table_bytes);
  auto* hdr = reinterpret_cast<des::BatchHeader*>(st);
  auto* bs = reinterpret_cast<des::BatchSet*>(st + kHeaderBytes);
  char* ws = static_cast<char*>(workspace);
  auto* tab = reinterpret_cast<des::SortJobs*>(st + L.jobs_off);
  des::SortJob* jobs = tab->jobs;
  int njobs = 0, sort_blocks = 0, sort_scan = 0;
  int64_t tile0 = 0, red = 0;
  for (int i = 0; i < n_sets; ++i) {
    const des_set& s = sets[i];
    des::Problem pr = make_problem(s.P, s.H, s.W);
    pr.N_face = s.N_face; pr.N_vertex = s.N_vertex;
    pr.normal = s.normal; pr.view_dir = s.view_dir; pr.albedo = s.albedo;
    pr.spec = s.specular_albedo; pr.rough = s.roughness; pr.envmap = s.envmap;
    pr.rotmat = s.light2obj_rotmat;
    pr.fvl = s.face_vertex_list; pr.tri = s.nearest_triangle_id; pr.bary = s.barycentric_coord;
    pr.vis = reinterpret_cast<const uint32_t*>(s.visibility);
    pr.sh = s.shading; pr.ds = s.diffuse_shading;
    pr.g_sh = s.grad_shading; pr.g_ds = s.grad_diffuse_shading;
    pr.gn = s.grad_normal; pr.gv = s.grad_view_dir; pr.ga = s.grad_albedo;
    pr.gs = s.grad_specular_albedo; pr.gr = s.grad_roughness; pr.gE = s.grad_envmap;
    pr.part = reinterpret_cast<float*>(ws + L.part_off[i]);
    des::BatchSet& b = bs[i];
    b.pr = pr;
    if (L.order_off[i] != 0) {
      des::SortJob job = des::sort_job(s.normal, s.P, ws + L.order_off[i]);
      job.blk0 = sort_blocks;
      job.scan0 = sort_scan;
      sort_blocks += job.nblk;
      sort_scan += des::sort_scan_groups(job);
      jobs[njobs++] = job;
      if (use_order(s.P, pr.T, false)) b.order_fwd = job.order;
      b.order_bwd = job.order;
    }
    b.tl_fwd = tiling_for(ntc_forward, pr.T, false);
    b.tl_bwd = tiling_for(ntc_backward, pr.T, true);
    const int64_t tiles = n_tiles(s.P);
    b.tile0 = tile0;
    b.red_item0 = red;
    tile0 += tiles;
    red += (3 * pr.T + 63) / 64;  // batched reduce: 64-column items
  }
  hdr->n_sets = n_sets;
  hdr->variant = variant;
  hdr->has_backward = with_backward ? 1 : 0;
  hdr->sig_fwd = des::batch_signature(variant, diff, spec, ntc_forward);
  hdr->sig_bwd = des::batch_signature(variant, diff, spec, ntc_backward);
  hdr->total_tiles = tile0;
  hdr->total_red = red;
  hdr->sort_jobs = reinterpret_cast<const des::SortJobs*>(ws + L.jobs_off);
  tab->njobs = njobs;
  tab->nblocks = sort_blocks;
  tab->nscan = sort_scan;
  hipError_t e = hipMemcpyAsync(workspace, staging, L.table_bytes, hipMemcpyHostToDevice,
                                static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_error(e, "descriptor upload failed");
  record_batch(workspace, PreparedBatch{variant, diff, spec, with_backward != 0, ntc_forward,
                                        ntc_backward});
  e = des::launch_point_order_jobs(&reinterpret_cast<des::BatchHeader*>(ws)->sort_jobs, num_cus(),
                                   static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_error(e, "point order failed");
  return ok();
}

DES_API int des_batch_sort(void* workspace, void* stream) {
  REQUIRE(workspace);
  {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    if (g_batches.find(workspace) == g_batches.end())
      return fail(DES_ERR_INVALID_ARGUMENT, "workspace was not prepared by des_batch_prepare");
  }
  hipError_t e = des::launch_point_order_jobs(
      &static_cast<des::BatchHeader*>(workspace)->sort_jobs, num_cus(),
      static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_error(e, "point order failed");
  return ok();
}

DES_API int des_batch_forward(int32_t variant, int enable_diffuse, int enable_specular,
                              int32_t ntc_forward, void* workspace, void* stream) {
  if (variant < des::kDiffuse || variant > des::kVis)
    return fail(DES_ERR_INVALID_ARGUMENT, "unknown variant %d", variant);
  if (int rc = check_ntc(ntc_forward, 8, "ntc_forward")) return rc;
  REQUIRE(workspace);
  const bool diff = variant == des::kDiffuse || variant == des::kV2 || variant == des::kVis || enable_diffuse;
  const bool spec = variant != des::kDiffuse && enable_specular;
  if (int rc = check_prepared(workspace, variant, diff, spec, true, ntc_forward)) return rc;
  const des::BatchLaunch bl = batch_launch(workspace, variant, diff, spec, true, ntc_forward);
  hipError_t e = des::launch_batch_forward(variant, diff, spec, bl, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_error(e, "batched forward launch failed");
  return ok();
}

DES_API int des_batch_backward(int32_t variant, int enable_diffuse, int enable_specular,
                               int32_t ntc_backward, void* workspace, void* stream) {
  if (variant < des::kDiffuse || variant > des::kVis)
    return fail(DES_ERR_INVALID_ARGUMENT, "unknown variant %d", variant);
  if (int rc = check_ntc(ntc_backward, 8, "ntc_backward")) return rc;
  REQUIRE(workspace);
  const bool diff = variant == des::kDiffuse || variant == des::kV2 || variant == des::kVis || enable_diffuse;
  const bool spec = variant != des::kDiffuse && enable_specular;
  if (int rc = check_prepared(workspace, variant, diff, spec, false, ntc_backward)) return rc;
  const des::BatchLaunch bl = batch_launch(workspace, variant, diff, spec, false, ntc_backward);
  hipError_t e = des::launch_batch_backward(variant, diff, spec, bl, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return launch_error(e, "batched backward launch failed");
  return ok();
}

}  // extern "C"
