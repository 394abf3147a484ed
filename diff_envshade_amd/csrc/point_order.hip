// point_order.hip -- normal-coherent point order for the shading kernels.
//
// A tile of 64 points whose normals point the same way leaves whole texels
// with no lit lane: every texel in the opposite hemisphere of all 64 normals
// has l.n <= 0 on every lane and contributes exactly nothing, so the shading
// kernels skip it.  With random point order a tile's normals cover the sphere
// and no texel is ever dark; sorted by normal direction about 45% of the
// (tile, texel) pairs of a 16x32 map are (at least a third in every tile).
//
// Key: the 10-bit Hilbert index of the normal's cell on a 32 x 32 octahedral
// map, so consecutive keys are neighbouring directions.  The order is a
// stable counting sort by key -- points with equal keys keep their index
// order -- computed without atomics on positions, so it depends only on the
// normals, and every output -- the per-point sums, grouped by the tile's
// waves, and the per-tile env-map-gradient partials -- is bitwise
// reproducible; against input order the outputs move at fp32 rounding level.
//
// Blocks of 1024 points, one per thread (small blocks spread the ranking over
// many CUs):
//   1. key of each point and its rank among the equal-key points of its block
//      that precede it (within a wave: a match mask built from one ballot per
//      key bit; across waves and rounds: per-key counts in LDS), plus the
//      block's per-key counts: up to kFusedBlocks blocks block-major as 16-bit
//      words, count[block][key], beyond that key-major, count[key][block];
//   2. (more than kFusedBlocks blocks only) per key, an exclusive scan of
//      count[key][.] over blocks in place, and the key totals;
//   3. key bases (an exclusive scan over keys of the totals, in LDS) and the
//      scatter: each point goes to base[key] + prefix[key][block] + rank.  Up
//      to kFusedBlocks blocks, each block reads every block's count row
//      itself (coalesced 2 KB rows) and step 2 is not launched.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "envshade_internal.h"

namespace des {

namespace {

constexpr int kThreads = 1024;
constexpr int kRounds = kOrderBlock / kThreads;  // 1
constexpr int kWaves = kThreads / 64;             // 16
constexpr int kKeys = kOrderKeys;                 // 1024
constexpr int kKeyBits = 10;
// Up to this many blocks each scatter block sums the count rows itself (that
// work grows with the square of the block count; 128 blocks are 512 KB read
// per block); beyond, a separate per-key scan runs first.
constexpr int kFusedBlocks = 128;
static_assert(kOrderBlock <= 0xffff, "a block's per-key count fits 16 bits");

// Hilbert-curve index of cell (x, y) of a 32 x 32 grid.  Consecutive indices
// are always neighbouring cells, so a run of 64 sorted points stays compact
// (a Morton order jumps across the grid at power-of-two boundaries, and a tile
// straddling such a jump has a wide normal cone and few dark texels).
__device__ __forceinline__ uint32_t hilbert5(uint32_t x, uint32_t y) {
  uint32_t d = 0u;
#pragma unroll
  for (uint32_t s = 16u; s > 0u; s >>= 1) {
    const uint32_t rx = (x & s) ? 1u : 0u;
    const uint32_t ry = (y & s) ? 1u : 0u;
    d += s * s * ((3u * rx) ^ ry);
    if (ry == 0u) {
      if (rx == 1u) {
        x = 31u - x;
        y = 31u - y;
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

// 10-bit key of direction (x, y, z): the Hilbert index of its cell on a
// 32 x 32 octahedral map.  Zero or non-finite normals get key 0.
__device__ __forceinline__ uint32_t normal_key(float x, float y, float z) {
  const float s = fabsf(x) + fabsf(y) + fabsf(z);
  if (!(s > 0.f) || !(s < INFINITY)) return 0u;
  const float is = __builtin_amdgcn_rcpf(s);  // any deterministic rounding will do
  float u = x * is, v = y * is;
  if (z < 0.f) {
    const float fu = (1.f - fabsf(v)) * (u < 0.f ? -1.f : 1.f);
    const float fv = (1.f - fabsf(u)) * (v < 0.f ? -1.f : 1.f);
    u = fu;
    v = fv;
  }
  const int iu = min(max((int)((u + 1.f) * 16.f), 0), 31);
  const int iv = min(max((int)((v + 1.f) * 16.f), 0), 31);
  return hilbert5((uint32_t)iu, (uint32_t)iv);
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// 1. keys, in-block ranks, per-block key counts
__device__ __forceinline__ void rank_block(const SortJob& job, int blk) {
  const float* __restrict__ normal = job.normal;
  const int64_t P = job.P;
  uint32_t* __restrict__ keyrank = job.keyrank;
  uint32_t* __restrict__ count = job.count;
  const int nblk = job.nblk;
  const int nblk_pad = job.nblk_pad;
  // wcount[k][w] = (round + 1) << 12 | n: first the number n of key-k points of
  // wave w in that round, then (after the column scan) the number of key-k
  // points of this block that precede wave w's.  Key-major, so that thread k
  // moves its key's 16 entries with two 16-byte LDS accesses.
  __shared__ __attribute__((aligned(16))) uint16_t wcount[kKeys * kWaves];  // 32 KB
  __shared__ uint16_t run[kKeys];  // key-k points of this block in earlier rounds
  static_assert(kKeys == kThreads && kWaves == 16, "thread k owns key k's 16 wave counts");
  const int tid = threadIdx.x;
  uint16_t* const col = wcount + tid * kWaves;  // 32 bytes, 16-byte aligned
  const int lane = tid & 63;
  const int wv = tid >> 6;
  // all four rounds' normals are requested before any is used: one memory
  // latency per block instead of one per round
  float nrm[kRounds][3];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int64_t p = (int64_t)blk * kOrderBlock + r * kThreads + tid;
    nrm[r][0] = nrm[r][1] = nrm[r][2] = 0.f;
    if (p < P) {
      nrm[r][0] = normal[3 * p];
      nrm[r][1] = normal[3 * p + 1];
      nrm[r][2] = normal[3 * p + 2];
    }
  }
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  __builtin_memcpy(col, &zero, 16);
  __builtin_memcpy(col + 8, &zero, 16);
  run[tid] = 0;
  __syncthreads();
  // ranks stay in registers until the end: a barrier waits for outstanding
  // global stores, so none are issued between barriers
  uint32_t kr[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int64_t p = (int64_t)blk * kOrderBlock + r * kThreads + tid;
    const bool valid = p < P;
    const uint32_t key = valid ? normal_key(nrm[r][0], nrm[r][1], nrm[r][2]) : 0u;
    // lanes of this wave holding the same key: intersect, over the key bits,
    // the ballot of lanes that agree with this lane on the bit
    uint64_t eq = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kKeyBits; ++b) {
      const bool bit = (key >> b) & 1u;
      const uint64_t m = __ballot(bit);
      eq &= bit ? m : ~m;
    }
    const uint32_t rank_w = (uint32_t)__popcll(eq & lanemask_lt(lane));
    const uint32_t tag = (uint32_t)(r + 1) << 12;
    if (valid && rank_w == 0u) wcount[key * kWaves + wv] = (uint16_t)(tag | (uint32_t)__popcll(eq));
    __syncthreads();
    // column scan of key tid over the waves: this round's counts -> prefixes
    // (entries of other rounds count 0; every entry is rewritten with this
    // round's tag, which the next round reads as stale)
    {
      uint4 c0, c1;
      __builtin_memcpy(&c0, col, 16);
      __builtin_memcpy(&c1, col + 8, 16);
      const uint32_t c[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
      uint32_t acc = run[tid];
      uint32_t out[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t lo = c[i] & 0xffffu, hi = c[i] >> 16;
        const uint32_t plo = acc;
        acc += ((lo & 0xf000u) == tag) ? (lo & 0x0fffu) : 0u;
        const uint32_t phi = acc;
        acc += ((hi & 0xf000u) == tag) ? (hi & 0x0fffu) : 0u;
        out[i] = (tag | plo) | ((tag | phi) << 16);
      }
      const uint4 o0 = make_uint4(out[0], out[1], out[2], out[3]);
      const uint4 o1 = make_uint4(out[4], out[5], out[6], out[7]);
      __builtin_memcpy(col, &o0, 16);
      __builtin_memcpy(col + 8, &o1, 16);
      run[tid] = (uint16_t)acc;
    }
    __syncthreads();
    kr[r] = valid ? ((key << 16) | ((uint32_t)(wcount[key * kWaves + wv] & 0x0fffu) + rank_w)) : 0u;
  }
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int64_t p = (int64_t)blk * kOrderBlock + r * kThreads + tid;
    if (p < P) keyrank[p] = kr[r];
  }
  if (job.fused) {
    // block-major 16-bit counts (at most kOrderBlock per key): the scatter reads whole rows
    reinterpret_cast<uint16_t*>(count)[(int64_t)blk * kKeys + tid] = run[tid];
  } else {
    count[(int64_t)tid * nblk_pad + blk] = run[tid];
  }
  (void)nblk;
}

// inclusive scan over the 64 lanes of a wave
__device__ __forceinline__ uint32_t wave_scan(uint32_t x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// 2. per key: exclusive scan of count[key][0..nblk) in place, total[key]
__device__ __forceinline__ void scan_group(const SortJob& job, int group) {
  uint32_t* __restrict__ count = job.count;
  uint32_t* __restrict__ total = job.total;
  const int nblk = job.nblk;
  const int nblk_pad = job.nblk_pad;
  const int lane = threadIdx.x & 63;
  const int key = group * 4 + (threadIdx.x >> 6);
  if (key >= kKeys) return;
  uint32_t* row = count + (int64_t)key * nblk_pad;
  uint32_t carry = 0u;
  for (int b0 = 0; b0 < nblk; b0 += 64) {
    const int b = b0 + lane;
    const uint32_t c = b < nblk ? row[b] : 0u;
    const uint32_t inc = wave_scan(c, lane);
    if (b < nblk) row[b] = carry + inc - c;
    carry += __shfl(inc, 63, 64);
  }
  if (lane == 0) total[key] = carry;
}

// 3. key bases, then scatter.  fused: count holds the raw per-block counts,
// block-major; otherwise the key-major exclusive prefixes of step 2 and
// total[] the key totals.
__device__ __forceinline__ void scatter_block(const SortJob& job, int blk) {
  const uint32_t* __restrict__ keyrank = job.keyrank;
  const uint32_t* __restrict__ count = job.count;
  const uint32_t* __restrict__ total = job.total;
  const int64_t P = job.P;
  const int nblk_pad = job.nblk_pad;
  const int fused = job.fused;
  int32_t* __restrict__ order = job.order;
  __shared__ uint32_t off[kKeys];
  __shared__ uint32_t wsum[kWaves];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  uint32_t kr[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int64_t p = (int64_t)blk * kOrderBlock + r * kThreads + tid;
    kr[r] = p < P ? keyrank[p] : 0u;
  }
  uint32_t tot = 0u, pre = 0u;
  if (fused) {
    // key tid's count in every block: one 2 KB row per block, read whole by
    // the workgroup (32 rows in flight: the reads are the scatter's latency)
    const uint16_t* col = reinterpret_cast<const uint16_t*>(count) + tid;
    const int nblk = job.nblk;
    int j = 0;
    for (; j + 32 <= nblk; j += 32) {
      uint32_t c[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) c[i] = col[(int64_t)(j + i) * kKeys];  // widened
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        tot += c[i];
        pre += (j + i < blk) ? c[i] : 0u;
      }
    }
    for (; j < nblk; ++j) {
      const uint32_t c = col[(int64_t)j * kKeys];  // widened
      tot += c;
      pre += (j < blk) ? c : 0u;
    }
  } else {
    tot = total[tid];
    pre = count[(int64_t)tid * nblk_pad + blk];
  }
  const uint32_t inc = wave_scan(tot, lane);
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  uint32_t o = 0u;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) o += (w < wv) ? wsum[w] : 0u;  // all 16 reads issued at once
  off[tid] = o + inc - tot + pre;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int64_t p = (int64_t)blk * kOrderBlock + r * kThreads + tid;
    if (p < P) order[off[kr[r] >> 16] + (kr[r] & 0xffffu)] = (int32_t)p;
  }
}

// One point set: the job by value, one workgroup per block (per 4 keys for
// the scan).
__global__ __launch_bounds__(kThreads) void order_rank_kernel(SortJob job) {
  rank_block(job, blockIdx.x);
}
__global__ __launch_bounds__(256) void order_scan_kernel(SortJob job) { scan_group(job, blockIdx.x); }
__global__ __launch_bounds__(kThreads) void order_scatter_kernel(SortJob job) {
  scatter_block(job, blockIdx.x);
}

// Many point sets in one launch (a batch): work item b belongs to the last
// job whose first item (blk0, or scan0 for the scan) is at most b.  The job
// table and the item totals live in device memory, so a fixed grid of
// persistent workgroups serves any batch (a captured graph replays for the
// next batch prepared into the same workspace).
template <bool SCAN>
__device__ __forceinline__ int find_job(const SortJob* jobs, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int first = SCAN ? jobs[mid].scan0 : jobs[mid].blk0;
    if (first <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kThreads) void order_rank_jobs_kernel(const SortJobs* const* tabp) {
  const SortJobs* tab = *tabp;
  const int njobs = tab->njobs, nitems = tab->nblocks;
  for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
    const SortJob job = tab->jobs[find_job<false>(tab->jobs, njobs, b)];
    __syncthreads();  // the previous item's LDS is free
    rank_block(job, b - job.blk0);
  }
}
__global__ __launch_bounds__(256) void order_scan_jobs_kernel(const SortJobs* const* tabp) {
  const SortJobs* tab = *tabp;
  const int njobs = tab->njobs, nitems = tab->nscan;
  for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
    const SortJob job = tab->jobs[find_job<true>(tab->jobs, njobs, b)];
    scan_group(job, b - job.scan0);
  }
}
__global__ __launch_bounds__(kThreads) void order_scatter_jobs_kernel(const SortJobs* const* tabp) {
  const SortJobs* tab = *tabp;
  const int njobs = tab->njobs, nitems = tab->nblocks;
  for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
    const SortJob job = tab->jobs[find_job<false>(tab->jobs, njobs, b)];
    __syncthreads();
    scatter_block(job, b - job.blk0);
  }
}

}  // namespace

OrderLayout order_layout(int64_t P) {
  OrderLayout L{0, 0, 0, 0, 0};
  if (P <= 0) return L;
  const int64_t nblk = (P + kOrderBlock - 1) / kOrderBlock;
  const int64_t nblk_pad = (nblk + 3) & ~int64_t(3);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  L.order_off = 0;
  L.keyrank_off = al(size_t(P) * 4);
  L.count_off = L.keyrank_off + al(size_t(P) * 4);
  L.total_off = L.count_off + al(size_t(kKeys) * size_t(nblk_pad) * 4);
  L.total = L.total_off + al(size_t(kKeys) * 4);
  return L;
}

SortJob sort_job(const float* normal, int64_t P, void* ws) {
  SortJob job{};
  if (P <= 0) return job;
  const OrderLayout L = order_layout(P);
  char* base = static_cast<char*>(ws);
  job.normal = normal;
  job.P = P;
  job.order = reinterpret_cast<int32_t*>(base + L.order_off);
  job.keyrank = reinterpret_cast<uint32_t*>(base + L.keyrank_off);
  job.count = reinterpret_cast<uint32_t*>(base + L.count_off);
  job.total = reinterpret_cast<uint32_t*>(base + L.total_off);
  job.nblk = (int)((P + kOrderBlock - 1) / kOrderBlock);
  job.nblk_pad = (job.nblk + 3) & ~3;
  job.fused = job.nblk <= kFusedBlocks ? 1 : 0;
  return job;
}

int sort_scan_groups(const SortJob& job) { return job.fused ? 0 : (kKeys + 3) / 4; }

hipError_t launch_point_order(const float* normal, int64_t P, void* ws, hipStream_t stream) {
  if (P <= 0) return hipSuccess;
  const SortJob job = sort_job(normal, P, ws);
  hipLaunchKernelGGL(order_rank_kernel, dim3(job.nblk), dim3(kThreads), 0, stream, job);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (!job.fused) {
    hipLaunchKernelGGL(order_scan_kernel, dim3(sort_scan_groups(job)), dim3(256), 0, stream, job);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(order_scatter_kernel, dim3(job.nblk), dim3(kThreads), 0, stream, job);
  return hipGetLastError();
}

hipError_t launch_point_order_jobs(const SortJobs* const* tabp, int max_groups,
                                   hipStream_t stream) {
  hipLaunchKernelGGL(order_rank_jobs_kernel, dim3(max_groups), dim3(kThreads), 0, stream, tabp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(order_scan_jobs_kernel, dim3(4 * max_groups), dim3(256), 0, stream, tabp);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(order_scatter_jobs_kernel, dim3(max_groups), dim3(kThreads), 0, stream, tabp);
  return hipGetLastError();
}

}  // namespace des
