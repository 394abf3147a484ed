// mesh_kernels.hip -- the shader's mesh-side inputs on gfx950: nearest mesh
// points (the producer of nearest_triangle_id / barycentric_coord), the
// per-point closest point on a given triangle, and the visibility bitfield
// decoder.  Semantics follow zjumsj/diff_envshade etc.cu (line numbers below).
//
// The closest-point geometry runs in FP64 with the reference's region rule and
// a fixed sequence of IEEE operations -- adds, multiplies, fused multiply-adds
// where written as fma(), one correctly rounded division -- with no other
// contraction.  A result is then a function of the inputs alone, the same on
// any IEEE machine, so the face chosen among faces at equal distance --
// common, since a query's closest point often lies on an edge or vertex shared
// by several faces -- is reproducible and checkable exactly.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "mesh_internal.h"

#pragma clang fp contract(off)

namespace des {
namespace {

struct D3 {
  double x, y, z;
};

__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return fma(a.z, b.z, fma(a.y, b.y, a.x * b.x)); }
// a NaN passes through, as through the reference's clamp (etc.cu:63-67)
__device__ __forceinline__ double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// Closest point of a triangle to q (etc.cu:302-393): x0 + s e0 + t e1 with
// e0 = x1 - x0, e1 = x2 - x0; type = the region the closest point lies in
// (0 interior, 1 edge x0-x1, 2 edge x0-x2, 3 edge x1-x2).  Each region needs
// one quotient num / den, taken as num * (1 / den): interior (S, T) / det,
// edge x0-x1 -d / a, edge x0-x2 -e / c, edge x1-x2 (c + e - b - d) /
// (a - 2b + c).  Selecting num and den by region first keeps the wave on one
// path: a single reciprocal per face, whatever mix of regions its lanes hit.
struct TriGeo {
  D3 e0, e1;
  double a, b, c, d, e, det, S, T, inv, u;
  int type;
};

__device__ __forceinline__ TriGeo tri_geo(D3 x0, D3 x1, D3 x2, D3 q) {
  TriGeo g;
  g.e0 = dsub(x1, x0);
  g.e1 = dsub(x2, x0);
  const D3 v0 = dsub(x0, q);
  g.a = ddot(g.e0, g.e0);
  g.b = ddot(g.e0, g.e1);
  g.c = ddot(g.e1, g.e1);
  g.d = ddot(g.e0, v0);
  g.e = ddot(g.e1, v0);
  g.det = fma(g.a, g.c, -(g.b * g.b));
  g.S = fma(g.b, g.e, -(g.c * g.d));
  g.T = fma(g.b, g.d, -(g.a * g.e));
  // region rule, etc.cu:316-354
  if (g.S + g.T < g.det) {
    if (g.S < 0.0) g.type = (g.T < 0.0 && g.d < 0.0) ? 1 : 2;
    else g.type = (g.T < 0.0) ? 1 : 0;
  } else {
    if (g.S < 0.0) g.type = (g.c + g.e > g.b + g.d) ? 3 : 2;
    else if (g.T < 0.0) g.type = (g.a + g.d > g.b + g.e) ? 3 : 1;
    else g.type = 3;
  }
  const bool t0 = g.type == 0, t1 = g.type == 1, t2 = g.type == 2;
  const double den = t0 ? g.det : (t1 ? g.a : (t2 ? g.c : (g.a - 2.0 * g.b) + g.c));
  const double num = t0 ? g.S : (t1 ? -g.d : (t2 ? -g.e : ((g.c + g.e) - g.b) - g.d));
  g.inv = 1.0 / den;
  g.u = num * g.inv;
  return g;
}

// etc.cu:356-392
__device__ __forceinline__ void closest_st(const TriGeo& g, double& s, double& t) {
  const double cu = clamp01(g.u);
  s = g.type == 0 ? g.u : (g.type == 2 ? 0.0 : cu);
  t = g.type == 0 ? g.T * g.inv : (g.type == 1 ? 0.0 : (g.type == 2 ? cu : 1.0 - cu));
}

// etc.cu:205-300: gradient of the query point from (grad_s, grad_t); a clamp
// passes the gradient only where its argument lies in [0, 1].
__device__ __forceinline__ D3 closest_st_vjp(const TriGeo& g, double gs, double gt) {
  if (g.type == 0) {
    const double u = fma(gs, g.c, -(gt * g.b)) * g.inv, w = fma(gt, g.a, -(gs * g.b)) * g.inv;
    return {fma(g.e1.x, w, g.e0.x * u), fma(g.e1.y, w, g.e0.y * u), fma(g.e1.z, w, g.e0.z * u)};
  }
  if (!(g.u >= 0.0 && g.u <= 1.0)) return {0.0, 0.0, 0.0};
  if (g.type == 1) {
    const double k = gs * g.inv;
    return {g.e0.x * k, g.e0.y * k, g.e0.z * k};
  }
  if (g.type == 2) {
    const double k = gt * g.inv;
    return {g.e1.x * k, g.e1.y * k, g.e1.z * k};
  }
  const double k = (gs - gt) * g.inv;
  return {(g.e0.x - g.e1.x) * k, (g.e0.y - g.e1.y) * k, (g.e0.z - g.e1.z) * k};
}

// x0 + s e0 + t e1
__device__ __forceinline__ D3 tri_point(D3 x0, const TriGeo& g, double s, double t) {
  return {fma(t, g.e1.x, fma(s, g.e0.x, x0.x)), fma(t, g.e1.y, fma(s, g.e0.y, x0.y)),
          fma(t, g.e1.z, fma(s, g.e0.z, x0.z))};
}

// Vertices of face f (face_vertex_list 3 x N_face, vertex_pos 3 x N_vertex);
// false if the face or one of its vertices is out of range.  The three vertex
// ids are loaded together, then the nine coordinates; offsets are 32-bit
// (N_face, N_vertex < 2^31) from per-plane bases, which the loads take as
// scalar base + vector offset.
__device__ __forceinline__ bool face_vertices(const MeshArgs& m, int f, D3 (&x)[3]) {
  if (f < 0 || f >= m.N_face) return false;
  const uint32_t uf = (uint32_t)f;
  const int32_t* fv1 = m.fvl + m.N_face;
  const int32_t* fv2 = fv1 + m.N_face;
  const int v0 = m.fvl[uf], v1 = fv1[uf], v2 = fv2[uf];
  const uint32_t nv = (uint32_t)m.N_vertex;
  if ((uint32_t)v0 >= nv || (uint32_t)v1 >= nv || (uint32_t)v2 >= nv) return false;
  const float* py = m.vpos + m.N_vertex;
  const float* pz = py + m.N_vertex;
  const uint32_t u0 = (uint32_t)v0, u1 = (uint32_t)v1, u2 = (uint32_t)v2;
  x[0] = {(double)m.vpos[u0], (double)py[u0], (double)pz[u0]};
  x[1] = {(double)m.vpos[u1], (double)py[u1], (double)pz[u1]};
  x[2] = {(double)m.vpos[u2], (double)py[u2], (double)pz[u2]};
  return true;
}

// LDS writes of a wave made visible to the wave's other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kNearestWindow = 64;  // candidate faces staged per query at a time
constexpr int kStageBatch = 8;      // adjacency entries a lane loads at once

// GetNearestMeshPoints_Forward (etc.cu:779-862), G lanes per query point and
// 64 / G points per wave.  The candidate faces -- those around the point's K
// nearest vertices, vertex by vertex in idxs order, each vertex's faces in
// adjacency_list order -- are numbered in that order.  The G lanes first take
// G vertices each: a segmented scan of their face counts places each vertex's
// faces in the numbering, and each lane copies its faces into the point's
// window of kNearestWindow staged entries in LDS.  The lanes then shade the
// window's candidates, lane l taking l, l + G, ...; each lane keeps its
// smallest squared distance (the first, in its own increasing numbering) and
// the group reduces to the smallest distance, the lowest candidate number
// among equals.  That is the face the reference's deduplicating slot scan
// keeps (the first strictly nearest face in first-occurrence order), without
// its limit of 128 distinct faces.  Loop bounds are uniform over the wave.
template <int G>
__global__ __launch_bounds__(256) void nearest_forward_kernel(MeshArgs m) {
  constexpr int QW = 64 / G;  // points per wave
  constexpr int CAP = kNearestWindow;
  __shared__ int32_t staged[4][QW][CAP];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gi = lane / G, gl = lane % G;
  const int64_t p = ((int64_t)blockIdx.x * 4 + wv) * QW + gi;
  const bool live = p < m.P;  // a group past the end loads nothing but joins the shuffles
  int32_t* buf = staged[wv][gi];
  D3 q{0.0, 0.0, 0.0};
  if (live) q = {(double)m.query[p], (double)m.query[m.P + p], (double)m.query[2 * m.P + p]};
  double best = 3.4028234663852886e38;  // FLT_MAX, as the reference starts from
  long long best_i = LLONG_MAX;
  double best_s = 0.0, best_t = 0.0;
  int best_f = -1;
  long long base = 0;  // candidates numbered before this chunk of vertices
  for (int k0 = 0; k0 < m.K; k0 += G) {
    // lane gl: the faces around vertex idxs[p][k0 + gl], adjacency_list[off, off + cnt)
    int cnt = 0, off = 0;  // counts and offsets stay below 2^26, a group's total below 2^30
    const int k = k0 + gl;
    if (live && k < m.K) {
      const int64_t v = m.idxs[p * m.K + k];
      if (v >= 0 && v < m.N_vertex) {
        const int n = m.adj_head[v], end = m.adj_head[(uint32_t)m.N_vertex + (uint32_t)v];
        if (n > 0 && end - n >= 0 && end <= m.n_adj) {
          cnt = n;
          off = end - n;
        }
      }
    }
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
      const int y = __shfl_up(incl, d, G);
      if (gl >= d) incl += y;
    }
    const int start = incl - cnt;
    const int total = __shfl(incl, G - 1, G);
    int wmax = total;  // the wave's largest count bounds its loops
#pragma unroll
    for (int d = G; d < 64; d <<= 1) wmax = max(wmax, __shfl_xor(wmax, d, 64));
    for (int w0 = 0; w0 < wmax; w0 += CAP) {
      // stage this lane's faces numbered within [w0, w0 + CAP)
      const int j0 = max(0, w0 - start), j1 = min(cnt, w0 + CAP - start);
      for (int j = j0; j < j1; j += kStageBatch) {  // a vertex's faces, loads in flight together
        int fj[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) fj[u] = j + u < j1 ? m.adj_list[(uint32_t)(off + j + u)] : 0;
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u)
          if (j + u < j1) buf[start + j + u - w0] = fj[u];
      }
      wave_lds_sync();
      const int nwin = min(CAP, total - w0);  // <= 0 once this point is done
      const int rounds = (min(CAP, wmax - w0) + G - 1) / G;
      for (int r = 0; r < rounds; ++r) {
        const int i = gl + r * G;
        if (i < nwin) {
          const int f = buf[i];
          D3 x[3];
          if (face_vertices(m, f, x)) {
            const TriGeo g = tri_geo(x[0], x[1], x[2], q);
            double s, t;
            closest_st(g, s, t);
            const D3 df = dsub(tri_point(x[0], g, s, t), q);
            const double d2 = ddot(df, df);
            if (d2 < best) {  // NaN never wins, as in the reference
              best = d2;
              best_i = base + w0 + i;
              best_s = s;
              best_t = t;
              best_f = f;
            }
          }
        }
      }
      wave_lds_sync();  // the window is read before the next one is staged
    }
    base += total;
  }
  // group argmin of (distance, candidate number), then the winner's (s, t, face)
  int src = lane;
#pragma unroll
  for (int d = G / 2; d >= 1; d >>= 1) {
    const double ob = __shfl_xor(best, d, 64);
    const long long oi = __shfl_xor(best_i, d, 64);
    const int os = __shfl_xor(src, d, 64);
    if (ob < best || (ob == best && oi < best_i)) {
      best = ob;
      best_i = oi;
      src = os;
    }
  }
  best_s = __shfl(best_s, src, 64);
  best_t = __shfl(best_t, src, 64);
  best_f = __shfl(best_f, src, 64);
  if (gl == 0 && live) {
    m.tri_out[p] = best_f;
    const bool found = best_f >= 0;
    m.bary_out[p] = found ? (float)(1.0 - best_s - best_t) : 0.f;
    m.bary_out[m.P + p] = found ? (float)best_s : 0.f;
    m.bary_out[2 * m.P + p] = found ? (float)best_t : 0.f;
  }
}

// GetNearestMeshPoints_Backward (etc.cu:864-920), the upstream gradient read
// from grad_barycentric_coord (3 x P); an invalid triangle gets zero gradient.
__global__ __launch_bounds__(256) void nearest_backward_kernel(MeshArgs m) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= m.P) return;
  D3 x[3];
  D3 r{0.0, 0.0, 0.0};
  const int64_t f = m.tri_in[p];
  if (f >= 0 && f < m.N_face && face_vertices(m, (int)f, x)) {
    const D3 q{(double)m.query[p], (double)m.query[m.P + p], (double)m.query[2 * m.P + p]};
    const double g0 = m.gbary[p], g1 = m.gbary[m.P + p], g2 = m.gbary[2 * m.P + p];
    r = closest_st_vjp(tri_geo(x[0], x[1], x[2], q), g1 - g0, g2 - g0);
  }
  m.gquery[p] = (float)r.x;
  m.gquery[m.P + p] = (float)r.y;
  m.gquery[2 * m.P + p] = (float)r.z;
}

// GetUVofTriangle_Forward / _Backward (etc.cu:707-777): the triangle of each
// point given directly, triangles P x 3 x 3, query P x 3, barycentric P x 3.
__device__ __forceinline__ void load_tri(const float* tri, int64_t p, D3 (&x)[3]) {
  const float* v = tri + 9 * p;
#pragma unroll
  for (int j = 0; j < 3; ++j) x[j] = {(double)v[3 * j], (double)v[3 * j + 1], (double)v[3 * j + 2]};
}

__global__ __launch_bounds__(256) void uv_forward_kernel(int64_t P, const float* __restrict__ tri,
                                                         const float* __restrict__ query,
                                                         float* __restrict__ bary) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  D3 x[3];
  load_tri(tri, p, x);
  const D3 q{(double)query[3 * p], (double)query[3 * p + 1], (double)query[3 * p + 2]};
  double s, t;
  closest_st(tri_geo(x[0], x[1], x[2], q), s, t);
  bary[3 * p] = (float)(1.0 - s - t);
  bary[3 * p + 1] = (float)s;
  bary[3 * p + 2] = (float)t;
}

__global__ __launch_bounds__(256) void uv_backward_kernel(int64_t P, const float* __restrict__ tri,
                                                          const float* __restrict__ query,
                                                          const float* __restrict__ gbary,
                                                          float* __restrict__ gquery) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  D3 x[3];
  load_tri(tri, p, x);
  const D3 q{(double)query[3 * p], (double)query[3 * p + 1], (double)query[3 * p + 2]};
  const double g0 = gbary[3 * p], g1 = gbary[3 * p + 1], g2 = gbary[3 * p + 2];
  const D3 r = closest_st_vjp(tri_geo(x[0], x[1], x[2], q), g1 - g0, g2 - g0);
  gquery[3 * p] = (float)r.x;
  gquery[3 * p + 1] = (float)r.y;
  gquery[3 * p + 2] = (float)r.z;
}

// ExtractBitfield_Forward (etc.cu:1021-1040): out[b][k] = bit k % 32 of word
// [k / 32][i_start + b].  A block takes 64 points: it stages their words in
// LDS, up to kBitWords per point at a time, reading each word slot as one
// coalesced row, then its waves write whole output rows (the 64 rows are one
// contiguous stretch of the output), a lane writing 4 consecutive floats per
// store where rows are 16-byte aligned, else 1.
constexpr int kBitPoints = 64, kBitWords = 16;

__global__ __launch_bounds__(256) void bitfield_kernel(const uint32_t* __restrict__ bits,
                                                       int64_t N, int N_bit, int64_t i_start,
                                                       int64_t nb, float* __restrict__ out,
                                                       bool vec) {
  __shared__ uint32_t words[kBitWords][kBitPoints + 1];  // padded: the rows' reads hit distinct banks
  const int nw = (N_bit + 31) >> 5;
  const int64_t b0 = (int64_t)blockIdx.x * kBitPoints;
  const int npts = (int)min((int64_t)kBitPoints, nb - b0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int w0 = 0; w0 < nw; w0 += kBitWords) {
    const int cw = min(kBitWords, nw - w0);
    for (int e = threadIdx.x; e < cw * kBitPoints; e += 256) {
      const int sl = e / kBitPoints, j = e % kBitPoints;
      if (j < npts) words[sl][j] = bits[(int64_t)(w0 + sl) * N + i_start + b0 + j];
    }
    __syncthreads();
    const int c0 = 32 * w0;
    const int nc = min(32 * cw, N_bit - c0);  // output columns of this word chunk
    for (int r = wv; r < npts; r += 4) {
      float* o = out + (b0 + r) * N_bit + c0;
      if (vec) {
        for (int c4 = lane; c4 < nc / 4; c4 += 64) {
          const uint32_t w = words[c4 >> 3][r] >> ((4 * c4) & 31);
          float4 v;
          v.x = (float)(w & 1u);
          v.y = (float)((w >> 1) & 1u);
          v.z = (float)((w >> 2) & 1u);
          v.w = (float)((w >> 3) & 1u);
          *reinterpret_cast<float4*>(o + 4 * c4) = v;
        }
      } else {
        for (int c = lane; c < nc; c += 64) o[c] = (float)((words[c >> 5][r] >> (c & 31)) & 1u);
      }
    }
    __syncthreads();  // the rows are written before the next chunk is staged
  }
}

}  // namespace

// Lanes per query point: enough that one pass over the staged window keeps
// each lane's loop short, few enough that a wave holds several points (their
// dependent gathers then overlap).
int nearest_group(int K) {
  if (K <= 2) return 4;
  if (K <= 4) return 8;
  return 16;
}

hipError_t launch_nearest_forward(const MeshArgs& m, hipStream_t stream) {
  if (m.P <= 0) return hipSuccess;
  const int G = nearest_group(m.K);
  const int64_t per_block = 4 * (64 / G);
  const dim3 grid((unsigned)((m.P + per_block - 1) / per_block));
  switch (G) {
    case 4: hipLaunchKernelGGL(nearest_forward_kernel<4>, grid, dim3(256), 0, stream, m); break;
    case 8: hipLaunchKernelGGL(nearest_forward_kernel<8>, grid, dim3(256), 0, stream, m); break;
    default: hipLaunchKernelGGL(nearest_forward_kernel<16>, grid, dim3(256), 0, stream, m); break;
  }
  return hipGetLastError();
}

hipError_t launch_nearest_backward(const MeshArgs& m, hipStream_t stream) {
  if (m.P <= 0) return hipSuccess;
  hipLaunchKernelGGL(nearest_backward_kernel, dim3((unsigned)((m.P + 255) / 256)), dim3(256), 0,
                     stream, m);
  return hipGetLastError();
}

hipError_t launch_uv_forward(int64_t P, const float* tri, const float* query, float* bary,
                             hipStream_t stream) {
  if (P <= 0) return hipSuccess;
  hipLaunchKernelGGL(uv_forward_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream, P,
                     tri, query, bary);
  return hipGetLastError();
}

hipError_t launch_uv_backward(int64_t P, const float* tri, const float* query, const float* gbary,
                              float* gquery, hipStream_t stream) {
  if (P <= 0) return hipSuccess;
  hipLaunchKernelGGL(uv_backward_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream,
                     P, tri, query, gbary, gquery);
  return hipGetLastError();
}

hipError_t launch_bitfield(const uint32_t* bits, int64_t N, int N_bit, int64_t i_start, int64_t nb,
                           float* out, hipStream_t stream) {
  if (nb <= 0 || N_bit <= 0) return hipSuccess;
  const bool vec = (N_bit & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  hipLaunchKernelGGL(bitfield_kernel, dim3((unsigned)((nb + kBitPoints - 1) / kBitPoints)),
                     dim3(256), 0, stream, bits, N, N_bit, i_start, nb, out, vec);
  return hipGetLastError();
}

}  // namespace des
