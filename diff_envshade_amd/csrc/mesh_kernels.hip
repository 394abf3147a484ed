// mesh_kernels.hip -- the shader's mesh-side inputs on gfx950: nearest mesh
// points (the producer of nearest_triangle_id / barycentric_coord), the
// per-point closest point on a given triangle, and the visibility bitfield
// decoder.  Semantics follow zjumsj/diff_envshade etc.cu (line numbers below).
//
// The closest-point geometry runs in FP64 with the reference's region rule and
// its order of operations, every operation a correctly rounded IEEE add,
// multiply or divide (no contraction into fused multiply-adds).  A result is
// then a function of the inputs alone, the same on any IEEE machine, so the
// face chosen among faces at equal distance -- common, since a query's closest
// point often lies on an edge or vertex shared by several faces -- is
// reproducible and checkable exactly.  These kernels are bound by their
// gathers, not by the FP64 arithmetic.
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
__device__ __forceinline__ double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// a NaN passes through, as through the reference's clamp
__device__ __forceinline__ double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// Closest point of a triangle to q (etc.cu:302-393): x0 + s e0 + t e1 with
// e0 = x1 - x0, e1 = x2 - x0; type = the region the closest point lies in
// (0 interior, 1 edge x0-x1, 2 edge x0-x2, 3 edge x1-x2).
struct TriGeo {
  D3 e0, e1;
  double a, b, c, d, e, det, s, t;
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
  g.det = g.a * g.c - g.b * g.b;
  g.s = g.b * g.e - g.c * g.d;
  g.t = g.b * g.d - g.a * g.e;
  // region rule, etc.cu:316-354
  if (g.s + g.t < g.det) {
    if (g.s < 0.0) g.type = (g.t < 0.0 && g.d < 0.0) ? 1 : 2;
    else g.type = (g.t < 0.0) ? 1 : 0;
  } else {
    if (g.s < 0.0) g.type = (g.c + g.e > g.b + g.d) ? 3 : 2;
    else if (g.t < 0.0) g.type = (g.a + g.d > g.b + g.e) ? 3 : 1;
    else g.type = 3;
  }
  return g;
}

// etc.cu:356-392
__device__ __forceinline__ void closest_st(const TriGeo& g, double& s, double& t) {
  if (g.type == 0) {
    const double inv = 1.0 / g.det;
    s = g.s * inv;
    t = g.t * inv;
  } else if (g.type == 1) {
    s = clamp01(-g.d / g.a);
    t = 0.0;
  } else if (g.type == 2) {
    s = 0.0;
    t = clamp01(-g.e / g.c);
  } else {
    s = clamp01((g.c + g.e - g.b - g.d) / (g.a - 2.0 * g.b + g.c));
    t = 1.0 - s;
  }
}

// etc.cu:205-300: gradient of the query point from (grad_s, grad_t); a clamp
// passes the gradient only where its argument lies in [0, 1].
__device__ __forceinline__ D3 closest_st_vjp(const TriGeo& g, double gs, double gt) {
  D3 r{0.0, 0.0, 0.0};
  if (g.type == 0) {
    const double inv = 1.0 / g.det;
    const double u = (gs * g.c - gt * g.b) * inv, w = (gt * g.a - gs * g.b) * inv;
    r = {g.e0.x * u + g.e1.x * w, g.e0.y * u + g.e1.y * w, g.e0.z * u + g.e1.z * w};
  } else if (g.type == 1) {
    const double v = -g.d / g.a;
    if (v >= 0.0 && v <= 1.0) r = {g.e0.x * gs / g.a, g.e0.y * gs / g.a, g.e0.z * gs / g.a};
  } else if (g.type == 2) {
    const double v = -g.e / g.c;
    if (v >= 0.0 && v <= 1.0) r = {g.e1.x * gt / g.c, g.e1.y * gt / g.c, g.e1.z * gt / g.c};
  } else {
    const double denom = g.a - 2.0 * g.b + g.c;
    const double v = (g.c + g.e - g.b - g.d) / denom;
    if (v >= 0.0 && v <= 1.0) {
      const double k = (gs - gt) / denom;
      r = {(g.e0.x - g.e1.x) * k, (g.e0.y - g.e1.y) * k, (g.e0.z - g.e1.z) * k};
    }
  }
  return r;
}

// Vertices of face f (face_vertex_list 3 x N_face, vertex_pos 3 x N_vertex);
// false if the face or one of its vertices is out of range.
__device__ __forceinline__ bool face_vertices(const MeshArgs& m, int64_t f, D3 (&x)[3]) {
  if (f < 0 || f >= m.N_face) return false;
  int v[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    v[j] = m.fvl[j * m.N_face + f];
    if (v[j] < 0 || v[j] >= m.N_vertex) return false;
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
    x[j] = {(double)m.vpos[v[j]], (double)m.vpos[(int64_t)m.N_vertex + v[j]],
            (double)m.vpos[2 * (int64_t)m.N_vertex + v[j]]};
  return true;
}

__device__ __forceinline__ long long wave_incl_scan(long long x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const long long y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// GetNearestMeshPoints_Forward (etc.cu:779-862), one wave per query point.
// The candidate faces -- those around the point's K nearest vertices, vertex by
// vertex in idxs order, each vertex's faces in adjacency_list order -- are
// numbered in that order and shaded 64 at a time, one per lane; the wave keeps
// the smallest squared distance and, among equal distances, the lowest
// candidate number.  That is the face the reference's deduplicating slot scan
// keeps (the first strictly nearest face in first-occurrence order), without
// its limit of 128 distinct faces.
__global__ __launch_bounds__(256) void nearest_forward_kernel(MeshArgs m) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= m.P) return;  // whole waves only
  const D3 q{(double)m.query[p], (double)m.query[m.P + p], (double)m.query[2 * m.P + p]};
  double best = 3.4028234663852886e38;  // FLT_MAX, as the reference starts from
  long long best_i = LLONG_MAX;
  double best_s = 0.0, best_t = 0.0;
  int64_t best_f = -1;
  long long base = 0;  // candidates numbered before this chunk of vertices
  for (int k0 = 0; k0 < m.K; k0 += 64) {
    // lane k: the faces around vertex idxs[p][k0 + k], adjacency_list[off, off + cnt)
    long long cnt = 0, off = 0;
    const int k = k0 + lane;
    if (k < m.K) {
      const int64_t v = m.idxs[p * m.K + k];
      if (v >= 0 && v < m.N_vertex) {
        const long long n = m.adj_head[v], end = m.adj_head[(int64_t)m.N_vertex + v];
        if (n > 0 && end - n >= 0 && end <= m.n_adj) {
          cnt = n;
          off = end - n;
        }
      }
    }
    const long long incl = wave_incl_scan(cnt, lane);
    const long long start = incl - cnt;
    const long long total = __shfl(incl, 63, 64);
    const int nk = min(64, m.K - k0);
    for (long long c0 = 0; c0 < total; c0 += 64) {
      const long long i = c0 + lane;  // this lane's candidate within the chunk
      // its vertex: the last lane whose start is <= i (lanes with no faces
      // share the start of the next lane and are passed over)
      int owner = 0;
      for (int j = 1; j < nk; ++j)
        if (__shfl(start, j, 64) <= i) owner = j;
      const long long ostart = __shfl(start, owner, 64), ooff = __shfl(off, owner, 64);
      if (i < total) {
        const int64_t f = m.adj_list[ooff + (i - ostart)];
        D3 x[3];
        if (face_vertices(m, f, x)) {
          const TriGeo g = tri_geo(x[0], x[1], x[2], q);
          double s, t;
          closest_st(g, s, t);
          const D3 cp{x[0].x + s * g.e0.x + t * g.e1.x, x[0].y + s * g.e0.y + t * g.e1.y,
                      x[0].z + s * g.e0.z + t * g.e1.z};
          const D3 df = dsub(cp, q);
          const double d2 = ddot(df, df);
          if (d2 < best) {  // NaN never wins, as in the reference
            best = d2;
            best_i = base + i;
            best_s = s;
            best_t = t;
            best_f = f;
          }
        }
      }
    }
    base += total;
  }
  // wave argmin of (distance, candidate number)
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const double ob = __shfl_xor(best, d, 64);
    const long long oi = __shfl_xor(best_i, d, 64);
    const double os = __shfl_xor(best_s, d, 64), ot = __shfl_xor(best_t, d, 64);
    const long long of = __shfl_xor((long long)best_f, d, 64);
    if (ob < best || (ob == best && oi < best_i)) {
      best = ob;
      best_i = oi;
      best_s = os;
      best_t = ot;
      best_f = of;
    }
  }
  if (lane == 0) {
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
  if (face_vertices(m, m.tri_in[p], x)) {
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
// [k / 32][i_start + b].  One thread per (point, word): a coalesced word load
// and up to 32 contiguous floats written as float4s.
__global__ __launch_bounds__(256) void bitfield_kernel(const uint32_t* __restrict__ bits,
                                                       int64_t N, int N_bit, int64_t i_start,
                                                       int64_t nb, float* __restrict__ out,
                                                       bool vec) {
  const int nw = (N_bit + 31) >> 5;
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= nb * nw) return;
  const int w = (int)(idx / nb);    // word: consecutive threads read consecutive points
  const int64_t b = idx - (int64_t)w * nb;
  const uint32_t word = bits[(int64_t)w * N + i_start + b];
  float* o = out + b * N_bit + 32 * w;
  const int nbits = min(32, N_bit - 32 * w);
  if (nbits == 32 && vec) {  // rows of a multiple of 4 floats on a 16-byte aligned output
#pragma unroll
    for (int j = 0; j < 32; j += 4) {
      float4 v;
      v.x = (float)((word >> j) & 1u);
      v.y = (float)((word >> (j + 1)) & 1u);
      v.z = (float)((word >> (j + 2)) & 1u);
      v.w = (float)((word >> (j + 3)) & 1u);
      *reinterpret_cast<float4*>(o + j) = v;
    }
  } else {
    for (int j = 0; j < nbits; ++j) o[j] = (float)((word >> j) & 1u);
  }
}

}  // namespace

hipError_t launch_nearest_forward(const MeshArgs& m, hipStream_t stream) {
  if (m.P <= 0) return hipSuccess;
  hipLaunchKernelGGL(nearest_forward_kernel, dim3((unsigned)((m.P + 3) / 4)), dim3(256), 0, stream,
                     m);
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
  const int64_t items = nb * ((N_bit + 31) / 32);
  if (items <= 0) return hipSuccess;
  const bool vec = (N_bit & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  hipLaunchKernelGGL(bitfield_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, stream,
                     bits, N, N_bit, i_start, nb, out, vec);
  return hipGetLastError();
}

}  // namespace des
