/* mesh_oracle.c -- FP64 restatement of the reference's mesh utilities
 * (zjumsj/diff_envshade etc.cu), TEST INFRASTRUCTURE ONLY: loaded by tests/ and
 * bench.py's CPU baseline through oracle/oracle.py, never by the product.
 *
 *   closest point on a triangle and its region rule   etc.cu:302-393
 *   the VJP of its (s, t) w.r.t. the query point       etc.cu:205-300
 *   get_uv_of_triangle forward / backward              etc.cu:707-777, 1180-1275
 *   get_nearest_mesh_points forward / backward         etc.cu:779-920, 1277-1365
 *   extract_bitfield forward                           etc.cu:1021-1071
 *
 * Deliberate differences from the reference, kept identical in the HIP path:
 *   - the nearest-point backward reads its upstream gradient from
 *     grad_barycentric_coord (the reference reads its own output buffer,
 *     etc.cu:884-888);
 *   - candidate faces are not capped at 128 and a point with no valid face gets
 *     triangle id -1 and zero barycentrics (the reference traps in both cases);
 *   - indices out of range (vertex, face, adjacency offsets) are skipped instead
 *     of read.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double x, y, z; } m3;

/* The FP64 evaluation is a fixed sequence of IEEE operations -- adds,
 * multiplies, fused multiply-adds where written as fma(), one correctly
 * rounded division per face -- with no other contraction (built with
 * -ffp-contract=off), so every result is a function of the inputs alone. */
static inline m3 msub(m3 a, m3 b) { m3 c = {a.x - b.x, a.y - b.y, a.z - b.z}; return c; }
static inline double mdot(m3 a, m3 b) { return fma(a.z, b.z, fma(a.y, b.y, a.x * b.x)); }
/* etc.cu:63-67: a NaN passes through */
static inline double mclamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

/* Region of the closest point (etc.cu:316-354): 0 interior, 1 edge x0-x1 (s),
 * 2 edge x0-x2 (t), 3 edge x1-x2 (s + t = 1). */
static int region(double a, double b, double c, double d, double e, double det, double s,
                  double t) {
  if (s + t < det) {
    if (s < 0.0) {
      if (t < 0.0) return d < 0.0 ? 1 : 2;
      return 2;
    }
    if (t < 0.0) return 1;
    return 0;
  }
  if (s < 0.0) return (c + e > b + d) ? 3 : 2;
  if (t < 0.0) return (a + d > b + e) ? 3 : 1;
  return 3;
}

/* Per face: the quantities of etc.cu:310-320 and the one quotient the region
 * needs, taken as num * (1 / den): interior (S, T) / det, edge x0-x1 -d / a,
 * edge x0-x2 -e / c, edge x1-x2 (c + e - b - d) / (a - 2b + c). */
typedef struct {
  m3 e0, e1;
  double a, b, c, d, e, det, S, T, inv, u;
  int type;
} tri_geo;

static tri_geo geometry(m3 x0, m3 x1, m3 x2, m3 q) {
  tri_geo g;
  g.e0 = msub(x1, x0);
  g.e1 = msub(x2, x0);
  m3 v0 = msub(x0, q);
  g.a = mdot(g.e0, g.e0);
  g.b = mdot(g.e0, g.e1);
  g.c = mdot(g.e1, g.e1);
  g.d = mdot(g.e0, v0);
  g.e = mdot(g.e1, v0);
  g.det = fma(g.a, g.c, -(g.b * g.b));
  g.S = fma(g.b, g.e, -(g.c * g.d));
  g.T = fma(g.b, g.d, -(g.a * g.e));
  g.type = region(g.a, g.b, g.c, g.d, g.e, g.det, g.S, g.T);
  double num, den;
  switch (g.type) {
    case 0: num = g.S; den = g.det; break;
    case 1: num = -g.d; den = g.a; break;
    case 2: num = -g.e; den = g.c; break;
    default: num = ((g.c + g.e) - g.b) - g.d; den = (g.a - 2.0 * g.b) + g.c; break;
  }
  g.inv = 1.0 / den;
  g.u = num * g.inv;
  return g;
}

/* etc.cu:356-392: (s, t) of the closest point x0 + s e0 + t e1. */
static void st_of(const tri_geo* g, double* os, double* ot) {
  switch (g->type) {
    case 0: *os = g->u; *ot = g->T * g->inv; break;
    case 1: *os = mclamp01(g->u); *ot = 0.0; break;
    case 2: *os = 0.0; *ot = mclamp01(g->u); break;
    default: *os = mclamp01(g->u); *ot = 1.0 - *os; break;
  }
}

static void closest_st(m3 x0, m3 x1, m3 x2, m3 q, double* os, double* ot) {
  tri_geo g = geometry(x0, x1, x2, q);
  st_of(&g, os, ot);
}

/* etc.cu:205-300: grad of the query point from grad_s, grad_t (the clamps pass
 * the gradient only where the unclamped value lies in [0, 1]). */
static m3 closest_st_vjp(m3 x0, m3 x1, m3 x2, m3 q, double gs, double gt) {
  tri_geo g = geometry(x0, x1, x2, q);
  m3 r = {0.0, 0.0, 0.0};
  if (g.type == 0) {
    double u = fma(gs, g.c, -(gt * g.b)) * g.inv, w = fma(gt, g.a, -(gs * g.b)) * g.inv;
    r.x = fma(g.e1.x, w, g.e0.x * u);
    r.y = fma(g.e1.y, w, g.e0.y * u);
    r.z = fma(g.e1.z, w, g.e0.z * u);
    return r;
  }
  if (!(g.u >= 0.0 && g.u <= 1.0)) return r;
  if (g.type == 1) {
    double k = gs * g.inv;
    r.x = g.e0.x * k; r.y = g.e0.y * k; r.z = g.e0.z * k;
  } else if (g.type == 2) {
    double k = gt * g.inv;
    r.x = g.e1.x * k; r.y = g.e1.y * k; r.z = g.e1.z * k;
  } else {
    double k = (gs - gt) * g.inv;
    r.x = (g.e0.x - g.e1.x) * k;
    r.y = (g.e0.y - g.e1.y) * k;
    r.z = (g.e0.z - g.e1.z) * k;
  }
  return r;
}

void des_oracle_closest_point(const double* x0, const double* x1, const double* x2,
                              const double* q, double* s, double* t) {
  m3 a = {x0[0], x0[1], x0[2]}, b = {x1[0], x1[1], x1[2]}, c = {x2[0], x2[1], x2[2]};
  m3 p = {q[0], q[1], q[2]};
  closest_st(a, b, c, p, s, t);
}

/* GetUVofTriangle_Forward (etc.cu:707-738): triangles P x 3 x 3, query P x 3,
 * barycentric P x 3 = (1 - s - t, s, t). */
int des_oracle_uv_of_triangle_forward(int64_t P, const double* tri, const double* q,
                                      double* bary) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < P; p++) {
    const double* v = tri + 9 * p;
    m3 x0 = {v[0], v[1], v[2]}, x1 = {v[3], v[4], v[5]}, x2 = {v[6], v[7], v[8]};
    m3 qp = {q[3 * p], q[3 * p + 1], q[3 * p + 2]};
    double s, t;
    closest_st(x0, x1, x2, qp, &s, &t);
    bary[3 * p] = 1.0 - s - t;
    bary[3 * p + 1] = s;
    bary[3 * p + 2] = t;
  }
  return 0;
}

/* GetUVofTriangle_Backward (etc.cu:740-777): grad_s = g1 - g0, grad_t = g2 - g0. */
int des_oracle_uv_of_triangle_backward(int64_t P, const double* tri, const double* q,
                                       const double* gbary, double* gq) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < P; p++) {
    const double* v = tri + 9 * p;
    m3 x0 = {v[0], v[1], v[2]}, x1 = {v[3], v[4], v[5]}, x2 = {v[6], v[7], v[8]};
    m3 qp = {q[3 * p], q[3 * p + 1], q[3 * p + 2]};
    const double* g = gbary + 3 * p;
    m3 r = closest_st_vjp(x0, x1, x2, qp, g[1] - g[0], g[2] - g[0]);
    gq[3 * p] = r.x;
    gq[3 * p + 1] = r.y;
    gq[3 * p + 2] = r.z;
  }
  return 0;
}

static int face_vertices(const int32_t* fvl, int64_t N_face, int N_vertex, int64_t f,
                         const double* vpos, m3* x) {
  if (f < 0 || f >= N_face) return 0;
  for (int j = 0; j < 3; j++) {
    int v = fvl[j * N_face + f];
    if (v < 0 || v >= N_vertex) return 0;
    x[j].x = vpos[v];
    x[j].y = vpos[(int64_t)N_vertex + v];
    x[j].z = vpos[2 * (int64_t)N_vertex + v];
  }
  return 1;
}

/* GetNearestMeshPoints_Forward (etc.cu:779-862).  adjacency_head 2 x N_vertex:
 * [0][v] faces around vertex v, [1][v] their inclusive prefix sum, so v's faces
 * are adjacency_list[head1 - head0 .. head1).  Candidates are visited in the
 * reference's order (vertex k = 0..K-1, then its list), each face once (its
 * first occurrence); the first strictly nearest closest point wins. */
int des_oracle_nearest_forward(int N_vertex, int64_t N_face, int64_t P, int K,
                               const int32_t* adj_head, const int32_t* adj_list, int64_t n_adj,
                               const int32_t* fvl, const double* vpos, const double* q,
                               const int64_t* idxs, int64_t* tri_out, double* bary) {
  int ok = 0;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t p = 0; p < P; p++) {
    m3 qp = {q[p], q[P + p], q[2 * P + p]};
    double best = FLT_MAX, bs = 0.0, bt = 0.0; /* the reference starts from FLT_MAX */
    int64_t bf = -1;
    int64_t seen_cap = 64, nseen = 0;
    int64_t* seen = (int64_t*)malloc(sizeof(int64_t) * seen_cap);
    if (!seen) { ok = -1; continue; }
    for (int k = 0; k < K; k++) {
      int64_t v = idxs[p * K + k];
      if (v < 0 || v >= N_vertex) continue;
      int64_t n = adj_head[v], end = adj_head[(int64_t)N_vertex + v];
      int64_t off = end - n;
      if (n <= 0 || off < 0 || end > n_adj) continue;
      for (int64_t j = off; j < end; j++) {
        int64_t f = adj_list[j];
        int dup = 0;
        for (int64_t i = 0; i < nseen; i++) if (seen[i] == f) { dup = 1; break; }
        if (dup) continue;
        if (nseen == seen_cap) {
          seen_cap *= 2;
          int64_t* grown = (int64_t*)realloc(seen, sizeof(int64_t) * seen_cap);
          if (!grown) { ok = -1; break; }
          seen = grown;
        }
        seen[nseen++] = f;
        m3 x[3];
        if (!face_vertices(fvl, N_face, N_vertex, f, vpos, x)) continue;
        tri_geo g = geometry(x[0], x[1], x[2], qp);
        double s, t;
        st_of(&g, &s, &t);
        m3 cp = {fma(t, g.e1.x, fma(s, g.e0.x, x[0].x)), fma(t, g.e1.y, fma(s, g.e0.y, x[0].y)),
                 fma(t, g.e1.z, fma(s, g.e0.z, x[0].z))};
        m3 df = msub(cp, qp);
        double d2 = mdot(df, df);
        if (d2 < best) {
          best = d2;
          bf = f;
          bs = s;
          bt = t;
        }
      }
    }
    free(seen);
    tri_out[p] = bf;
    bary[p] = bf >= 0 ? 1.0 - bs - bt : 0.0;
    bary[P + p] = bf >= 0 ? bs : 0.0;
    bary[2 * P + p] = bf >= 0 ? bt : 0.0;
  }
  return ok;
}

/* GetNearestMeshPoints_Backward (etc.cu:864-920), upstream gradient from
 * grad_barycentric_coord (3 x P); an invalid triangle gets zero gradient. */
int des_oracle_nearest_backward(int N_vertex, int64_t N_face, int64_t P, const int32_t* fvl,
                                const double* vpos, const double* q, const int64_t* tri,
                                const double* gbary, double* gq) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < P; p++) {
    m3 x[3];
    m3 r = {0.0, 0.0, 0.0};
    if (face_vertices(fvl, N_face, N_vertex, tri[p], vpos, x)) {
      m3 qp = {q[p], q[P + p], q[2 * P + p]};
      double g0 = gbary[p], g1 = gbary[P + p], g2 = gbary[2 * P + p];
      r = closest_st_vjp(x[0], x[1], x[2], qp, g1 - g0, g2 - g0);
    }
    gq[p] = r.x;
    gq[P + p] = r.y;
    gq[2 * P + p] = r.z;
  }
  return 0;
}

/* ExtractBitfield_Forward (etc.cu:1021-1040): out[b][k] = bit k % 32 of word
 * bits[k / 32][i_start + b], for b < i_end - i_start, k < N_bit. */
void des_oracle_extract_bitfield(const uint32_t* bits, int64_t N, int N_bit, int64_t i_start,
                                 int64_t i_end, double* out) {
  for (int64_t b = 0; b < i_end - i_start; b++)
    for (int k = 0; k < N_bit; k++)
      out[b * N_bit + k] = (double)((bits[(int64_t)(k / 32) * N + i_start + b] >> (k % 32)) & 1u);
}
