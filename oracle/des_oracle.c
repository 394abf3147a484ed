/*
 * des_oracle.c -- FP64 CPU restatement of diff_envshade's brute-force
 * environment-lighting shader (forward and backward, all five variants).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness checker and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product path (diff_envshade_amd) never calls
 * it and has no CPU fallback.
 *
 * Every quantity is computed in IEEE binary64 from the (exactly widened) fp32
 * inputs, in the same straight-line order as the reference kernels, so the
 * FP32 GPU path can be checked against it to ~1e-5 relative.
 *
 * Reference citations ("bs" = bruteforce_shader.cu, all in zjumsj/diff_envshade):
 *   texel geometry          bs:124-147 (getDeltaOmega, getLightDir), config.h:11 row flip
 *   Fresnel / G / D         bs:149-227 (+ their VJPs), constants F0 / thres bs:30-31
 *   normalize VJP           bs:62-71
 *   diffuse kernels         bs:231-530
 *   specular (v1) kernels   bs:532-1038
 *   clamp kernels           bs:1040-1552 (per-point gate n.v > 0)
 *   v2 kernels              bs:1554-2066 (+ diffuse_shading output)
 *   visibility v2 kernels   bs:2068-2683 (barycentric-weighted bitmask)
 *   host wrappers           bs:2687-3453 (zeros for disabled branches)
 *
 * Pinning (what checks this file, see tests/test_oracle_pins.py):
 *   closed forms (sum of solid angles = 4 pi, uniform-env Lambert -> 1,
 *   roughness 1 -> D = 1/pi, normal incidence F = F0), an angle-based
 *   textbook Cook-Torrance evaluation of single texels, one-hot env maps for
 *   row/column orientation and visibility bit order, cross-variant identities,
 *   the adjoint identity for grad_envmap and FP64 central finite differences
 *   for every gradient.  The reference holds no golden vectors (it has no
 *   tests) and cannot be built here (CUDA-only), so nothing is pinned against
 *   its own outputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DES_PI 3.14159265358979323846

/* variant ids shared with the Python wrapper (oracle/oracle.py) */
enum {
    ORC_DIFFUSE = 0,  /* bs:231-530   */
    ORC_SPECULAR = 1, /* bs:532-1038  */
    ORC_CLAMP = 2,    /* bs:1040-1552 */
    ORC_V2 = 3,       /* bs:1554-2066 */
    ORC_VIS = 4       /* bs:2068-2683 */
};

static const double F0 = 0.04;   /* bs:30 */
static const double THRES = (double)1e-9f; /* bs:31 (the fp32 constant, widened) */

typedef struct { double x, y, z; } v3;

static inline double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 add3(v3 a, v3 b) { v3 c = {a.x + b.x, a.y + b.y, a.z + b.z}; return c; }
static inline double saturate(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }
/* bs:49-51 */
static inline double grad_saturate(double x) { return (x >= 0.0 && x <= 1.0) ? 1.0 : 0.0; }

/* bs:53-60 */
static inline v3 normalize3(v3 a) {
    double L = sqrt(dot3(a, a));
    v3 o = {a.x / L, a.y / L, a.z / L};
    return o;
}

/* bs:62-71: (I |v|^2 - v v^T) g / |v|^3 */
static inline v3 grad_normalize3(v3 v, v3 g) {
    double v2 = dot3(v, v);
    double v15 = sqrt(v2) * v2;
    double gx = g.x * (v.y * v.y + v.z * v.z) - g.y * (v.x * v.y) - g.z * (v.x * v.z);
    double gy = -g.x * (v.x * v.y) + g.y * (v.x * v.x + v.z * v.z) - g.z * (v.y * v.z);
    double gz = -g.x * (v.x * v.z) - g.y * (v.y * v.z) + g.z * (v.x * v.x + v.y * v.y);
    v3 o = {gx / v15, gy / v15, gz / v15};
    return o;
}

/* bs:124-131: solid angle of lat-long row iy (difference of sines form). */
double des_oracle_delta_omega(int iy, int W, int H) {
    double dphi = 2.0 * DES_PI / W;
    double dtheta = DES_PI / H;
    double s1 = sin(dtheta * (iy + 1) - DES_PI * 0.5);
    double s0 = sin(dtheta * iy - DES_PI * 0.5);
    return (s1 - s0) * dphi;
}

/* bs:133-147: texel-centre direction, +y up. */
static v3 light_dir(int iy, int ix, int W, int H) {
    double v = ((double)iy + 0.5) / H;
    double u = ((double)ix + 0.5) / W;
    double phi = u * DES_PI * 2.0;
    double theta = (v - 0.5) * DES_PI;
    v3 c = {cos(theta) * cos(phi), sin(theta), cos(theta) * sin(phi)};
    return c;
}

/* Storage texel t -> rotated light direction and solid angle.
 * Row flip from config.h:11 (LEFT_TOP_AS_ORIGIN): iy = H-1-t/W (bs:299-301).
 * Rotation l = M d with M row-major (bs:305-307). */
static void texel(int t, int H, int W, const double* M, v3* l, double* domega) {
    int ix = t % W;
    int iy = H - 1 - t / W;
    v3 d = light_dir(iy, ix, W, H);
    *domega = des_oracle_delta_omega(iy, W, H);
    l->x = M[0] * d.x + M[1] * d.y + M[2] * d.z;
    l->y = M[3] * d.x + M[4] * d.y + M[5] * d.z;
    l->z = M[6] * d.x + M[7] * d.y + M[8] * d.z;
}

void des_oracle_texel_table(int H, int W, const double* M, double* l_out, double* domega_out) {
    for (int t = 0; t < H * W; t++) {
        v3 l; double dw;
        texel(t, H, W, M, &l, &dw);
        l_out[3 * t + 0] = l.x; l_out[3 * t + 1] = l.y; l_out[3 * t + 2] = l.z;
        domega_out[t] = dw;
    }
}

/* bs:149-153 */
static inline double fresnel(double R0, double c) {
    double a = 1.0 - saturate(c);
    double a2 = a * a;
    return (a2 * a2 * a) * (1.0 - R0) + R0;
}
/* bs:155-159 */
static inline double grad_fresnel(double R0, double c, double g) {
    double a = 1.0 - saturate(c);
    double a2 = a * a;
    return 5.0 * (R0 - 1.0) * g * (a2 * a2) * grad_saturate(c);
}
/* bs:161-171 */
static inline double g1_robust(v3 l, v3 v, v3 n, double k) {
    double dvn = dot3(v, n), dln = dot3(l, n);
    dvn = dvn > 0.0 ? dvn : 0.0;
    dln = dln > 0.0 ? dln : 0.0;
    double a = dln / (dln * (1.0 - k) + k);
    return a / (dvn * (1.0 - k) + k);
}
/* bs:173-201; gk_abs accumulates the magnitudes of the two terms added to gk */
static inline void grad_g1_robust(v3 l, v3 v, v3 n, double k, double g,
                                  v3* gl, v3* gv, v3* gn, double* gk, double* gk_abs) {
    double dvn = dot3(v, n), dln = dot3(l, n);
    dvn = dvn > 0.0 ? dvn : 0.0;
    dln = dln > 0.0 ? dln : 0.0;
    double a_det = dln * (1.0 - k) + k;
    double a = dln / a_det;
    double b = dvn * (1.0 - k) + k;
    double grad_a = g / b;
    double grad_b = -g * a / (b * b);
    double gdl = grad_a * k / (a_det * a_det);
    if (dln > 0.0) {
        gl->x += gdl * n.x; gl->y += gdl * n.y; gl->z += gdl * n.z;
        gn->x += gdl * l.x; gn->y += gdl * l.y; gn->z += gdl * l.z;
    }
    const double gk_l = grad_a * dln * (dln - 1.0) / (a_det * a_det);
    *gk += gk_l;
    double gdv = grad_b * (1.0 - k);
    if (dvn > 0.0) {
        gv->x += gdv * n.x; gv->y += gdv * n.y; gv->z += gdv * n.z;
        gn->x += gdv * v.x; gn->y += gdv * v.y; gn->z += gdv * v.z;
    }
    const double gk_v = grad_b * (1.0 - dvn);
    *gk += gk_v;
    *gk_abs += fabs(gk_l) + fabs(gk_v);
}
/* bs:203-210 */
static inline double ndf1(v3 n, v3 h, double alpha2) {
    double dnh = dot3(n, h);
    double t = dnh * dnh * (alpha2 - 1.0) + 1.0;
    double den = DES_PI * t * t;
    return alpha2 / (den > THRES ? den : THRES);
}
/* bs:212-227; ga2_abs accumulates the magnitudes of the terms added to ga2 */
static inline void grad_ndf1(v3 n, v3 h, double alpha2, double g, v3* gn, v3* gh, double* ga2,
                             double* ga2_abs) {
    double dnh = dot3(n, h);
    double t = dnh * dnh * (alpha2 - 1.0) + 1.0;
    double t2 = DES_PI * t * t;
    const double ga2_d = g / (t2 > THRES ? t2 : THRES);
    *ga2 += ga2_d;
    *ga2_abs += fabs(ga2_d);
    if (t2 > THRES) {
        double gt = (-2.0 * alpha2 * g) / (t2 * t);
        *ga2 += dnh * dnh * gt;
        *ga2_abs += fabs(dnh * dnh * gt);
        double gdnh = 2.0 * gt * (alpha2 - 1.0) * dnh;
        gn->x += gdnh * h.x; gn->y += gdnh * h.y; gn->z += gdnh * h.z;
        gh->x += gdnh * n.x; gh->y += gdnh * n.y; gh->z += gdnh * n.z;
    }
}

/* Per-point inputs gathered once (bs:2112-2152). */
typedef struct {
    v3 n, v;
    double a[3], s[3], r, alpha2, k;
    int vert[3];
    double bary[3];
    int culled; /* clamp variant: n.v <= 0 (bs:1111) */
} point_t;

typedef struct {
    int variant, diff, spec;
    int64_t P;
    int H, W, N_face, N_vertex;
    const double *normal, *view_dir, *albedo, *spec_albedo, *rough, *envmap, *M;
    const int32_t* fvl;
    const int64_t* tri;
    const double* bary;
    const uint32_t* vis;
    const v3* tl;        /* rotated texel directions, computed once per call */
    const double* tdw;   /* texel solid angles */
} prob_t;

static void load_point(const prob_t* q, int64_t p, point_t* o) {
    o->n.x = q->normal[3 * p]; o->n.y = q->normal[3 * p + 1]; o->n.z = q->normal[3 * p + 2];
    if (q->view_dir) {
        o->v.x = q->view_dir[3 * p]; o->v.y = q->view_dir[3 * p + 1]; o->v.z = q->view_dir[3 * p + 2];
    } else {
        o->v.x = o->v.y = o->v.z = 0.0;
    }
    for (int c = 0; c < 3; c++) {
        o->a[c] = q->albedo ? q->albedo[3 * p + c] : 0.0;
        o->s[c] = q->spec_albedo ? q->spec_albedo[3 * p + c] : 0.0;
    }
    o->r = q->rough ? q->rough[p] : 0.0;
    /* bs:2126-2129: alpha2 = r^4, k = (r+1)^2/8 */
    double a0 = o->r * o->r;
    o->alpha2 = a0 * a0;
    o->k = (o->r + 1.0) * (o->r + 1.0) / 8.0;
    o->culled = (q->variant == ORC_CLAMP) ? !(dot3(o->n, o->v) > 0.0) : 0;
    if (q->variant == ORC_VIS) {
        int64_t f = q->tri[p];
        for (int j = 0; j < 3; j++) {
            o->vert[j] = q->fvl[(int64_t)j * q->N_face + f];
            o->bary[j] = q->bary[(int64_t)j * q->P + p];
        }
    }
}

/* bs:2180-2190: vis = sum_j c_j * bit(visibility[(t/32)*N_vertex + v_j], t%32) */
static double vis_weight(const prob_t* q, const point_t* pt, int t) {
    if (q->variant != ORC_VIS) return 1.0;
    int64_t slot = t / 32;
    uint32_t bit = 1u << (t % 32);
    double w = 0.0;
    for (int j = 0; j < 3; j++) {
        uint32_t word = q->vis[slot * q->N_vertex + pt->vert[j]];
        w += ((word & bit) ? 1.0 : 0.0) * pt->bary[j];
    }
    return w;
}

static int has_dout(int variant) { return variant == ORC_V2 || variant == ORC_VIS; }
static int diffuse_on(const prob_t* q) {
    /* v2 / visibility variants always shade diffuse (bs:1554 ff); diffuse-only always. */
    if (q->variant == ORC_DIFFUSE || has_dout(q->variant)) return 1;
    return q->diff;
}
static int specular_on(const prob_t* q) { return q->variant == ORC_DIFFUSE ? 0 : q->spec; }

/* Specular lobe value F*G*D/4 for one (point, light) pair (bs:2218-2226). */
static double spec_tmp(v3 l, const point_t* pt) {
    v3 h = normalize3(add3(l, pt->v));
    double ch = dot3(l, h);
    double F = fresnel(F0, ch);
    double G = g1_robust(l, pt->v, pt->n, pt->k);
    double D = ndf1(pt->n, h, pt->alpha2);
    return (F * G * D) / 4.0;
}

static void forward_point(const prob_t* q, int64_t p, double* sh, double* ds) {
    point_t pt;
    load_point(q, p, &pt);
    double acc[3] = {0, 0, 0}, dacc[3] = {0, 0, 0};
    int T = q->H * q->W;
    int don = diffuse_on(q), son = specular_on(q);
    if (!pt.culled) {
        for (int t = 0; t < T; t++) {
            v3 l = q->tl[t];
            double dw = q->tdw[t];
            double vis = vis_weight(q, &pt, t);
            double E[3];
            for (int c = 0; c < 3; c++) E[c] = q->envmap[(int64_t)c * T + t];
            if (don) {
                double cs = dot3(l, pt.n);
                if (cs > 0.0) {
                    for (int c = 0; c < 3; c++) {
                        double w = E[c] * (cs / DES_PI) * dw * vis;
                        dacc[c] += w;
                        acc[c] += w * pt.a[c];
                    }
                }
            }
            if (son) {
                double tmp = spec_tmp(l, &pt);
                if (tmp > 0.0)
                    for (int c = 0; c < 3; c++) acc[c] += E[c] * pt.s[c] * (tmp * dw * vis);
            }
        }
    }
    for (int c = 0; c < 3; c++) {
        sh[3 * p + c] = acc[c];
        if (ds) ds[3 * p + c] = dacc[c];
    }
}

/*
 * Per-point backward (bs:2248-2506 and its v1/clamp/diffuse siblings) plus
 * the per-(point, texel) env-map gradient contribution (bs:2508-2683),
 * accumulated into gE_acc (3 x T).  gra_o, when given, receives the scale of
 * grad_roughness: the magnitudes of the terms the reference adds into the
 * alpha^2 and k gradients (two each per texel, bs:173-227), through the
 * same chain factors as the gradient.  A finite-precision evaluation of
 * those sums of terms of both signs errs in proportion to it.
 */
static void backward_point(const prob_t* q, int64_t p, const double* gsh, const double* gds,
                           double* gn_o, double* gv_o, double* ga_o, double* gs_o, double* gr_o,
                           double* gra_o, double* gE_acc) {
    point_t pt;
    load_point(q, p, &pt);
    int T = q->H * q->W;
    int don = diffuse_on(q), son = specular_on(q), dout = has_dout(q->variant);
    double gs[3], gd[3];
    for (int c = 0; c < 3; c++) {
        gs[c] = gsh[3 * p + c];
        gd[c] = (dout && gds) ? gds[3 * p + c] : 0.0;
    }
    v3 gn = {0, 0, 0}, gv = {0, 0, 0};
    double ga[3] = {0, 0, 0}, gsa[3] = {0, 0, 0};
    double galpha2 = 0.0, gk = 0.0, ga2_abs = 0.0, gk_abs = 0.0;
    const double dr_da2 = 4.0 * pt.r * pt.r * pt.r, dr_dk = (1.0 + pt.r) / 4.0; /* bs:2501-2502 */
    if (!pt.culled) {
        for (int t = 0; t < T; t++) {
            v3 l = q->tl[t];
            double dw = q->tdw[t];
            double vis = vis_weight(q, &pt, t);
            double E[3];
            for (int c = 0; c < 3; c++) E[c] = q->envmap[(int64_t)c * T + t];
            double gE_pt[3] = {0, 0, 0};
            if (don) {
                double cs = dot3(l, pt.n);
                if (cs > 0.0) {
                    double gdot = 0.0;
                    for (int c = 0; c < 3; c++) {
                        double I = E[c] * vis;
                        ga[c] += gs[c] * I * (cs / DES_PI) * dw;
                        /* v2/vis: (gs a + gd) (bs:2432); v1/clamp/diffuse: gs a (bs:816, 421) */
                        double up = gs[c] * pt.a[c] + gd[c];
                        gdot += up * I * dw / DES_PI;
                        gE_pt[c] += up * (cs / DES_PI) * dw * vis;
                    }
                    gn.x += gdot * l.x; gn.y += gdot * l.y; gn.z += gdot * l.z;
                }
            }
            if (son) {
                v3 hb = add3(l, pt.v);
                v3 h = normalize3(hb);
                double ch = dot3(l, h);
                double F = fresnel(F0, ch);
                double G = g1_robust(l, pt.v, pt.n, pt.k);
                double D = ndf1(pt.n, h, pt.alpha2);
                double tmp = (F * G * D) / 4.0;
                if (tmp > 0.0) {
                    double gtmp = 0.0;
                    for (int c = 0; c < 3; c++) {
                        double I = E[c] * vis;
                        gsa[c] += gs[c] * I * (tmp * dw);
                        gtmp += gs[c] * I * pt.s[c] * dw;
                        gE_pt[c] += gs[c] * pt.s[c] * tmp * dw * vis;
                    }
                    double gF = gtmp * (G * D) / 4.0;
                    double gG = gtmp * (F * D) / 4.0;
                    double gD = gtmp * (F * G) / 4.0;
                    double gch = grad_fresnel(F0, ch, gF);
                    v3 gh = {gch * l.x, gch * l.y, gch * l.z};
                    v3 gl = {gch * h.x, gch * h.y, gch * h.z}; /* discarded, as bs:2457 */
                    v3 gvp = {0, 0, 0}, gnp = {0, 0, 0};
                    grad_g1_robust(l, pt.v, pt.n, pt.k, gG, &gl, &gvp, &gnp, &gk, &gk_abs);
                    grad_ndf1(pt.n, h, pt.alpha2, gD, &gnp, &gh, &galpha2, &ga2_abs);
                    gh = grad_normalize3(hb, gh);
                    gvp = add3(gvp, gh);
                    gv = add3(gv, gvp);
                    gn = add3(gn, gnp);
                }
            }
            if (gE_acc)
                for (int c = 0; c < 3; c++) gE_acc[(int64_t)c * T + t] += gE_pt[c];
        }
    }
    gn_o[3 * p] = gn.x; gn_o[3 * p + 1] = gn.y; gn_o[3 * p + 2] = gn.z;
    if (don && ga_o)
        for (int c = 0; c < 3; c++) ga_o[3 * p + c] = ga[c];
    if (son) {
        if (gv_o) { gv_o[3 * p] = gv.x; gv_o[3 * p + 1] = gv.y; gv_o[3 * p + 2] = gv.z; }
        if (gs_o)
            for (int c = 0; c < 3; c++) gs_o[3 * p + c] = gsa[c];
        if (gr_o) gr_o[p] = galpha2 * dr_da2 + gk * dr_dk;
        if (gra_o) gra_o[p] = ga2_abs * dr_da2 + gk_abs * dr_dk;
    }
}

/* Texel table of one call: the geometry depends only on (H, W, M), so it is
 * evaluated once (bs:124-147, 299-307) instead of once per (point, texel). */
static int make_table(prob_t* q, v3** tl, double** tdw) {
    int T = q->H * q->W;
    *tl = (v3*)malloc(sizeof(v3) * (size_t)T);
    *tdw = (double*)malloc(sizeof(double) * (size_t)T);
    if (!*tl || !*tdw) { free(*tl); free(*tdw); return -5; }
    for (int t = 0; t < T; t++) texel(t, q->H, q->W, q->M, &(*tl)[t], &(*tdw)[t]);
    q->tl = *tl;
    q->tdw = *tdw;
    return 0;
}

static int check_problem(const prob_t* q) {
    if (q->P < 0 || q->H <= 0 || q->W <= 0) return -1;
    if (q->variant < ORC_DIFFUSE || q->variant > ORC_VIS) return -2;
    if (!q->normal || !q->envmap || !q->M) return -3;
    if (q->variant == ORC_VIS && (!q->fvl || !q->tri || !q->bary || !q->vis)) return -4;
    return 0;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int des_oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/*
 * Forward: shading (P x 3) and, for the v2 / visibility variants,
 * diffuse_shading (P x 3).  Returns 0 or a negative error code.
 */
int des_oracle_forward(int variant, int enable_diffuse, int enable_specular, int64_t P, int H,
                       int W, int N_face, int N_vertex, const double* normal,
                       const double* view_dir, const double* albedo, const double* spec_albedo,
                       const double* rough, const double* envmap, const double* rotmat,
                       const int32_t* fvl, const int64_t* tri, const double* bary,
                       const uint32_t* vis, double* shading, double* diffuse_shading,
                       int nthreads) {
    prob_t q = {variant, enable_diffuse, enable_specular, P, H, W, N_face, N_vertex,
                normal, view_dir, albedo, spec_albedo, rough, envmap, rotmat, fvl, tri, bary, vis,
                NULL, NULL};
    int rc = check_problem(&q);
    if (rc) return rc;
    v3* tl; double* tdw;
    if ((rc = make_table(&q, &tl, &tdw))) return rc;
    set_threads(nthreads);
    double* ds = has_dout(variant) ? diffuse_shading : NULL;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < P; p++) forward_point(&q, p, shading, ds);
    free(tl);
    free(tdw);
    return 0;
}

/*
 * Backward: grad_normal, grad_view_dir, grad_albedo, grad_specular_albedo
 * (P x 3 each), grad_roughness (P) and grad_envmap (3 x T).  Every output is
 * zero-filled first, so disabled branches read back as zeros (bs:2889-2905,
 * bs:3205-3214).  Per-thread env-map partials are summed in thread order.
 */
int des_oracle_backward_ex(int variant, int enable_diffuse, int enable_specular, int64_t P,
                           int H, int W, int N_face, int N_vertex, const double* grad_shading,
                           const double* grad_diffuse_shading, const double* normal,
                           const double* view_dir, const double* albedo, const double* spec_albedo,
                           const double* rough, const double* envmap, const double* rotmat,
                           const int32_t* fvl, const int64_t* tri, const double* bary,
                           const uint32_t* vis, double* gn, double* gv, double* ga, double* gs,
                           double* gr, double* gE, double* gr_abs, int nthreads) {
    prob_t q = {variant, enable_diffuse, enable_specular, P, H, W, N_face, N_vertex,
                normal, view_dir, albedo, spec_albedo, rough, envmap, rotmat, fvl, tri, bary, vis,
                NULL, NULL};
    int rc = check_problem(&q);
    if (rc) return rc;
    if (!grad_shading || !gn) return -3;
    v3* tl; double* tdw;
    if ((rc = make_table(&q, &tl, &tdw))) return rc;
    set_threads(nthreads);
    int T = H * W;
    memset(gn, 0, sizeof(double) * 3 * (size_t)P);
    if (gv) memset(gv, 0, sizeof(double) * 3 * (size_t)P);
    if (ga) memset(ga, 0, sizeof(double) * 3 * (size_t)P);
    if (gs) memset(gs, 0, sizeof(double) * 3 * (size_t)P);
    if (gr) memset(gr, 0, sizeof(double) * (size_t)P);
    if (gr_abs) memset(gr_abs, 0, sizeof(double) * (size_t)P);
    if (gE) memset(gE, 0, sizeof(double) * 3 * (size_t)T);
    int nt = des_oracle_max_threads();
    double* part = gE ? (double*)calloc((size_t)nt * 3 * T, sizeof(double)) : NULL;
    if (gE && !part) { free(tl); free(tdw); return -5; }
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* acc = part ? part + (size_t)tid * 3 * T : NULL;
#pragma omp for schedule(static)
        for (int64_t p = 0; p < P; p++)
            backward_point(&q, p, grad_shading, grad_diffuse_shading, gn, gv, ga, gs, gr, gr_abs,
                           acc);
    }
    if (gE) {
        for (int i = 0; i < nt; i++)
            for (int64_t j = 0; j < 3 * (int64_t)T; j++) gE[j] += part[(size_t)i * 3 * T + j];
        free(part);
    }
    free(tl);
    free(tdw);
    return 0;
}

/* The backward without the conditioning scale of grad_roughness. */
int des_oracle_backward(int variant, int enable_diffuse, int enable_specular, int64_t P, int H,
                        int W, int N_face, int N_vertex, const double* grad_shading,
                        const double* grad_diffuse_shading, const double* normal,
                        const double* view_dir, const double* albedo, const double* spec_albedo,
                        const double* rough, const double* envmap, const double* rotmat,
                        const int32_t* fvl, const int64_t* tri, const double* bary,
                        const uint32_t* vis, double* gn, double* gv, double* ga, double* gs,
                        double* gr, double* gE, int nthreads) {
    return des_oracle_backward_ex(variant, enable_diffuse, enable_specular, P, H, W, N_face,
                                  N_vertex, grad_shading, grad_diffuse_shading, normal, view_dir,
                                  albedo, spec_albedo, rough, envmap, rotmat, fvl, tri, bary, vis,
                                  gn, gv, ga, gs, gr, gE, NULL, nthreads);
}
