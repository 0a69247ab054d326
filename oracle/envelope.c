/*
 * envelope.c -- oracle: triangle sampling, point-triangle distance and the
 * envelope decisions.  TEST INFRASTRUCTURE ONLY.
 *
 * sampleTriangle restates src/tetwild/Common.cpp:143-228 operation by
 * operation, with geogram's vec3 semantics for length2 / length / distance /
 * dot / cross / normalize (geogram is not vendored in the reference, so
 * ulp-level agreement with it is unpinned).
 *
 * The point-triangle distance is Eberly's region method ("Distance Between
 * Point and Triangle in 3D", Geometric Tools), the method behind
 * GEO::Geom::point_triangle_squared_distance called from
 * src/tetwild/DistanceQuery.h:20-39 and src/tetwild/geogram/mesh_AABB.cpp:153-171.
 *
 * Envelope decision.  src/tetwild/LocalOperations.cpp:1046-1109 returns "out"
 * iff some sample s has, after the hint + AABB search of
 * src/tetwild/geogram/mesh_AABB.cpp:482-548, sq_dist > eps_2.  That search only
 * skips boxes whose squared distance exceeds eps_2, so the decision is
 *      out(s)  <=>  min over all facets f of d2(s, f) > eps_2,
 * independent of facet order, traversal order and hints.  This file computes
 * exactly that function: a brute-force minimum (orc_*_brute) and an implicit
 * AABB tree over the facets (root 1, children 2n / 2n+1, split at
 * b + (e-b)/2 as in mesh_AABB.cpp:356-379) whose pruning threshold carries a
 * margin above eps_2 large enough that a facet with d2 <= eps_2 is never
 * pruned, so tree and brute force agree bit for bit.
 *
 * orc_nearest restates MeshFacetsAABBWithEps::nearest_facet
 * (mesh_AABB.h:130-141, mesh_AABB.cpp:381-480) including its traversal
 * order, so its result depends on the facet order of the tree.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- geogram-style vec3 helpers --------------------------------------- */
typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 ld(const double* p) { return mk(p[0], p[1], p[2]); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 mulv(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 divv(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double len2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double len(v3 a) { return sqrt(len2(a)); }
static inline double dist(v3 a, v3 b) { return len(sub(a, b)); }
static inline v3 cross(v3 a, v3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 normalize(v3 v)
{
    double s = len(v);
    if (s > 1e-30) s = 1.0 / s;
    return mulv(v, s);
}

/* ---- sampleTriangle ---------------------------------------------------- */
typedef struct {
    double* ps;
    int64_t cap, n;
} sink;

static inline void push(sink* k, v3 p)
{
    if (k->n < k->cap) {
        k->ps[3 * k->n + 0] = p.x;
        k->ps[3 * k->n + 1] = p.y;
        k->ps[3 * k->n + 2] = p.z;
    }
    k->n++;
}

int64_t orc_sample_triangle(const double* a, const double* b, const double* c, double d,
                            double* ps, int64_t cap)
{
    sink k = {ps, cap, 0};
    v3 vs[3] = {ld(a), ld(b), ld(c)};
    double sqrt3_2 = sqrt(3.0) / 2;
    double ls[3];
    for (int i = 0; i < 3; ++i) ls[i] = len2(sub(vs[i], vs[(i + 1) % 3]));
    /* std::minmax_element: first smallest, last largest */
    int max_i = 0;
    for (int i = 1; i < 3; ++i)
        if (!(ls[i] < ls[max_i])) max_i = i;
    double N = sqrt(ls[max_i]) / d;
    if (N <= 1) {
        for (int i = 0; i < 3; ++i) push(&k, vs[i]);
        return k.n;
    }
    if (N == (int)N) N -= 1;

    v3 v0 = vs[max_i], v1 = vs[(max_i + 1) % 3], v2 = vs[(max_i + 2) % 3];
    v3 n_v0v1 = normalize(sub(v1, v0));
    for (int n = 0; n <= N; n++) push(&k, add(v0, mulv(mulv(n_v0v1, d), (double)n)));
    push(&k, v1);

    v3 e01 = sub(v1, v0);
    double h = dist(add(divv(mulv(e01, dot(sub(v2, v0), e01)), ls[max_i]), v0), v2);
    int M = (int)(h / (sqrt3_2 * d));
    if (M < 1) {
        push(&k, v2);
        return k.n;
    }

    v3 n_v0v2 = normalize(sub(v2, v0));
    v3 n_v1v2 = normalize(sub(v2, v1));
    double sin_v0 = len(cross(sub(v2, v0), sub(v1, v0))) / (dist(v0, v2) * dist(v0, v1));
    double tan_v0 = len(cross(sub(v2, v0), sub(v1, v0))) / dot(sub(v2, v0), sub(v1, v0));
    double tan_v1 = len(cross(sub(v2, v1), sub(v0, v1))) / dot(sub(v2, v1), sub(v0, v1));
    double sin_v1 = len(cross(sub(v2, v1), sub(v0, v1))) / (dist(v1, v2) * dist(v0, v1));
    (void)tan_v1;

    for (int m = 1; m <= M; m++) {
        int n = (int)(sqrt3_2 / tan_v0 * m + 0.5);
        int n1 = (int)(sqrt3_2 / tan_v0 * m);
        if (m % 2 == 0 && n == n1) n += 1;
        v3 v0_m = add(v0, mulv(n_v0v2, m * sqrt3_2 * d / sin_v0));
        v3 v1_m = add(v1, mulv(n_v1v2, m * sqrt3_2 * d / sin_v1));
        if (dist(v0_m, v1_m) <= d) break;
        double delta_d = ((n + (m % 2) / 2.0) - m * sqrt3_2 / tan_v0) * d;
        v3 v = add(v0_m, mulv(n_v0v1, delta_d));
        int N1 = (int)(dist(v, v1_m) / d);
        for (int i = 0; i <= N1; i++) push(&k, add(v, mulv(mulv(n_v0v1, (double)i), d)));
    }
    push(&k, v2);

    N = sqrt(ls[(max_i + 1) % 3]) / d;
    if (N > 1) {
        if (N == (int)N) N -= 1;
        v3 nn = normalize(sub(v2, v1));
        for (int n = 1; n <= N; n++) push(&k, add(v1, mulv(mulv(nn, d), (double)n)));
    }
    N = sqrt(ls[(max_i + 2) % 3]) / d;
    if (N > 1) {
        if (N == (int)N) N -= 1;
        v3 nn = normalize(sub(v0, v2));
        for (int n = 1; n <= N; n++) push(&k, add(v2, mulv(mulv(nn, d), (double)n)));
    }
    return k.n;
}

/* ---- point / segment / triangle distances ----------------------------- */
static double point_seg_d2(v3 p, v3 a, v3 b, v3* cp)
{
    double l2 = len2(sub(a, b));
    double t = dot(sub(p, a), sub(b, a));
    if (t <= 0.0 || l2 == 0.0) {
        *cp = a;
        return len2(sub(p, a));
    }
    if (t > l2) {
        *cp = b;
        return len2(sub(p, b));
    }
    double l1 = t / l2;
    double l0 = 1.0 - l1;
    *cp = add(mulv(a, l0), mulv(b, l1));
    return len2(sub(p, *cp));
}

static double tri_d2(v3 p, v3 V0, v3 V1, v3 V2, v3* cp)
{
    v3 diff = sub(V0, p);
    v3 e0 = sub(V1, V0);
    v3 e1 = sub(V2, V0);
    double a00 = len2(e0);
    double a01 = dot(e0, e1);
    double a11 = len2(e1);
    double b0 = dot(diff, e0);
    double b1 = dot(diff, e1);
    double c = len2(diff);
    double det = fabs(a00 * a11 - a01 * a01);
    double s = a01 * b1 - a11 * b0;
    double t = a01 * b0 - a00 * b1;
    double d2;

    if (det < 1e-30) {
        v3 q;
        double best = point_seg_d2(p, V0, V1, cp);
        double cur = point_seg_d2(p, V0, V2, &q);
        if (cur < best) { best = cur; *cp = q; }
        cur = point_seg_d2(p, V1, V2, &q);
        if (cur < best) { best = cur; *cp = q; }
        return best;
    }

    if (s + t <= det) {
        if (s < 0.0) {
            if (t < 0.0) { /* region 4 */
                if (b0 < 0.0) {
                    t = 0.0;
                    if (-b0 >= a00) { s = 1.0; d2 = a00 + 2.0 * b0 + c; }
                    else { s = -b0 / a00; d2 = b0 * s + c; }
                } else {
                    s = 0.0;
                    if (b1 >= 0.0) { t = 0.0; d2 = c; }
                    else if (-b1 >= a11) { t = 1.0; d2 = a11 + 2.0 * b1 + c; }
                    else { t = -b1 / a11; d2 = b1 * t + c; }
                }
            } else { /* region 3 */
                s = 0.0;
                if (b1 >= 0.0) { t = 0.0; d2 = c; }
                else if (-b1 >= a11) { t = 1.0; d2 = a11 + 2.0 * b1 + c; }
                else { t = -b1 / a11; d2 = b1 * t + c; }
            }
        } else if (t < 0.0) { /* region 5 */
            t = 0.0;
            if (b0 >= 0.0) { s = 0.0; d2 = c; }
            else if (-b0 >= a00) { s = 1.0; d2 = a00 + 2.0 * b0 + c; }
            else { s = -b0 / a00; d2 = b0 * s + c; }
        } else { /* region 0: interior */
            double inv = 1.0 / det;
            s *= inv;
            t *= inv;
            d2 = s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c;
        }
    } else {
        if (s < 0.0) { /* region 2 */
            double tmp0 = a01 + b0;
            double tmp1 = a11 + b1;
            if (tmp1 > tmp0) {
                double numer = tmp1 - tmp0;
                double denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) { s = 1.0; t = 0.0; d2 = a00 + 2.0 * b0 + c; }
                else {
                    s = numer / denom;
                    t = 1.0 - s;
                    d2 = s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c;
                }
            } else {
                s = 0.0;
                if (tmp1 <= 0.0) { t = 1.0; d2 = a11 + 2.0 * b1 + c; }
                else if (b1 >= 0.0) { t = 0.0; d2 = c; }
                else { t = -b1 / a11; d2 = b1 * t + c; }
            }
        } else if (t < 0.0) { /* region 6 */
            double tmp0 = a01 + b1;
            double tmp1 = a00 + b0;
            if (tmp1 > tmp0) {
                double numer = tmp1 - tmp0;
                double denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) { t = 1.0; s = 0.0; d2 = a11 + 2.0 * b1 + c; }
                else {
                    t = numer / denom;
                    s = 1.0 - t;
                    d2 = s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c;
                }
            } else {
                t = 0.0;
                if (tmp1 <= 0.0) { s = 1.0; d2 = a00 + 2.0 * b0 + c; }
                else if (b0 >= 0.0) { s = 0.0; d2 = c; }
                else { s = -b0 / a00; d2 = b0 * s + c; }
            }
        } else { /* region 1 */
            double numer = a11 + b1 - a01 - b0;
            if (numer <= 0.0) { s = 0.0; t = 1.0; d2 = a11 + 2.0 * b1 + c; }
            else {
                double denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) { s = 1.0; t = 0.0; d2 = a00 + 2.0 * b0 + c; }
                else {
                    s = numer / denom;
                    t = 1.0 - s;
                    d2 = s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c;
                }
            }
        }
    }
    if (d2 < 0.0) d2 = 0.0;
    *cp = add(add(V0, mulv(e0, s)), mulv(e1, t));
    return d2;
}

double orc_point_tri_d2(const double* p, const double* a, const double* b, const double* c,
                        double* cp)
{
    v3 q;
    double d2 = tri_d2(ld(p), ld(a), ld(b), ld(c), &q);
    if (cp) { cp[0] = q.x; cp[1] = q.y; cp[2] = q.z; }
    return d2;
}

/* ---- surface + implicit AABB tree ------------------------------------- */
struct orc_surface {
    int64_t nf;
    double* tri;   /* nf x 9, facet corners in tree order */
    int64_t* fid;  /* tree position -> caller's facet index */
    double* box;   /* (max_node+1) x 6: min xyz, max xyz */
    int64_t nnode;
    double rmax;   /* longest facet edge */
};

static int64_t max_node(int64_t n, int64_t b, int64_t e)
{
    if (b + 1 == e) return n;
    int64_t m = b + (e - b) / 2;
    int64_t l = max_node(2 * n, b, m), r = max_node(2 * n + 1, m, e);
    return l > r ? l : r;
}

static void build(orc_surface* s, int64_t n, int64_t b, int64_t e)
{
    double* B = s->box + 6 * n;
    if (b + 1 == e) {
        const double* t = s->tri + 9 * b;
        for (int c = 0; c < 3; ++c) {
            B[c] = t[c];
            B[3 + c] = t[c];
        }
        for (int k = 1; k < 3; ++k)
            for (int c = 0; c < 3; ++c) {
                if (t[3 * k + c] < B[c]) B[c] = t[3 * k + c];
                if (t[3 * k + c] > B[3 + c]) B[3 + c] = t[3 * k + c];
            }
        return;
    }
    int64_t m = b + (e - b) / 2;
    build(s, 2 * n, b, m);
    build(s, 2 * n + 1, m, e);
    const double* L = s->box + 6 * (2 * n);
    const double* R = s->box + 6 * (2 * n + 1);
    for (int c = 0; c < 3; ++c) {
        B[c] = L[c] < R[c] ? L[c] : R[c];
        B[3 + c] = L[3 + c] > R[3 + c] ? L[3 + c] : R[3 + c];
    }
}

orc_surface* orc_surface_create(const double* V, int64_t nv, const int32_t* F, int64_t nf)
{
    (void)nv;
    if (nf <= 0) return NULL;
    orc_surface* s = (orc_surface*)calloc(1, sizeof(orc_surface));
    s->nf = nf;
    s->tri = (double*)malloc(sizeof(double) * 9 * nf);
    s->fid = (int64_t*)malloc(sizeof(int64_t) * nf);
    s->rmax = 0.0;
    for (int64_t f = 0; f < nf; ++f) {
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) s->tri[9 * f + 3 * k + c] = V[3 * F[3 * f + k] + c];
        s->fid[f] = f;
        for (int k = 0; k < 3; ++k) {
            double l = dist(ld(s->tri + 9 * f + 3 * k), ld(s->tri + 9 * f + 3 * ((k + 1) % 3)));
            if (l > s->rmax) s->rmax = l;
        }
    }
    s->nnode = max_node(1, 0, nf) + 1;
    s->box = (double*)calloc((size_t)s->nnode * 6, sizeof(double));
    build(s, 1, 0, nf);
    return s;
}

void orc_surface_destroy(orc_surface* s)
{
    if (!s) return;
    free(s->tri);
    free(s->fid);
    free(s->box);
    free(s);
}

/* point_box_signed_squared_distance, mesh_AABB.cpp:180-220 */
static double box_sd2(v3 p, const double* B)
{
    double pc[3] = {p.x, p.y, p.z};
    int inside = 1;
    double r = 0.0;
    for (int c = 0; c < 3; ++c) {
        if (pc[c] < B[c]) {
            inside = 0;
            double d = pc[c] - B[c];
            r += d * d;
        } else if (pc[c] > B[3 + c]) {
            inside = 0;
            double d = pc[c] - B[3 + c];
            r += d * d;
        }
    }
    if (inside) {
        double d = pc[0] - B[0];
        double in = d * d;
        d = pc[0] - B[3];
        if (d * d < in) in = d * d;
        for (int c = 1; c < 3; ++c) {
            d = pc[c] - B[c];
            if (d * d < in) in = d * d;
            d = pc[c] - B[3 + c];
            if (d * d < in) in = d * d;
        }
        r = -in;
    }
    return r;
}

static inline double facet_d2(const orc_surface* s, int64_t f, v3 p, v3* cp)
{
    const double* t = s->tri + 9 * f;
    return tri_d2(p, ld(t), ld(t + 3), ld(t + 6), cp);
}

/* pruning threshold: eps_2 plus a margin that dominates the rounding error of
 * the box and facet distances (a few ulps of (rmax + eps)^2 and of eps_2) */
static inline double prune_threshold(const orc_surface* s, double eps2)
{
    double r = s->rmax + sqrt(eps2);
    return eps2 + eps2 * 0x1p-20 + 0x1p-44 * (r * r);
}

static int in_rec(const orc_surface* s, v3 p, double eps2, double T, int64_t n, int64_t b,
                  int64_t e)
{
    if (b + 1 == e) {
        v3 cp;
        return facet_d2(s, b, p, &cp) <= eps2;
    }
    int64_t m = b + (e - b) / 2;
    if (box_sd2(p, s->box + 6 * (2 * n)) <= T && in_rec(s, p, eps2, T, 2 * n, b, m)) return 1;
    if (box_sd2(p, s->box + 6 * (2 * n + 1)) <= T && in_rec(s, p, eps2, T, 2 * n + 1, m, e))
        return 1;
    return 0;
}

int orc_point_out(const orc_surface* s, const double* p, double eps2)
{
    double T = prune_threshold(s, eps2);
    v3 q = ld(p);
    if (box_sd2(q, s->box + 6) > T) return 1;
    return !in_rec(s, q, eps2, T, 1, 0, s->nf);
}

double orc_min_d2_brute(const orc_surface* s, const double* p)
{
    v3 q = ld(p), cp;
    double best = INFINITY;
    for (int64_t f = 0; f < s->nf; ++f) {
        double d = facet_d2(s, f, q, &cp);
        if (d < best) best = d;
    }
    return best;
}

int orc_point_out_brute(const orc_surface* s, const double* p, double eps2)
{
    v3 q = ld(p), cp;
    for (int64_t f = 0; f < s->nf; ++f)
        if (facet_d2(s, f, q, &cp) <= eps2) return 0;
    return 1;
}

int orc_face_out(const orc_surface* s, const double* tri9, double eps2, double d,
                 int64_t* n_samples)
{
    if (orc_collinear3(tri9, tri9 + 3, tri9 + 6)) {
        if (n_samples) *n_samples = 0;
        return 0;
    }
    {   /* guard shared with the device: longest edge > 65536 d -> out, unsampled */
        v3 a = ld(tri9), b = ld(tri9 + 3), c = ld(tri9 + 6);
        double l = fmax(len2(sub(a, b)), fmax(len2(sub(b, c)), len2(sub(c, a))));
        if (!(sqrt(l) / d <= 65536.0)) {
            if (n_samples) *n_samples = 0;
            return 1;
        }
    }
    double local[3 * 256];
    int64_t n = orc_sample_triangle(tri9, tri9 + 3, tri9 + 6, d, local, 256);
    double* ps = local;
    if (n > 256) {
        ps = (double*)malloc(sizeof(double) * 3 * n);
        orc_sample_triangle(tri9, tri9 + 3, tri9 + 6, d, ps, n);
    }
    if (n_samples) *n_samples = n;
    int out = 0;
    for (int64_t i = 0; i < n && !out; ++i) out = orc_point_out(s, ps + 3 * i, eps2);
    if (ps != local) free(ps);
    return out;
}

void orc_face_out_batch(const orc_surface* s, const double* tri9, int64_t n, double eps2,
                        double d, uint8_t* out, int64_t* total_samples)
{
    int64_t tot = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t k = 0;
        out[i] = (uint8_t)orc_face_out(s, tri9 + 9 * i, eps2, d, &k);
        tot += k;
    }
    if (total_samples) *total_samples = tot;
}

void orc_point_out_batch(const orc_surface* s, const double* p3, int64_t n, double eps2,
                         uint8_t* out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)orc_point_out(s, p3 + 3 * i, eps2);
}

/* ---- nearest_facet (reference traversal order) ------------------------ */
static double center_d2(v3 p, const double* B)
{
    double pc[3] = {p.x, p.y, p.z};
    double r = 0.0;
    for (int c = 0; c < 3; ++c) {
        double d = pc[c] - 0.5 * (B[c] + B[3 + c]);
        r += d * d;
    }
    return r;
}

static void nearest_rec(const orc_surface* s, v3 p, int64_t* nf, v3* np, double* sq, int64_t n,
                        int64_t b, int64_t e)
{
    if (b + 1 == e) {
        v3 cp;
        double d = facet_d2(s, b, p, &cp);
        if (d < *sq) {
            *nf = b;
            *np = cp;
            *sq = d;
        }
        return;
    }
    int64_t m = b + (e - b) / 2;
    double dl = box_sd2(p, s->box + 6 * (2 * n));
    double dr = box_sd2(p, s->box + 6 * (2 * n + 1));
    if (dl < dr) {
        if (dl < *sq) nearest_rec(s, p, nf, np, sq, 2 * n, b, m);
        if (dr < *sq) nearest_rec(s, p, nf, np, sq, 2 * n + 1, m, e);
    } else {
        if (dr < *sq) nearest_rec(s, p, nf, np, sq, 2 * n + 1, m, e);
        if (dl < *sq) nearest_rec(s, p, nf, np, sq, 2 * n, b, m);
    }
}

int64_t orc_nearest(const orc_surface* s, const double* pin, double* cp, double* d2)
{
    v3 p = ld(pin);
    int64_t b = 0, e = s->nf, n = 1;
    while (e != b + 1) {
        int64_t m = b + (e - b) / 2;
        if (center_d2(p, s->box + 6 * (2 * n)) < center_d2(p, s->box + 6 * (2 * n + 1))) {
            e = m;
            n = 2 * n;
        } else {
            b = m;
            n = 2 * n + 1;
        }
    }
    int64_t nf = b;
    v3 np = ld(s->tri + 9 * b);
    v3 dd = sub(np, p);
    double sq = len2(dd);
    nearest_rec(s, p, &nf, &np, &sq, 1, 0, s->nf);
    if (cp) { cp[0] = np.x; cp[1] = np.y; cp[2] = np.z; }
    if (d2) *d2 = sq;
    return nf;
}
