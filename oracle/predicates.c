/*
 * predicates.c -- oracle: exact orientation and collinearity of double
 * coordinates.  TEST INFRASTRUCTURE ONLY.
 *
 * Replaces CGAL::orientation (src/tetwild/LocalOperations.cpp:535-553, 695-773,
 * 862-884) and Triangle_3f::is_degenerate (CGAL::collinear,
 * src/tetwild/LocalOperations.cpp:1048) on Epick points: the result is the
 * sign of an exact determinant of doubles, which is unique, so any exact method
 * returns the same answer.  Method: a floating-point filter with a static
 * relative error bound, then exact expansion arithmetic (Shewchuk, "Adaptive
 * Precision Floating-Point Arithmetic and Fast Robust Geometric Predicates",
 * 1997: Two-Sum, Two-Product, Grow-Expansion, Scale-Expansion), written here
 * from the paper.  Sign convention as CGAL: +1 (POSITIVE) when
 * det[q-p; r-p; s-p] > 0; the unit tet (0,0,0),(1,0,0),(0,1,0),(0,0,1) is +1.
 */
#include "oracle.h"

#include <math.h>

#define EPS 0x1p-53

static inline void two_sum(double a, double b, double* x, double* y)
{
    double s = a + b;
    double bv = s - a;
    double av = s - bv;
    double br = b - bv;
    double ar = a - av;
    *x = s;
    *y = ar + br;
}

static inline void two_diff(double a, double b, double* x, double* y)
{
    double s = a - b;
    double bv = a - s;
    double av = s + bv;
    double br = bv - b;
    double ar = a - av;
    *x = s;
    *y = ar + br;
}

static inline void two_prod(double a, double b, double* x, double* y)
{
    double p = a * b;
    *x = p;
    *y = fma(a, b, -p);
}

/* h = e + b; e nonoverlapping, increasing magnitude; zero components dropped */
static int grow(int elen, const double* e, double b, double* h)
{
    double q = b;
    int k = 0;
    for (int i = 0; i < elen; ++i) {
        double s, t;
        two_sum(q, e[i], &s, &t);
        if (t != 0.0) h[k++] = t;
        q = s;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

/* h = e + f (h must not alias e or f); tmp has room for elen+flen */
static int add(int elen, const double* e, int flen, const double* f, double* h, double* tmp)
{
    int n = elen;
    for (int i = 0; i < elen; ++i) h[i] = e[i];
    for (int j = 0; j < flen; ++j) {
        int m = grow(n, h, f[j], tmp);
        for (int i = 0; i < m; ++i) h[i] = tmp[i];
        n = m;
    }
    return n;
}

/* h = b * e */
static int scale(int elen, const double* e, double b, double* h)
{
    int k = 0;
    double q, t;
    two_prod(e[0], b, &q, &t);
    if (t != 0.0) h[k++] = t;
    for (int i = 1; i < elen; ++i) {
        double p1, p0, s, u;
        two_prod(e[i], b, &p1, &p0);
        two_sum(q, p0, &s, &u);
        if (u != 0.0) h[k++] = u;
        two_sum(p1, s, &q, &u);
        if (u != 0.0) h[k++] = u;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

/* h = e * f; h (also used as scratch) needs room for 2*elen*flen terms */
static int mul(int elen, const double* e, int flen, const double* f, double* h)
{
    double part[64], acc[128], tmp[128];
    int n = scale(elen, e, f[0], acc);
    for (int j = 1; j < flen; ++j) {
        int m = scale(elen, e, f[j], part);
        n = add(n, acc, m, part, tmp, h);
        for (int i = 0; i < n; ++i) acc[i] = tmp[i];
    }
    for (int i = 0; i < n; ++i) h[i] = acc[i];
    return n;
}

static inline int neg(int n, double* e)
{
    for (int i = 0; i < n; ++i) e[i] = -e[i];
    return n;
}

static inline int sign_of(int n, const double* e)
{
    double top = e[n - 1];
    return (top > 0.0) - (top < 0.0);
}

/* difference a - b as a 1- or 2-component expansion */
static inline int diff_exp(double a, double b, double* h)
{
    double x, y;
    two_diff(a, b, &x, &y);
    if (y != 0.0) { h[0] = y; h[1] = x; return 2; }
    h[0] = x;
    return 1;
}

/* exact a*d - b*c with 2-term inputs */
static int minor2(int na, const double* a, int nb, const double* b, int nc, const double* c,
                  int nd, const double* d, double* h)
{
    double p[16], q[16], tmp[32];
    int np_ = mul(na, a, nd, d, p);
    int nq = mul(nb, b, nc, c, q);
    neg(nq, q);
    return add(np_, p, nq, q, h, tmp);
}

static int orient3d_exact(const double* p, const double* q, const double* r, const double* s)
{
    double ax[2], ay[2], az[2], bx[2], by[2], bz[2], cx[2], cy[2], cz[2];
    int nax = diff_exp(q[0], p[0], ax), nay = diff_exp(q[1], p[1], ay), naz = diff_exp(q[2], p[2], az);
    int nbx = diff_exp(r[0], p[0], bx), nby = diff_exp(r[1], p[1], by), nbz = diff_exp(r[2], p[2], bz);
    int ncx = diff_exp(s[0], p[0], cx), ncy = diff_exp(s[1], p[1], cy), ncz = diff_exp(s[2], p[2], cz);
    double m1[16], m2[16], m3[16];
    int n1 = minor2(nby, by, nbz, bz, ncy, cy, ncz, cz, m1); /* by*cz - bz*cy */
    int n2 = minor2(nbz, bz, nbx, bx, ncz, cz, ncx, cx, m2); /* bz*cx - bx*cz */
    int n3 = minor2(nbx, bx, nby, by, ncx, cx, ncy, cy, m3); /* bx*cy - by*cx */
    double t1[64], t2[64], t3[64], s12[128], all[192], tmp[192];
    int k1 = mul(n1, m1, nax, ax, t1);
    int k2 = mul(n2, m2, nay, ay, t2);
    int k3 = mul(n3, m3, naz, az, t3);
    int k12 = add(k1, t1, k2, t2, s12, tmp);
    int k = add(k12, s12, k3, t3, all, tmp);
    return sign_of(k, all);
}

int orc_orient3d(const double* p, const double* q, const double* r, const double* s)
{
    double ax = q[0] - p[0], ay = q[1] - p[1], az = q[2] - p[2];
    double bx = r[0] - p[0], by = r[1] - p[1], bz = r[2] - p[2];
    double cx = s[0] - p[0], cy = s[1] - p[1], cz = s[2] - p[2];
    double m1 = by * cz - bz * cy;
    double m2 = bz * cx - bx * cz;
    double m3 = bx * cy - by * cx;
    double det = ax * m1 + ay * m2 + az * m3;
    double perm = fabs(ax) * (fabs(by * cz) + fabs(bz * cy)) +
                  fabs(ay) * (fabs(bz * cx) + fabs(bx * cz)) +
                  fabs(az) * (fabs(bx * cy) + fabs(by * cx));
    double bound = 8.0 * EPS * perm;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return orient3d_exact(p, q, r, s);
}

/* exact zero test of a_u*b_v - a_v*b_u for the (u,v) plane */
static int minor_is_zero(const double* p, const double* q, const double* r, int u, int v)
{
    double au = q[u] - p[u], av = q[v] - p[v], bu = r[u] - p[u], bv = r[v] - p[v];
    double m = au * bv - av * bu;
    double perm = fabs(au * bv) + fabs(av * bu);
    if (fabs(m) > 4.0 * EPS * perm) return 0;
    double eau[2], eav[2], ebu[2], ebv[2], h[32];
    int nau = diff_exp(q[u], p[u], eau), nav = diff_exp(q[v], p[v], eav);
    int nbu = diff_exp(r[u], p[u], ebu), nbv = diff_exp(r[v], p[v], ebv);
    int n = minor2(nau, eau, nav, eav, nbu, ebu, nbv, ebv, h);
    return sign_of(n, h) == 0;
}

int orc_collinear3(const double* p, const double* q, const double* r)
{
    return minor_is_zero(p, q, r, 0, 1) && minor_is_zero(p, q, r, 1, 2) &&
           minor_is_zero(p, q, r, 2, 0);
}
