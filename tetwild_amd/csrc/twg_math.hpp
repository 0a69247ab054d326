// twg_math.hpp -- device (and host) arithmetic of the mesh-optimisation loop.
//
// FP64 contract: every expression below is written so that its evaluation
// order is fixed (no contraction: the library is built with
// -ffp-contract=off), and uses only IEEE +, -, *, / and sqrt, which are
// correctly rounded on gfx950 and on the host.  The results are therefore
// identical bit for bit to a host evaluation of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define TWG_HD __host__ __device__ __forceinline__

namespace twg {

constexpr double kMaxEnergy = 1e50;  // State::MAX_ENERGY, src/tetwild/State.h:29
constexpr double kTwoThirds = 2.0 / 3.0;
constexpr double kOneThird = 1.0 / 3.0;

TWG_HD uint64_t f2u(double x) { return __builtin_bit_cast(uint64_t, x); }
TWG_HD double u2f(uint64_t u) { return __builtin_bit_cast(double, u); }
TWG_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
TWG_HD float u2f32(uint32_t u) { return __builtin_bit_cast(float, u); }

TWG_HD bool is_finite(double x) { return (f2u(x) & 0x7FF0000000000000ull) != 0x7FF0000000000000ull; }
TWG_HD bool is_inf(double x) { return (f2u(x) & 0x7FFFFFFFFFFFFFFFull) == 0x7FF0000000000000ull; }
TWG_HD bool is_nan(double x) { return x != x; }
TWG_HD double fabs_(double x) { return u2f(f2u(x) & 0x7FFFFFFFFFFFFFFFull); }
TWG_HD float fabsf_(float x) { return u2f32(f2u(x) & 0x7FFFFFFFu); }

// x^(-1/3): exponent-bit seed (relative error < 3.5%), then four Newton steps
// y <- y + y (1 - x y^3) / 3.  Max error ~1.6 ulp over the normal range.
TWG_HD double rcbrt(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? __builtin_inf() : __builtin_nan("");
    if (x == __builtin_inf()) return 0.0;
    double scale = 1.0;
    if (x < 0x1p-1000) { x = x * 0x1p300; scale = 0x1p100; }
    else if (x > 0x1p1000) { x = x * 0x1p-300; scale = 0x1p-100; }
    uint64_t i = f2u(x);
    double y = u2f(0x553EF0FF289DD796ull - i / 3u);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double t = y * y;
        t = t * y;
        t = x * t;
        double r = 1.0 - t;
        r = r * kOneThird;
        y = y + y * r;
    }
    return y * scale;
}

// float x^(-1/3) for the mixed path: same scheme, three Newton steps; relative
// error below 4 * 2^-24 for normal x (checked on the host over all exponents).
TWG_HD float rcbrtf(float x)
{
    if (!(x > 0.0f)) return (x == 0.0f) ? __builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return 0.0f;
    float scale = 1.0f;
    if (x < 0x1p-100f) { x = x * 0x1p60f; scale = 0x1p20f; }
    else if (x > 0x1p100f) { x = x * 0x1p-60f; scale = 0x1p-20f; }
    uint32_t i = f2u(x);
    float y = u2f32(0x54A1F1E1u - i / 3u);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float t = y * y;
        t = t * y;
        t = x * t;
        float r = 1.0f - t;
        r = r * (1.0f / 3.0f);
        y = y + y * r;
    }
    return y * scale;
}

struct Tet {
    double p[4][3];
};

// ---- AMIPS, FP64 ------------------------------------------------------
// Quantity of src/tetwild/LocalOperations.cpp:28-291: with d_i = v_i - v0,
// Q = 2 sum|d_i|^2 - |d1+d2+d3|^2/2, V6 = det[d1 d2 d3],
// E = Q (2 V6^2)^(-1/3); gradient/Hessian w.r.t. v0:
// g = E u, u = (2/3) n/V6 - s/Q with s = d1+d2+d3, n = (v2-v1) x (v3-v1);
// H = E (u u^T - (s/Q)(s/Q)^T + (2/3)(n/V6)(n/V6)^T) + 3 (E/Q) I.
struct AmipsCore {
    double E, w, Q, V6, sx, sy, sz;
};

TWG_HD AmipsCore amips_core(const Tet& T)
{
    double d1x = T.p[1][0] - T.p[0][0], d1y = T.p[1][1] - T.p[0][1], d1z = T.p[1][2] - T.p[0][2];
    double d2x = T.p[2][0] - T.p[0][0], d2y = T.p[2][1] - T.p[0][1], d2z = T.p[2][2] - T.p[0][2];
    double d3x = T.p[3][0] - T.p[0][0], d3y = T.p[3][1] - T.p[0][1], d3z = T.p[3][2] - T.p[0][2];
    double cx = d2y * d3z - d2z * d3y;
    double cy = d2z * d3x - d2x * d3z;
    double cz = d2x * d3y - d2y * d3x;
    double V6 = d1x * cx + d1y * cy + d1z * cz;
    double sx = d1x + d2x + d3x;
    double sy = d1y + d2y + d3y;
    double sz = d1z + d2z + d3z;
    double S1 = d1x * d1x + d1y * d1y + d1z * d1z;
    double S2 = d2x * d2x + d2y * d2y + d2z * d2z;
    double S3 = d3x * d3x + d3y * d3y + d3z * d3z;
    double S = S1 + S2 + S3;
    double ss = sx * sx + sy * sy + sz * sz;
    double Q = 2.0 * S - 0.5 * ss;
    double w = rcbrt(2.0 * (V6 * V6));
    AmipsCore c;
    c.E = Q * w;
    c.w = w;
    c.Q = Q;
    c.V6 = V6;
    c.sx = sx;
    c.sy = sy;
    c.sz = sz;
    return c;
}

TWG_HD double amips_energy(const Tet& T) { return amips_core(T).E; }

// g[3]; u, al, be are the intermediate vectors the Hessian needs
TWG_HD void amips_grad(const Tet& T, const AmipsCore& c, double* g, double* u, double* al,
                       double* be)
{
    double e21x = T.p[2][0] - T.p[1][0], e21y = T.p[2][1] - T.p[1][1], e21z = T.p[2][2] - T.p[1][2];
    double e31x = T.p[3][0] - T.p[1][0], e31y = T.p[3][1] - T.p[1][1], e31z = T.p[3][2] - T.p[1][2];
    double nx = e21y * e31z - e21z * e31y;
    double ny = e21z * e31x - e21x * e31z;
    double nz = e21x * e31y - e21y * e31x;
    double rQ = 1.0 / c.Q;
    double rV = 1.0 / c.V6;
    al[0] = c.sx * rQ;
    al[1] = c.sy * rQ;
    al[2] = c.sz * rQ;
    be[0] = nx * rV;
    be[1] = ny * rV;
    be[2] = nz * rV;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u[k] = kTwoThirds * be[k] - al[k];
        g[k] = c.E * u[k];
    }
}

// h[9] row-major, symmetric
TWG_HD void amips_hess(const AmipsCore& c, const double* u, const double* al, const double* be,
                       double* h)
{
    double D = 3.0 * c.w;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = i; j < 3; ++j) {
            double t = u[i] * u[j];
            t = t - al[i] * al[j];
            t = t + kTwoThirds * (be[i] * be[j]);
            double v = c.E * t;
            if (i == j) v = v + D;
            h[i * 3 + j] = v;
            h[j * 3 + i] = v;
        }
    }
}

// ---- AMIPS, FP32 with a forward error bound -----------------------------
// Inputs are the FP64 edge vectors d_i rounded to FP32.  Relative error of the
// returned energy, including that rounding: |E32 - E| <= rel * E with
// rel = 2^-24 (64 + 16 kappa), kappa = perm / |V6| the condition number of
// the determinant (perm = |d1| . (|d2| x |d3|) summed in absolute values).
// Q = 2S - |s|^2/2 >= S/2 bounds its cancellation by a constant, so the bound
// is dominated by the determinant's conditioning.
struct AmipsCore32 {
    float E, w, Q, V6, sx, sy, sz, rel;
};

TWG_HD AmipsCore32 amips_core32(const float d[3][3])
{
    float d1x = d[0][0], d1y = d[0][1], d1z = d[0][2];
    float d2x = d[1][0], d2y = d[1][1], d2z = d[1][2];
    float d3x = d[2][0], d3y = d[2][1], d3z = d[2][2];
    float m1 = d2y * d3z, m2 = d2z * d3y;
    float m3 = d2z * d3x, m4 = d2x * d3z;
    float m5 = d2x * d3y, m6 = d2y * d3x;
    float cx = m1 - m2, cy = m3 - m4, cz = m5 - m6;
    float V6 = d1x * cx + d1y * cy + d1z * cz;
    float perm = fabsf_(d1x) * (fabsf_(m1) + fabsf_(m2)) + fabsf_(d1y) * (fabsf_(m3) + fabsf_(m4)) +
                 fabsf_(d1z) * (fabsf_(m5) + fabsf_(m6));
    float sx = d1x + d2x + d3x, sy = d1y + d2y + d3y, sz = d1z + d2z + d3z;
    float S = (d1x * d1x + d1y * d1y + d1z * d1z) + (d2x * d2x + d2y * d2y + d2z * d2z) +
              (d3x * d3x + d3y * d3y + d3z * d3z);
    float ss = sx * sx + sy * sy + sz * sz;
    float Q = 2.0f * S - 0.5f * ss;
    float w = rcbrtf(2.0f * (V6 * V6));
    AmipsCore32 c;
    c.E = Q * w;
    c.w = w;
    c.Q = Q;
    c.V6 = V6;
    c.sx = sx;
    c.sy = sy;
    c.sz = sz;
    c.rel = 0x1p-24f * (64.0f + 16.0f * (perm / fabsf_(V6)));
    return c;
}

// ---- exact predicates (expansion arithmetic) ---------------------------
// Two-Sum / Two-Product / Grow-Expansion / Scale-Expansion after Shewchuk,
// "Adaptive Precision Floating-Point Arithmetic and Fast Robust Geometric
// Predicates" (1997).  Only reached when the floating-point filter cannot
// decide, so the code favours a small footprint over speed.
TWG_HD void two_sum(double a, double b, double& x, double& y)
{
    double s = a + b;
    double bv = s - a;
    double av = s - bv;
    double br = b - bv;
    double ar = a - av;
    x = s;
    y = ar + br;
}

TWG_HD void two_diff(double a, double b, double& x, double& y)
{
    double s = a - b;
    double bv = a - s;
    double av = s + bv;
    double br = bv - b;
    double ar = a - av;
    x = s;
    y = ar + br;
}

TWG_HD void two_prod(double a, double b, double& x, double& y)
{
    double p = a * b;
    x = p;
    y = __builtin_fma(a, b, -p);
}

TWG_HD int exp_grow(int elen, const double* e, double b, double* h)
{
    double q = b;
    int k = 0;
    for (int i = 0; i < elen; ++i) {
        double s, t;
        two_sum(q, e[i], s, t);
        if (t != 0.0) h[k++] = t;
        q = s;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

// h = e + f; tmp needs elen + flen slots
TWG_HD int exp_add(int elen, const double* e, int flen, const double* f, double* h, double* tmp)
{
    int n = elen;
    for (int i = 0; i < elen; ++i) h[i] = e[i];
    for (int j = 0; j < flen; ++j) {
        int m = exp_grow(n, h, f[j], tmp);
        for (int i = 0; i < m; ++i) h[i] = tmp[i];
        n = m;
    }
    return n;
}

TWG_HD int exp_scale(int elen, const double* e, double b, double* h)
{
    int k = 0;
    double q, t;
    two_prod(e[0], b, q, t);
    if (t != 0.0) h[k++] = t;
    for (int i = 1; i < elen; ++i) {
        double p1, p0, s, u;
        two_prod(e[i], b, p1, p0);
        two_sum(q, p0, s, u);
        if (u != 0.0) h[k++] = u;
        two_sum(p1, s, q, u);
        if (u != 0.0) h[k++] = u;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

// h = e * f; h needs 2*elen*flen slots (also used as scratch), elen*2 <= 32
TWG_HD int exp_mul(int elen, const double* e, int flen, const double* f, double* h)
{
    double part[32], acc[64], tmp[64];
    int n = exp_scale(elen, e, f[0], acc);
    for (int j = 1; j < flen; ++j) {
        int m = exp_scale(elen, e, f[j], part);
        n = exp_add(n, acc, m, part, tmp, h);
        for (int i = 0; i < n; ++i) acc[i] = tmp[i];
    }
    for (int i = 0; i < n; ++i) h[i] = acc[i];
    return n;
}

TWG_HD int exp_diff(double a, double b, double* h)
{
    double x, y;
    two_diff(a, b, x, y);
    if (y != 0.0) {
        h[0] = y;
        h[1] = x;
        return 2;
    }
    h[0] = x;
    return 1;
}

// exact a*d - b*c for <= 2-term inputs (<= 16 terms out)
TWG_HD int exp_minor2(int na, const double* a, int nb, const double* b, int nc, const double* c,
                      int nd, const double* d, double* h)
{
    double p[16], q[16], tmp[32];
    int np_ = exp_mul(na, a, nd, d, p);
    int nq = exp_mul(nb, b, nc, c, q);
    for (int i = 0; i < nq; ++i) q[i] = -q[i];
    return exp_add(np_, p, nq, q, h, tmp);
}

TWG_HD int exp_sign(int n, const double* e)
{
    double top = e[n - 1];
    return (top > 0.0) - (top < 0.0);
}

// sign of det[q-p; r-p; s-p], exact
__host__ __device__ __attribute__((noinline)) inline int orient3d_exact(const double* p,
                                                                        const double* q,
                                                                        const double* r,
                                                                        const double* s)
{
    double a[3][2], b[3][2], c[3][2];
    int na[3], nb[3], nc[3];
    for (int k = 0; k < 3; ++k) {
        na[k] = exp_diff(q[k], p[k], a[k]);
        nb[k] = exp_diff(r[k], p[k], b[k]);
        nc[k] = exp_diff(s[k], p[k], c[k]);
    }
    double m[16], t[64], acc[192], acc2[192], tmp[192];
    // by*cz - bz*cy
    int nm = exp_minor2(nb[1], b[1], nb[2], b[2], nc[1], c[1], nc[2], c[2], m);
    int nacc = exp_mul(nm, m, na[0], a[0], acc);
    // bz*cx - bx*cz
    nm = exp_minor2(nb[2], b[2], nb[0], b[0], nc[2], c[2], nc[0], c[0], m);
    int nt = exp_mul(nm, m, na[1], a[1], t);
    nacc = exp_add(nacc, acc, nt, t, acc2, tmp);
    // bx*cy - by*cx
    nm = exp_minor2(nb[0], b[0], nb[1], b[1], nc[0], c[0], nc[1], c[1], m);
    nt = exp_mul(nm, m, na[2], a[2], t);
    nacc = exp_add(nacc, acc2, nt, t, acc, tmp);
    return exp_sign(nacc, acc);
}

// The floating-point filter of orient3d alone: +1 / -1 when it decides the
// sign of det[q-p; r-p; s-p], kUndecided otherwise.  Kernels that call this
// instead of orient3d keep the exact predicate (and its registers and stack)
// out of their hot loop and finish the undecided cases in a separate pass.
constexpr int kUndecided = 2;
TWG_HD int orient3d_filter(const double* p, const double* q, const double* r, const double* s)
{
    double ax = q[0] - p[0], ay = q[1] - p[1], az = q[2] - p[2];
    double bx = r[0] - p[0], by = r[1] - p[1], bz = r[2] - p[2];
    double cx = s[0] - p[0], cy = s[1] - p[1], cz = s[2] - p[2];
    double m1 = by * cz - bz * cy;
    double m2 = bz * cx - bx * cz;
    double m3 = bx * cy - by * cx;
    double det = ax * m1 + ay * m2 + az * m3;
    double perm = fabs_(ax) * (fabs_(by * cz) + fabs_(bz * cy)) +
                  fabs_(ay) * (fabs_(bz * cx) + fabs_(bx * cz)) +
                  fabs_(az) * (fabs_(bx * cy) + fabs_(by * cx));
    double bound = 0x1p-50 * perm;  // > (7 + 56 u) u perm, u = 2^-53
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return kUndecided;
}

// CGAL::orientation(p, q, r, s) on doubles: +1 POSITIVE, 0, -1
TWG_HD int orient3d(const double* p, const double* q, const double* r, const double* s)
{
    int f = orient3d_filter(p, q, r, s);
    return f != kUndecided ? f : orient3d_exact(p, q, r, s);
}

__host__ __device__ __attribute__((noinline)) inline int minor_zero_exact(const double* p,
                                                                          const double* q,
                                                                          const double* r, int u,
                                                                          int v)
{
    double au[2], av[2], bu[2], bv[2], h[32];
    int nau = exp_diff(q[u], p[u], au), nav = exp_diff(q[v], p[v], av);
    int nbu = exp_diff(r[u], p[u], bu), nbv = exp_diff(r[v], p[v], bv);
    int n = exp_minor2(nau, au, nav, av, nbu, bu, nbv, bv, h);
    return exp_sign(n, h) == 0;
}

TWG_HD bool minor_zero(const double* p, const double* q, const double* r, int u, int v)
{
    double au = q[u] - p[u], av = q[v] - p[v], bu = r[u] - p[u], bv = r[v] - p[v];
    double m = au * bv - av * bu;
    double perm = fabs_(au * bv) + fabs_(av * bu);
    if (fabs_(m) > 0x1p-51 * perm) return false;
    return minor_zero_exact(p, q, r, u, v);
}

// CGAL::collinear(p, q, r) on doubles (Triangle_3f::is_degenerate)
TWG_HD bool collinear3(const double* p, const double* q, const double* r)
{
    return minor_zero(p, q, r, 0, 1) && minor_zero(p, q, r, 1, 2) && minor_zero(p, q, r, 2, 0);
}

// Filter only: true when some 2x2 minor of (q-p) x (r-p) is certainly nonzero,
// i.e. the triangle is certainly not degenerate.  false = only the exact test
// can tell (never for a triangle with a clearly nonzero area).
TWG_HD bool certainly_not_collinear(const double* p, const double* q, const double* r)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int u = k, v = (k + 1) % 3;
        double au = q[u] - p[u], av = q[v] - p[v], bu = r[u] - p[u], bv = r[v] - p[v];
        double m = au * bv - av * bu;
        double perm = fabs_(au * bv) + fabs_(av * bu);
        if (fabs_(m) > 0x1p-51 * perm) return true;
    }
    return false;
}

// XCD-aware block order: blocks are dispatched round-robin over the 8 XCDs of
// an MI355X, so block b runs on XCD b % 8.  Remapping b -> (b % 8) * (G / 8) +
// b / 8 gives every XCD a contiguous range of the work, so neighbouring items
// (which share vertices or tree nodes) meet in the same L2.  Blocks beyond the
// last full round of 8 keep their index.
__device__ __forceinline__ int64_t xcd_block()
{
    const unsigned g = gridDim.x, b = blockIdx.x;
    const unsigned per = g / 8;
    const unsigned k = b / 8;
    if (k >= per) return b;
    return (int64_t)(b % 8) * per + k;
}

}  // namespace twg
