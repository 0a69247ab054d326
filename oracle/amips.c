/*
 * amips.c -- oracle: conformal AMIPS energy, gradient and Hessian of one tet
 * w.r.t. its first vertex, and calTetQualities.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows the quantity computed by src/tetwild/LocalOperations.cpp:28-291
 * (comformalAMIPSEnergy_new / Jacobian_new / Hessian_new): with J the map of a
 * regular tet onto (v0..v3), E = ||J||_F^2 / |det J|^(2/3).  The reference
 * evaluates that with code-generated straight-line expressions; this
 * restatement evaluates the same quantity in closed form:
 *
 *   d_i = v_i - v0,  Q = 2 sum|d_i|^2 - |d1+d2+d3|^2 / 2  (= half the sum of
 *   the six squared edge lengths), V6 = det[d1 d2 d3],
 *   E  = Q * (2 V6^2)^(-1/3),
 *   dE/dv0 = E u,  u = a/Q - (2/3) b/V6,  a = dQ/dv0 = -(d1+d2+d3),
 *            b = dV6/dv0 = -(v2-v1) x (v3-v1),
 *   H  = E (u u^T - a a^T/Q^2 + (2/3) b b^T/V6^2) + 3 (E/Q) I.
 *
 * The reference's own functions agree with this to ~1e-13 relative on
 * well-shaped tets (tests/golden/amips_golden.npz holds their outputs).
 * The operation order below is the contract the HIP kernels reproduce bit for
 * bit; the cube root is a deterministic Newton iteration built only from IEEE
 * +,-,*,/ so that host and device give identical bits.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

static const double TWO_THIRDS = 2.0 / 3.0;
static const double ONE_THIRD = 1.0 / 3.0;

static inline uint64_t f2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double u2f(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* x^(-1/3): seed from the exponent bits, then four Newton steps on y^-3 = x.
 * 0 -> +inf, +inf -> 0, negative or NaN -> NaN. */
double orc_rcbrt(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? INFINITY : NAN;
    if (x == INFINITY) return 0.0;
    double scale = 1.0;
    if (x < 0x1p-1000) { x = x * 0x1p300; scale = 0x1p100; }
    else if (x > 0x1p1000) { x = x * 0x1p-300; scale = 0x1p-100; }
    uint64_t i = f2u(x);
    double y = u2f(UINT64_C(0x553EF0FF289DD796) - i / 3u);
    for (int k = 0; k < 4; ++k) {
        double t = y * y;
        t = t * y;
        t = x * t;
        double r = 1.0 - t;
        r = r * ONE_THIRD;
        y = y + y * r;
    }
    return y * scale;
}

typedef struct {
    double E, w, Q, V6;
    double sx, sy, sz; /* d1+d2+d3 */
} amips_core;

static inline void core(const double* T, amips_core* c)
{
    double d1x = T[3] - T[0], d1y = T[4] - T[1], d1z = T[5] - T[2];
    double d2x = T[6] - T[0], d2y = T[7] - T[1], d2z = T[8] - T[2];
    double d3x = T[9] - T[0], d3y = T[10] - T[1], d3z = T[11] - T[2];
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
    double w = orc_rcbrt(2.0 * (V6 * V6));
    c->E = Q * w;
    c->w = w;
    c->Q = Q;
    c->V6 = V6;
    c->sx = sx; c->sy = sy; c->sz = sz;
}

double orc_amips_energy(const double* T)
{
    amips_core c;
    core(T, &c);
    return c.E;
}

/* gradient part; fills u (3) and the scaled vectors al = s/Q, be = n/V6 */
static inline void grad(const double* T, const amips_core* c, double* g, double* u, double* al,
                        double* be)
{
    double e21x = T[6] - T[3], e21y = T[7] - T[4], e21z = T[8] - T[5];
    double e31x = T[9] - T[3], e31y = T[10] - T[4], e31z = T[11] - T[5];
    double nx = e21y * e31z - e21z * e31y;
    double ny = e21z * e31x - e21x * e31z;
    double nz = e21x * e31y - e21y * e31x;
    double rQ = 1.0 / c->Q;
    double rV = 1.0 / c->V6;
    al[0] = c->sx * rQ; al[1] = c->sy * rQ; al[2] = c->sz * rQ;
    be[0] = nx * rV; be[1] = ny * rV; be[2] = nz * rV;
    for (int k = 0; k < 3; ++k) {
        u[k] = TWO_THIRDS * be[k] - al[k];
        g[k] = c->E * u[k];
    }
}

static inline void hess(const amips_core* c, const double* u, const double* al, const double* be,
                        double* h)
{
    double D = 3.0 * c->w;
    for (int i = 0; i < 3; ++i) {
        for (int j = i; j < 3; ++j) {
            double t = u[i] * u[j];
            t = t - al[i] * al[j];
            t = t + TWO_THIRDS * (be[i] * be[j]);
            double v = c->E * t;
            if (i == j) v = v + D;
            h[i * 3 + j] = v;
            h[j * 3 + i] = v;
        }
    }
}

void orc_amips_jacobian(const double* T, double* g3)
{
    amips_core c;
    double u[3], al[3], be[3];
    core(T, &c);
    grad(T, &c, g3, u, al, be);
}

void orc_amips_hessian(const double* T, double* h9)
{
    amips_core c;
    double g[3], u[3], al[3], be[3];
    core(T, &c);
    grad(T, &c, g, u, al, be);
    hess(&c, u, al, be, h9);
}

double orc_amips_ejh(const double* T, double* g3, double* h9)
{
    amips_core c;
    double u[3], al[3], be[3];
    core(T, &c);
    grad(T, &c, g3, u, al, be);
    hess(&c, u, al, be, h9);
    return c.E;
}

static inline void gather(const int32_t* t, const double* x, const double* y, const double* z,
                          double* T)
{
    for (int j = 0; j < 4; ++j) {
        T[3 * j + 0] = x[t[j]];
        T[3 * j + 1] = y[t[j]];
        T[3 * j + 2] = z[t[j]];
    }
}

/* mode 0: E; 1: E+grad; 2: E+grad+Hessian.  G and H are per-tet rows:
 * G[3*i + k], H[9*i + k] with k the row-major component index. */
void orc_amips_batch(const int32_t* tets, int64_t n, const double* x, const double* y,
                     const double* z, int mode, double* E, double* G, double* H)
{
    for (int64_t i = 0; i < n; ++i) {
        double T[12], g[3], h[9];
        gather(tets + 4 * i, x, y, z, T);
        if (mode == 0) {
            E[i] = orc_amips_energy(T);
        } else if (mode == 1) {
            amips_core c;
            double u[3], al[3], be[3];
            core(T, &c);
            grad(T, &c, g, u, al, be);
            E[i] = c.E;
            for (int k = 0; k < 3; ++k) G[3 * i + k] = g[k];
        } else {
            E[i] = orc_amips_ejh(T, g, h);
            for (int k = 0; k < 3; ++k) G[3 * i + k] = g[k];
            for (int k = 0; k < 9; ++k) H[9 * i + k] = h[k];
        }
    }
}

/* calTetQuality_AMIPS, src/tetwild/LocalOperations.cpp:862-884 (the scalar
 * branch of calTetQualities, :768-772): non-POSITIVE orientation -> MAX_ENERGY,
 * else the energy, with inf / nan / <= 0 mapped to MAX_ENERGY. */
void orc_tet_qualities(const int32_t* tets, int64_t n, const double* x, const double* y,
                       const double* z, double* slim_energy)
{
    for (int64_t i = 0; i < n; ++i) {
        double T[12];
        gather(tets + 4 * i, x, y, z, T);
        double e;
        if (orc_orient3d(T, T + 3, T + 6, T + 9) != 1) {
            e = ORC_MAX_ENERGY;
        } else {
            e = orc_amips_energy(T);
            if (isinf(e) || isnan(e)) e = ORC_MAX_ENERGY;
        }
        if (isinf(e) || isnan(e) || e <= 0) e = ORC_MAX_ENERGY;
        slim_energy[i] = e;
    }
}

/* isTetFlip on rounded vertices, src/tetwild/LocalOperations.cpp:535-553 */
void orc_tet_flip(const int32_t* tets, int64_t n, const double* x, const double* y,
                  const double* z, uint8_t* flipped)
{
    for (int64_t i = 0; i < n; ++i) {
        double T[12];
        gather(tets + 4 * i, x, y, z, T);
        flipped[i] = (uint8_t)(orc_orient3d(T, T + 3, T + 6, T + 9) != 1);
    }
}
