/*
 * smooth.c -- oracle: the VertexSmoother Newton step on one frozen one-ring.
 * TEST INFRASTRUCTURE ONLY.
 *
 * NewtonsMethod  src/tetwild/VertexSmoother.cpp:465-542
 * getNewEnergy   src/tetwild/VertexSmoother.cpp:544-625 (scalar branch)
 * NewtonsUpdate  src/tetwild/VertexSmoother.cpp:627-702 (scalar branch)
 * isFlip         src/tetwild/LocalOperations.cpp:535-573 (rounded vertices)
 *
 * The linear solve H X = H X0 - a J is a Householder QR with column pivoting
 * (Businger-Golub, as in Golub & Van Loan 5.4.1) with the rank rule of Eigen's
 * ColPivHouseholderQR (pivot squared norm below (max column norm * 2^-52)^2 /
 * rows * (rows - k) ends the nonzero pivots; the remaining unknowns are 0).
 * Eigen is not vendored in the reference, so ulp-level agreement with it is
 * unpinned.  Sums over the ring run in ring order, as in the reference.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

#define MAX_STEP 15
#define MAX_IT 20

static inline void coords(const double* x, const double* y, const double* z, int32_t v, double* p)
{
    p[0] = x[v];
    p[1] = y[v];
    p[2] = z[v];
}

/* position override for the smoothed vertex */
typedef struct {
    const double *x, *y, *z;
    int32_t v;
    double p[3];
} view;

static inline void vget(const view* w, int32_t v, double* p)
{
    if (v == w->v) { p[0] = w->p[0]; p[1] = w->p[1]; p[2] = w->p[2]; }
    else coords(w->x, w->y, w->z, v, p);
}

static int ring_flip(const int32_t* rt, int m, const view* w)
{
    for (int i = 0; i < m; ++i) {
        double T[12];
        for (int j = 0; j < 4; ++j) vget(w, rt[4 * i + j], T + 3 * j);
        if (orc_orient3d(T, T + 3, T + 6, T + 9) != 1) return 1;
    }
    return 0;
}

static double ring_energy(const int32_t* rt, int m, const view* w)
{
    double s = 0.0;
    for (int i = 0; i < m; ++i) {
        double T[12];
        for (int j = 0; j < 4; ++j) vget(w, rt[4 * i + j], T + 3 * j);
        s += orc_amips_energy(T);
    }
    if (isinf(s) || isnan(s) || s <= 0 || s > ORC_MAX_ENERGY) s = ORC_MAX_ENERGY;
    return s;
}

static int ring_update(const int32_t* rt, int m, const view* w, double* E, double* J, double* H,
                       double* X0)
{
    double e = 0.0;
    for (int k = 0; k < 3; ++k) { J[k] = 0.0; X0[k] = w->p[k]; }
    for (int k = 0; k < 9; ++k) H[k] = 0.0;
    for (int i = 0; i < m; ++i) {
        const int32_t* t = rt + 4 * i;
        int start = 0;
        for (int j = 0; j < 4; ++j)
            if (t[j] == w->v) { start = j; break; }
        double T[12], g[3], h[9];
        for (int j = 0; j < 4; ++j) vget(w, t[(start + j) % 4], T + 3 * j);
        e += orc_amips_ejh(T, g, h);
        for (int k = 0; k < 3; ++k) J[k] += g[k];
        for (int k = 0; k < 9; ++k) H[k] += h[k];
    }
    if (isinf(e)) e = ORC_MAX_ENERGY;
    *E = e;
    if (isnan(e) || e <= 0) return 0;
    for (int k = 0; k < 3; ++k)
        if (!isfinite(J[k])) return 0;
    for (int k = 0; k < 9; ++k)
        if (!isfinite(H[k])) return 0;
    return 1;
}

int orc_solve3_qrcp(const double* Ain, const double* b, double* x)
{
    double A[3][3];
    int perm[3] = {0, 1, 2};
    double tau[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = Ain[3 * i + j];

    double norm[3];
    double maxn = 0.0;
    for (int j = 0; j < 3; ++j) {
        norm[j] = sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
        if (norm[j] > maxn) maxn = norm[j];
    }
    double thr = (maxn * 0x1p-52) * (maxn * 0x1p-52) / 3.0;
    int rank = 3;
    for (int k = 0; k < 3; ++k) {
        int big = k;
        for (int j = k + 1; j < 3; ++j)
            if (norm[j] > norm[big]) big = j;
        if (rank == 3 && norm[big] * norm[big] < thr * (double)(3 - k)) rank = k;
        if (big != k) {
            for (int i = 0; i < 3; ++i) { double t = A[i][k]; A[i][k] = A[i][big]; A[i][big] = t; }
            double t = norm[k]; norm[k] = norm[big]; norm[big] = t;
            int p = perm[k]; perm[k] = perm[big]; perm[big] = p;
        }
        /* Householder vector for A[k..2][k]: v = (1, ess), H = I - tau v v^T */
        double tail = 0.0;
        for (int i = k + 1; i < 3; ++i) tail += A[i][k] * A[i][k];
        double c0 = A[k][k];
        double beta;
        if (tail <= 0x1p-1022) {
            tau[k] = 0.0;
            beta = c0;
            for (int i = k + 1; i < 3; ++i) A[i][k] = 0.0;
        } else {
            beta = sqrt(c0 * c0 + tail);
            if (c0 >= 0.0) beta = -beta;
            for (int i = k + 1; i < 3; ++i) A[i][k] = A[i][k] / (c0 - beta);
            tau[k] = (beta - c0) / beta;
        }
        A[k][k] = beta;
        /* apply to the remaining columns */
        for (int j = k + 1; j < 3; ++j) {
            if (tau[k] == 0.0) break;
            double w = A[k][j];
            for (int i = k + 1; i < 3; ++i) w += A[i][k] * A[i][j];
            A[k][j] -= tau[k] * w;
            for (int i = k + 1; i < 3; ++i) A[i][j] -= tau[k] * A[i][k] * w;
        }
        /* column norms of the trailing block, recomputed directly */
        for (int j = k + 1; j < 3; ++j) {
            double s = 0.0;
            for (int i = k + 1; i < 3; ++i) s += A[i][j] * A[i][j];
            norm[j] = sqrt(s);
        }
    }
    if (rank == 0) {
        x[0] = x[1] = x[2] = 0.0;
        return 0;
    }
    double c[3] = {b[0], b[1], b[2]};
    for (int k = 0; k < rank; ++k) {
        if (tau[k] == 0.0) continue;
        double w = c[k];
        for (int i = k + 1; i < 3; ++i) w += A[i][k] * c[i];
        c[k] -= tau[k] * w;
        for (int i = k + 1; i < 3; ++i) c[i] -= tau[k] * A[i][k] * w;
    }
    for (int i = rank - 1; i >= 0; --i) {
        double s = c[i];
        for (int j = i + 1; j < rank; ++j) s -= A[i][j] * c[j];
        c[i] = s / A[i][i];
    }
    double out[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < rank; ++i) out[perm[i]] = c[i];
    x[0] = out[0];
    x[1] = out[1];
    x[2] = out[2];
    return rank == 3;
}

int orc_newtons_update(const int32_t* ring_tets, int m, int32_t v_id, const double* x,
                       const double* y, const double* z, double* E, double* J, double* H,
                       double* X0)
{
    view w = {x, y, z, v_id, {x[v_id], y[v_id], z[v_id]}};
    return ring_update(ring_tets, m, &w, E, J, H, X0);
}

double orc_new_energy(const int32_t* ring_tets, int m, const double* x, const double* y,
                      const double* z)
{
    view w = {x, y, z, -1, {0, 0, 0}};
    return ring_energy(ring_tets, m, &w);
}

/* A mesh with exact positions: posf (x, y, z), pos (px, py, pz) and
 * is_rounded (r) per vertex; see orc_smooth_vertex. */
typedef struct {
    double *x, *y, *z, *px, *py, *pz;
    uint8_t* r;
} emesh;

static void eget(const emesh* M, int32_t u, int exact, double* p)
{
    if (exact) { p[0] = M->px[u]; p[1] = M->py[u]; p[2] = M->pz[u]; }
    else { p[0] = M->x[u]; p[1] = M->y[u]; p[2] = M->z[u]; }
}

static void eset(double* a, double* b, double* c, int32_t v, const double* p)
{
    a[v] = p[0];
    b[v] = p[1];
    c[v] = p[2];
}

static void moved_back(const emesh* M, int32_t v, const double* pf, const double* pos)
{
    eset(M->x, M->y, M->z, v, pf);
    eset(M->px, M->py, M->pz, v, pos);
}

static int e_is_flip(const emesh* M, const int32_t* rt, int m)
{
    for (int i = 0; i < m; ++i) {
        const int32_t* t = rt + 4 * i;
        int all = 1;
        for (int j = 0; j < 4; ++j) all = all && M->r[t[j]];
        double T[12];
        for (int j = 0; j < 4; ++j) eget(M, t[j], !all, T + 3 * j);
        if (orc_orient3d(T, T + 3, T + 6, T + 9) != 1) return 1;
    }
    return 0;
}

/* Near-tie accounting for the line search (risk may be NULL).  The reference
 * sums a ring in unordered_set order (src/tetwild/VertexSmoother.cpp:145-150),
 * so its energies may differ from these ring-order sums by up to the
 * recursive-summation bound (m - 1) u sum|E_i| each (u = 2^-53; the terms
 * are positive, so sum|E_i| is the sum itself).  A comparison whose operands
 * lie within the sum of both bounds of a tie is one the reference could
 * decide the other way.
 *   risk[0] energy comparisons new >= old (finite operands)
 *   risk[1] of them within the band: |new - old| <= 2 (m - 1) u (new + old)
 *   risk[2] |dE| < 1e-5 tests of an accepted step
 *   risk[3] of them within the band: ||dE| - 1e-5| <= 2 (m - 1) u (new + old) */
static void note_energy_compare(int64_t* risk, int m, double new_e, double old_e)
{
    if (!risk || !isfinite(new_e) || !isfinite(old_e)) return;
    double band = 2.0 * (double)(m > 1 ? m - 1 : 1) * 0x1p-53 * (fabs(new_e) + fabs(old_e));
    risk[0]++;
    if (fabs(new_e - old_e) <= band) risk[1]++;
}

static void note_threshold(int64_t* risk, int m, double new_e, double old_e)
{
    if (!risk) return;
    double band = 2.0 * (double)(m > 1 ? m - 1 : 1) * 0x1p-53 * (fabs(new_e) + fabs(old_e));
    risk[2]++;
    if (fabs(fabs(new_e - old_e) - 1e-5) <= band) risk[3]++;
}

/* NewtonsMethod.  M == NULL: every vertex rounded, flips on posf with the
 * vertex at w.p.  M != NULL: the reference's rounding rules -- posf and pos
 * move together (VertexSmoother.cpp:499-500), isTetFlip per tet on posf or
 * pos, both restored at the end (:538-539). */
static int newtons(const int32_t* rt, int m, int32_t v_id, const double* x, const double* y,
                   const double* z, double* p_out, int64_t* risk, const emesh* M)
{
    view w = {x, y, z, v_id, {x[v_id], y[v_id], z[v_id]}};
    const double pf0[3] = {x[v_id], y[v_id], z[v_id]};
    double pos0[3] = {0, 0, 0};
    if (M) { pos0[0] = M->px[v_id]; pos0[1] = M->py[v_id]; pos0[2] = M->pz[v_id]; }
    int is_moved = 0;
    double old_energy = 0.0;
    double J[3], H[9], X0[3];
    for (int step = 0; step < MAX_STEP; ++step) {
        if (!ring_update(rt, m, &w, &old_energy, J, H, X0)) break;
        double old_p[3] = {w.p[0], w.p[1], w.p[2]};
        double old_pos[3] = {0, 0, 0};
        if (M) { old_pos[0] = M->px[v_id]; old_pos[1] = M->py[v_id]; old_pos[2] = M->pz[v_id]; }
        double a = 1.0;
        int step_taken = 0;
        double new_energy = 0.0;
        for (int it = 0; it < MAX_IT; ++it) {
            double rhs[3], X[3];
            for (int i = 0; i < 3; ++i) {
                double hx = H[3 * i + 0] * X0[0] + H[3 * i + 1] * X0[1] + H[3 * i + 2] * X0[2];
                rhs[i] = hx - a * J[i];
            }
            orc_solve3_qrcp(H, rhs, X);
            if (!(isfinite(X[0]) && isfinite(X[1]) && isfinite(X[2]))) {
                a /= 2.0;
                continue;
            }
            w.p[0] = X[0]; w.p[1] = X[1]; w.p[2] = X[2];
            if (M) {
                eset(M->x, M->y, M->z, v_id, X);
                eset(M->px, M->py, M->pz, v_id, X);
            }
            if (M ? e_is_flip(M, rt, m) : ring_flip(rt, m, &w)) {
                memcpy(w.p, old_p, sizeof old_p);
                if (M) moved_back(M, v_id, old_p, old_pos);
                a /= 2.0;
                continue;
            }
            new_energy = ring_energy(rt, m, &w);
            note_energy_compare(risk, m, new_energy, old_energy);
            if (new_energy >= old_energy || isinf(new_energy) || isnan(new_energy)) {
                memcpy(w.p, old_p, sizeof old_p);
                if (M) moved_back(M, v_id, old_p, old_pos);
                a /= 2.0;
                continue;
            }
            step_taken = 1;
            break;
        }
        if (step_taken) note_threshold(risk, m, new_energy, old_energy);
        if (step_taken && fabs(new_energy - old_energy) < 1e-5) step_taken = 0;
        if (!step_taken) {
            is_moved = step != 0;
            break;
        }
        is_moved = 1;
    }
    p_out[0] = w.p[0];
    p_out[1] = w.p[1];
    p_out[2] = w.p[2];
    if (M) moved_back(M, v_id, pf0, pos0);
    return is_moved;
}

int orc_newtons_method(const int32_t* rt, int m, int32_t v_id, const double* x, const double* y,
                       const double* z, double* p_out)
{
    return newtons(rt, m, v_id, x, y, z, p_out, NULL, NULL);
}

/* ---- one vertex of smoothSingle / smoothSurface on a frozen mesh -------- */
/* smoothSingle body  src/tetwild/VertexSmoother.cpp:145-212 (rounded vertices)
 * smoothSurface body src/tetwild/VertexSmoother.cpp:282-455
 * The caller has applied the reference's vertex filters (removed, bbox,
 * surface, locked, isolated) and supplies the boundary flag through kind. */
static double tet_quality_at(const view* w, const int32_t* t)
{
    double T[12];
    for (int j = 0; j < 4; ++j) vget(w, t[j], T + 3 * j);
    double e;
    if (orc_orient3d(T, T + 3, T + 6, T + 9) != 1) {
        e = ORC_MAX_ENERGY;
    } else {
        e = orc_amips_energy(T);
        if (isinf(e) || isnan(e)) e = ORC_MAX_ENERGY;
    }
    if (isinf(e) || isnan(e) || e <= 0) e = ORC_MAX_ENERGY;
    return e;
}

int orc_smooth_problem(int kind, int32_t v, const int32_t* rt, int m, const double* x,
                       const double* y, const double* z, const double* old_q,
                       const int32_t* tri_ab, int ntri, const orc_surface* sf,
                       const orc_surface* bt, double eps2, double d, double* p_out,
                       double* new_q)
{
    return orc_smooth_problem_risk(kind, v, rt, m, x, y, z, old_q, tri_ab, ntri, sf, bt, eps2, d,
                                   p_out, new_q, NULL);
}

int orc_smooth_problem_risk(int kind, int32_t v, const int32_t* rt, int m, const double* x,
                            const double* y, const double* z, const double* old_q,
                            const int32_t* tri_ab, int ntri, const orc_surface* sf,
                            const orc_surface* bt, double eps2, double d, double* p_out,
                            double* new_q, int64_t* risk)
{
    view w = {x, y, z, v, {x[v], y[v], z[v]}};
    double p0[3] = {x[v], y[v], z[v]};
    p_out[0] = p0[0]; p_out[1] = p0[1]; p_out[2] = p0[2];
    if (kind == 3)
        return newtons(rt, m, v, x, y, z, p_out, risk, NULL) ? ORC_SMOOTH_MOVED
                                                       : ORC_SMOOTH_NOT_MOVED;
    if (ring_flip(rt, m, &w)) return ORC_SMOOTH_INVALID;
    double p[3];
    if (!newtons(rt, m, v, x, y, z, p, risk, NULL)) return ORC_SMOOTH_NOT_MOVED;
    if (kind == 0) {
        memcpy(w.p, p, sizeof p);
        if (ring_flip(rt, m, &w)) return ORC_SMOOTH_FLIP_END;
        memcpy(p_out, p, sizeof p);
        return ORC_SMOOTH_MOVED;
    }
    double pf[3], d2;
    orc_nearest(kind == 2 ? bt : sf, p, pf, &d2);
    memcpy(w.p, pf, sizeof pf);
    if (ring_flip(rt, m, &w)) return ORC_SMOOTH_REJECT_FLIP;
    double old_max = 0.0, new_max = 0.0;
    for (int i = 0; i < m; ++i)
        if (old_q[i] > old_max) old_max = old_q[i];
    for (int i = 0; i < m; ++i) {
        new_q[i] = tet_quality_at(&w, rt + 4 * i);
        if (new_q[i] > new_max) new_max = new_q[i];
    }
    if (!(new_max < old_max)) return ORC_SMOOTH_REJECT_QUALITY;
    for (int i = 0; i < ntri; ++i) {
        double tri[9];
        memcpy(tri, pf, sizeof pf);
        coords(x, y, z, tri_ab[2 * i], tri + 3);
        coords(x, y, z, tri_ab[2 * i + 1], tri + 6);
        if (orc_collinear3(tri, tri + 3, tri + 6)) continue;
        if (orc_face_out(sf, tri, eps2, d, NULL)) return ORC_SMOOTH_REJECT_ENVELOPE;
    }
    memcpy(p_out, pf, sizeof pf);
    return ORC_SMOOTH_MOVED;
}

/* ---- the per-vertex bodies with the reference's rounding rules ---------- */
/* Each vertex has posf (x, y, z), an exact position pos (px, py, pz: doubles
 * in these inputs) and is_rounded (rounded[]).  Everything follows the
 * reference literally: isTetFlip orients on posf when the tet's four
 * vertices are rounded and on pos otherwise (src/tetwild/LocalOperations.cpp:
 * 535-553); NewtonsMethod moves posf and pos together (VertexSmoother.cpp:
 * 499-500) and restores both; smoothSingle tries to round the vertex first
 * (:152-161), validates the ring on posf (:163-172) and commits pos, posf
 * and is_rounded = true unless the final isFlip fails (:196-211);
 * smoothSurface (:272-455) likewise, rounding the vertex at its end. */
int orc_smooth_vertex(int kind, int32_t v, const int32_t* rt, int m, double* x, double* y,
                      double* z, double* px, double* py, double* pz, uint8_t* rounded,
                      const double* old_q, const int32_t* tri_ab, int ntri,
                      const orc_surface* sf, const orc_surface* bt, double eps2, double d,
                      double* new_q)
{
    emesh M = {x, y, z, px, py, pz, rounded};
    /* try to round the vertex */
    if (!rounded[v]) {
        double old_p[3] = {px[v], py[v], pz[v]}, pf[3] = {x[v], y[v], z[v]};
        eset(px, py, pz, v, pf);
        if (e_is_flip(&M, rt, m)) eset(px, py, pz, v, old_p);
        else rounded[v] = 1;
    }
    /* validity on posf */
    for (int i = 0; i < m; ++i) {
        double T[12];
        for (int j = 0; j < 4; ++j) eget(&M, rt[4 * i + j], 0, T + 3 * j);
        if (orc_orient3d(T, T + 3, T + 6, T + 9) != 1) return ORC_SMOOTH_INVALID;
    }
    double pn[3];
    if (!newtons(rt, m, v, x, y, z, pn, NULL, &M)) return ORC_SMOOTH_NOT_MOVED;
    double old_pf[3] = {x[v], y[v], z[v]}, old_p[3] = {px[v], py[v], pz[v]};
    const uint8_t old_r = rounded[v];
    if (kind == 0) {
        eset(px, py, pz, v, pn);
        eset(x, y, z, v, pn);
        rounded[v] = 1;
        if (e_is_flip(&M, rt, m)) {
            eset(px, py, pz, v, old_p);
            eset(x, y, z, v, old_pf);
            rounded[v] = old_r;
            return ORC_SMOOTH_FLIP_END;
        }
        return ORC_SMOOTH_MOVED;
    }
    double pf[3], d2;
    orc_nearest(kind == 2 ? bt : sf, pn, pf, &d2);
    eset(x, y, z, v, pf);
    eset(px, py, pz, v, pf);
    if (e_is_flip(&M, rt, m)) {
        eset(px, py, pz, v, old_p);
        eset(x, y, z, v, old_pf);
        return ORC_SMOOTH_REJECT_FLIP;
    }
    double om = 0.0, nm = 0.0;
    for (int i = 0; i < m; ++i)
        if (old_q[i] > om) om = old_q[i];
    view w = {x, y, z, -1, {0, 0, 0}};
    for (int i = 0; i < m; ++i) {
        new_q[i] = tet_quality_at(&w, rt + 4 * i);
        if (new_q[i] > nm) nm = new_q[i];
    }
    if (!(nm < om)) {
        eset(px, py, pz, v, old_p);
        eset(x, y, z, v, old_pf);
        return ORC_SMOOTH_REJECT_QUALITY;
    }
    for (int i = 0; i < ntri; ++i) {
        double tri[9];
        memcpy(tri, pf, sizeof pf);
        coords(x, y, z, tri_ab[2 * i], tri + 3);
        coords(x, y, z, tri_ab[2 * i + 1], tri + 6);
        if (orc_collinear3(tri, tri + 3, tri + 6)) continue;
        if (orc_face_out(sf, tri, eps2, d, NULL)) {
            eset(px, py, pz, v, old_p);
            eset(x, y, z, v, old_pf);
            return ORC_SMOOTH_REJECT_ENVELOPE;
        }
    }
    if (!rounded[v]) {
        eset(px, py, pz, v, pf);
        if (e_is_flip(&M, rt, m)) {
            eset(px, py, pz, v, old_p);
            rounded[v] = 0;
        } else {
            rounded[v] = 1;
        }
    }
    return ORC_SMOOTH_MOVED;
}
