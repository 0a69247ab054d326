/*
 * baseline.c -- the oracle's batch loops spread over host threads (OpenMP).
 * TEST INFRASTRUCTURE ONLY: bench.py times these as the multi-core CPU
 * baseline.  Every item (tet, query, one-ring, candidate) is independent and
 * computed by the same single-item function the sequential loops call, so the
 * results are the sequential loops' bit for bit; only the wall clock changes.
 * The near-tie accounting of the envelope decision lives here too.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

static int threads(int nthreads) { return nthreads > 0 ? nthreads : omp_get_max_threads(); }

int orc_max_threads(void) { return omp_get_max_threads(); }

/* chunks of 4096 items, each through the sequential loop */
void orc_amips_batch_mt(const int32_t* tets, int64_t n, const double* x, const double* y,
                        const double* z, int mode, double* E, double* G, double* H,
                        int nthreads)
{
    const int64_t C = 4096;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads(nthreads))
    for (int64_t s = 0; s < n; s += C) {
        int64_t k = n - s < C ? n - s : C;
        orc_amips_batch(tets + 4 * s, k, x, y, z, mode, E + s, G ? G + 3 * s : NULL,
                        H ? H + 9 * s : NULL);
    }
}

void orc_tet_qualities_mt(const int32_t* tets, int64_t n, const double* x, const double* y,
                          const double* z, double* slim_energy, int nthreads)
{
    const int64_t C = 4096;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads(nthreads))
    for (int64_t s = 0; s < n; s += C) {
        int64_t k = n - s < C ? n - s : C;
        orc_tet_qualities(tets + 4 * s, k, x, y, z, slim_energy + s);
    }
}

void orc_face_out_batch_mt(const orc_surface* s, const double* tri9, int64_t n, double eps2,
                           double d, uint8_t* out, int nthreads)
{
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads(nthreads))
    for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)orc_face_out(s, tri9 + 9 * i, eps2, d, NULL);
}

void orc_point_out_batch_mt(const orc_surface* s, const double* p3, int64_t n, double eps2,
                            uint8_t* out, int nthreads)
{
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads(nthreads))
    for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)orc_point_out(s, p3 + 3 * i, eps2);
}

void orc_smooth_batch_mt(int64_t n, const uint8_t* kind, const int32_t* center,
                         const int32_t* ring_off, const int32_t* ring_tets, const double* x,
                         const double* y, const double* z, const double* old_q,
                         const int32_t* tri_off, const int32_t* tri_ab, const orc_surface* sf,
                         const orc_surface* bt, double eps2, double d, double* p_out,
                         uint8_t* status, double* new_q, int64_t* risk, int nthreads)
{
    int mmax = 1;
    for (int64_t p = 0; p < n; ++p)
        if (ring_off[p + 1] - ring_off[p] > mmax) mmax = ring_off[p + 1] - ring_off[p];
    int64_t r0s = 0, r1s = 0, r2s = 0, r3s = 0;
    static const int32_t no_tri[2] = {0, 0};
#pragma omp parallel num_threads(threads(nthreads)) reduction(+ : r0s, r1s, r2s, r3s)
    {
        double* nq = (double*)malloc(sizeof(double) * (size_t)mmax);
        int64_t rk[4] = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 16)
        for (int64_t p = 0; p < n; ++p) {
            const int r0 = ring_off[p], m = ring_off[p + 1] - r0;
            const int t0 = tri_off ? tri_off[p] : 0;
            const int nt = tri_off ? tri_off[p + 1] - t0 : 0;
            const int k = kind ? kind[p] : 0;
            int s = orc_smooth_problem_risk(k, center[p], ring_tets + 4 * (int64_t)r0, m, x, y, z,
                                            old_q + r0, nt > 0 ? tri_ab + 2 * (int64_t)t0 : no_tri,
                                            nt, sf, bt, eps2, d, p_out + 3 * p, nq,
                                            risk ? rk : NULL);
            status[p] = (uint8_t)s;
            if (s == ORC_SMOOTH_MOVED && (k == 1 || k == 2))
                memcpy(new_q + r0, nq, sizeof(double) * (size_t)m);
        }
        free(nq);
        r0s += rk[0];
        r1s += rk[1];
        r2s += rk[2];
        r3s += rk[3];
    }
    if (risk) {
        risk[0] += r0s;
        risk[1] += r1s;
        risk[2] += r2s;
        risk[3] += r3s;
    }
}

void orc_collapse_batch_mt(int64_t n, const int32_t* v1, const int32_t* v2, const uint8_t* flags,
                           const int32_t* ring_off, const int32_t* ring_tets, const double* x,
                           const double* y, const double* z, const double* old_q,
                           const int32_t* tri_off, const int32_t* tri_ids,
                           const orc_surface* sf, const orc_surface* bt, double eps2, double d,
                           int check_quality, int check_eps, int is_soft, double soft_energy,
                           uint8_t* status, double* new_q, int nthreads)
{
    static const int32_t none[4] = {0, 0, 0, 0};
    static const double zero = 0.0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads(nthreads))
    for (int64_t c = 0; c < n; ++c) {
        const int r0 = ring_off[c], m = ring_off[c + 1] - r0;
        const int t0 = tri_off[c], nt = tri_off[c + 1] - t0;
        double one = 0.0;
        double* nq = m > 0 ? new_q + r0 : &one;
        if (m > 0) memset(nq, 0, sizeof(double) * (size_t)m);
        status[c] = (uint8_t)orc_collapse_check(
            v1[c], v2[c], flags[c], m > 0 ? ring_tets + 4 * (int64_t)r0 : none, m, x, y, z,
            m > 0 ? old_q + r0 : &zero, nt > 0 ? tri_ids + 3 * (int64_t)t0 : none, nt, sf, bt,
            eps2, d, check_quality, check_eps, is_soft, soft_energy, nq);
    }
}

/* sampleTriangle's samples of one triangle, as orc_face_out draws them */
static int64_t samples_of(const double* tri9, double d, double** ps, double* local, int64_t cap)
{
    int64_t n = orc_sample_triangle(tri9, tri9 + 3, tri9 + 6, d, local, cap);
    *ps = local;
    if (n > cap) {
        *ps = (double*)malloc(sizeof(double) * 3 * (size_t)n);
        orc_sample_triangle(tri9, tri9 + 3, tri9 + 6, d, *ps, n);
    }
    return n;
}

void orc_face_risk_batch_mt(const orc_surface* s, const double* tri9, int64_t n, double eps2,
                            double d, double band, int64_t* counts, int nthreads)
{
    int64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads(nthreads)) \
    reduction(+ : c0, c1, c2, c3)
    for (int64_t i = 0; i < n; ++i) {
        const double* t = tri9 + 9 * i;
        /* the triangles orc_face_out decides without sampling carry no risk */
        if (orc_collinear3(t, t + 3, t + 6)) continue;
        double l = 0.0;
        for (int e = 0; e < 3; ++e) {
            const double* a = t + 3 * e;
            const double* b = t + 3 * ((e + 1) % 3);
            double q = 0.0;
            for (int k = 0; k < 3; ++k) q += (a[k] - b[k]) * (a[k] - b[k]);
            l = q > l ? q : l;
        }
        if (!(sqrt(l) / d <= 65536.0)) continue;
        double local[3 * 256], *ps;
        int64_t m = samples_of(t, d, &ps, local, 256);
        int64_t nb = 0, clear_out = 0;
        for (int64_t k = 0; k < m; ++k) {
            double cp[3], d2;
            orc_nearest(s, ps + 3 * k, cp, &d2);
            if (fabs(d2 - eps2) <= band) nb++;
            else if (d2 > eps2) clear_out++;
        }
        if (ps != local) free(ps);
        c0 += m;
        c1 += nb;
        c2 += nb > 0;
        c3 += nb > 0 && clear_out == 0;
    }
    counts[0] += c0;
    counts[1] += c1;
    counts[2] += c2;
    counts[3] += c3;
}

void orc_point_risk_batch_mt(const orc_surface* s, const double* p3, int64_t n, double eps2,
                             double band, int64_t* counts, int nthreads)
{
    int64_t c = 0;
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads(nthreads)) reduction(+ : c)
    for (int64_t i = 0; i < n; ++i) {
        double cp[3], d2;
        orc_nearest(s, p3 + 3 * i, cp, &d2);
        c += fabs(d2 - eps2) <= band;
    }
    counts[0] += n;
    counts[1] += c;
}
