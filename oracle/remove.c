/*
 * remove.c -- oracle: the decision of EdgeRemover for one edge, i.e. which
 * of removeAnEdge_32 / _44 / _56 succeeds on a frozen mesh, with the new
 * tets and qualities it would write.  TEST INFRASTRUCTURE ONLY.
 *
 * EdgeRemover::swap's order  src/tetwild/EdgeRemover.cpp:97-106
 * removeAnEdge_32            src/tetwild/EdgeRemover.cpp:127-174
 * removeAnEdge_44            src/tetwild/EdgeRemover.cpp:276-368
 * removeAnEdge_56            src/tetwild/EdgeRemover.cpp:458-617 (new tets :612-625)
 * isBetterThan / isBetterOrEqualThan (AMIPS): src/tetwild/TetmeshElements.h:90-109
 * getCheckQuality (use_energy_max): src/tetwild/LocalOperations.cpp:575-603
 */
#include "oracle.h"

#include <string.h>

/* isFlip over m tets */
static int flips(const int32_t* t, int m, const double* x, const double* y, const double* z)
{
    uint8_t f[8];
    orc_tet_flip(t, m, x, y, z, f);
    for (int i = 0; i < m; ++i)
        if (f[i]) return 1;
    return 0;
}

static double quals(const int32_t* t, int m, const double* x, const double* y, const double* z,
                    double* q)
{
    orc_tet_qualities(t, m, x, y, z, q);
    double mx = 0;
    for (int i = 0; i < m; ++i)
        if (q[i] > mx) mx = q[i];
    return mx;
}

static void replace(int32_t* t, int32_t from, int32_t to)
{
    for (int j = 0; j < 4; ++j)
        if (t[j] == from) {
            t[j] = to;
            return;
        }
}

static int has(const int32_t* t, int32_t v)
{
    return t[0] == v || t[1] == v || t[2] == v || t[3] == v;
}

/* new quality acceptable?  equal_buget allows equality while it lasts */
static int better(double new_q, double old_q, int* equal_buget)
{
    if (*equal_buget > 0) {
        (*equal_buget)--;
        return new_q <= old_q;
    }
    return new_q < old_q;
}

/* the ring's "n12" vertices (the two vertices of each tet other than v1, v2)
 * chained around the edge, and the tets in chain order */
static void chain(const int32_t* ring, int m, int32_t v1, int32_t v2, int32_t* nv, int* nt)
{
    int32_t es[5][3];
    for (int i = 0; i < m; ++i) {
        int c = 0;
        for (int j = 0; j < 4; ++j)
            if (ring[4 * i + j] != v1 && ring[4 * i + j] != v2) es[i][c++] = ring[4 * i + j];
        es[i][2] = i;
    }
    int visited[5] = {1, 0, 0, 0, 0};
    int nvn = 0, ntn = 0;
    nv[nvn++] = es[0][0];
    nv[nvn++] = es[0][1];
    nt[ntn++] = 0;
    for (int i = 0; i < m - 2; ++i)
        for (int j = 0; j < m; ++j) {
            if (visited[j]) continue;
            if (es[j][0] == nv[nvn - 1]) {
                visited[j] = 1;
                nv[nvn++] = es[j][1];
            } else if (es[j][1] == nv[nvn - 1]) {
                visited[j] = 1;
                nv[nvn++] = es[j][0];
            }
            if (visited[j]) {
                nt[ntn++] = j;
                break;
            }
        }
    for (int j = 0; j < m; ++j)
        if (!visited[j]) {
            nt[ntn++] = j;
            break;
        }
}

int orc_remove_edge(int32_t v1, int32_t v2, const int32_t* ring, int m, const double* old_q,
                    const double* x, const double* y, const double* z, int* equal_buget,
                    int32_t* new_tets, double* new_q, int* n_new)
{
    *n_new = 0;
    double old_max = 0;
    for (int i = 0; i < m; ++i)
        if (old_q[i] > old_max) old_max = old_q[i];
    if (m == 3) {
        int32_t vi[2], c = 0;
        for (int j = 0; j < 4; ++j)
            if (ring[j] != v1 && ring[j] != v2) vi[c++] = ring[j];
        int32_t t[8];
        int a = has(ring + 4, vi[0]) ? 1 : 2, b = 3 - a;
        memcpy(t, ring + 4 * a, 16);
        memcpy(t + 4, ring + 4 * b, 16);
        replace(t, v1, vi[1]);
        replace(t + 4, v2, vi[0]);
        if (flips(t, 2, x, y, z)) return 0;
        double q[2];
        double nm = quals(t, 2, x, y, z, q);
        if (!better(nm, old_max, equal_buget)) return 0;
        memcpy(new_tets, t, sizeof t);
        memcpy(new_q, q, sizeof q);
        *n_new = 2;
        return 3;
    }
    if (m == 4) {
        int32_t nv[5];
        int nt[5];
        chain(ring, 4, v1, v2, nv, nt);
        int valid = 0;
        int32_t best_t[16];
        double best_q[4];
        for (int i = 0; i < 2; ++i) {
            const int32_t a = nv[i], b = nv[2 + i];
            int32_t t[16];
            for (int j = 0; j < 4; ++j) {
                memcpy(t + 4 * j, ring + 4 * j, 16);
                if (has(t + 4 * j, a)) replace(t + 4 * j, v2, b);
                else replace(t + 4 * j, v1, a);
            }
            if (flips(t, 4, x, y, z)) continue;
            double q[4];
            double nm = quals(t, 4, x, y, z, q);
            if (!better(nm, old_max, equal_buget)) return 0;
            valid = 1;
            old_max = nm;
            memcpy(best_t, t, sizeof t);
            memcpy(best_q, q, sizeof q);
        }
        if (!valid) return 0;
        memcpy(new_tets, best_t, sizeof best_t);
        memcpy(new_q, best_q, sizeof best_q);
        *n_new = 4;
        return 4;
    }
    if (m == 5) {
        int32_t nv[5];
        int nt[5];
        chain(ring, 5, v1, v2, nv, nt);
        int ok[5] = {1, 1, 1, 1, 1}, have[5] = {0, 0, 0, 0, 0};
        int32_t pt[10][8];
        double pq[10][2];
        for (int i = 0; i < 5; ++i) {
            if (!ok[(i + 1) % 5] && !ok[(i + 4) % 5]) continue;
            int32_t* t = pt[i];
            memcpy(t, ring + 4 * nt[i], 16);
            memcpy(t + 4, ring + 4 * nt[i], 16);
            replace(t, v1, nv[(i + 4) % 5]);
            replace(t + 4, v2, nv[(i + 4) % 5]);
            if (flips(t, 2, x, y, z)) {
                ok[(i + 1) % 5] = 0;
                ok[(i + 4) % 5] = 0;
                continue;
            }
            quals(t, 2, x, y, z, pq[i]);
            have[i] = 1;
        }
        if (!ok[0] && !ok[1] && !ok[2] && !ok[3] && !ok[4]) return 0;
        int sel = -1;
        for (int i = 0; i < 5; ++i) {
            if (!ok[i]) continue;
            int32_t t[8];
            memcpy(t, ring + 4 * nt[(i + 2) % 5], 16);
            memcpy(t + 4, ring + 4 * nt[(i + 2) % 5], 16);
            replace(t, v1, nv[i]);
            replace(t + 4, v2, nv[i]);
            if (flips(t, 2, x, y, z)) continue;
            double q[2];
            quals(t, 2, x, y, z, q);
            const int l = (i + 1) % 5, r = (i + 4) % 5;
            if (!have[l] || !have[r]) return -1; /* cannot happen (see the reference's loop) */
            double six[6] = {q[0], q[1], pq[l][0], pq[r][0], pq[l][1], pq[r][1]};
            double nm = 0;
            for (int k = 0; k < 6; ++k)
                if (six[k] > nm) nm = six[k];
            if (!better(nm, old_max, equal_buget)) continue;
            old_max = nm;
            sel = i;
            memcpy(pt[5 + i], t, sizeof t);
            memcpy(pq[5 + i], q, sizeof q);
        }
        if (sel < 0) return 0;
        const int src[3] = {(sel + 1) % 5, (sel + 4) % 5, sel + 5};
        for (int s = 0; s < 3; ++s) {
            memcpy(new_tets + 8 * s, pt[src[s]], 32);
            new_q[2 * s] = pq[src[s]][0];
            new_q[2 * s + 1] = pq[src[s]][1];
        }
        *n_new = 6;
        return 5;
    }
    return 0;
}
