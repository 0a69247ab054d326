/*
 * reduce.c -- oracle: the energy reductions of LocalOperations and the group
 * evaluation of EdgeRemover's configurations.  TEST INFRASTRUCTURE ONLY.
 *
 * getAvgMaxEnergy     src/tetwild/LocalOperations.cpp:605-622
 * getMaxEnergy        src/tetwild/LocalOperations.cpp:624-635
 * getSecondMaxEnergy  src/tetwild/LocalOperations.cpp:637-650
 * getFilterEnergy     src/tetwild/LocalOperations.cpp:652-693
 * isTetLocked_ui      src/tetwild/LocalOperations.cpp:1377-1384
 * isFlip + calTetQualities + getCheckQuality over a configuration
 *                     src/tetwild/EdgeRemover.cpp:138-174, 327-368, 501-575
 *
 * Sums run in index order, as the reference's loops do.
 */
#include "oracle.h"

#include <math.h>

static int locked(const int32_t* tets, const uint8_t* v_locked, int64_t i)
{
    if (!tets || !v_locked) return 0;
    for (int j = 0; j < 4; ++j)
        if (v_locked[tets[4 * i + j]]) return 1;
    return 0;
}

void orc_energy_stats(const double* e, int64_t n, const uint8_t* removed, const int32_t* tets,
                      const uint8_t* v_locked, double thres, double* sum, double* max,
                      double* second_max, int64_t* count, int64_t* buckets)
{
    double s = 0, m = 0, m2 = 0;
    int64_t cnt = 0;
    for (int k = 0; k < 11; ++k) buckets[k] = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (removed && removed[i]) continue;
        /* getFilterEnergy's histogram */
        if (e[i] > thres - 1 + 1e10) {
            buckets[10]++;
        } else {
            for (int j = 0; j < 10; ++j)
                if (e[i] > thres - 1 + pow(10, j) && e[i] <= thres - 1 + pow(10, j + 1)) {
                    buckets[j]++;
                    break;
                }
        }
        if (locked(tets, v_locked, i)) continue;
        if (e[i] > m) m = e[i];
        if (e[i] != ORC_MAX_ENERGY && e[i] > m2) m2 = e[i];
        s += e[i];
        cnt++;
    }
    *sum = s;
    *max = m;
    *second_max = m2;
    *count = cnt;
}

double orc_filter_energy(const int64_t* buckets, double thres, int* is_clean_up)
{
    int64_t t1[10], t2[10];
    for (int i = 0; i < 10; ++i) {
        t1[i] = 0;
        t2[i] = 0;
        for (int k = 0; k <= i; ++k) t1[i] += buckets[k];
        for (int k = i + 1; k < 11; ++k) t2[i] += buckets[k];
    }
    if (t1[0] >= t2[0]) {
        *is_clean_up = t2[5] > 0;
        return 8;
    }
    if (t1[8] <= t2[8]) return 1e11;
    for (int i = 0; i < 8; ++i)
        if (t1[i] < t2[i] && t1[i + 1] > t2[i + 1]) return thres - 1 + 5 * pow(10, i + 1);
    return 8;
}

void orc_tet_groups(const int32_t* tets, const int32_t* off, int64_t n_groups, const double* x,
                    const double* y, const double* z, uint8_t* flipped, double* max_q, double* q)
{
    for (int64_t g = 0; g < n_groups; ++g) {
        int flip = 0;
        for (int64_t k = off[g]; k < off[g + 1] && !flip; ++k) {
            double T[12];
            for (int j = 0; j < 4; ++j) {
                T[3 * j] = x[tets[4 * k + j]];
                T[3 * j + 1] = y[tets[4 * k + j]];
                T[3 * j + 2] = z[tets[4 * k + j]];
            }
            flip = orc_orient3d(T, T + 3, T + 6, T + 9) != 1;
        }
        flipped[g] = (uint8_t)flip;
        if (flip) {
            max_q[g] = NAN;
            continue;
        }
        double m = 0;
        for (int64_t k = off[g]; k < off[g + 1]; ++k) {
            double v;
            orc_tet_qualities(tets + 4 * k, 1, x, y, z, &v);
            if (q) q[k] = v;
            if (v > m) m = v;
        }
        max_q[g] = m;
    }
}
