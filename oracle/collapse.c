/*
 * collapse.c -- oracle: the checks EdgeCollapser::collapseAnEdge runs for one
 * (v1 -> v2) candidate on a frozen mesh.  TEST INFRASTRUCTURE ONLY.
 *
 * collapseAnEdge         src/tetwild/EdgeCollapser.cpp:284-446 (everything
 *                        before the "real update")
 * isCollapsable_epsilon  src/tetwild/EdgeCollapser.cpp:727-775
 * getCheckQuality        src/tetwild/LocalOperations.cpp:575-603
 *                        (State::use_energy_max is the constant true,
 *                        src/tetwild/State.h:53)
 * isBetterOrEqualThan    src/tetwild/TetmeshElements.h:100-109 (AMIPS energy)
 * isBoundarySlide        returns false on its first line
 *                        (src/tetwild/LocalOperations.cpp:1123-1124), so
 *                        the reference's "check 2.5" never rejects.
 *
 * The caller supplies what the reference derives from the mesh topology:
 * v1's one-ring (conn_tets order), the stored qualities of those tets, the
 * vertex flags after the isolated / boundary-point updates (:298-307), and
 * the epsilon-check triangles (unique sorted one-ring surface faces of v1
 * that do not contain v2, with v1 replaced by v2 in place, :742-764).
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>

static inline void pt(const double* x, const double* y, const double* z, int32_t v, double* p)
{
    p[0] = x[v];
    p[1] = y[v];
    p[2] = z[v];
}

static double quality(const double* T)
{
    /* calTetQualities + calTetQuality_AMIPS, src/tetwild/LocalOperations.cpp:695-773, 862-884 */
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

int orc_collapse_check(int32_t v1, int32_t v2, int flags, const int32_t* rt, int m,
                       const double* x, const double* y, const double* z, const double* old_q,
                       const int32_t* tri_ids, int ntri, const orc_surface* sf,
                       const orc_surface* bt, double eps2, double d, int check_quality,
                       int check_eps, int is_soft, double soft_energy, double* new_q)
{
    const int v1_sf = (flags & ORC_CV1_SURFACE) != 0, v1_bd = (flags & ORC_CV1_BOUNDARY) != 0;
    const int v2_sf = (flags & ORC_CV2_SURFACE) != 0, v2_bd = (flags & ORC_CV2_BOUNDARY) != 0;
    double p1[3], p2[3];
    pt(x, y, z, v1, p1);
    pt(x, y, z, v2, p2);
    /* sqrt(squared_distance) == 0 <=> the squared distance is 0 */
    const double dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
    const int degenerate = (dx * dx + dy * dy + dz * dz) == 0.0;

    if (v1_bd && !v2_bd && !degenerate && orc_point_out(bt, p2, eps2)) return ORC_COLLAPSE_ENVELOP;
    if (v1_sf && !v2_sf && !degenerate && orc_point_out(sf, p2, eps2)) return ORC_COLLAPSE_ENVELOP;

    /* new_tets: ring tets without v2, v1 replaced by v2 in place */
    for (int i = 0; i < m; ++i) {
        const int32_t* t = rt + 4 * i;
        int has_v2 = t[0] == v2 || t[1] == v2 || t[2] == v2 || t[3] == v2;
        if (has_v2) continue;
        double T[12];
        for (int j = 0; j < 4; ++j) pt(x, y, z, t[j] == v1 ? v2 : t[j], T + 3 * j);
        if (orc_orient3d(T, T + 3, T + 6, T + 9) != 1) return ORC_COLLAPSE_FLIP;
    }
    double new_max = 0.0, old_max = 0.0;
    for (int i = 0; i < m; ++i) {
        const int32_t* t = rt + 4 * i;
        int has_v2 = t[0] == v2 || t[1] == v2 || t[2] == v2 || t[3] == v2;
        if (old_q[i] > old_max) old_max = old_q[i];
        if (has_v2) {
            new_q[i] = NAN;
            continue;
        }
        double T[12];
        for (int j = 0; j < 4; ++j) pt(x, y, z, t[j] == v1 ? v2 : t[j], T + 3 * j);
        new_q[i] = quality(T);
        if (new_q[i] > new_max) new_max = new_q[i];
    }
    if (check_quality) {
        if (is_soft) old_max = soft_energy;
        if (flags & ORC_CV1_UNROUNDED) new_max = 0.0;
        if (!degenerate && !(new_max <= old_max)) return ORC_COLLAPSE_QUALITY;
    }
    if (check_eps && v1_sf) {
        if (!degenerate) {
            for (int i = 0; i < ntri; ++i) {
                double tri[9];
                for (int k = 0; k < 3; ++k) pt(x, y, z, tri_ids[3 * i + k], tri + 3 * k);
                if (orc_face_out(sf, tri, eps2, d, NULL)) return ORC_COLLAPSE_ENVELOP;
            }
        }
        return ORC_COLLAPSE_ENVELOP_SUC;
    }
    return ORC_COLLAPSE_SUCCESS;
}
