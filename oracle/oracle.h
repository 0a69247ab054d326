/*
 * oracle.h -- CPU restatement of TetWild's mesh-optimisation inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * tetwild_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Arithmetic contract: plain IEEE binary64, built with -ffp-contract=off, no
 * fast-math.  The only fused multiply-adds are explicit fma() calls inside the
 * error-free product of the exact predicates, where fusion is the point.
 *
 * Pinning status per function (what independent evidence fixes each one):
 *   orc_amips_*           golden vectors produced by the reference's own
 *                         comformalAMIPS{Energy,Jacobian,Hessian}_new compiled
 *                         from src/tetwild/LocalOperations.cpp:28-291
 *                         (tests/golden/amips_golden.npz), plus 50-digit
 *                         mpmath closed form and invariants.
 *   orc_orient3d/collinear  exact; pinned against fractions.Fraction.
 *   orc_point_tri_d2      Eberly's region method; pinned against an exact
 *                         rational brute force on small inputs.  Bit-level
 *                         agreement with geogram's point_triangle_squared_distance
 *                         (not vendored in the reference): parity unpinned.
 *   orc_sample_triangle   an independent 50-digit model of the hexagonal
 *                         lattice (rows, even-row bump, break rule, side
 *                         edges, integer-ratio and tie rules), sample by
 *                         sample (tests/test_oracle_sampling.py, with planted
 *                         mistakes), plus invariants.  Bit-level agreement
 *                         with geogram vec3 arithmetic: parity unpinned.
 *   orc_envelope_*        brute force over all facets (exact-rational decisions
 *                         away from the eps^2 rounding band).
 *   orc_collapse_check    composition of the pinned parts above (orient3d,
 *                         qualities, point / face envelope) in the reference's
 *                         check order; the order and the early exits are pinned
 *                         by constructed candidates whose outcome geometry
 *                         fixes (tests/test_oracle_collapse.py).
 *   orc_energy_stats, orc_filter_energy  constructed histograms and values on
 *                         the bucket limits (tests/test_oracle_reduce.py);
 *                         the sum checked against exact rational sums.
 *   orc_tet_groups        composition of orient3d and tet qualities, checked
 *                         against them on constructed configurations.
 *   orc_remove_edge       constructed rings of 3, 4 and 5 tets around an edge:
 *                         the new tets must be the geometric re-triangulation
 *                         (face / diagonal / pentagon fan coned to v1 and v2)
 *                         and the decision the quality comparison
 *                         (tests/test_oracle_remove.py).
 *   orc_newton_*          invariants (monotone energy, no flips, step bounds,
 *                         is_moved semantics, X = X0 - a H^-1 J); MAX_STEP,
 *                         MAX_IT and the 1e-5 stop by constructed rings
 *                         (tests/test_oracle_newton.py, with planted
 *                         mistakes).  Bit-level agreement with Eigen's
 *                         ColPivHouseholderQR: parity unpinned.
 *   orc_smooth_vertex     the rounding rules: equal to orc_smooth_problem when
 *                         every vertex is rounded, constructed try-to-round
 *                         and inverted-exact-tet cases
 *                         (tests/test_oracle_smooth.py).
 *   orc_*_mt, risk counts the sequential loops' results bit for bit; the
 *                         band counts on constructed distances
 *                         (tests/test_oracle_baseline.py).
 */
#ifndef TWG_ORACLE_H
#define TWG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_ENERGY 1e50 /* State::MAX_ENERGY, src/tetwild/State.h:29 */

/* ---- (a) AMIPS, src/tetwild/LocalOperations.cpp:28-291 ---------------- */
/* T = 12 doubles v0xyz v1xyz v2xyz v3xyz, as the reference's const double* T */
double orc_amips_energy(const double* T);
void orc_amips_jacobian(const double* T, double* g3);
void orc_amips_hessian(const double* T, double* h9);
/* E, grad (3) and Hessian (9) in one pass; same bits as the three above */
double orc_amips_ejh(const double* T, double* g3, double* h9);
double orc_rcbrt(double x);

/* batch helpers over SoA vertex arrays and int32 tets (n x 4) */
void orc_amips_batch(const int32_t* tets, int64_t n, const double* x, const double* y,
                     const double* z, int mode, double* E, double* G, double* H);

/* ---- exact predicates ------------------------------------------------- */
/* sign of det[q-p; r-p; s-p]: +1 POSITIVE, 0, -1 (CGAL::orientation) */
int orc_orient3d(const double* p, const double* q, const double* r, const double* s);
/* 1 if p, q, r are exactly collinear (CGAL::collinear / Triangle_3::is_degenerate) */
int orc_collinear3(const double* p, const double* q, const double* r);

/* calTetQualities, src/tetwild/LocalOperations.cpp:695-773, 862-884 */
void orc_tet_qualities(const int32_t* tets, int64_t n, const double* x, const double* y,
                       const double* z, double* slim_energy);
/* isTetFlip per tet, src/tetwild/LocalOperations.cpp:535-553 */
void orc_tet_flip(const int32_t* tets, int64_t n, const double* x, const double* y,
                  const double* z, uint8_t* flipped);

/* ---- (b) envelope ------------------------------------------------------ */
/* sampleTriangle, src/tetwild/Common.cpp:143-228.  Returns sample count; writes
 * at most cap samples (xyz triples) into ps.  A negative return never happens. */
int64_t orc_sample_triangle(const double* v0, const double* v1, const double* v2,
                            double d, double* ps, int64_t cap);
/* squared distance point -> triangle (Eberly regions); closest point into cp */
double orc_point_tri_d2(const double* p, const double* a, const double* b, const double* c,
                        double* cp);

/* Surface for envelope queries: V (nv x 3), F (nf x 3).  The oracle's decision
 * for one point is: out <=> min over ALL facets f of d2(p, f) > eps2. */
typedef struct orc_surface orc_surface;
orc_surface* orc_surface_create(const double* V, int64_t nv, const int32_t* F, int64_t nf);
void orc_surface_destroy(orc_surface* s);
/* 1 if the point is out of the envelope (isPointOutEnvelop semantics) */
int orc_point_out(const orc_surface* s, const double* p, double eps2);
/* isFaceOutEnvelop_sampling semantics; returns 1 (out) / 0 (in).  n_samples
 * receives the sample count (0 for a degenerate triangle). */
int orc_face_out(const orc_surface* s, const double* tri9, double eps2, double d,
                 int64_t* n_samples);
void orc_face_out_batch(const orc_surface* s, const double* tri9, int64_t n, double eps2,
                        double d, uint8_t* out, int64_t* total_samples);
void orc_point_out_batch(const orc_surface* s, const double* p3, int64_t n, double eps2,
                         uint8_t* out);
/* nearest_facet: closest point on the surface, squared distance and facet id */
int64_t orc_nearest(const orc_surface* s, const double* p, double* cp, double* d2);
/* brute-force versions (O(nf) per point) used to pin the tree traversal */
int orc_point_out_brute(const orc_surface* s, const double* p, double eps2);
double orc_min_d2_brute(const orc_surface* s, const double* p);

/* ---- (c) smoothing, src/tetwild/VertexSmoother.cpp:114-702 ------------- */
/* One frozen one-ring problem.  ring_tets: m tets (int32 x4, indices into
 * x/y/z), v_id the smoothed vertex.  Writes the new position (p_out) and returns
 * is_moved (NewtonsMethod's return value). */
int orc_newtons_method(const int32_t* ring_tets, int m, int32_t v_id, const double* x,
                       const double* y, const double* z, double* p_out);
/* NewtonsUpdate: E, J (3), H (9 row-major), X0.  Returns 0 when the reference
 * returns false. */
int orc_newtons_update(const int32_t* ring_tets, int m, int32_t v_id, const double* x,
                       const double* y, const double* z, double* E, double* J, double* H,
                       double* X0);
/* getNewEnergy over the ring in the given order */
double orc_new_energy(const int32_t* ring_tets, int m, const double* x, const double* y,
                      const double* z);
/* 3x3 solve by Householder QR with column pivoting: returns 0 if rank < 3 */
int orc_solve3_qrcp(const double* A, const double* b, double* x);

/* One vertex of smoothSingle (kind 0) or smoothSurface (kind 1: surface tree
 * projection, kind 2: boundary tree projection) on a frozen mesh.  ring_tets:
 * the m tets around v in conn_tets order; old_q: their stored slim energies;
 * tri_ab: the other two vertices of each one-ring surface face, in the
 * reference's rotation order.  Writes the final position (the old one unless
 * moved) and, for a moved surface vertex, the ring's new qualities. */
#define ORC_SMOOTH_MOVED 0
#define ORC_SMOOTH_INVALID 1
#define ORC_SMOOTH_NOT_MOVED 2
#define ORC_SMOOTH_FLIP_END 3
#define ORC_SMOOTH_REJECT_FLIP 4
#define ORC_SMOOTH_REJECT_QUALITY 5
#define ORC_SMOOTH_REJECT_ENVELOPE 6
int orc_smooth_problem(int kind, int32_t v, const int32_t* ring_tets, int m, const double* x,
                       const double* y, const double* z, const double* old_q,
                       const int32_t* tri_ab, int ntri, const orc_surface* sf,
                       const orc_surface* bt, double eps2, double d, double* p_out,
                       double* new_q);

/* orc_smooth_problem with near-tie accounting of its line search: risk[0..3]
 * are incremented (energy comparisons, those within the summation band of a
 * tie, |dE| < 1e-5 tests, those within the band of 1e-5; smooth.c states the
 * band).  risk may be NULL. */
int orc_smooth_problem_risk(int kind, int32_t v, const int32_t* ring_tets, int m,
                            const double* x, const double* y, const double* z,
                            const double* old_q, const int32_t* tri_ab, int ntri,
                            const orc_surface* sf, const orc_surface* bt, double eps2, double d,
                            double* p_out, double* new_q, int64_t* risk);

/* One vertex of smoothSingle (kind 0) or smoothSurface (kind 1 = surface,
 * 2 = boundary) with the reference's rounding rules, on a mesh whose vertices
 * carry posf (x, y, z), an exact position pos (px, py, pz; doubles here) and
 * is_rounded (rounded[]): the try-to-round step, isTetFlip on posf or pos
 * per tet, NewtonsMethod moving both, and the reference's commits.  Updates
 * the vertex's posf, pos and rounded flag in place as the reference leaves
 * them; new_q as orc_smooth_problem.  Returns ORC_SMOOTH_*. */
int orc_smooth_vertex(int kind, int32_t v, const int32_t* ring_tets, int m, double* x, double* y,
                      double* z, double* px, double* py, double* pz, uint8_t* rounded,
                      const double* old_q, const int32_t* tri_ab, int ntri,
                      const orc_surface* sf, const orc_surface* bt, double eps2, double d,
                      double* new_q);

/* ---- edge-collapse candidates, src/tetwild/EdgeCollapser.cpp:284-446 ---- */
/* return codes: the reference's (src/tetwild/EdgeCollapser.h:52-56) */
#define ORC_COLLAPSE_SUCCESS 0
#define ORC_COLLAPSE_FLIP 1
#define ORC_COLLAPSE_QUALITY 2
#define ORC_COLLAPSE_ENVELOP 3
#define ORC_COLLAPSE_ENVELOP_SUC 4
/* candidate flags (vertex state after the isolated / boundary-point updates) */
#define ORC_CV1_SURFACE 1
#define ORC_CV1_BOUNDARY 2
#define ORC_CV2_SURFACE 4
#define ORC_CV2_BOUNDARY 8
#define ORC_CV1_UNROUNDED 16
/* One candidate v1 -> v2.  ring_tets: v1's m one-ring tets (int32 x4);
 * old_q: their stored slim energies; tri_ids: ntri epsilon-check triangles
 * (vertex ids, v2 already substituted).  new_q (m) receives the new tets'
 * qualities (NaN for ring tets containing v2) once the quality stage is
 * reached.  Returns the reference's code. */
int orc_collapse_check(int32_t v1, int32_t v2, int flags, const int32_t* ring_tets, int m,
                       const double* x, const double* y, const double* z, const double* old_q,
                       const int32_t* tri_ids, int ntri, const orc_surface* sf,
                       const orc_surface* bt, double eps2, double d, int check_quality,
                       int check_eps, int is_soft, double soft_energy, double* new_q);

/* ---- energy reductions, src/tetwild/LocalOperations.cpp:605-693 ----------- */
/* sum / max / second max / count over tets neither removed nor locked (sum in
 * index order), and getFilterEnergy's 11-bucket histogram over non-removed
 * tets.  removed, tets + v_locked may be NULL. */
void orc_energy_stats(const double* e, int64_t n, const uint8_t* removed, const int32_t* tets,
                      const uint8_t* v_locked, double thres, double* sum, double* max,
                      double* second_max, int64_t* count, int64_t* buckets);
/* getFilterEnergy's decision from the histogram; sets *is_clean_up only where
 * the reference does */
double orc_filter_energy(const int64_t* buckets, double thres, int* is_clean_up);

/* ---- configurations (EdgeRemover), src/tetwild/EdgeRemover.cpp ------------ */
/* per group: isFlip; when not flipped the qualities (q may be NULL) and their
 * max from 0; NaN max when flipped */
void orc_tet_groups(const int32_t* tets, const int32_t* off, int64_t n_groups, const double* x,
                    const double* y, const double* z, uint8_t* flipped, double* max_q, double* q);

/* EdgeRemover::swap for one edge (v1, v2): ring = the m tets around the edge
 * (int32 x4, the reference's old_t_ids order), old_q their stored qualities.
 * Returns 3, 4 or 5 when removeAnEdge_32 / _44 / _56 succeeds (0 when none
 * does), with the new tets (up to 6) and their qualities in the order the
 * reference writes them; *equal_buget is updated as the reference updates it. */
int orc_remove_edge(int32_t v1, int32_t v2, const int32_t* ring, int m, const double* old_q,
                    const double* x, const double* y, const double* z, int* equal_buget,
                    int32_t* new_tets, double* new_q, int* n_new);

/* ---- CPU baseline: the batch loops above over host threads (baseline.c) --
 * OpenMP, schedule(dynamic); nthreads <= 0 uses every thread OpenMP offers.
 * Results are identical to the single-threaded loops (every item is
 * independent and computed by the same code). */
int orc_max_threads(void);
void orc_amips_batch_mt(const int32_t* tets, int64_t n, const double* x, const double* y,
                        const double* z, int mode, double* E, double* G, double* H,
                        int nthreads);
void orc_tet_qualities_mt(const int32_t* tets, int64_t n, const double* x, const double* y,
                          const double* z, double* slim_energy, int nthreads);
void orc_face_out_batch_mt(const orc_surface* s, const double* tri9, int64_t n, double eps2,
                           double d, uint8_t* out, int nthreads);
void orc_point_out_batch_mt(const orc_surface* s, const double* p3, int64_t n, double eps2,
                            uint8_t* out, int nthreads);
/* n one-ring problems (workloads.smoothing_batch layout); new_q is written
 * for the ring of every surface problem that moved; risk (4 counters, may be
 * NULL) sums orc_smooth_problem_risk's */
void orc_smooth_batch_mt(int64_t n, const uint8_t* kind, const int32_t* center,
                         const int32_t* ring_off, const int32_t* ring_tets, const double* x,
                         const double* y, const double* z, const double* old_q,
                         const int32_t* tri_off, const int32_t* tri_ab, const orc_surface* sf,
                         const orc_surface* bt, double eps2, double d, double* p_out,
                         uint8_t* status, double* new_q, int64_t* risk, int nthreads);
/* n collapse candidates (workloads.collapse_batch layout) */
void orc_collapse_batch_mt(int64_t n, const int32_t* v1, const int32_t* v2, const uint8_t* flags,
                           const int32_t* ring_off, const int32_t* ring_tets, const double* x,
                           const double* y, const double* z, const double* old_q,
                           const int32_t* tri_off, const int32_t* tri_ids,
                           const orc_surface* sf, const orc_surface* bt, double eps2, double d,
                           int check_quality, int check_eps, int is_soft, double soft_energy,
                           uint8_t* status, double* new_q, int nthreads);
/* Near-tie accounting for isFaceOutEnvelop over n triangles: every sample's
 * nearest squared distance (a full nearest search) against eps2 +- band.
 * counts[0] samples, [1] samples with |d2 - eps2| <= band, [2] triangles with
 * such a sample, [3] triangles whose decision rests on the band (no sample
 * with d2 > eps2 + band, and one inside the band). */
void orc_face_risk_batch_mt(const orc_surface* s, const double* tri9, int64_t n, double eps2,
                            double d, double band, int64_t* counts, int nthreads);
/* the same for isPointOutEnvelop: counts[0] points, [1] points with
 * |d2 - eps2| <= band (each such decision rests on the band) */
void orc_point_risk_batch_mt(const orc_surface* s, const double* p3, int64_t n, double eps2,
                             double band, int64_t* counts, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
