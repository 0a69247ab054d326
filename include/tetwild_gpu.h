/*
 * tetwild_gpu.h -- C-ABI of the MI355X mesh-optimisation inner loop
 * (libtetwild_amd.so).
 *
 * Plain pointers and sizes only.  Unless a function says otherwise, array
 * arguments are DEVICE pointers (HIP global memory) and the call is enqueued
 * on the context's stream and returns before the work finishes; call
 * twg_ctx_sync() (or synchronise the stream) before reading results.  Every
 * function returns TWG_OK (0) or a nonzero status; twg_last_error() gives the
 * message of the calling thread's last failure.  A context is not
 * thread-safe: callers serialise the calls they make on one context.
 *
 * Layouts (all little-endian, naturally aligned):
 *   vertices  three float64 arrays x[nv], y[nv], z[nv]          (SoA)
 *             a tet whose vertex ids fall outside [0, nv) gets NaN outputs
 *             (flip test: flipped) instead of an out-of-bounds read
 *   tets      int32 [n][4] vertex ids, 16-byte aligned          (AoS, 16 B)
 *   gradients float64 [n][3]                                    (per-tet rows)
 *   Hessians  float64 [n][9] row-major 3x3                      (per-tet rows)
 *   triangles float64 [n][9] = v0xyz v1xyz v2xyz                (AoS)
 *   points    float64 [n][3]                                    (AoS)
 *
 * Reference interfaces replaced (paths relative to the TetWild sources):
 *   see each declaration.
 */
#ifndef TETWILD_GPU_H
#define TETWILD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWG_OK 0
#define TWG_EINVAL 1   /* bad argument (null pointer, negative size, ...) */
#define TWG_EDEVICE 2  /* HIP runtime error */
#define TWG_ENOMEM 3   /* device allocation failed */

#define TWG_FP64 0     /* FP64 arithmetic, bit-exact with the CPU restatement */
#define TWG_MIXED 1    /* FP32 bulk with certified bounds, FP64 fallback */

#define TWG_AMIPS_E 0   /* energy */
#define TWG_AMIPS_EG 1  /* energy + gradient w.r.t. vertex 0 */
#define TWG_AMIPS_EGH 2 /* energy + gradient + Hessian w.r.t. vertex 0 */
/* as TWG_AMIPS_EGH with H[n][6]: the Hessian's upper triangle row by row
 * (xx, xy, xz, yy, yz, zz); TWG_AMIPS_EGH's H is symmetric bit for bit, so
 * these are its values, in 96 instead of 120 output bytes per tet */
#define TWG_AMIPS_EGH_SYM 3

#define TWG_MAX_ENERGY 1e50 /* State::MAX_ENERGY, src/tetwild/State.h:29 */

typedef struct twg_ctx twg_ctx;
typedef struct twg_bvh twg_bvh;

/* ---- context, memory ------------------------------------------------- */
/* Creates a context on HIP device `device` with its own (default-priority)
 * stream.  No reference counterpart: the reference runs on the host thread. */
int twg_ctx_create(int device, twg_ctx** out);
int twg_ctx_destroy(twg_ctx* ctx);
/* Enqueue later calls on an external HIP stream (hipStream_t; NULL is the HIP
 * null stream).  twg_ctx_stream() returns the stream in use (initially the
 * context's own).  The context's scratch is shared by all its calls, so the
 * new stream is made to wait (an event, no host wait) for the work already
 * enqueued on the old one -- unless either stream is capturing a graph, in
 * which case ordering the two is the caller's. */
int twg_ctx_set_stream(twg_ctx* ctx, void* hip_stream);
void* twg_ctx_stream(twg_ctx* ctx);
int twg_ctx_sync(twg_ctx* ctx);
const char* twg_last_error(void);
/* Device memory helpers for hosts that do not link HIP themselves.  The
 * copies are synchronous with respect to the host. */
int twg_malloc(twg_ctx* ctx, size_t bytes, void** dptr);
int twg_free(twg_ctx* ctx, void* dptr);
int twg_memcpy_h2d(twg_ctx* ctx, void* dst, const void* src, size_t bytes);
int twg_memcpy_d2h(twg_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- (a) AMIPS energy, gradient, Hessian ------------------------------ */
/* Per tet i (vertex ids tets[i][0..3]): conformal AMIPS energy E[i], its
 * gradient G and Hessian H w.r.t. vertex tets[i][0].
 * Replaces LocalOperations::comformalAMIPSEnergy_new / Jacobian_new /
 * Hessian_new (src/tetwild/LocalOperations.h:109-111,
 * src/tetwild/LocalOperations.cpp:28-291) and energy_ispc
 * (src/ispc/energy.ispc:7-65), batched.  mode: TWG_AMIPS_E / _EG / _EGH /
 * _EGH_SYM; G and H may be NULL when the mode does not write them.  FP64, bit-exact with the
 * CPU restatement; within ~1e-12 relative of the reference's code-generated
 * expressions on well-shaped tets. */
int twg_amips_eval(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                   const double* y, const double* z, int64_t nv, int mode, double* E, double* G,
                   double* H);

/* Mixed-precision variant: FP32 arithmetic from FP64 coordinates, FP32
 * outputs for G and H, and an energy E[i] (float64) that is either the FP32
 * value, certified to relative error <= tol by a forward error bound, or, where
 * the bound exceeds tol (ill-conditioned tets), the FP64 value.  fallback[i]
 * (may be NULL) is 1 where the FP64 value was used.  E is finite-or-inf exactly
 * where the FP64 energy is.  mode TWG_AMIPS_EGH_SYM: H[n][6], the upper
 * triangle (the FP32 Hessian is symmetric too). */
int twg_amips_eval_mixed(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                         const double* y, const double* z, int64_t nv, int mode, double tol,
                         double* E, float* G, float* H, uint8_t* fallback);

/* calTetQualities (src/tetwild/LocalOperations.h:69,
 * src/tetwild/LocalOperations.cpp:695-773 with calTetQuality_AMIPS :862-884):
 * slim_energy[i] = MAX_ENERGY if the exact orientation of the tet is not
 * POSITIVE, else the AMIPS energy with inf / nan / <= 0 mapped to MAX_ENERGY.
 * precision TWG_MIXED computes energies as twg_amips_eval_mixed with tol. */
int twg_tet_qualities(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                      const double* y, const double* z, int64_t nv, int precision, double tol,
                      double* slim_energy, uint8_t* fallback);

/* isTetFlip on rounded vertices (src/tetwild/LocalOperations.cpp:535-553):
 * flipped[i] = 1 unless the exact orientation of tet i is POSITIVE. */
int twg_tet_flip(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                 const double* y, const double* z, int64_t nv, uint8_t* flipped);

/* ---- (b) envelope ------------------------------------------------------ */
/* Builds the device AABB tree of a triangle surface.  V (nv x 3 float64) and
 * F (nf x 3 int32) are HOST arrays; facets are Morton-ordered and the tree
 * uses MeshFacetsAABBWithEps's implicit layout (root 1, children 2n / 2n+1,
 * split at b + (e-b)/2; src/tetwild/geogram/mesh_AABB.cpp:356-379, the
 * constructor of GEO::MeshFacetsAABBWithEps, which MeshRefinement::refine
 * builds at src/tetwild/MeshRefinement.cpp:209-213).  Synchronous.  Facet ids
 * out of [0, nv) are rejected. */
int twg_bvh_build(twg_ctx* ctx, const double* V, int64_t nv, const int32_t* F, int64_t nf,
                  twg_bvh** out);
int twg_bvh_destroy(twg_bvh* bvh);
int twg_bvh_info(const twg_bvh* bvh, int64_t* nf, int64_t* nnode, int32_t* depth, double* rmax);
/* order[b] = caller's facet id at tree position b (HOST array of nf) */
int twg_bvh_facet_order(const twg_bvh* bvh, int32_t* order);

/* isPointOutEnvelop / isPointOutBoundaryEnvelop
 * (src/tetwild/LocalOperations.cpp:1034-1044, 1111-1121): out[i] = 1 iff the
 * squared distance from point i to the surface exceeds eps2.  precision
 * TWG_MIXED traverses FP32 boxes (rounded outward) with conservative pruning;
 * decisions are identical to TWG_FP64.  stats (device, 4 x uint64, may be NULL)
 * accumulates samples, full traversals, boxes tested and facets tested. */
int twg_envelope_points(twg_ctx* ctx, const twg_bvh* bvh, const double* p3, int64_t n,
                        double eps2, int precision, uint8_t* out, unsigned long long* stats);

/* isFaceOutEnvelop / isFaceOutEnvelop_sampling
 * (src/tetwild/LocalOperations.cpp:967-975, 1046-1109): out[i] = 0 for a
 * degenerate (exactly collinear) triangle, else 1 iff some sampleTriangle
 * sample (spacing sampling_dist) lies farther than sqrt(eps2) from the
 * surface.  Guard: a triangle whose longest edge exceeds 65536 * sampling_dist
 * is reported out without sampling (its sample count would be ~10^9). */
int twg_envelope_faces(twg_ctx* ctx, const twg_bvh* bvh, const double* tri9, int64_t n,
                       double eps2, double sampling_dist, int precision, uint8_t* out,
                       unsigned long long* stats);

/* Debug counters of the dense envelope batches (no counterpart in the
 * reference): while enabled, the LDS-list kernels of twg_envelope_faces and
 * twg_envelope_points count their waves that tested the wave's facet list
 * (out[0]) and those that walked the tree instead (out[1]: the list
 * overflowed, or a coordinate is not finite).  enable = 1 (re)starts the
 * counts at zero (16 bytes of device memory, allocated once per context);
 * enable = 0 stops counting, waits for the context's stream and writes the
 * counts to out (host, may be NULL).  Decisions do not depend on it. */
int twg_ctx_list_stats(twg_ctx* ctx, int enable, unsigned long long* out);

/* MeshFacetsAABBWithEps::nearest_facet (src/tetwild/geogram/mesh_AABB.h:130-141,
 * mesh_AABB.cpp:381-480) with the reference's traversal order: closest point
 * cp3[i], squared distance d2[i] and facet id (caller's numbering after
 * twg_bvh_facet_order: facet[i] is a TREE POSITION). */
int twg_nearest_points(twg_ctx* ctx, const twg_bvh* bvh, const double* p3, int64_t n,
                       double* cp3, double* d2, int32_t* facet);

/* sampleTriangle (src/tetwild/Common.cpp:143-228), batched in two passes:
 * counts[i] = number of samples of triangle i; then, with offsets[i] the
 * exclusive prefix sum of counts and total their sum, pts[offsets[i] + k]
 * (float64 x 3) = k-th sample of triangle i in the reference's order. */
int twg_sample_count(twg_ctx* ctx, const double* tri9, int64_t n, double d, int64_t* counts);
int twg_sample_write(twg_ctx* ctx, const double* tri9, int64_t n, double d,
                     const int64_t* offsets, int64_t total, double* pts);

/* ---- (c) vertex smoothing ----------------------------------------------- */
/* Outcome of one one-ring problem (the `continue` points of the reference's
 * per-vertex loop bodies). */
#define TWG_SMOOTH_MOVED 0           /* moved to p_out */
#define TWG_SMOOTH_INVALID 1         /* a ring tet is not positively oriented */
#define TWG_SMOOTH_NOT_MOVED 2       /* NewtonsMethod returned false */
#define TWG_SMOOTH_FLIP_END 3        /* smoothSingle: flip at the final position */
#define TWG_SMOOTH_REJECT_FLIP 4     /* smoothSurface: flip after projection */
#define TWG_SMOOTH_REJECT_QUALITY 5  /* smoothSurface: max energy did not improve */
#define TWG_SMOOTH_REJECT_ENVELOPE 6 /* smoothSurface: a one-ring face left the envelope */
#define TWG_SMOOTH_BAD_INPUT 7       /* an index of this problem is out of range */

/* A batch of independent one-ring problems on a frozen mesh snapshot.  All
 * arrays are device arrays.  Problem p smooths vertex center[p] over the tets
 * ring_tets[ring_off[p] .. ring_off[p+1]) (conn_tets order, as the reference
 * gathers them, src/tetwild/VertexSmoother.cpp:145-150).  kind[p]: 0 = interior
 * vertex (smoothSingle body, :145-212), 1 = surface vertex (smoothSurface body,
 * :282-455, projection on sf_tree), 2 = surface vertex on the boundary
 * (projection on b_tree), 3 = NewtonsMethod alone (:465-542: status MOVED /
 * NOT_MOVED = its is_moved, p_out = its returned position).  For kinds 1 and 2: old_q holds the stored
 * slim_energy of each ring tet, tri_ab[tri_off[p] .. tri_off[p+1]) the other
 * two vertex ids of each one-ring surface face in the reference's order
 * (:395-402), and new_q receives the ring's new qualities when the vertex
 * moves.  Flip tests are taken on x/y/z: exact for a ring whose vertices
 * other than center[p] are rounded (pos == posf).  A ring with another
 * unrounded vertex needs exact rational orientations the device does not
 * have: the caller runs it through twg_newton_step with its own tests. */
typedef struct twg_smooth_desc {
    const double* x;
    const double* y;
    const double* z;
    int64_t nv;
    int64_t n;
    const int32_t* center;    /* [n] */
    const int32_t* ring_off;  /* [n + 1] */
    const int32_t* ring_tets; /* [n_ring][4], 16-byte aligned */
    int64_t n_ring;
    const uint8_t* kind;      /* [n] or NULL (all interior) */
    const double* old_q;      /* [n_ring] */
    const int32_t* tri_off;   /* [n + 1] */
    const int32_t* tri_ab;    /* [n_tri][2] */
    int64_t n_tri;
    const twg_bvh* sf_tree;
    const twg_bvh* b_tree;
    double eps2;
    double sampling_dist;
    double* p_out;            /* [n][3] final position (unchanged unless MOVED) */
    uint8_t* status;          /* [n] TWG_SMOOTH_* */
    double* new_q;            /* [n_ring] */
} twg_smooth_desc;

/* Runs every problem of the batch in one persistent launch: the per-vertex
 * bodies of VertexSmoother::smoothSingle / smoothSurface with NewtonsMethod,
 * NewtonsUpdate and getNewEnergy (src/tetwild/VertexSmoother.h:23-30,
 * src/tetwild/VertexSmoother.cpp:114-702).  Results are bit-identical to the
 * CPU restatement for any tile (16, 32, 64 lanes per problem, or
 * TWG_SMOOTH_TILE_SPLIT: 64 lanes for each problem of kind 1 or 2, which
 * the launch takes first, and 32 for the others; 0 = tuned or default) and
 * chunk (problems per queue pop; 0 = tuned or default).
 * precision TWG_MIXED screens line-search energies in FP32; results are
 * identical to TWG_FP64. */
#define TWG_SMOOTH_TILE_SPLIT 1
int twg_smooth_batch(twg_ctx* ctx, const twg_smooth_desc* desc, int precision, int tile,
                     int chunk);
/* twg_smooth_batch without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results as twg_smooth_batch; its only scratch, the persistent
 * kernel's two work-queue counters, lives in a 256-byte file-scope device array
 * (one per device, held until the process ends) that every launch clears
 * with a kernel first.  Limit: n below 2^31 - 2^24 problems, as for
 * twg_smooth_batch.  All calls on one GPU share that array: calls that
 * overlap in time on one GPU (several streams or contexts) race; serialise
 * them. */
int twg_smooth_batch_noalloc(twg_ctx* ctx, const twg_smooth_desc* desc, int precision, int tile,
                             int chunk);
/* Times every (tile, chunk) pair on this batch and remembers the fastest for
 * batches of a similar size on this device.  The outputs are written (by the
 * timing runs) exactly as twg_smooth_batch would write them. */
int twg_smooth_autotune(twg_ctx* ctx, const twg_smooth_desc* desc, int precision, int* tile,
                        int* chunk, float* ms);
/* The same launch captured once into a hipGraph (counter reset + persistent
 * kernel) and replayed by twg_smooth_plan_run on the context's stream; the
 * descriptor's arrays must stay valid while the plan lives. */
typedef struct twg_smooth_plan twg_smooth_plan;
int twg_smooth_plan_create(twg_ctx* ctx, const twg_smooth_desc* desc, int precision, int tile,
                           int chunk, twg_smooth_plan** out);
int twg_smooth_plan_run(twg_smooth_plan* plan);
int twg_smooth_plan_info(const twg_smooth_plan* plan, int* tile, int* chunk);
int twg_smooth_plan_destroy(twg_smooth_plan* plan);

/* For a caller that keeps the mesh on the device and runs batches of
 * independent problems one after another (no two problems of a batch share a
 * tet: the levels of the reference's vertex order, or colour classes), so
 * that no batch waits on the host.  No counterpart in the reference, whose
 * positions and qualities are host arrays updated between vertices.
 * twg_smooth_gather_q: desc->old_q[r] = q[ring_tid[r]] for the batch's
 * n_ring entries (NaN for an id outside [0, n_q)); desc->old_q must be
 * writable device memory.
 * twg_smooth_apply, after twg_smooth_batch on the same descriptor: for each
 * problem of kind 0, 1 or 2 with status TWG_SMOOTH_MOVED, writes p_out into
 * x, y, z at its centre and, for kinds 1 and 2 (q non-NULL), its new_q into
 * q[ring_tid[r]] over its ring -- the reference's updates of posf and
 * tet_qualities after a move (src/tetwild/VertexSmoother.cpp:196-211,
 * 441-453).  Enqueued on the context's stream like the batch. */
int twg_smooth_gather_q(twg_ctx* ctx, const twg_smooth_desc* desc, const int32_t* ring_tid,
                        const double* q, int64_t n_q);
int twg_smooth_apply(twg_ctx* ctx, const twg_smooth_desc* desc, const int32_t* ring_tid,
                     double* x, double* y, double* z, double* q, int64_t n_q);

/* One step of NewtonsMethod (src/tetwild/VertexSmoother.cpp:478-528) laid
 * out for a caller that decides some of its flip tests itself: the tets with
 * an unrounded vertex other than the moving one, which the reference orients
 * on exact rational positions (isTetFlip, src/tetwild/LocalOperations.cpp:
 * 535-553).  With vertex v at (px, py, pz): NewtonsUpdate over the m ring
 * tets (ok, energy, J, H: the sums twg_smooth_batch forms, bit for bit; ok =
 * -1 when an index is out of range), then every line-search candidate
 * X[k] = solve(H, H X0 - 2^-k J), k = 0..19, with finite[k], flip[k] (some
 * ring tet j with skip[j] == 0 is not positively oriented with v at X[k];
 * skip may be NULL) and new_energy[k] (getNewEnergy at X[k]; 0 when X[k] is
 * not finite).  The caller walks k in order as the halving loop does and
 * accepts the first finite candidate that no flip test rejects and that
 * lowers the energy.  Arrays are device arrays; out points to device memory. */
typedef struct twg_newton_step_out {
    double energy;
    double J[3];
    double H[9];
    double X[20][3];
    double new_energy[20];
    int32_t ok;
    uint8_t finite[20];
    uint8_t flip[20];
} twg_newton_step_out;
int twg_newton_step(twg_ctx* ctx, const double* x, const double* y, const double* z, int64_t nv,
                    const int32_t* ring_tets, int32_t m, int32_t v, double px, double py,
                    double pz, const uint8_t* skip, twg_newton_step_out* out);

/* ---- edge-collapse candidates -------------------------------------------- */
/* Return codes of EdgeCollapser::collapseAnEdge (src/tetwild/EdgeCollapser.h:52-56) */
#define TWG_COLLAPSE_SUCCESS 0
#define TWG_COLLAPSE_FLIP 1
#define TWG_COLLAPSE_QUALITY 2
#define TWG_COLLAPSE_ENVELOP 3
#define TWG_COLLAPSE_ENVELOP_SUC 4
#define TWG_COLLAPSE_BAD_INPUT 7 /* an index of this candidate is out of range */
/* Candidate flags: the vertices' state after collapseAnEdge's isolated and
 * boundary-point updates (src/tetwild/EdgeCollapser.cpp:298-307). */
#define TWG_CV1_SURFACE 1
#define TWG_CV1_BOUNDARY 2
#define TWG_CV2_SURFACE 4
#define TWG_CV2_BOUNDARY 8
#define TWG_CV1_UNROUNDED 16 /* v1 not rounded: its new quality counts as 0 (:408) */

/* A batch of independent collapse candidates v1 -> v2 on a frozen mesh.
 * ring_tets holds v1's one-ring (conn_tets, as 4 vertex ids); the tets that
 * contain v2 are the ones the collapse removes, the others become new tets
 * with v1 replaced by v2 in place.  tri_ids holds, per candidate, the
 * triangles isCollapsable_epsilon tests (src/tetwild/EdgeCollapser.cpp:742-764:
 * unique sorted one-ring surface faces of v1 without v2, v1 replaced by v2).
 * Every vertex of the new tets and triangles must be rounded (posf). */
typedef struct twg_collapse_desc {
    const double* x;
    const double* y;
    const double* z;
    int64_t nv;
    int64_t n;
    const int32_t* v1;        /* [n] */
    const int32_t* v2;        /* [n] */
    const uint8_t* flags;     /* [n] TWG_CV* bits, or NULL (all interior, rounded) */
    const int32_t* ring_off;  /* [n + 1] */
    const int32_t* ring_tets; /* [n_ring][4], 16-byte aligned */
    int64_t n_ring;
    const double* old_q;      /* [n_ring] stored slim_energy of the ring tets */
    const int32_t* tri_off;   /* [n + 1], or NULL (no epsilon triangles) */
    const int32_t* tri_ids;   /* [n_tri][3] */
    int64_t n_tri;
    const twg_bvh* sf_tree;   /* needed when a candidate has TWG_CV1_SURFACE */
    const twg_bvh* b_tree;    /* needed when a candidate has TWG_CV1_BOUNDARY */
    double eps2;
    double sampling_dist;
    int check_quality;        /* energy_type != ENERGY_NA && is_check_quality */
    int check_eps;            /* state.eps is neither EPSILON_NA nor EPSILON_INFINITE */
    int is_soft;              /* EdgeCollapser::is_soft */
    double soft_energy;       /* EdgeCollapser::soft_energy */
    uint8_t* status;          /* [n] TWG_COLLAPSE_* */
    double* new_q;            /* [n_ring] new slim_energy per ring tet (NaN for the
                                 removed ones); what the caller stores on success */
} twg_collapse_desc;

/* Evaluates every candidate of the batch: the checks of collapseAnEdge up to
 * its "real update" (src/tetwild/EdgeCollapser.cpp:284-446) -- boundary and
 * surface point envelope of v2, isFlip and calTetQualities of the new tets,
 * getCheckQuality / isBetterOrEqualThan, isCollapsable_epsilon -- with the
 * reference's return code.  The one-ring checks run in one persistent launch
 * (tile = 16, 32 or 64 lanes per candidate, chunk = candidates per queue pop;
 * 0 = tuned or default).  Codes and new_q are identical to the CPU
 * restatement for any tile and chunk.  precision TWG_MIXED walks the envelope
 * trees with FP32 boxes (identical decisions).  new_q is complete when the
 * status is SUCCESS or ENVELOP_SUC. */
int twg_collapse_batch(twg_ctx* ctx, const twg_collapse_desc* desc, int precision, int tile,
                       int chunk);
/* Times every (tile, chunk) pair on this batch and remembers the fastest for
 * batches of a similar size on this device; the outputs are written as
 * twg_collapse_batch would write them. */
int twg_collapse_autotune(twg_ctx* ctx, const twg_collapse_desc* desc, int precision, int* tile,
                          int* chunk, float* ms);
/* The batch's launches captured once into a hipGraph (queue reset, persistent
 * ring pass, exact ring pass, point, face, exact face and final passes) and
 * replayed by twg_collapse_plan_run on the context's stream.  The plan owns
 * its scratch; the descriptor's arrays must stay valid while the plan lives. */
typedef struct twg_collapse_plan twg_collapse_plan;
int twg_collapse_plan_create(twg_ctx* ctx, const twg_collapse_desc* desc, int precision, int tile,
                             int chunk, twg_collapse_plan** out);
int twg_collapse_plan_run(twg_collapse_plan* plan);
int twg_collapse_plan_info(const twg_collapse_plan* plan, int* tile, int* chunk);
int twg_collapse_plan_destroy(twg_collapse_plan* plan);

/* ---- configurations of a few tets (edge removal) ------------------------ */
#define TWG_GROUP_BAD_INPUT 2 /* flipped[g]: an index of the group is out of range */

/* For every group g of tets (tets[group_off[g] .. group_off[g+1]), int32 x4):
 * flipped[g] = isFlip(group) (src/tetwild/LocalOperations.cpp:555-573, exact
 * orientation on posf); when not flipped, q[k] = calTetQualities of each tet
 * (:695-773, 862-884; q may be NULL) and max_q[g] = getCheckQuality's max
 * (:575-588, max over the group starting from 0), else max_q[g] = NaN.  The
 * evaluation step of EdgeRemover::removeAnEdge_32 / _44 / _56
 * (src/tetwild/EdgeRemover.cpp:138-174, 327-368, 501-575): one group per
 * configuration; the caller keeps the selection logic.  One thread per group:
 * meant for small groups (a few tets).  Identical to the CPU restatement. */
int twg_tet_groups(twg_ctx* ctx, const int32_t* tets, const int32_t* group_off, int64_t n_groups,
                   int64_t n_tets, const double* x, const double* y, const double* z, int64_t nv,
                   uint8_t* flipped, double* max_q, double* q);

/* ---- energy reductions ---------------------------------------------------- */
typedef struct twg_energy_summary {
    double sum;          /* sum of slim_energy over counted tets */
    double max;          /* getMaxEnergy: max over counted tets (from 0) */
    double second_max;   /* getSecondMaxEnergy: same, excluding MAX_ENERGY */
    int64_t count;       /* tets neither removed nor locked */
    int64_t buckets[11]; /* getFilterEnergy's histogram (removed tets skipped,
                            locks ignored, as in the reference) */
} twg_energy_summary;

/* One pass over the per-tet slim energies (device arrays): the quantities of
 * LocalOperations::getAvgMaxEnergy, getMaxEnergy, getSecondMaxEnergy and
 * getFilterEnergy (src/tetwild/LocalOperations.cpp:605-693) that drive pass
 * termination in MeshRefinement (src/tetwild/MeshRefinement.cpp:185-206,
 * 285-292).  t_is_removed (uint8, may be NULL), tets + v_is_locked (uint8 per
 * vertex; both NULL = nothing locked) give isTetLocked_ui (:1377-1384).  The
 * bucket limits are filter_energy_thres - 1 + 10^j as the reference forms
 * them.  avg = sum / count (inf -> MAX_ENERGY) is the caller's.  Max, second
 * max, count and buckets are exact; sum is a deterministic reduction (within
 * n 2^-53 sum|e| of the reference's sequential sum).  Synchronous: *out is a
 * HOST struct, filled when the call returns. */
int twg_energy_stats(twg_ctx* ctx, const double* slim_energy, int64_t n,
                     const uint8_t* t_is_removed, const int32_t* tets, const uint8_t* v_is_locked,
                     int64_t nv, double filter_energy_thres, twg_energy_summary* out);

#ifdef __cplusplus
}
#endif
#endif
