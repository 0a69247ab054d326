// collapse.hip -- batched evaluation of edge-collapse candidates: the checks
// EdgeCollapser::collapseAnEdge (src/tetwild/EdgeCollapser.cpp:284-446) runs
// before its "real update", for many (v1 -> v2) candidates on a frozen mesh.
//
// The reference returns at the first failing check, in the order: boundary
// point envelope of v2, surface point envelope of v2, isFlip of the new tets,
// the getCheckQuality comparison, isCollapsable_epsilon.  Every check is a
// pure function of the frozen mesh, so the batch evaluates them in four
// launches, each uniform in its own work, and combines the outcomes in that
// order:
//   1. ring pass   a tile of TILE lanes per candidate (lanes take the ring's
//                  tets): input validation, the degenerate-edge test, isFlip,
//                  calTetQualities (written to new_q), the quality decision;
//                  marks the point checks and the epsilon check still open and
//                  maps the candidate's triangles to it.  Orientation comes
//                  from the floating-point filter; a candidate with no tet
//                  certainly flipped but some tet undecided is listed, and
//   1b.            the exact ring pass redoes the listed candidates with the
//                  exact predicate (the same code, so the same results).
//   2. point pass  one thread per candidate: the open point checks.
//   3. face pass   a wave per triangle of a candidate whose epsilon check is
//                  still open (its samples spread over the lanes):
//                  isFaceOutEnvelop; triangles whose collinearity the filter
//                  cannot decide are marked, and
//   3b.            the exact face pass finishes them.
//   4. final pass  one thread per candidate: the reference's return code.
// The exact predicates stay out of passes 1 and 3, which keeps their register
// count and stack small.  Candidates rejected early still pay for the checks of
// pass 1 (flip and quality of a ring), never for envelope checks they would
// not reach.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

#include "twg_bvh.hpp"
#include "twg_internal.hpp"

namespace twg {
namespace {

// per-candidate work bits (context scratch, one u32 per candidate)
constexpr unsigned kNeedBoundaryPoint = 1u;
constexpr unsigned kNeedSurfacePoint = 2u;
constexpr unsigned kFacesOpen = 4u;
constexpr unsigned kFaceOut = 8u;
constexpr unsigned kNeedExact = 16u;  // the ring pass must be redone exactly

struct CollapseArgs {
    const double* __restrict__ x;
    const double* __restrict__ y;
    const double* __restrict__ z;
    int64_t nv;
    int64_t n;
    const int32_t* __restrict__ v1;
    const int32_t* __restrict__ v2;
    const uint8_t* __restrict__ flags;
    const int32_t* __restrict__ ring_off;
    const int4* __restrict__ ring;
    int64_t n_ring;
    const double* __restrict__ old_q;
    const int32_t* __restrict__ tri_off;
    const int32_t* __restrict__ tri_ids;
    int64_t n_tri;
    BvhView sf, bt;
    bool has_sf, has_bt;
    double eps2, d;
    double T_sf, sqrtT_sf, T_bt, sqrtT_bt;
    bool check_quality, check_eps, is_soft;
    double soft_energy;
    uint8_t* __restrict__ status;
    double* __restrict__ new_q;
    unsigned* __restrict__ work;
    int32_t* __restrict__ tri_cand;
    unsigned* counter;     // [0] the ring pass's queue, [kExactCount] exact-list length
    int32_t* exact_list;   // candidates the ring pass left to the exact pass
    unsigned chunk;        // candidates per queue pop
};

constexpr int kExactCount = 16;  // counter slot (64 bytes after the queue)

__device__ __forceinline__ V3<double> vtx(const CollapseArgs& A, int id)
{
    return mk3(A.x[id], A.y[id], A.z[id]);
}

template <int TILE>
__device__ __forceinline__ bool tile_any(int wbase, bool pred)
{
    unsigned long long b = __ballot(pred);
    unsigned long long m = TILE == 64 ? ~0ull : ((1ull << (TILE & 63)) - 1ull);
    return ((b >> wbase) & m) != 0ull;
}

template <int TILE>
__device__ __forceinline__ double tile_max(double v)
{
    for (int off = TILE / 2; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, TILE);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ bool has_vertex(int4 t, int v)
{
    return t.x == v || t.y == v || t.z == v || t.w == v;
}

// the new tet: t with v1 replaced by v2 (at position p2), in t's vertex order
__device__ __forceinline__ void new_tet(const CollapseArgs& A, int4 t, int v1, V3<double> p2, Tet& T)
{
    int id[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        V3<double> q = id[j] == v1 ? p2 : vtx(A, id[j]);
        T.p[j][0] = q.x;
        T.p[j][1] = q.y;
        T.p[j][2] = q.z;
    }
}

// calTetQualities + calTetQuality_AMIPS (src/tetwild/LocalOperations.cpp:695-773, 862-884)
// for a tet known to be positively oriented (isFlip of the new tets passed)
__device__ __forceinline__ double tet_quality_positive(const Tet& T)
{
    double e = amips_energy(T);
    if (is_inf(e) || is_nan(e) || e <= 0) e = kMaxEnergy;
    return e;
}

// The common case of ring_candidate (at most TILE ring tets, so lane tl owns
// tet tl): one sweep validates the lane's tet, then orients it and evaluates
// its quality; the qualities are written only when no tet is flipped or
// undecided, as in the sweep-per-check order.  Same outputs as the general
// path below.
template <int TILE, bool EXACT>
__device__ void ring_candidate_fused(const CollapseArgs& A, int64_t c, int tl, int wbase, int v1,
                                     int v2, unsigned fl, int m, const int4* rt, int64_t r0,
                                     int64_t t0, int64_t t1)
{
    const bool v1_sf = fl & TWG_CV1_SURFACE, v1_bd = fl & TWG_CV1_BOUNDARY;
    const bool v2_sf = fl & TWG_CV2_SURFACE, v2_bd = fl & TWG_CV2_BOUNDARY;
    const bool mine = tl < m;
    const int4 t = mine ? rt[tl] : make_int4(0, 0, 0, 0);
    bool bid = mine && (t.x < 0 || t.x >= A.nv || t.y < 0 || t.y >= A.nv || t.z < 0 ||
                        t.z >= A.nv || t.w < 0 || t.w >= A.nv || !has_vertex(t, v1));
    for (int64_t k = t0 + tl; k < t1; k += TILE) {
        int a = A.tri_ids[3 * k], b = A.tri_ids[3 * k + 1], e = A.tri_ids[3 * k + 2];
        bid = bid || a < 0 || a >= A.nv || b < 0 || b >= A.nv || e < 0 || e >= A.nv;
    }
    unsigned work = 0;
    uint8_t st = TWG_COLLAPSE_SUCCESS;
    const bool bad = tile_any<TILE>(wbase, bid);
    if (bad) {
        st = TWG_COLLAPSE_BAD_INPUT;
    } else {
        const V3<double> p1 = vtx(A, v1), p2 = vtx(A, v2);
        const double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
        const bool degenerate = (dx * dx + dy * dy + dz * dz) == 0.0;
        if (v1_bd && !v2_bd && !degenerate) work |= kNeedBoundaryPoint;
        if (v1_sf && !v2_sf && !degenerate) work |= kNeedSurfacePoint;
        const bool removed = mine && has_vertex(t, v2);
        bool flip = false, undecided = false;
        double q = __builtin_nan("");
        if (mine && !removed) {
            Tet T;
            new_tet(A, t, v1, p2, T);
            const int o = EXACT ? orient3d(T.p[0], T.p[1], T.p[2], T.p[3])
                                : orient3d_filter(T.p[0], T.p[1], T.p[2], T.p[3]);
            if (o == kUndecided) undecided = true;
            else flip = o != 1;
            q = tet_quality_positive(T);
        }
        if (tile_any<TILE>(wbase, flip)) {
            st = TWG_COLLAPSE_FLIP;
        } else if (!EXACT && tile_any<TILE>(wbase, undecided)) {
            work = kNeedExact;
        } else {
            double om = mine ? A.old_q[r0 + tl] : 0.0;
            om = om > 0.0 ? om : 0.0;
            double nm = mine && !removed ? (q > 0.0 ? q : 0.0) : 0.0;
            if (mine) A.new_q[r0 + tl] = q;
            om = tile_max<TILE>(om);
            nm = tile_max<TILE>(nm);
            if (A.check_quality) {
                if (A.is_soft) om = A.soft_energy;
                if (fl & TWG_CV1_UNROUNDED) nm = 0.0;
            }
            if (A.check_quality && !degenerate && !(nm <= om)) {
                st = TWG_COLLAPSE_QUALITY;
            } else if (A.check_eps && v1_sf) {
                st = TWG_COLLAPSE_ENVELOP_SUC;
                if (!degenerate && t1 > t0) work |= kFacesOpen;
            }
        }
    }
    if (!bad)
        for (int64_t k = t0 + tl; k < t1; k += TILE)
            A.tri_cand[k] = (work & kFacesOpen) ? (int32_t)c : -1;
    if (tl == 0) {
        A.status[c] = st;
        A.work[c] = work;
        if (!EXACT && (work & kNeedExact)) {
            const unsigned k = atomicAdd(A.counter + kExactCount, 1u);
            A.exact_list[k] = (int32_t)c;
        }
    }
}

// EXACT: orientations from the exact predicate; otherwise from the filter,
// with undecided rings marked kNeedExact
template <int TILE, bool EXACT>
__device__ void ring_candidate(const CollapseArgs& A, int64_t c, int tl, int wbase)
{
    const int v1 = A.v1[c], v2 = A.v2[c];
    const unsigned fl = A.flags ? A.flags[c] : 0u;
    const int64_t r0 = A.ring_off[c], r1 = A.ring_off[c + 1];
    int64_t t0 = 0, t1 = 0;
    if (A.tri_off) {
        t0 = A.tri_off[c];
        t1 = A.tri_off[c + 1];
    }
    const bool v1_sf = fl & TWG_CV1_SURFACE, v1_bd = fl & TWG_CV1_BOUNDARY;
    const bool v2_sf = fl & TWG_CV2_SURFACE, v2_bd = fl & TWG_CV2_BOUNDARY;
    // input validation: every index read below stays inside its array
    bool bad = v1 < 0 || v1 >= A.nv || v2 < 0 || v2 >= A.nv || v1 == v2 || r0 < 0 || r1 < r0 ||
               r1 > A.n_ring || t0 < 0 || t1 < t0 || t1 > A.n_tri ||
               (v1_sf && !A.has_sf) || (v1_bd && !A.has_bt);
    const int m = bad ? 0 : (int)(r1 - r0);
    const int4* rt = A.ring + r0;
    if (!bad && m <= TILE) {
        ring_candidate_fused<TILE, EXACT>(A, c, tl, wbase, v1, v2, fl, m, rt, r0, t0, t1);
        return;
    }
    if (!bad) {
        bool bid = false;
        for (int j = tl; j < m; j += TILE) {
            int4 t = rt[j];
            bid = bid || t.x < 0 || t.x >= A.nv || t.y < 0 || t.y >= A.nv || t.z < 0 ||
                  t.z >= A.nv || t.w < 0 || t.w >= A.nv || !has_vertex(t, v1);
        }
        for (int64_t k = t0 + tl; k < t1; k += TILE) {
            int a = A.tri_ids[3 * k], b = A.tri_ids[3 * k + 1], e = A.tri_ids[3 * k + 2];
            bid = bid || a < 0 || a >= A.nv || b < 0 || b >= A.nv || e < 0 || e >= A.nv;
        }
        bad = tile_any<TILE>(wbase, bid);
    }
    unsigned work = 0;
    uint8_t st = TWG_COLLAPSE_SUCCESS;
    if (bad) {
        st = TWG_COLLAPSE_BAD_INPUT;
    } else {
        const V3<double> p1 = vtx(A, v1), p2 = vtx(A, v2);
        const double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
        const bool degenerate = (dx * dx + dy * dy + dz * dz) == 0.0;
        if (v1_bd && !v2_bd && !degenerate) work |= kNeedBoundaryPoint;
        if (v1_sf && !v2_sf && !degenerate) work |= kNeedSurfacePoint;
        // isFlip(new_tets): a certain flip decides; otherwise an undecided tet
        // sends the candidate to the exact pass
        bool flip = false, undecided = false;
        for (int j = tl; j < m && !flip; j += TILE) {
            int4 t = rt[j];
            if (has_vertex(t, v2)) continue;
            Tet T;
            new_tet(A, t, v1, p2, T);
            const int o = EXACT ? orient3d(T.p[0], T.p[1], T.p[2], T.p[3])
                                : orient3d_filter(T.p[0], T.p[1], T.p[2], T.p[3]);
            if (o == kUndecided) undecided = true;
            else flip = o != 1;
        }
        if (tile_any<TILE>(wbase, flip)) {
            st = TWG_COLLAPSE_FLIP;
        } else if (!EXACT && tile_any<TILE>(wbase, undecided)) {
            work = kNeedExact;
        } else {
            // calTetQualities(new_tets); getCheckQuality of old and new
            double om = 0.0, nm = 0.0;
            for (int j = tl; j < m; j += TILE) {
                int4 t = rt[j];
                double o = A.old_q[r0 + j];
                om = o > om ? o : om;
                double q;
                if (has_vertex(t, v2)) {
                    q = __builtin_nan("");
                } else {
                    Tet T;
                    new_tet(A, t, v1, p2, T);
                    q = tet_quality_positive(T);
                    nm = q > nm ? q : nm;
                }
                A.new_q[r0 + j] = q;
            }
            om = tile_max<TILE>(om);
            nm = tile_max<TILE>(nm);
            if (A.check_quality) {
                if (A.is_soft) om = A.soft_energy;
                if (fl & TWG_CV1_UNROUNDED) nm = 0.0;
            }
            if (A.check_quality && !degenerate && !(nm <= om)) {
                st = TWG_COLLAPSE_QUALITY;
            } else if (A.check_eps && v1_sf) {
                st = TWG_COLLAPSE_ENVELOP_SUC;
                if (!degenerate && t1 > t0) work |= kFacesOpen;
            }
        }
    }
    // triangles of this candidate -> candidate id while its epsilon check is open
    if (!bad)
        for (int64_t k = t0 + tl; k < t1; k += TILE)
            A.tri_cand[k] = (work & kFacesOpen) ? (int32_t)c : -1;
    if (tl == 0) {
        A.status[c] = st;
        A.work[c] = work;
        if (!EXACT && (work & kNeedExact)) {
            const unsigned k = atomicAdd(A.counter + kExactCount, 1u);
            A.exact_list[k] = (int32_t)c;
        }
    }
}

// Persistent: tiles pop chunks of candidates from an atomic queue, so rings
// of 4..200 tets balance across the machine (the queue start is uniform
// across a tile).
template <int TILE>
__global__ __launch_bounds__(kBlock) void collapse_ring_kernel(CollapseArgs A)
{
    const int tl = threadIdx.x % TILE;
    const int wbase = (threadIdx.x % 64) - tl;
    for (;;) {
        unsigned start = 0;
        if (tl == 0) start = atomicAdd(A.counter, A.chunk);
        start = __shfl(start, 0, TILE);
        if ((int64_t)start >= A.n) break;
        int64_t end = (int64_t)start + A.chunk;
        if (end > A.n) end = A.n;
        for (int64_t c = start; c < end; ++c) ring_candidate<TILE, false>(A, c, tl, wbase);
    }
}

// the candidates the filter pass listed, with the exact predicate; a tile per
// candidate, grid-stride over the list (the loop bound is uniform across a tile)
template <int TILE>
__global__ __launch_bounds__(kBlock) void collapse_ring_exact_kernel(CollapseArgs A)
{
    const int tl = threadIdx.x % TILE;
    const int wbase = (threadIdx.x % 64) - tl;
    const int64_t stride = (int64_t)gridDim.x * (kBlock / TILE);
    const int64_t listed = A.counter[kExactCount];
    for (int64_t k = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / TILE; k < listed; k += stride)
        ring_candidate<TILE, true>(A, A.exact_list[k], tl, wbase);
}

__global__ void reset_queue(unsigned* c)
{
    c[0] = 0u;
    c[kExactCount] = 0u;
}

template <bool F32>
__global__ __launch_bounds__(kBlock) void collapse_point_kernel(CollapseArgs A)
{
    const int64_t c = xcd_block() * (int64_t)kBlock + threadIdx.x;
    if (c >= A.n) return;
    const unsigned w = A.work[c];
    if (!(w & (kNeedBoundaryPoint | kNeedSurfacePoint))) return;
    const V3<double> p2 = vtx(A, A.v2[c]);
    Stats st;
    bool out = false;
    if (w & kNeedBoundaryPoint) {
        int hint = -1;
        out = !in_envelope<F32, false>(walk(A.bt), p2, A.eps2, A.T_bt, A.sqrtT_bt, hint, st);
    }
    if (!out && (w & kNeedSurfacePoint)) {
        int hint = -1;
        out = !in_envelope<F32, false>(walk(A.sf), p2, A.eps2, A.T_sf, A.sqrtT_sf, hint, st);
    }
    if (out) {
        A.status[c] = TWG_COLLAPSE_ENVELOP;
        A.work[c] = 0u;  // nothing after the point checks runs
    }
}

// lanes per triangle in the face pass (samples spread over the group)
constexpr int kFaceGroup = 64;
// facets listed per wave in the face pass (a triangle's neighbourhood)
constexpr int kFaceListCap = 64;
constexpr int kFaceListBytes = kFaceListCap * (32 + 8 + 72);

template <bool F32>
__global__ __launch_bounds__(kBlock) void collapse_face_kernel(CollapseArgs A)
{
    __shared__ __align__(16) unsigned char face_lists[(kBlock / 64) * kFaceListBytes];
    const int r = threadIdx.x % kFaceGroup;
    const int gbase = (threadIdx.x % 64) - r;
    const int64_t k = (xcd_block() * (int64_t)kBlock + threadIdx.x) / kFaceGroup;
    if (k >= A.n_tri) return;
    const int64_t c = A.tri_cand[k];
    // stale entries (from earlier batches) fail the range test
    if (c < 0 || c >= A.n) return;
    if (k < A.tri_off[c] || k >= A.tri_off[c + 1]) return;
    unsigned w = __hip_atomic_load(A.work + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((w & (kFacesOpen | kFaceOut)) != kFacesOpen) return;
    double t9[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        V3<double> p = vtx(A, A.tri_ids[3 * k + j]);
        t9[3 * j] = p.x;
        t9[3 * j + 1] = p.y;
        t9[3 * j + 2] = p.z;
    }
    if (!certainly_not_collinear(t9, t9 + 3, t9 + 6)) {
        if (r == 0) A.tri_cand[k] = -2 - (int32_t)c;  // for the exact face pass
        return;
    }
    // the facets near the triangle, listed in the wave's LDS (collect_facets);
    // the tree walk when they do not fit (without the subtree step after a
    // hint miss: on the 1M-candidate sheet workload it ran 6% slower)
    const BvhWalk W = walk(A.sf);
    unsigned char* mine = face_lists + (threadIdx.x / 64) * kFaceListBytes;
    LdsNode32W* lst = (LdsNode32W*)mine;
    LdsU32* fr = (LdsU32*)(mine + kFaceListCap * 32);
    LdsF64* tri = (LdsF64*)(mine + kFaceListCap * 40);
    double lo[3], hi[3], amax = 0.0;
    bool fin = true;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        lo[j] = fmin(t9[j], fmin(t9[3 + j], t9[6 + j]));
        hi[j] = fmax(t9[j], fmax(t9[3 + j], t9[6 + j]));
        fin = fin && is_finite(t9[j]) && is_finite(t9[3 + j]) && is_finite(t9[6 + j]);
        amax = fmax(amax, fmax(fabs(lo[j]), fabs(hi[j])));
    }
    const double pad = 1.0001 * A.sqrtT_sf + 1e-9 * amax;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        lo[j] -= pad;
        hi[j] += pad;
    }
    const int nl = fin ? collect_facets<64>(W, lo, hi, lst, fr, kFaceListCap, r, 0) : -1;
    for (int e = r; e < nl; e += 64) {
        const double* f = W.tri + 9 * (int64_t)lst[e].leaf;
#pragma unroll
        for (int q = 0; q < 9; ++q) tri[9 * e + q] = f[q];
    }
    wave_sync();
    if (face_out_group<F32, kFaceGroup, false>(W, ld3(t9), ld3(t9 + 3), ld3(t9 + 6), A.eps2, A.d,
                                               A.T_sf, A.sqrtT_sf, r, gbase, lst, tri, nl) &&
        r == 0)
        atomicOr(A.work + c, kFaceOut);
}

// triangles the collinearity filter left undecided (tri_cand = -2 - candidate)
template <bool F32>
__global__ __launch_bounds__(kBlock) void collapse_face_exact_kernel(CollapseArgs A)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < A.n_tri; k += stride) {
        const int64_t c = -2 - (int64_t)A.tri_cand[k];
        if (c < 0 || c >= A.n) continue;
        if (k < A.tri_off[c] || k >= A.tri_off[c + 1]) continue;
        unsigned w = __hip_atomic_load(A.work + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((w & (kFacesOpen | kFaceOut)) != kFacesOpen) continue;
        double t9[9];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            V3<double> p = vtx(A, A.tri_ids[3 * k + j]);
            t9[3 * j] = p.x;
            t9[3 * j + 1] = p.y;
            t9[3 * j + 2] = p.z;
        }
        Stats st;
        if (face_out<F32, false>(walk(A.sf), t9, A.eps2, A.d, A.T_sf, A.sqrtT_sf, st))
            atomicOr(A.work + c, kFaceOut);
    }
}

__global__ __launch_bounds__(kBlock) void collapse_final_kernel(CollapseArgs A)
{
    const int64_t c = xcd_block() * (int64_t)kBlock + threadIdx.x;
    if (c >= A.n) return;
    const unsigned w = A.work[c];
    if (w & kFacesOpen) A.status[c] = (w & kFaceOut) ? TWG_COLLAPSE_ENVELOP : TWG_COLLAPSE_ENVELOP_SUC;
}

int make_args(twg_ctx* ctx, const twg_collapse_desc* d, CollapseArgs& A)
{
    TWG_REQUIRE(d != nullptr, "collapse: NULL descriptor");
    TWG_REQUIRE(d->n >= 0 && d->n < (int64_t(1) << 31) - (int64_t(1) << 24),
                "collapse: n out of range");
    TWG_REQUIRE(d->nv >= 0 && d->n_ring >= 0 && d->n_tri >= 0 && d->n_ring < (int64_t(1) << 31) &&
                    d->n_tri < (int64_t(1) << 31),
                "collapse: size out of range");
    TWG_REQUIRE(d->n == 0 || (d->x && d->y && d->z && d->v1 && d->v2 && d->ring_off &&
                              d->ring_tets && d->old_q && d->status && d->new_q),
                "collapse: NULL array");
    TWG_REQUIRE(d->n_tri == 0 || (d->tri_off && d->tri_ids), "collapse: NULL triangle array");
    TWG_REQUIRE(((uintptr_t)d->ring_tets & 15) == 0, "collapse: ring_tets must be 16-byte aligned");
    TWG_REQUIRE(!d->sf_tree || d->sf_tree->device == ctx->device,
                "collapse: sf_tree on another device");
    TWG_REQUIRE(!d->b_tree || d->b_tree->device == ctx->device, "collapse: b_tree on another device");
    TWG_REQUIRE(!d->check_eps || d->n_tri == 0 || d->sampling_dist > 0,
                "collapse: sampling_dist must be positive");
    A.x = d->x;
    A.y = d->y;
    A.z = d->z;
    A.nv = d->nv;
    A.n = d->n;
    A.v1 = d->v1;
    A.v2 = d->v2;
    A.flags = d->flags;
    A.ring_off = d->ring_off;
    A.ring = reinterpret_cast<const int4*>(d->ring_tets);
    A.n_ring = d->n_ring;
    A.old_q = d->old_q;
    A.tri_off = d->n_tri ? d->tri_off : nullptr;
    A.tri_ids = d->tri_ids;
    A.n_tri = d->tri_off ? d->n_tri : 0;
    A.has_sf = d->sf_tree != nullptr;
    A.has_bt = d->b_tree != nullptr;
    BvhView none = no_bvh();
    A.sf = d->sf_tree ? bvh_view(d->sf_tree) : none;
    A.bt = d->b_tree ? bvh_view(d->b_tree) : none;
    A.eps2 = d->eps2;
    A.d = d->sampling_dist;
    A.T_sf = d->sf_tree ? bvh_prune_threshold(d->sf_tree, d->eps2) : 0.0;
    A.sqrtT_sf = std::sqrt(A.T_sf);
    A.T_bt = d->b_tree ? bvh_prune_threshold(d->b_tree, d->eps2) : 0.0;
    A.sqrtT_bt = std::sqrt(A.T_bt);
    A.check_quality = d->check_quality != 0;
    A.check_eps = d->check_eps != 0;
    A.is_soft = d->is_soft != 0;
    A.soft_energy = d->soft_energy;
    A.status = d->status;
    A.new_q = d->new_q;
    return TWG_OK;
}

// lanes per candidate from the mean ring size (when not tuned)
int default_tile(const CollapseArgs& A)
{
    double mean = A.n ? (double)A.n_ring / (double)A.n : 0.0;
    return mean <= 12.0 ? 16 : (mean <= 40.0 ? 32 : 64);
}

// bytes of per-batch scratch: work[n], tri_cand[n_tri], the counters (256
// bytes), exact_list[n]
size_t scratch_bytes(const CollapseArgs& A)
{
    const size_t w = ((size_t)std::max<int64_t>(A.n, 1) * 4 + 255) & ~(size_t)255;
    return 2 * w + (((size_t)std::max<int64_t>(A.n_tri, 1) * 4 + 255) & ~(size_t)255) + 256;
}

void bind_scratch(CollapseArgs& A, void* base)
{
    char* b = static_cast<char*>(base);
    const size_t w = ((size_t)std::max<int64_t>(A.n, 1) * 4 + 255) & ~(size_t)255;
    const size_t t = ((size_t)std::max<int64_t>(A.n_tri, 1) * 4 + 255) & ~(size_t)255;
    A.work = reinterpret_cast<unsigned*>(b);
    A.tri_cand = reinterpret_cast<int32_t*>(b + w);
    A.counter = reinterpret_cast<unsigned*>(b + w + t);
    A.exact_list = reinterpret_cast<int32_t*>(b + w + t + 256);
}

using RingFn = void (*)(CollapseArgs);

RingFn ring_fn(int tile, bool exact)
{
    if (tile == 16) return exact ? collapse_ring_exact_kernel<16> : collapse_ring_kernel<16>;
    if (tile == 32) return exact ? collapse_ring_exact_kernel<32> : collapse_ring_kernel<32>;
    return exact ? collapse_ring_exact_kernel<64> : collapse_ring_kernel<64>;
}

// every launch of one batch, in order, on stream s (graph-capturable: the
// queue counter is cleared by a kernel)
int enqueue(twg_ctx* ctx, const CollapseArgs& A, int tile, bool f32, hipStream_t s)
{
    RingFn ring = ring_fn(tile, false);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(ring),
                                                     kBlock, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    int64_t need = (A.n * tile + kBlock - 1) / kBlock;
    const unsigned gring = (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)ctx->num_cus * per_cu));
    const unsigned gexact = (unsigned)std::min<int64_t>(std::max<int64_t>(need, 1), 2048);
    hipLaunchKernelGGL(reset_queue, dim3(1), dim3(1), 0, s, A.counter);
    TWG_LAUNCHED();
    hipLaunchKernelGGL(ring, dim3(gring), dim3(kBlock), 0, s, A);
    TWG_LAUNCHED();
    hipLaunchKernelGGL(ring_fn(tile, true), dim3(gexact), dim3(kBlock), 0, s, A);
    TWG_LAUNCHED();
    if (A.has_sf || A.has_bt) {
        if (f32) hipLaunchKernelGGL(collapse_point_kernel<true>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A);
        else hipLaunchKernelGGL(collapse_point_kernel<false>, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A);
        TWG_LAUNCHED();
    }
    if (A.n_tri > 0 && A.check_eps) {
        const unsigned gtri = grid_for(A.n_tri * kFaceGroup);
        const unsigned gtri_exact = std::min<unsigned>(grid_for(A.n_tri), 1024u);
        if (f32) {
            hipLaunchKernelGGL(collapse_face_kernel<true>, dim3(gtri), dim3(kBlock), 0, s, A);
            TWG_LAUNCHED();
            hipLaunchKernelGGL(collapse_face_exact_kernel<true>, dim3(gtri_exact), dim3(kBlock), 0, s, A);
        } else {
            hipLaunchKernelGGL(collapse_face_kernel<false>, dim3(gtri), dim3(kBlock), 0, s, A);
            TWG_LAUNCHED();
            hipLaunchKernelGGL(collapse_face_exact_kernel<false>, dim3(gtri_exact), dim3(kBlock), 0, s, A);
        }
        TWG_LAUNCHED();
        hipLaunchKernelGGL(collapse_final_kernel, dim3(grid_for(A.n)), dim3(kBlock), 0, s, A);
        TWG_LAUNCHED();
    }
    return TWG_OK;
}

// autotuner cache: (device, f32, size class) -> (tile, chunk)
TuneCache g_tuned;

int size_class(int64_t n) { return n < 4096 ? 0 : (n < 65536 ? 1 : 2); }

void pick(const twg_ctx* ctx, const CollapseArgs& A, bool f32, int tile, int chunk, int& t,
          unsigned& ch)
{
    int tt = default_tile(A);
    unsigned tc = 1u;
    g_tuned.get(ctx->device, f32, size_class(A.n), tt, tc);
    t = tile ? tile : tt;
    ch = chunk ? (unsigned)chunk : tc;
}

int check_launch_args(int precision, int tile, int chunk, const char* fn)
{
    TWG_REQUIRE(precision == TWG_FP64 || precision == TWG_MIXED, "%s: bad precision", fn);
    TWG_REQUIRE(tile == 0 || tile == 16 || tile == 32 || tile == 64,
                "%s: tile must be 0, 16, 32 or 64", fn);
    TWG_REQUIRE(chunk >= 0 && chunk <= 1024, "%s: chunk out of range", fn);
    return TWG_OK;
}

}  // namespace
}  // namespace twg

using namespace twg;

struct twg_collapse_plan {
    twg_ctx* ctx = nullptr;
    CollapseArgs args;
    void* scratch = nullptr;
    int tile = 32;
    unsigned chunk = 1;
    bool f32 = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
};

extern "C" {

int twg_collapse_batch(twg_ctx* ctx, const twg_collapse_desc* d, int precision, int tile, int chunk)
{
    TWG_REQUIRE(ctx != nullptr, "collapse: NULL context");
    if (int r = check_launch_args(precision, tile, chunk, "twg_collapse_batch")) return r;
    CollapseArgs A;
    int r = make_args(ctx, d, A);
    if (r) return r;
    if (A.n == 0) return TWG_OK;
    if ((r = bind(ctx))) return r;
    if ((r = ensure_scratch(ctx, scratch_bytes(A)))) return r;
    bind_scratch(A, ctx->scratch);
    const bool f32 = precision == TWG_MIXED;
    int t;
    pick(ctx, A, f32, tile, chunk, t, A.chunk);
    return enqueue(ctx, A, t, f32, ctx->cur);
}

int twg_collapse_autotune(twg_ctx* ctx, const twg_collapse_desc* d, int precision, int* tile_out,
                          int* chunk_out, float* ms_out)
{
    TWG_REQUIRE(ctx != nullptr, "twg_collapse_autotune: NULL context");
    if (int r = check_launch_args(precision, 0, 0, "twg_collapse_autotune")) return r;
    CollapseArgs A;
    int r = make_args(ctx, d, A);
    if (r) return r;
    if ((r = bind(ctx))) return r;
    if ((r = ensure_scratch(ctx, scratch_bytes(A)))) return r;
    bind_scratch(A, ctx->scratch);
    const bool f32 = precision == TWG_MIXED;
    const int tiles[3] = {16, 32, 64};
    const unsigned chunks[3] = {1, 4, 16};
    float best = 1e30f;
    int bt = default_tile(A);
    unsigned bc = 1;
    if (A.n > 0) {
        EventPair ev;
        TWG_HIP(ev.create());
        for (int t : tiles)
            for (unsigned ch : chunks) {
                A.chunk = ch;
                if ((r = enqueue(ctx, A, t, f32, ctx->cur))) return r;  // warm
                TWG_HIP(hipEventRecord(ev.e0, ctx->cur));
                for (int rep = 0; rep < 2; ++rep)
                    if ((r = enqueue(ctx, A, t, f32, ctx->cur))) return r;
                TWG_HIP(hipEventRecord(ev.e1, ctx->cur));
                TWG_HIP(hipEventSynchronize(ev.e1));
                float ms = 0;
                TWG_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
                ms /= 2;
                if (ms < best) {
                    best = ms;
                    bt = t;
                    bc = ch;
                }
            }
        g_tuned.put(ctx->device, f32, size_class(A.n), bt, bc);
    }
    if (tile_out) *tile_out = bt;
    if (chunk_out) *chunk_out = (int)bc;
    if (ms_out) *ms_out = A.n > 0 ? best : 0.0f;
    return TWG_OK;
}

int twg_collapse_plan_create(twg_ctx* ctx, const twg_collapse_desc* d, int precision, int tile,
                             int chunk, twg_collapse_plan** out)
{
    TWG_REQUIRE(ctx && d && out, "twg_collapse_plan_create: NULL argument");
    if (int r = check_launch_args(precision, tile, chunk, "twg_collapse_plan_create")) return r;
    if (int r = bind(ctx)) return r;
    twg_collapse_plan* P = new twg_collapse_plan();
    P->ctx = ctx;
    P->f32 = precision == TWG_MIXED;
    if (int r = make_args(ctx, d, P->args)) {
        delete P;
        return r;
    }
    pick(ctx, P->args, P->f32, tile, chunk, P->tile, P->chunk);
    P->args.chunk = P->chunk;
    // the plan owns its scratch: a graph keeps the addresses it was captured with
    hipError_t e = hipMalloc(&P->scratch, scratch_bytes(P->args));
    if (e != hipSuccess) {
        delete P;
        return hip_fail(e, "twg_collapse_plan_create", __FILE__, __LINE__);
    }
    bind_scratch(P->args, P->scratch);
    int r = TWG_OK;
    if (P->args.n > 0) {
        e = hipStreamCreateWithFlags(&P->capture_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(P->capture_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) r = enqueue(ctx, P->args, P->tile, P->f32, P->capture_stream);
        if (P->capture_stream) {
            hipError_t e2 = hipStreamEndCapture(P->capture_stream, &P->graph);
            if (e == hipSuccess) e = e2;
        }
        if (e == hipSuccess && r == TWG_OK) e = hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0);
    }
    if (e != hipSuccess || r != TWG_OK) {
        if (P->graph) (void)hipGraphDestroy(P->graph);
        if (P->capture_stream) (void)hipStreamDestroy(P->capture_stream);
        (void)hipFree(P->scratch);
        delete P;
        return r ? r : hip_fail(e, "twg_collapse_plan_create capture", __FILE__, __LINE__);
    }
    *out = P;
    return TWG_OK;
}

int twg_collapse_plan_run(twg_collapse_plan* P)
{
    TWG_REQUIRE(P != nullptr, "twg_collapse_plan_run: NULL plan");
    if (!P->exec) return TWG_OK;  // empty batch
    if (int r = bind(P->ctx)) return r;
    TWG_HIP(hipGraphLaunch(P->exec, P->ctx->cur));
    return TWG_OK;
}

int twg_collapse_plan_info(const twg_collapse_plan* P, int* tile, int* chunk)
{
    TWG_REQUIRE(P != nullptr, "twg_collapse_plan_info: NULL plan");
    if (tile) *tile = P->tile;
    if (chunk) *chunk = (int)P->chunk;
    return TWG_OK;
}

int twg_collapse_plan_destroy(twg_collapse_plan* P)
{
    if (!P) return TWG_OK;
    (void)hipSetDevice(P->ctx->device);
    (void)hipStreamSynchronize(P->ctx->cur);
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->capture_stream) (void)hipStreamDestroy(P->capture_stream);
    if (P->scratch) (void)hipFree(P->scratch);
    delete P;
    return TWG_OK;
}

}  // extern "C"
