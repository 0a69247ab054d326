// twg_bvh.hpp -- device AABB tree of the envelope surface and its traversals
// (shared by the envelope kernels and the smoothing kernel).
//
// Tree: MeshFacetsAABBWithEps's implicit layout (src/tetwild/geogram/
// mesh_AABB.cpp:356-379): root 1, children 2n and 2n+1, facet range split at
// b + (e-b)/2, one facet per leaf, facets in Morton order.  Node n is a 64-byte
// record (FP64 box + leaf facet id) so that the two children of a node -- the
// pair every traversal step reads -- are one 128-byte line; the mixed path
// reads a 32-byte FP32 copy of the boxes (rounded outward) instead.
//
// Decision: a sample p is in the envelope iff some facet f has
// d2(p, f) <= eps_2 (d2 = the FP64 Eberly distance, bit-identical to the host
// restatement).  The traversal prunes a box only when its squared distance
// exceeds eps_2 by a margin larger than the rounding error of any facet
// distance inside it, so the result does not depend on traversal order, on
// the facet-hint shortcut (src/tetwild/LocalOperations.cpp:1078-1098), or on
// FP32 boxes.  The traversal is stackless (a bit per tree level marks a
// pending far child).
#pragma once

#include <vector>

#include "twg_geom.hpp"

struct BvhNode64 {
    double lo[3];
    double hi[3];
    int32_t leaf;  // facet position for leaves, -1 internal
    int32_t pad[3];
};
struct BvhNode32 {
    float lo[3];
    float hi[3];
    int32_t leaf;
    int32_t pad;
};
static_assert(sizeof(BvhNode64) == 64, "node64");
static_assert(sizeof(BvhNode32) == 32, "node32");

struct twg_bvh {
    int device = 0;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // root box
    int64_t nf = 0, nnode = 0;
    int depth = 0;
    double rmax = 0.0;
    BvhNode64* node64 = nullptr;
    BvhNode32* node32 = nullptr;
    double* tri = nullptr;  // nf x 9, tree order
    // nf x 3, tree order: the facet across edge k (vertices k, k+1), -1 none
    int32_t* adj = nullptr;
    std::vector<int32_t> order;  // tree position -> caller's facet id
};

namespace twg {

// node records in LDS (address space 3): typed so that the compiler emits LDS
// reads for them (a generic pointer would turn every node read into a flat load)
typedef __attribute__((address_space(3))) const BvhNode64 LdsNode64;
typedef __attribute__((address_space(3))) const BvhNode32 LdsNode32;

// field-wise copies out of LDS (a whole-struct assignment across address
// spaces does not compile)
__device__ __forceinline__ BvhNode32 from_lds(LdsNode32* p)
{
    BvhNode32 r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = p->lo[k];
        r.hi[k] = p->hi[k];
    }
    r.leaf = p->leaf;
    return r;
}
__device__ __forceinline__ BvhNode64 from_lds(LdsNode64* p)
{
    BvhNode64 r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = p->lo[k];
        r.hi[k] = p->hi[k];
    }
    r.leaf = p->leaf;
    return r;
}

// What kernels receive (by value, as a kernel argument or inside one): global
// pointers and sizes only, so that the host and device layouts agree.
struct BvhView {
    const BvhNode64* __restrict__ n64;
    const BvhNode32* __restrict__ n32;
    const double* __restrict__ tri;
    int64_t nf;
    int64_t nnode;  // node slots (array length)
    const int32_t* __restrict__ adj;  // nf x 3 edge neighbours (tree positions)
};
static_assert(sizeof(BvhView) == 48, "BvhView: kernel-argument layout");

// What the traversals read: the view plus an optional copy of the top of the
// tree in LDS (node slots [0, ncache), the levels every traversal walks
// first; ncache = 0: none).  LDS pointers are 32-bit in device code and 64-bit
// in the host pass, so a walker is made inside a kernel and never crosses into
// a kernel argument.
struct BvhWalk {
    const BvhNode64* __restrict__ n64;
    const BvhNode32* __restrict__ n32;
    const double* __restrict__ tri;
    int64_t nf;
    int64_t nnode;
    const int32_t* __restrict__ adj;
    LdsNode64* c64;
    LdsNode32* c32;
    int64_t ncache;
    // which copy holds the cached slots (an LDS pointer cannot say: offset 0,
    // where the copy may sit, is also its null value)
    bool top64;
};

// a walker that reads every node from global memory
__device__ __forceinline__ BvhWalk walk(const BvhView& B)
{
    BvhWalk w;
    w.n64 = B.n64;
    w.n32 = B.n32;
    w.tri = B.tri;
    w.nf = B.nf;
    w.nnode = B.nnode;
    w.adj = B.adj;
    w.c64 = nullptr;
    w.c32 = nullptr;
    w.ncache = 0;
    w.top64 = false;
    return w;
}

// bytes of the per-block LDS copy of the top of the tree: 1024 FP32 nodes
// (10 levels) or 512 FP64 nodes (9 levels)
constexpr int kTreeCacheBytes = 32768;

// Every thread of the block calls this once, before any traversal: copies
// node slots [0, ncache) of the tree (FP32 or FP64 records, BYTES bytes at
// most) into cache and returns a walker that reads them from there.
template <bool F32, int BYTES = kTreeCacheBytes>
__device__ __forceinline__ BvhWalk cache_tree_top(const BvhView& B, void* cache)
{
    static_assert(BYTES % 128 == 0, "cache size: an even number of FP64 nodes");
    const int64_t per = BYTES / (F32 ? (int64_t)sizeof(BvhNode32) : (int64_t)sizeof(BvhNode64));
    const int64_t nc = B.nnode < per ? B.nnode : per;
    const uint4* src = F32 ? reinterpret_cast<const uint4*>(B.n32) : reinterpret_cast<const uint4*>(B.n64);
    uint4* dst = static_cast<uint4*>(cache);
    const int64_t chunks = nc * (F32 ? (int64_t)sizeof(BvhNode32) : (int64_t)sizeof(BvhNode64)) / 16;
    for (int64_t k = threadIdx.x; k < chunks; k += blockDim.x) dst[k] = src[k];
    __syncthreads();
    BvhWalk w = walk(B);
    if (F32) w.c32 = (LdsNode32*)cache;
    else w.c64 = (LdsNode64*)cache;
    w.ncache = nc;
    w.top64 = !F32;
    return w;
}

// the two children c, c + 1 (same level: both cached or neither)
__device__ __forceinline__ void load_pair(const BvhWalk& B, int64_t c, BvhNode32& L, BvhNode32& R)
{
    if (c < B.ncache) {
        L = from_lds(B.c32 + c);
        R = from_lds(B.c32 + c + 1);
    } else {
        L = B.n32[c];
        R = B.n32[c + 1];
    }
}

__device__ __forceinline__ void load_pair(const BvhWalk& B, int64_t c, BvhNode64& L, BvhNode64& R)
{
    if (c < B.ncache) {
        L = from_lds(B.c64 + c);
        R = from_lds(B.c64 + c + 1);
    } else {
        L = B.n64[c];
        R = B.n64[c + 1];
    }
}

__device__ __forceinline__ V3<double> ld3(const double* p) { return mk3(p[0], p[1], p[2]); }

__device__ __forceinline__ double facet_d2(const BvhWalk& B, int f, V3<double> p)
{
    const double* t = B.tri + 9 * (int64_t)f;
    V3<double> cp;
    return point_tri_d2<double, false>(p, ld3(t), ld3(t + 3), ld3(t + 6), cp);
}

struct Stats {
    unsigned long long samples = 0, traversals = 0, boxes = 0, leaves = 0;
};

// Tf: FP32 pruning threshold, conservative for the rounding of p to FP32 and
// of the FP32 box distance (see the header comment of this file).
__device__ __forceinline__ float f32_threshold(V3<double> p, double sqrtT)
{
    double pm = fmax(fabs(p.x), fmax(fabs(p.y), fabs(p.z)));
    double ep = 2.0e-7 * pm;
    double t = (sqrtT + ep) * (sqrtT + ep) * (1.0 + 0x1p-19);
    float tf = (float)t;
    if ((double)tf < t) tf = u2f32(f2u(tf) + 1u);  // t > 0: next float up
    return tf;
}

// Levels above the hint facet's leaf whose subtree is searched first when
// the hint itself misses: consecutive samples of a triangle usually move to a
// nearby facet, which lies in that subtree.  Measured on 20 M query triangles
// against 1 M facets: 3 / 5 / 7 / 10 / 12 levels gave 500 / 513 / 525 / 542 /
// 539 M checks/s.
constexpr int kHintLevels = 10;

// the implicit tree's node holding facet position b: the median splits of
// the tree's build (b + (e - b) / 2), integer work only
__device__ __forceinline__ int64_t leaf_node(int64_t nf, int64_t b)
{
    int64_t n = 1, lo = 0, hi = nf;
    while (hi - lo > 1) {
        const int64_t m = lo + (hi - lo) / 2;
        if (b < m) {
            n = 2 * n;
            hi = m;
        } else {
            n = 2 * n + 1;
            lo = m;
        }
    }
    return n;
}

// SUBTREE: after a hint miss, search the hint's ancestor subtree first
template <bool F32, bool STATS, bool SUBTREE = true>
__device__ bool in_envelope(const BvhWalk& B, V3<double> p, double eps2, double T, double sqrtT,
                            int& hint, Stats& st)
{
    int64_t first = 1;  // the subtree searched first
    if (hint >= 0) {
        if (STATS) st.leaves++;
        if (facet_d2(B, hint, p) <= eps2) return true;
        if (SUBTREE) {
            const int64_t a = leaf_node(B.nf, hint) >> kHintLevels;
            first = a > 1 ? a : 1;
        }
    }
    if (STATS) st.traversals++;
    if (B.nf == 1) {
        if (STATS) st.leaves++;
        bool in = facet_d2(B, 0, p) <= eps2;
        if (in) hint = 0;
        return in;
    }
    V3<float> pf = mk3((float)p.x, (float)p.y, (float)p.z);
    float Tf = F32 ? f32_threshold(p, sqrtT) : 0.f;
    // the subtree of `first`, then (if it held no facet within eps) the whole
    // tree; the decision -- some facet within eps -- is the whole tree's
    for (int64_t root = first;; root = 1) {
        uint32_t trail = 0;
        int depth = 0;
        int64_t n = root;
        bool ok;
        int root_leaf;
        if (F32) {
            BvhNode32 r;
            if (root < B.ncache) r = from_lds(B.c32 + root);
            else r = B.n32[root];
            ok = box_d2(pf, r.lo, r.hi) <= Tf;
            root_leaf = r.leaf;
        } else {
            BvhNode64 r;
            if (root < B.ncache) r = from_lds(B.c64 + root);
            else r = B.n64[root];
            ok = box_d2(p, r.lo, r.hi) <= T;
            root_leaf = r.leaf;
        }
        if (STATS) st.boxes++;
        bool found = false;
        if (ok && root_leaf >= 0) {
            if (STATS) st.leaves++;
            if (facet_d2(B, root_leaf, p) <= eps2) {
                hint = root_leaf;
                found = true;
            }
        } else if (ok) {
            for (;;) {
                // n is internal here: look at both children
                int64_t c = 2 * n;
                double dl, dr;
                bool okl, okr;
                int leafl, leafr;
                if (F32) {
                    BvhNode32 L, R;
                    load_pair(B, c, L, R);
                    float fl = box_d2(pf, L.lo, L.hi), fr = box_d2(pf, R.lo, R.hi);
                    okl = fl <= Tf;
                    okr = fr <= Tf;
                    dl = fl;
                    dr = fr;
                    leafl = L.leaf;
                    leafr = R.leaf;
                } else {
                    BvhNode64 L, R;
                    load_pair(B, c, L, R);
                    dl = box_d2(p, L.lo, L.hi);
                    dr = box_d2(p, R.lo, R.hi);
                    okl = dl <= T;
                    okr = dr <= T;
                    leafl = L.leaf;
                    leafr = R.leaf;
                }
                if (STATS) st.boxes += 2;
                bool descend = okl || okr;
                if (descend) {
                    bool left = okl && (!okr || dl <= dr);
                    if (okl && okr) trail |= 1u << (depth + 1);
                    n = left ? c : c + 1;
                    depth++;
                    int leaf = left ? leafl : leafr;
                    if (leaf < 0) continue;
                    if (STATS) st.leaves++;
                    if (facet_d2(B, leaf, p) <= eps2) {
                        hint = leaf;
                        found = true;
                        break;
                    }
                }
                // backtrack to the deepest pending far child (depths are
                // counted from the subtree's root, which has no sibling here)
                bool done = false;
                for (;;) {
                    if (trail == 0) {
                        done = true;
                        break;
                    }
                    int d = 31 - __clz(trail);
                    trail &= ~(1u << d);
                    n = (n >> (depth - d)) ^ 1;
                    depth = d;
                    int leaf;
                    if (F32) leaf = n < B.ncache ? B.c32[n].leaf : B.n32[n].leaf;
                    else leaf = n < B.ncache ? B.c64[n].leaf : B.n64[n].leaf;
                    if (leaf < 0) break;
                    if (STATS) st.leaves++;
                    if (facet_d2(B, leaf, p) <= eps2) {
                        hint = leaf;
                        found = true;
                        done = true;
                        break;
                    }
                }
                if (done) break;
            }
        }
        if (found) return true;
        if (root == 1) return false;
    }
}


// ---- the facets near a small region, in LDS --------------------------------
// writable LDS records
typedef __attribute__((address_space(3))) BvhNode32 LdsNode32W;
typedef __attribute__((address_space(3))) uint32_t LdsU32;

// the wave's LDS writes so far, visible to its other lanes
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x rounded to FP32 downward / upward
__device__ __forceinline__ float f32_down(double x)
{
    float f = (float)x;
    if ((double)f > x) f = f > 0.f ? u2f32(f2u(f) - 1u) : (f < 0.f ? u2f32(f2u(f) + 1u) : -0x1p-149f);
    return f;
}
__device__ __forceinline__ float f32_up(double x) { return -f32_down(-x); }

template <typename R, typename S>
__device__ __forceinline__ bool boxes_meet(const R* alo, const R* ahi, const S* blo, const S* bhi)
{
    return alo[0] <= bhi[0] && alo[1] <= bhi[1] && alo[2] <= bhi[2] && ahi[0] >= blo[0] &&
           ahi[1] >= blo[1] && ahi[2] >= blo[2];
}

__device__ __forceinline__ void put_lds(LdsNode32W* d, const BvhNode32& s)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d->lo[k] = s.lo[k];
        d->hi[k] = s.hi[k];
    }
    d->leaf = s.leaf;
    d->pad = -1;  // no neighbours recorded (list_neighbours fills them)
}

// The facets whose boxes meet the box [lo, hi], found by the G lanes of a
// group (G a power of two up to 64, lane gl, the group's first lane gbase in
// the wave) expanding a work list of internal nodes that meet the box (fr:
// two buffers of cap entries).  A lane expands one node: its children from
// the walker's FP64 LDS copy when they are there, else its children and
// grandchildren from the FP32 records in global memory, read together (one
// memory latency for two levels; a grandchild's box lies inside its
// parent's, and FP32 boxes are rounded outward, so no facet is lost).
// Ballots place what meets the box -- leaves into lst as FP32 records,
// internal nodes into the next work list.  Returns the number of facets in
// lst, or -1 (nothing usable in lst) when more than cap facets, or more than
// cap nodes in one work list, meet the box, when the box is not finite, or
// for a one-facet tree.  Every lane of the group calls it.
template <int G>
__device__ int collect_facets(const BvhWalk& W, const double lo[3], const double hi[3],
                              LdsNode32W* lst, LdsU32* fr, int cap, int gl, int gbase)
{
    bool finite = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) finite = finite && is_finite(lo[k]) && is_finite(hi[k]);
    if (!finite || W.nf < 2) return -1;
    float lof[3], hif[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lof[k] = f32_down(lo[k]);
        hif[k] = f32_up(hi[k]);
    }
    const unsigned long long gm = (G == 64 ? ~0ull : ((1ull << (G & 63)) - 1ull)) << gbase;
    const unsigned long long below = ((1ull << (gbase + gl)) - 1ull) & gm;
    LdsU32* cur = fr;
    LdsU32* nxt = fr + cap;
    if (gl == 0) cur[0] = 1u;
    wave_sync();
    int ncur = 1, nl = 0;
    while (ncur > 0) {
        int nn = 0;
        for (int i0 = 0; i0 < ncur; i0 += G) {
            const int i = i0 + gl;
            // up to four candidates: slot k holds node id[k] (0: none)
            BvhNode32 cand[4];
            int64_t id[4] = {0, 0, 0, 0};
            bool ok[4] = {false, false, false, false};
            if (i < ncur) {
                const int64_t ch = 2 * (int64_t)cur[i];
                if (ch < W.ncache && W.top64) {
                    const BvhNode64 L64 = from_lds(W.c64 + ch), R64 = from_lds(W.c64 + ch + 1);
                    ok[0] = boxes_meet(L64.lo, L64.hi, lo, hi);
                    ok[1] = boxes_meet(R64.lo, R64.hi, lo, hi);
                    id[0] = ch;
                    id[1] = ch + 1;
                    cand[0].leaf = L64.leaf;
                    cand[1].leaf = R64.leaf;
                    if (ok[0] && L64.leaf >= 0) cand[0] = W.n32[ch];
                    if (ok[1] && R64.leaf >= 0) cand[1] = W.n32[ch + 1];
                } else {
                    BvhNode32 C[2], GC[4];
                    if (ch < W.ncache) {
                        C[0] = from_lds(W.c32 + ch);
                        C[1] = from_lds(W.c32 + ch + 1);
                    } else {
                        C[0] = W.n32[ch];
                        C[1] = W.n32[ch + 1];
                    }
                    // grandchildren slots exist for an internal child (read
                    // before the child is known to be internal; used only then)
                    const int64_t g0 = 2 * ch;
                    if (g0 + 1 < W.nnode) {
                        GC[0] = W.n32[g0];
                        GC[1] = W.n32[g0 + 1];
                    }
                    if (g0 + 3 < W.nnode) {
                        GC[2] = W.n32[g0 + 2];
                        GC[3] = W.n32[g0 + 3];
                    }
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (!boxes_meet(C[h].lo, C[h].hi, lof, hif)) continue;
                        if (C[h].leaf >= 0) {
                            cand[2 * h] = C[h];
                            id[2 * h] = ch + h;
                            ok[2 * h] = true;
                        } else {
#pragma unroll
                            for (int e = 0; e < 2; ++e) {
                                cand[2 * h + e] = GC[2 * h + e];
                                id[2 * h + e] = 2 * (ch + h) + e;
                                ok[2 * h + e] = boxes_meet(GC[2 * h + e].lo, GC[2 * h + e].hi, lof, hif);
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const bool leaf = ok[k] && cand[k].leaf >= 0, inner = ok[k] && cand[k].leaf < 0;
                const unsigned long long ml = __ballot(leaf) & gm, mi = __ballot(inner) & gm;
                const int pl = nl + __popcll(ml & below), qi = nn + __popcll(mi & below);
                if (leaf && pl < cap) put_lds(lst + pl, cand[k]);
                if (inner && qi < cap) nxt[qi] = (uint32_t)id[k];
                nl += __popcll(ml);
                nn += __popcll(mi);
            }
        }
        if (nl > cap || nn > cap) return -1;
        wave_sync();
        LdsU32* t = cur;
        cur = nxt;
        nxt = t;
        ncur = nn;
    }
    return nl;
}

typedef __attribute__((address_space(3))) double LdsF64;
typedef __attribute__((address_space(3))) const double LdsF64c;

// For each entry of a collect_facets list (lanes gl, gl + G, ...), the list
// positions of its facet's three edge neighbours (the tree's adjacency),
// packed 8 bits each into the entry's pad field; 0xff where the neighbour is
// not in the list or the tree has no adjacency.  Lists of up to 255 entries.
// Every lane of the group calls it; it ends with the group's writes visible.
template <int G>
__device__ void list_neighbours(const BvhWalk& W, LdsNode32W* lst, int nl, int gl)
{
    if (W.adj && nl > 0 && nl < 256) {
        for (int e = gl; e < nl; e += G) {
            const int64_t f = lst[e].leaf;
            const int a0 = W.adj[3 * f], a1 = W.adj[3 * f + 1], a2 = W.adj[3 * f + 2];
            uint32_t q0 = 0xffu, q1 = 0xffu, q2 = 0xffu;
            for (int i = 0; i < nl; ++i) {
                const int li = lst[i].leaf;
                q0 = li == a0 ? (uint32_t)i : q0;
                q1 = li == a1 ? (uint32_t)i : q1;
                q2 = li == a2 ? (uint32_t)i : q2;
            }
            lst[e].pad = (int32_t)(q0 | (q1 << 8) | (q2 << 16) | 0xff000000u);
        }
    }
    wave_sync();
}

// facet_d2 of a triangle held in LDS (9 doubles): the same computation on
// the same values
__device__ __forceinline__ double tri_d2(LdsF64c* t, V3<double> p)
{
    V3<double> cp;
    return point_tri_d2<double, false>(p, mk3(t[0], t[1], t[2]), mk3(t[3], t[4], t[5]),
                                       mk3(t[6], t[7], t[8]), cp);
}

// in_envelope's decision for a point p of the box that a collect_facets list
// was made for, with that box padded by at least sqrt(T) beyond every point
// tested (then the list holds every facet the tree walk could accept): the
// facet at list position hl first (the one that held the previous point, -1:
// none), then the others, 64 positions at a time: the FP32 box tests of the
// window first, into a bit mask, then the triangles of the boxes that pass.
// (Box tests and triangle distances interleaved would make a wave pay a
// triangle distance at every position where some lane's box passes.)  tri
// holds the listed facets' triangles, 9 doubles per position.  On a hit hl
// becomes the facet's position.
template <bool NB = false>
__device__ __forceinline__ bool in_envelope_list(LdsNode32W* lst, LdsF64c* tri, int nl,
                                                 V3<double> p, double eps2, double sqrtT, int& hl)
{
    int nb[3] = {-1, -1, -1};
    if (hl >= 0) {
        if (tri_d2(tri + 9 * hl, p) <= eps2) return true;
        // a sample that left its predecessor's facet is usually within eps of
        // one of that facet's edge neighbours (NB: the list records them)
        const uint32_t pk = NB ? (uint32_t)lst[hl].pad : 0xffffffffu;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t q = (pk >> (8 * k)) & 0xffu;
            if (q != 0xffu) {
                if (tri_d2(tri + 9 * q, p) <= eps2) {
                    hl = (int)q;
                    return true;
                }
                nb[k] = (int)q;
            }
        }
    }
    const V3<float> pf = mk3((float)p.x, (float)p.y, (float)p.z);
    const float Tf = f32_threshold(p, sqrtT);
    for (int base = 0; base < nl; base += 64) {
        const int n = nl - base < 64 ? nl - base : 64;
        uint64_t m = 0;
        for (int k0 = 0; k0 < n; k0 += 8) {  // eight records read before their tests
            BvhNode32 e[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k0 + u < n) e[u] = from_lds(lst + base + k0 + u);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k0 + u < n) m |= (uint64_t)(box_d2(pf, e[u].lo, e[u].hi) <= Tf) << (k0 + u);
        }
        if (hl >= base && hl < base + n) m &= ~(1ull << (hl - base));
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (nb[k] >= base && nb[k] < base + n) m &= ~(1ull << (nb[k] - base));
        while (m) {
            const int i = base + __builtin_ctzll(m);
            m &= m - 1;
            if (tri_d2(tri + 9 * i, p) <= eps2) {
                hl = i;
                return true;
            }
        }
    }
    return false;
}

// The positions of a collect_facets list (nl <= 128 entries) whose FP32 boxes
// meet the box of triangle (a, b, c) padded as the list's own box is (by
// 1.0001 sqrt(T) + 1e-9 max |coordinate|): every facet within sqrt(T) of a
// point of the triangle -- or of a sample of it, up to its rounding -- is at
// a set position.  mk: two 64-bit words.
// the FP32 box face_list_mask compares the list's boxes with
__device__ __forceinline__ void face_mask_box(V3<double> a, V3<double> b, V3<double> c,
                                              double sqrtT, float lof[3], float hif[3])
{
    double lo[3] = {a.x, a.y, a.z}, hi[3] = {a.x, a.y, a.z};
    const double q[6] = {b.x, b.y, b.z, c.x, c.y, c.z};
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        lo[j % 3] = fmin(lo[j % 3], q[j]);
        hi[j % 3] = fmax(hi[j % 3], q[j]);
    }
    double amax = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) amax = fmax(amax, fmax(fabs(lo[k]), fabs(hi[k])));
    const double pad = 1.0001 * sqrtT + 1e-9 * amax;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lof[k] = f32_down(lo[k] - pad);
        hif[k] = f32_up(hi[k] + pad);
    }
}

__device__ __forceinline__ bool box_meets(const BvhNode32& r, const float lof[3], const float hif[3])
{
    return r.lo[0] <= hif[0] && r.hi[0] >= lof[0] && r.lo[1] <= hif[1] && r.hi[1] >= lof[1] &&
           r.lo[2] <= hif[2] && r.hi[2] >= lof[2];
}

template <class L>
__device__ __forceinline__ void face_list_mask(L lst, int nl, V3<double> a, V3<double> b,
                                               V3<double> c, double sqrtT, uint64_t mk[2])
{
    float lof[3], hif[3];
    face_mask_box(a, b, c, sqrtT, lof, hif);
    // the two words in registers (an array indexed at run time would go to
    // scratch memory), eight records read before their tests
    const int n = nl < 128 ? nl : 128;
    uint64_t w[2] = {0ull, 0ull};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int e1 = n - 64 * h < 64 ? n - 64 * h : 64;
        uint64_t m = 0ull;
        for (int e0 = 0; e0 < e1; e0 += 8) {
            BvhNode32 r[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + u < e1) r[u] = from_lds(lst + 64 * h + e0 + u);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool meet = e0 + u < e1 && box_meets(r[u], lof, hif);
                m |= (uint64_t)meet << (e0 + u);
            }
        }
        w[h] = m;
    }
    mk[0] = w[0];
    mk[1] = w[1];
}

// in_envelope_list over the list positions set in mask (two 64-bit words;
// lists of at most 128 entries), for a p whose caller guarantees that no
// facet at an unset position lies within sqrt(T) of p: the facet at hl
// first, then the set positions' FP32 box tests, then the triangles of the
// boxes that pass, in ascending position as in_envelope_list.  Every facet
// that could hold p is at a set position, so the decision and the facet left
// in hl are in_envelope_list's.  BATCH: the set positions' boxes are read
// four at a time before their tests (one LDS round trip per four; the
// smoothing kernels gain from it, the envelope list kernel, short of
// registers, loses).
template <bool NB = false, bool BATCH = false, class M>
__device__ __forceinline__ bool in_envelope_masked(LdsNode32W* lst, LdsF64c* tri, int nl, M mask,
                                                   V3<double> p, double eps2, double sqrtT, int& hl)
{
    int nb[3] = {-1, -1, -1};
    if (hl >= 0) {
        if (tri_d2(tri + 9 * hl, p) <= eps2) return true;
        // NB: the facet's edge neighbours next, as in_envelope_list<true>
        const uint32_t pk = NB ? (uint32_t)lst[hl].pad : 0xffffffffu;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t q = (pk >> (8 * k)) & 0xffu;
            if (q != 0xffu) {
                if (tri_d2(tri + 9 * q, p) <= eps2) {
                    hl = (int)q;
                    return true;
                }
                nb[k] = (int)q;
            }
        }
    }
    const V3<float> pf = mk3((float)p.x, (float)p.y, (float)p.z);
    const float Tf = f32_threshold(p, sqrtT);
    for (int w = 0; w < 2 && 64 * w < nl; ++w) {
        uint64_t mm = mask[w];
        if (hl >= 64 * w && hl < 64 * w + 64) mm &= ~(1ull << (hl - 64 * w));
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (nb[k] >= 64 * w && nb[k] < 64 * w + 64) mm &= ~(1ull << (nb[k] - 64 * w));
        uint64_t m = 0;
        while (!BATCH && mm) {
            const int b = __builtin_ctzll(mm);
            mm &= mm - 1;
            const BvhNode32 e = from_lds(lst + 64 * w + b);
            if (box_d2(pf, e.lo, e.hi) <= Tf) m |= 1ull << b;
        }
        while (BATCH && mm) {
            // four set positions at a time, their boxes read before their
            // tests (past the last set bit, the first is read again)
            int bq[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                bq[u] = mm ? __builtin_ctzll(mm) : -1;
                mm &= mm - 1;
            }
            float lo[4][3], hi[4][3];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const LdsNode32W* e = lst + 64 * w + (bq[u] >= 0 ? bq[u] : bq[0]);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    lo[u][k] = e->lo[k];
                    hi[u][k] = e->hi[k];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (bq[u] >= 0 && box_d2(pf, lo[u], hi[u]) <= Tf) m |= 1ull << bq[u];
        }
        while (m) {
            const int i = 64 * w + __builtin_ctzll(m);
            m &= m - 1;
            if (tri_d2(tri + 9 * i, p) <= eps2) {
                hl = i;
                return true;
            }
        }
    }
    return false;
}

// nearest_facet with the reference's traversal order
// (src/tetwild/geogram/mesh_AABB.h:130-141, mesh_AABB.cpp:381-480):
// get_nearest_facet_hint, then depth-first search visiting the nearer child
// first (the right one on ties) and pruning boxes whose signed distance is not
// below the best so far.  Stackless: one bit per level marks a far child still
// to visit; its box distance is recomputed on the way back (the same value the
// recursion keeps on its stack), so the walk needs no per-lane memory.
// node k of the FP64 tree, from the walker's LDS copy when it holds it
__device__ __forceinline__ BvhNode64 node64(const BvhWalk& B, int64_t k)
{
    if (k < B.ncache) return from_lds(B.c64 + k);
    return B.n64[k];
}

// get_nearest_facet_hint: down from the root to a leaf, into the child whose
// box center is nearer
__device__ __forceinline__ int nearest_hint(const BvhWalk& B, V3<double> p)
{
    int64_t nd = 1;
    while (node64(B, nd).leaf < 0) {
        const BvhNode64 L = node64(B, 2 * nd), R = node64(B, 2 * nd + 1);
        nd = box_center_d2(p, L.lo, L.hi) < box_center_d2(p, R.lo, R.hi) ? 2 * nd : 2 * nd + 1;
    }
    return node64(B, nd).leaf;
}

__device__ __forceinline__ int nearest_facet(const BvhWalk& B, V3<double> p, V3<double>& np,
                                             double& sq)
{
    int best = nearest_hint(B, p);
    np = ld3(B.tri + 9 * (int64_t)best);
    sq = len2(np - p);
    auto visit = [&](int leaf) {
        const double* t = B.tri + 9 * (int64_t)leaf;
        V3<double> cp;
        double dd = point_tri_d2<double, true>(p, ld3(t), ld3(t + 3), ld3(t + 6), cp);
        if (dd < sq) {
            best = leaf;
            np = cp;
            sq = dd;
        }
    };
    {
        const int root_leaf = node64(B, 1).leaf;
        if (root_leaf >= 0) {  // a one-facet tree: the root is the leaf
            visit(root_leaf);
            return best;
        }
    }
    int64_t n = 1;
    int depth = 0;
    uint32_t trail = 0;  // bit d: the far child at depth d is still to visit
    for (;;) {
        // n is internal: the nearer child first
        const int64_t c = 2 * n;
        BvhNode64 L, R;
        load_pair(B, c, L, R);
        const double dl = box_signed_d2(p, L.lo, L.hi);
        const double dr = box_signed_d2(p, R.lo, R.hi);
        const bool left = dl < dr;
        // the far child is never nearer, so a pruned near child prunes both
        if ((left ? dl : dr) < sq) {
            trail |= 1u << (depth + 1);
            n = left ? c : c + 1;
            depth++;
            const int leaf = left ? L.leaf : R.leaf;
            if (leaf < 0) continue;
            visit(leaf);
        }
        // back to the deepest pending far child that can still hold a nearer facet
        for (;;) {
            if (trail == 0) return best;
            const int d = 31 - __clz(trail);
            trail &= ~(1u << d);
            const int64_t far = (n >> (depth - d)) ^ 1;
            n = far;
            depth = d;
            const BvhNode64 F = node64(B, far);
            if (!(box_signed_d2(p, F.lo, F.hi) < sq)) continue;
            const int leaf = F.leaf;
            if (leaf < 0) break;
            visit(leaf);
        }
    }
}

// nearest_facet's result (facet, np, sq -- bit for bit) computed by the G
// lanes of a group (lane gl, first lane gbase in the wave; G a power of two up
// to 64), or by nearest_facet itself where the group cannot be sure to agree.
// The walk starts from the hint facet h with sq = s0 = |v0(h) - p|^2 and
// takes a facet only when its distance is strictly below the best so far.
// The group lists the facets whose boxes meet a cube around p (collect_facets)
// -- half-width rr plus more than the rounding of a facet distance, so the
// list holds every facet at distance rr or less -- and computes their
// distances; rr starts at the hint leaf's box diagonal (about one facet) and
// doubles, up to sqrt(s0), until some listed facet lies within rr.  None
// below s0 with rr = sqrt(s0): the walk keeps (h, v0(h)).  Else let f be the
// facet of least distance D: when no other facet has distance D and every
// node on f's path from the root has a box distance strictly below D, the
// walk cannot prune that path (it prunes a box only at or above the best so
// far, never below D) and nothing it visits later replaces f, so it returns
// f -- taken here.  Otherwise (ties, a box at or above D, a list too long for
// cap, a query that is not finite, three cubes without a facet) nearest_facet
// walks.  Every lane of the group calls it with the same p.
template <int G>
__device__ int nearest_facet_group(const BvhWalk& B, V3<double> p, V3<double>& np, double& sq,
                                   LdsNode32W* lst, LdsU32* fr, int cap, int gl, int gbase)
{
    if (B.nf < 2) return nearest_facet(B, p, np, sq);
    int64_t na = 1;
    while (node64(B, na).leaf < 0) {
        const BvhNode64 L = node64(B, 2 * na), R = node64(B, 2 * na + 1);
        na = box_center_d2(p, L.lo, L.hi) < box_center_d2(p, R.lo, R.hi) ? 2 * na : 2 * na + 1;
    }
    const BvhNode64 HL = node64(B, na);
    const int h = HL.leaf;
    const V3<double> v0 = ld3(B.tri + 9 * (int64_t)h);
    const double s0 = len2(v0 - p), sr0 = __builtin_sqrt(s0);
    if (!(s0 <= 1.0e300)) return nearest_facet(B, p, np, sq);
    // more than the rounding of a facet distance: ~1e-7 (|p| + d + 2 edge) in
    // distance units (the root box's diagonal bounds every edge)
    const BvhNode64 root = node64(B, 1);
    const double diag = __builtin_sqrt(len2(mk3(root.hi[0] - root.lo[0], root.hi[1] - root.lo[1],
                                                root.hi[2] - root.lo[2])));
    const double eabs = 0x1p-20 * (diag + fmax(fabs(p.x), fmax(fabs(p.y), fabs(p.z))));
    double r = __builtin_sqrt(len2(mk3(HL.hi[0] - HL.lo[0], HL.hi[1] - HL.lo[1], HL.hi[2] - HL.lo[2])));
    const unsigned long long gm = (G == 64 ? ~0ull : ((1ull << (G & 63)) - 1ull)) << gbase;
    for (int attempt = 0; attempt < 3; ++attempt, r *= 2.0) {
        const bool last = !(r < sr0);  // the cube holds every facet below s0
        const double rr = last ? sr0 : r;
        const double half = rr + 0x1p-20 * rr + eabs;
        const double lo[3] = {p.x - half, p.y - half, p.z - half};
        const double hi[3] = {p.x + half, p.y + half, p.z + half};
        // listed from the global tree (~1 % faster than through an LDS copy
        // of its top; the descents above read the copy)
        BvhWalk GW = B;
        GW.ncache = 0;
        GW.top64 = false;
        const int n = collect_facets<G>(GW, lo, hi, lst, fr, cap, gl, gbase);
        if (n < 0) break;
        // the lane's least distance below s0, its facet and how many reach it
        double bd = s0;
        int bf = -1, nb = 0;
        V3<double> bcp = v0;
        for (int i = gl; i < n; i += G) {
            const int f = lst[i].leaf;
            const double* t = B.tri + 9 * (int64_t)f;
            V3<double> cp;
            const double dd = point_tri_d2<double, true>(p, ld3(t), ld3(t + 3), ld3(t + 6), cp);
            if (dd < bd) {
                bd = dd;
                bf = f;
                bcp = cp;
                nb = 1;
            } else if (dd == bd && bf >= 0) {
                nb++;
            }
        }
        double m = bd;
        for (int off = G / 2; off > 0; off >>= 1) {
            const double o = __shfl_xor(m, off, G);
            m = o < m ? o : m;
        }
        const unsigned long long at = __ballot(bf >= 0 && bd == m) & gm;
        if (at == 0ull) {
            if (!last) continue;
            np = v0;  // no facet below s0
            sq = s0;
            return h;
        }
        if (!last && !(m <= rr * rr)) continue;  // a nearer facet may lie outside
        const int wl = __builtin_ctzll(at) - gbase;
        const int wf = __shfl(bf, wl, G);
        bool sure = (at & (at - 1)) == 0ull && __shfl(nb, wl, G) == 1;
        if (sure) {
            // every box on the path from the root to wf's leaf strictly below m
            const int64_t leaf = leaf_node(B.nf, wf);
            const int depth = 63 - __builtin_clzll((unsigned long long)leaf);
            bool below = true;
            for (int k = 1 + gl; k <= depth; k += G) {
                const BvhNode64 N = node64(B, leaf >> (depth - k));
                below = below && box_signed_d2(p, N.lo, N.hi) < m;
            }
            sure = (__ballot(!below) & gm) == 0ull;
        }
        if (!sure) break;
        np = mk3(__shfl(bcp.x, wl, G), __shfl(bcp.y, wl, G), __shfl(bcp.z, wl, G));
        sq = m;
        return wf;
    }
    return nearest_facet(B, p, np, sq);
}

// Triangles whose longest edge exceeds kMaxEdgeSamples sampling distances are
// reported out without sampling: their sample count (quadratic in that ratio)
// would keep one lane busy for seconds.  TetWild's faces are orders of
// magnitude below this.
constexpr double kMaxEdgeSamples = 65536.0;

// isFaceOutEnvelop_sampling for a triangle known not to be degenerate
template <bool F32, bool STATS>
__device__ bool face_out_sampled(const BvhWalk& B, const double* t9, double eps2, double d,
                                 double T, double sqrtT, Stats& st)
{
    const V3<double> a = ld3(t9), b = ld3(t9 + 3), c = ld3(t9 + 6);
    {
        double l = fmax(len2(a - b), fmax(len2(b - c), len2(c - a)));
        if (!(__builtin_sqrt(l) / d <= kMaxEdgeSamples)) return true;
    }
    int hint = -1;
    return for_each_sample(a, b, c, d, [&](V3<double> s) {
        if (STATS) st.samples++;
        return !in_envelope<F32, STATS>(B, s, eps2, T, sqrtT, hint, st);
    });
}

// face_out_sampled's decision with the facets from a collect_facets list
// (lst, tri; see in_envelope_list) whose box holds the triangle padded by
// more than sqrt(T): every sample lies in that box, so the list holds every
// facet the tree walk could accept.  nl < 0: the tree walk.
template <bool F32>
__device__ bool face_out_listed(const BvhWalk& B, const double* t9, double eps2, double d,
                                double T, double sqrtT, LdsNode32W* lst, LdsF64c* tri, int nl)
{
    if (nl < 0) {
        Stats st;
        return face_out_sampled<F32, false>(B, t9, eps2, d, T, sqrtT, st);
    }
    const V3<double> a = ld3(t9), b = ld3(t9 + 3), c = ld3(t9 + 6);
    {
        double l = fmax(len2(a - b), fmax(len2(b - c), len2(c - a)));
        if (!(__builtin_sqrt(l) / d <= kMaxEdgeSamples)) return true;
    }
    int hl = -1;
    if (nl <= 128) {
        // only the listed facets that can lie near this triangle
        uint64_t mk[2];
        face_list_mask(lst, nl, a, b, c, sqrtT, mk);
        return for_each_sample(a, b, c, d, [&](V3<double> s) {
            return !in_envelope_masked<true>(lst, tri, nl, mk, s, eps2, sqrtT, hl);
        });
    }
    return for_each_sample(a, b, c, d, [&](V3<double> s) {
        return !in_envelope_list<true>(lst, tri, nl, s, eps2, sqrtT, hl);
    });
}

// face_out_sampled's decision computed by a group of G lanes (G a power of
// two up to 64, groups aligned in the wave; every lane of a group passes the
// same triangle), against a collect_facets list (lst, tri, nl >= 0; its box
// holding the triangle padded by more than sqrt(T)) or the tree (nl < 0).  The sampler's sequence of `total` samples is cut into G
// consecutive blocks of K = ceil(total / G); lane r moves to sample r K
// (SampleGen::skip passes whole rows by their headers) and tests its block
// in order, one sample per step, and after each step a ballot over the group
// stops it at the first sample found outside.  The decision -- some sample
// has no facet within eps -- does not depend on which lane tests which
// sample, so it equals face_out_sampled's.  For batches of few triangles
// with many samples each, where one lane per triangle leaves most of the
// machine idle.
template <bool F32, int G, bool SUBTREE = true, bool NB = false>
__device__ bool face_out_group(const BvhWalk& B, V3<double> a, V3<double> b, V3<double> c,
                               double eps2, double d, double T, double sqrtT, int r, int gbase,
                               LdsNode32W* lst = nullptr, LdsF64c* tri = nullptr, int nl = -1)
{
    static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group size");
    {
        double l = fmax(len2(a - b), fmax(len2(b - c), len2(c - a)));
        if (!(__builtin_sqrt(l) / d <= kMaxEdgeSamples)) return true;
    }
    int64_t total;
    {
        SampleGen cnt;
        cnt.init(a, b, c, d);
        total = cnt.skip(INT64_MAX);
    }
    const int64_t K = (total + G - 1) / G;
    const int64_t first = (int64_t)r * K;
    SampleGen g;
    g.init(a, b, c, d);
    g.skip(first);
    int hint = -1, hl = -1;
    Stats st;
    const unsigned long long gmask = G == 64 ? ~0ull : (1ull << (G & 63)) - 1ull;
    for (int64_t j = 0; j < K; ++j) {
        V3<double> q = a;
        const bool have = first + j < total && g.next(q);
        const bool out = have && !(nl >= 0 ? in_envelope_list<NB>(lst, tri, nl, q, eps2, sqrtT, hl)
                                           : in_envelope<F32, false, SUBTREE>(B, q, eps2, T, sqrtT, hint, st));
        if ((__ballot(out) >> gbase) & gmask) return true;
    }
    return false;
}

// isFaceOutEnvelop_sampling for one triangle (false for a collinear one)
template <bool F32, bool STATS>
__device__ bool face_out(const BvhWalk& B, const double* t9, double eps2, double d, double T,
                         double sqrtT, Stats& st)
{
    if (collinear3(t9, t9 + 3, t9 + 6)) return false;
    return face_out_sampled<F32, STATS>(B, t9, eps2, d, T, sqrtT, st);
}

inline BvhView bvh_view(const twg_bvh* b)
{
    BvhView v;
    v.n64 = b->node64;
    v.n32 = b->node32;
    v.tri = b->tri;
    v.nf = b->nf;
    v.nnode = b->nnode;
    v.adj = b->adj;
    return v;
}

inline BvhView no_bvh()
{
    BvhView v = {nullptr, nullptr, nullptr, 0, 0, nullptr};
    return v;
}

inline double bvh_prune_threshold(const twg_bvh* b, double eps2)
{
    double r = b->rmax + __builtin_sqrt(eps2);
    return eps2 + eps2 * 0x1p-20 + 0x1p-44 * (r * r);
}

}  // namespace twg
