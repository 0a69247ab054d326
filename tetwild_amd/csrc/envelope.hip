// envelope.hip -- (b) the envelope test: isFaceOutEnvelop_sampling,
// isPointOutEnvelop, nearest_facet and sampleTriangle, over a flattened
// implicit AABB tree resident in HBM.
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
// FP32 boxes.  One thread per query; the traversal is stackless (a bit per
// tree level marks a pending far child).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "twg_geom.hpp"
#include "twg_internal.hpp"

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
    int64_t nf = 0, nnode = 0;
    int depth = 0;
    double rmax = 0.0;
    BvhNode64* node64 = nullptr;
    BvhNode32* node32 = nullptr;
    double* tri = nullptr;  // nf x 9, tree order
    std::vector<int32_t> order;  // tree position -> caller's facet id
};

namespace twg {
namespace {

struct BvhView {
    const BvhNode64* __restrict__ n64;
    const BvhNode32* __restrict__ n32;
    const double* __restrict__ tri;
    int64_t nf;
};

__device__ __forceinline__ V3<double> ld3(const double* p) { return mk3(p[0], p[1], p[2]); }

__device__ __forceinline__ double facet_d2(const BvhView& B, int f, V3<double> p)
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

template <bool F32, bool STATS>
__device__ bool in_envelope(const BvhView& B, V3<double> p, double eps2, double T, double sqrtT,
                            int& hint, Stats& st)
{
    if (hint >= 0) {
        if (STATS) st.leaves++;
        if (facet_d2(B, hint, p) <= eps2) return true;
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
    uint32_t trail = 0;
    int depth = 0;
    int64_t n = 1;
    // the root box
    bool ok;
    if (F32) ok = box_d2(pf, B.n32[1].lo, B.n32[1].hi) <= Tf;
    else ok = box_d2(p, B.n64[1].lo, B.n64[1].hi) <= T;
    if (!ok) return false;
    for (;;) {
        // n is internal here: look at both children
        int64_t c = 2 * n;
        double dl, dr;
        bool okl, okr;
        int leafl, leafr;
        if (F32) {
            const BvhNode32 L = B.n32[c], R = B.n32[c + 1];
            float fl = box_d2(pf, L.lo, L.hi), fr = box_d2(pf, R.lo, R.hi);
            okl = fl <= Tf;
            okr = fr <= Tf;
            dl = fl;
            dr = fr;
            leafl = L.leaf;
            leafr = R.leaf;
        } else {
            const BvhNode64 L = B.n64[c], R = B.n64[c + 1];
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
                return true;
            }
        }
        // backtrack to the deepest pending far child
        for (;;) {
            if (trail == 0) return false;
            int d = 31 - __clz(trail);
            trail &= ~(1u << d);
            n = (n >> (depth - d)) ^ 1;
            depth = d;
            int leaf = F32 ? B.n32[n].leaf : B.n64[n].leaf;
            if (leaf < 0) break;
            if (STATS) st.leaves++;
            if (facet_d2(B, leaf, p) <= eps2) {
                hint = leaf;
                return true;
            }
        }
    }
}

// XCD-aware block order: consecutive logical blocks (neighbouring queries)
// land on the same XCD, so each XCD's L2 sees a compact region of the tree.
__device__ __forceinline__ int64_t logical_block()
{
    const unsigned g = gridDim.x, b = blockIdx.x;
    const unsigned per = g / 8, rem = g % 8;
    const unsigned xcd = b % 8, k = b / 8;
    if (k >= per) return b;  // tail beyond the last full round
    (void)rem;
    return (int64_t)xcd * per + k;
}

__device__ __forceinline__ void flush_stats(const Stats& st, unsigned long long* out)
{
    unsigned long long v[4] = {st.samples, st.traversals, st.boxes, st.leaves};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        unsigned long long s = v[k];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if ((threadIdx.x & 63) == 0 && s) atomicAdd(&out[k], s);
    }
}

template <bool F32, bool STATS>
__global__ __launch_bounds__(kBlock) void points_kernel(BvhView B, const double* __restrict__ p3,
                                                        int64_t n, double eps2, double T,
                                                        double sqrtT, uint8_t* __restrict__ out,
                                                        unsigned long long* stats)
{
    int64_t i = logical_block() * kBlock + threadIdx.x;
    Stats st;
    if (i < n) {
        int hint = -1;
        V3<double> p = ld3(p3 + 3 * i);
        if (STATS) st.samples++;
        out[i] = in_envelope<F32, STATS>(B, p, eps2, T, sqrtT, hint, st) ? 0 : 1;
    }
    if (STATS) flush_stats(st, stats);
}

template <bool F32, bool STATS>
__global__ __launch_bounds__(kBlock) void faces_kernel(BvhView B, const double* __restrict__ tri9,
                                                       int64_t n, double eps2, double d, double T,
                                                       double sqrtT, uint8_t* __restrict__ out,
                                                       unsigned long long* stats)
{
    int64_t i = logical_block() * kBlock + threadIdx.x;
    Stats st;
    if (i < n) {
        const double* t = tri9 + 9 * i;
        V3<double> vs[3] = {ld3(t), ld3(t + 3), ld3(t + 6)};
        uint8_t res = 0;
        if (!collinear3(t, t + 3, t + 6)) {
            int hint = -1;
            bool stopped = for_each_sample(vs, d, [&](V3<double> s) {
                if (STATS) st.samples++;
                return !in_envelope<F32, STATS>(B, s, eps2, T, sqrtT, hint, st);
            });
            res = stopped ? 1 : 0;
        }
        out[i] = res;
    }
    if (STATS) flush_stats(st, stats);
}

// nearest_facet with the reference's traversal order
// (src/tetwild/geogram/mesh_AABB.h:130-141, mesh_AABB.cpp:381-480)
__global__ __launch_bounds__(kBlock) void nearest_kernel(BvhView B, const double* __restrict__ p3,
                                                         int64_t n, double* __restrict__ cp3,
                                                         double* __restrict__ d2out,
                                                         int32_t* __restrict__ fout)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    V3<double> p = ld3(p3 + 3 * i);
    // get_nearest_facet_hint: descend towards the nearer box centre
    int64_t nd = 1;
    while (B.n64[nd].leaf < 0) {
        const BvhNode64& L = B.n64[2 * nd];
        const BvhNode64& R = B.n64[2 * nd + 1];
        nd = box_center_d2(p, L.lo, L.hi) < box_center_d2(p, R.lo, R.hi) ? 2 * nd : 2 * nd + 1;
    }
    int best = B.n64[nd].leaf;
    V3<double> np = ld3(B.tri + 9 * (int64_t)best);
    double sq = len2(np - p);
    int64_t stk_n[64];
    double stk_d[64];
    int sp = 0;
    stk_n[sp] = 1;
    stk_d[sp] = -__builtin_inf();
    sp++;
    while (sp > 0) {
        --sp;
        int64_t nn = stk_n[sp];
        double dn = stk_d[sp];
        if (!(dn < sq)) continue;
        int leaf = B.n64[nn].leaf;
        if (leaf >= 0) {
            const double* t = B.tri + 9 * (int64_t)leaf;
            V3<double> cp;
            double dd = point_tri_d2<double, true>(p, ld3(t), ld3(t + 3), ld3(t + 6), cp);
            if (dd < sq) {
                best = leaf;
                np = cp;
                sq = dd;
            }
            continue;
        }
        const BvhNode64& L = B.n64[2 * nn];
        const BvhNode64& R = B.n64[2 * nn + 1];
        double dl = box_signed_d2(p, L.lo, L.hi);
        double dr = box_signed_d2(p, R.lo, R.hi);
        if (sp + 2 > 64) break;  // unreachable: depth <= 30 is enforced at build
        if (dl < dr) {
            stk_n[sp] = 2 * nn + 1; stk_d[sp] = dr; sp++;
            stk_n[sp] = 2 * nn; stk_d[sp] = dl; sp++;
        } else {
            stk_n[sp] = 2 * nn; stk_d[sp] = dl; sp++;
            stk_n[sp] = 2 * nn + 1; stk_d[sp] = dr; sp++;
        }
    }
    cp3[3 * i + 0] = np.x;
    cp3[3 * i + 1] = np.y;
    cp3[3 * i + 2] = np.z;
    d2out[i] = sq;
    fout[i] = best;
}

__global__ __launch_bounds__(kBlock) void sample_count_kernel(const double* __restrict__ tri9,
                                                              int64_t n, double d,
                                                              int64_t* __restrict__ counts)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double* t = tri9 + 9 * i;
    V3<double> vs[3] = {ld3(t), ld3(t + 3), ld3(t + 6)};
    int64_t k = 0;
    for_each_sample(vs, d, [&](V3<double>) {
        ++k;
        return false;
    });
    counts[i] = k;
}

__global__ __launch_bounds__(kBlock) void sample_write_kernel(const double* __restrict__ tri9,
                                                              int64_t n, double d,
                                                              const int64_t* __restrict__ offsets,
                                                              int64_t total,
                                                              double* __restrict__ pts)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double* t = tri9 + 9 * i;
    V3<double> vs[3] = {ld3(t), ld3(t + 3), ld3(t + 6)};
    int64_t k = offsets[i];
    for_each_sample(vs, d, [&](V3<double> s) {
        if (k >= 0 && k < total) {
            pts[3 * k + 0] = s.x;
            pts[3 * k + 1] = s.y;
            pts[3 * k + 2] = s.z;
        }
        ++k;
        return false;
    });
}

// ---- host-side build --------------------------------------------------
inline uint64_t spread3(uint64_t v)
{
    v &= 0x1FFFFF;
    v = (v | v << 32) & 0x1F00000000FFFFull;
    v = (v | v << 16) & 0x1F0000FF0000FFull;
    v = (v | v << 8) & 0x100F00F00F00F00Full;
    v = (v | v << 4) & 0x10C30C30C30C30C3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

int64_t max_node(int64_t n, int64_t b, int64_t e)
{
    if (b + 1 == e) return n;
    int64_t m = b + (e - b) / 2;
    return std::max(max_node(2 * n, b, m), max_node(2 * n + 1, m, e));
}

void fill(std::vector<BvhNode64>& nodes, const std::vector<double>& tri, int64_t n, int64_t b,
          int64_t e, int depth, int& maxdepth)
{
    maxdepth = std::max(maxdepth, depth);
    BvhNode64& N = nodes[n];
    if (b + 1 == e) {
        const double* t = &tri[9 * b];
        for (int c = 0; c < 3; ++c) {
            N.lo[c] = N.hi[c] = t[c];
            for (int k = 1; k < 3; ++k) {
                N.lo[c] = std::min(N.lo[c], t[3 * k + c]);
                N.hi[c] = std::max(N.hi[c], t[3 * k + c]);
            }
        }
        N.leaf = (int32_t)b;
        return;
    }
    int64_t m = b + (e - b) / 2;
    fill(nodes, tri, 2 * n, b, m, depth + 1, maxdepth);
    fill(nodes, tri, 2 * n + 1, m, e, depth + 1, maxdepth);
    const BvhNode64& L = nodes[2 * n];
    const BvhNode64& R = nodes[2 * n + 1];
    for (int c = 0; c < 3; ++c) {
        N.lo[c] = std::min(L.lo[c], R.lo[c]);
        N.hi[c] = std::max(L.hi[c], R.hi[c]);
    }
    N.leaf = -1;
}

float round_down(double v)
{
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}
float round_up(double v)
{
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, INFINITY);
    return f;
}

BvhView view(const twg_bvh* b)
{
    BvhView v;
    v.n64 = b->node64;
    v.n32 = b->node32;
    v.tri = b->tri;
    v.nf = b->nf;
    return v;
}

double prune_threshold(const twg_bvh* b, double eps2)
{
    double r = b->rmax + std::sqrt(eps2);
    return eps2 + eps2 * 0x1p-20 + 0x1p-44 * (r * r);
}

}  // namespace
}  // namespace twg

using namespace twg;

extern "C" {

int twg_bvh_build(twg_ctx* ctx, const double* V, int64_t nv, const int32_t* F, int64_t nf,
                  twg_bvh** out)
{
    TWG_REQUIRE(ctx && out, "twg_bvh_build: NULL argument");
    TWG_REQUIRE(nf >= 1 && nv >= 1 && V && F, "twg_bvh_build: empty surface");
    TWG_REQUIRE(nf < (int64_t(1) << 30), "twg_bvh_build: too many facets (%lld)", (long long)nf);
    for (int64_t f = 0; f < 3 * nf; ++f)
        TWG_REQUIRE(F[f] >= 0 && F[f] < nv, "twg_bvh_build: facet %lld has vertex id %d outside [0, %lld)",
                    (long long)(f / 3), F[f], (long long)nv);
    // Morton order of facet centroids
    std::vector<double> cen(3 * nf);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) {
            double v = (V[3 * F[3 * f] + c] + V[3 * F[3 * f + 1] + c] + V[3 * F[3 * f + 2] + c]) / 3.0;
            cen[3 * f + c] = v;
            lo[c] = std::min(lo[c], v);
            hi[c] = std::max(hi[c], v);
        }
    std::vector<std::pair<uint64_t, int32_t>> key(nf);
    for (int64_t f = 0; f < nf; ++f) {
        uint64_t k = 0;
        for (int c = 0; c < 3; ++c) {
            double span = hi[c] - lo[c];
            double q = span > 0 ? (cen[3 * f + c] - lo[c]) / span * 2097152.0 : 0.0;
            uint64_t qi = q <= 0 ? 0 : (q >= 2097151.0 ? 2097151 : (uint64_t)q);
            k |= spread3(qi) << c;
        }
        key[f] = {k, (int32_t)f};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, int32_t>& a, const std::pair<uint64_t, int32_t>& b) {
                         return a.first < b.first;
                     });
    twg_bvh* B = new twg_bvh();
    B->device = ctx->device;
    B->nf = nf;
    B->order.resize(nf);
    std::vector<double> tri(9 * nf);
    double rmax = 0.0;
    for (int64_t b = 0; b < nf; ++b) {
        int32_t f = key[b].second;
        B->order[b] = f;
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) tri[9 * b + 3 * k + c] = V[3 * F[3 * f + k] + c];
        for (int k = 0; k < 3; ++k) {
            const double* a = &tri[9 * b + 3 * k];
            const double* e = &tri[9 * b + 3 * ((k + 1) % 3)];
            double l = std::sqrt((a[0] - e[0]) * (a[0] - e[0]) + (a[1] - e[1]) * (a[1] - e[1]) +
                                 (a[2] - e[2]) * (a[2] - e[2]));
            rmax = std::max(rmax, l);
        }
    }
    B->rmax = rmax;
    B->nnode = max_node(1, 0, nf) + 1;
    std::vector<BvhNode64> nodes(B->nnode);
    for (auto& N : nodes) {
        for (int c = 0; c < 3; ++c) {
            N.lo[c] = INFINITY;
            N.hi[c] = -INFINITY;
        }
        N.leaf = -2;
        N.pad[0] = N.pad[1] = N.pad[2] = 0;
    }
    int depth = 0;
    fill(nodes, tri, 1, 0, nf, 0, depth);
    B->depth = depth;
    if (depth > 30) {
        delete B;
        set_error("twg_bvh_build: tree depth %d exceeds 30", depth);
        return TWG_EINVAL;
    }
    std::vector<BvhNode32> n32(B->nnode);
    for (int64_t i = 0; i < B->nnode; ++i) {
        for (int c = 0; c < 3; ++c) {
            n32[i].lo[c] = round_down(nodes[i].lo[c]);
            n32[i].hi[c] = round_up(nodes[i].hi[c]);
        }
        n32[i].leaf = nodes[i].leaf;
        n32[i].pad = 0;
    }
    int r = bind(ctx);
    hipError_t e = hipSuccess;
    if (!r) e = hipMalloc(&B->node64, sizeof(BvhNode64) * B->nnode);
    if (!r && e == hipSuccess) e = hipMalloc(&B->node32, sizeof(BvhNode32) * B->nnode);
    if (!r && e == hipSuccess) e = hipMalloc(&B->tri, sizeof(double) * 9 * nf);
    if (!r && e == hipSuccess)
        e = hipMemcpy(B->node64, nodes.data(), sizeof(BvhNode64) * B->nnode, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess)
        e = hipMemcpy(B->node32, n32.data(), sizeof(BvhNode32) * B->nnode, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess)
        e = hipMemcpy(B->tri, tri.data(), sizeof(double) * 9 * nf, hipMemcpyHostToDevice);
    if (r || e != hipSuccess) {
        if (B->node64) (void)hipFree(B->node64);
        if (B->node32) (void)hipFree(B->node32);
        if (B->tri) (void)hipFree(B->tri);
        delete B;
        return r ? r : hip_fail(e, "twg_bvh_build upload", __FILE__, __LINE__);
    }
    *out = B;
    return TWG_OK;
}

int twg_bvh_destroy(twg_bvh* B)
{
    if (!B) return TWG_OK;
    (void)hipSetDevice(B->device);
    (void)hipDeviceSynchronize();
    if (B->node64) (void)hipFree(B->node64);
    if (B->node32) (void)hipFree(B->node32);
    if (B->tri) (void)hipFree(B->tri);
    delete B;
    return TWG_OK;
}

int twg_bvh_info(const twg_bvh* B, int64_t* nf, int64_t* nnode, int32_t* depth, double* rmax)
{
    TWG_REQUIRE(B != nullptr, "twg_bvh_info: NULL bvh");
    if (nf) *nf = B->nf;
    if (nnode) *nnode = B->nnode;
    if (depth) *depth = B->depth;
    if (rmax) *rmax = B->rmax;
    return TWG_OK;
}

int twg_bvh_facet_order(const twg_bvh* B, int32_t* order)
{
    TWG_REQUIRE(B && order, "twg_bvh_facet_order: NULL argument");
    std::memcpy(order, B->order.data(), sizeof(int32_t) * B->nf);
    return TWG_OK;
}

static int check_env(twg_ctx* ctx, const twg_bvh* B, const void* in, int64_t n, const void* out,
                     double eps2, int precision, const char* fn)
{
    TWG_REQUIRE(ctx && B, "%s: NULL ctx or bvh", fn);
    TWG_REQUIRE(n >= 0, "%s: negative size", fn);
    TWG_REQUIRE(n == 0 || (in && out), "%s: NULL array", fn);
    TWG_REQUIRE(eps2 >= 0 && std::isfinite(eps2), "%s: bad eps2", fn);
    TWG_REQUIRE(precision == TWG_FP64 || precision == TWG_MIXED, "%s: bad precision %d", fn,
                precision);
    TWG_REQUIRE(B->device == ctx->device, "%s: bvh lives on device %d, ctx on %d", fn, B->device,
                ctx->device);
    return TWG_OK;
}

int twg_envelope_points(twg_ctx* ctx, const twg_bvh* B, const double* p3, int64_t n, double eps2,
                        int precision, uint8_t* out, unsigned long long* stats)
{
    if (int r = check_env(ctx, B, p3, n, out, eps2, precision, "twg_envelope_points")) return r;
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    double T = prune_threshold(B, eps2), sT = std::sqrt(T);
    dim3 g(grid_for(n)), b(kBlock);
    BvhView v = view(B);
    if (precision == TWG_FP64) {
        if (stats) hipLaunchKernelGGL((points_kernel<false, true>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, stats);
        else hipLaunchKernelGGL((points_kernel<false, false>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, stats);
    } else {
        if (stats) hipLaunchKernelGGL((points_kernel<true, true>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, stats);
        else hipLaunchKernelGGL((points_kernel<true, false>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, stats);
    }
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_envelope_faces(twg_ctx* ctx, const twg_bvh* B, const double* tri9, int64_t n, double eps2,
                       double sampling_dist, int precision, uint8_t* out,
                       unsigned long long* stats)
{
    if (int r = check_env(ctx, B, tri9, n, out, eps2, precision, "twg_envelope_faces")) return r;
    TWG_REQUIRE(sampling_dist > 0 && std::isfinite(sampling_dist),
                "twg_envelope_faces: sampling_dist must be positive");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    double T = prune_threshold(B, eps2), sT = std::sqrt(T);
    dim3 g(grid_for(n)), b(kBlock);
    BvhView v = view(B);
    if (precision == TWG_FP64) {
        if (stats) hipLaunchKernelGGL((faces_kernel<false, true>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, stats);
        else hipLaunchKernelGGL((faces_kernel<false, false>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, stats);
    } else {
        if (stats) hipLaunchKernelGGL((faces_kernel<true, true>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, stats);
        else hipLaunchKernelGGL((faces_kernel<true, false>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, stats);
    }
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_nearest_points(twg_ctx* ctx, const twg_bvh* B, const double* p3, int64_t n, double* cp3,
                       double* d2, int32_t* facet)
{
    if (int r = check_env(ctx, B, p3, n, cp3, 0.0, TWG_FP64, "twg_nearest_points")) return r;
    TWG_REQUIRE(n == 0 || (d2 && facet), "twg_nearest_points: NULL output");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    hipLaunchKernelGGL(nearest_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->cur, view(B), p3,
                       n, cp3, d2, facet);
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_sample_count(twg_ctx* ctx, const double* tri9, int64_t n, double d, int64_t* counts)
{
    TWG_REQUIRE(ctx != nullptr, "twg_sample_count: ctx is NULL");
    TWG_REQUIRE(n >= 0 && (n == 0 || (tri9 && counts)), "twg_sample_count: bad arrays");
    TWG_REQUIRE(d > 0 && std::isfinite(d), "twg_sample_count: d must be positive");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    hipLaunchKernelGGL(sample_count_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->cur, tri9, n,
                       d, counts);
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_sample_write(twg_ctx* ctx, const double* tri9, int64_t n, double d,
                     const int64_t* offsets, int64_t total, double* pts)
{
    TWG_REQUIRE(ctx != nullptr, "twg_sample_write: ctx is NULL");
    TWG_REQUIRE(n >= 0 && total >= 0 && (n == 0 || (tri9 && offsets && (total == 0 || pts))),
                "twg_sample_write: bad arrays");
    TWG_REQUIRE(d > 0 && std::isfinite(d), "twg_sample_write: d must be positive");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    hipLaunchKernelGGL(sample_write_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->cur, tri9, n,
                       d, offsets, total, pts);
    TWG_LAUNCHED();
    return TWG_OK;
}

}  // extern "C"
