// envelope.hip -- (b) the envelope test: isFaceOutEnvelop_sampling,
// isPointOutEnvelop, nearest_facet and sampleTriangle, over a flattened
// implicit AABB tree resident in HBM.
//
// The tree layout and the decision rule are described in twg_bvh.hpp.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "twg_bvh.hpp"
#include "twg_internal.hpp"

namespace twg {
namespace {

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

// LDS copy of the tree top for the point kernel (its registers allow more
// resident blocks than 32 KB of LDS per block would)
constexpr int kPointCacheBytes = 16384;

template <bool F32, bool STATS>
__global__ __launch_bounds__(kBlock) void points_kernel(BvhView B, const double* __restrict__ p3,
                                                        int64_t n, double eps2, double T,
                                                        double sqrtT, uint8_t* __restrict__ out,
                                                        unsigned long long* stats,
                                                        const int32_t* __restrict__ order)
{
    __shared__ __align__(16) unsigned char cache[kPointCacheBytes];
    const BvhWalk W = cache_tree_top<F32, kPointCacheBytes>(B, cache);
    int64_t i = xcd_block() * kBlock + threadIdx.x;
    Stats st;
    if (i < n) {
        if (order) i = order[i];
        int hint = -1;
        V3<double> p = ld3(p3 + 3 * i);
        if (STATS) st.samples++;
        out[i] = in_envelope<F32, STATS>(W, p, eps2, T, sqrtT, hint, st) ? 0 : 1;
    }
    if (STATS) flush_stats(st, stats);
}

// points_kernel for batches walked in Morton order with many points per
// facet: the wave's 64 consecutive points lie close together, so the wave
// lists once (collect_facets, LDS) the facets whose boxes meet the box of its
// points padded by more than sqrt(T), with their triangles, and each lane
// tests its point against the list (in_envelope_list).  A wave whose list
// exceeds CAP entries, or that holds a coordinate that is not finite, walks
// the tree (from global memory: the LDS holds the lists).
template <bool F32, int CAP>
__global__ __launch_bounds__(kBlock) void points_list_kernel(BvhView B, const double* __restrict__ p3,
                                                             int64_t n, double eps2, double T,
                                                             double sqrtT, uint8_t* __restrict__ out,
                                                             const int32_t* __restrict__ order,
                                                             unsigned long long* lstat)
{
    constexpr int kWaveBytes = CAP * (32 + 8 + 72);
    __shared__ __align__(16) unsigned char lists[(kBlock / 64) * kWaveBytes];
    const BvhWalk W = walk(B);
    const int lane = threadIdx.x % 64;
    unsigned char* mine = lists + (threadIdx.x / 64) * kWaveBytes;
    LdsNode32W* lst = (LdsNode32W*)mine;
    LdsU32* fr = (LdsU32*)(mine + CAP * 32);
    LdsF64* tri = (LdsF64*)(mine + CAP * 40);
    int64_t i = xcd_block() * kBlock + threadIdx.x;
    const bool act = i < n;
    V3<double> p = mk3(0.0, 0.0, 0.0);
    if (act) {
        if (order) i = order[i];
        p = ld3(p3 + 3 * i);
    }
    double lo[3] = {act ? p.x : INFINITY, act ? p.y : INFINITY, act ? p.z : INFINITY};
    double hi[3] = {act ? p.x : -INFINITY, act ? p.y : -INFINITY, act ? p.z : -INFINITY};
    double amax = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            const double ol = __shfl_xor(lo[k], off, 64), oh = __shfl_xor(hi[k], off, 64);
            lo[k] = ol < lo[k] ? ol : lo[k];
            hi[k] = oh > hi[k] ? oh : hi[k];
        }
        amax = fmax(amax, fmax(fabs(lo[k]), fabs(hi[k])));
    }
    const double pad = 1.0001 * sqrtT + 1e-9 * amax;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] -= pad;
        hi[k] += pad;
    }
    // a coordinate that is not finite can leave the lanes with different
    // boxes (min / max ignore NaN): then the whole wave walks the tree
    const bool fin = !act || (is_finite(p.x) && is_finite(p.y) && is_finite(p.z));
    const int nl = __ballot(!fin) ? -1 : collect_facets<64>(W, lo, hi, lst, fr, CAP, lane, 0);
    if (lstat && lane == 0) atomicAdd(lstat + (nl >= 0 ? 0 : 1), 1ull);
    for (int e = lane; e < nl; e += 64) {
        const double* f = B.tri + 9 * (int64_t)lst[e].leaf;
#pragma unroll
        for (int k = 0; k < 9; ++k) tri[9 * e + k] = f[k];
    }
    wave_sync();
    if (act) {
        bool in;
        if (nl >= 0) {
            int hl = -1;
            in = in_envelope_list(lst, tri, nl, p, eps2, sqrtT, hl);
        } else {
            int hint = -1;
            Stats st;
            in = in_envelope<F32, false>(W, p, eps2, T, sqrtT, hint, st);
        }
        out[i] = in ? 0 : 1;
    }
}

// Triangles the degeneracy filter cannot settle are queued (pending[],
// *n_pending) for faces_exact_kernel, which keeps the exact predicate -- and its
// register footprint -- out of this kernel.  Capped at 128 VGPRs (4 waves
// per SIMD, a few spills): more waves in flight hide more of the traversal's
// load latency (+18% checks/s on 20 M queries against 1 M facets).
template <bool F32, bool STATS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void faces_kernel(
    BvhView B, const double* __restrict__ tri9, int64_t n, double eps2, double d, double T,
    double sqrtT, uint8_t* __restrict__ out, unsigned long long* stats, unsigned* n_pending,
    int64_t* pending, const int32_t* __restrict__ order)
{
    __shared__ __align__(16) unsigned char cache[kTreeCacheBytes];
    const BvhWalk W = cache_tree_top<F32>(B, cache);
    int64_t i = xcd_block() * kBlock + threadIdx.x;
    Stats st;
    if (i < n) {
        if (order) i = order[i];
        const double* t = tri9 + 9 * i;
        if (certainly_not_collinear(t, t + 3, t + 6)) {
            out[i] = face_out_sampled<F32, STATS>(W, t, eps2, d, T, sqrtT, st) ? 1 : 0;
        } else {
            unsigned k = atomicAdd(n_pending, 1u);
            pending[k] = i;
        }
    }
    if (STATS) flush_stats(st, stats);
}

// face_out_listed for the wave's 64 triangles at once (a list of nl <= 128
// facets, nl >= 0; act: the lane has a triangle, t9 its corners).  Every lane
// tests its triangle's first sample; most triangles outside the envelope stop
// there.  The other samples of the triangles still open are then spread over
// the wave: the concatenation of their sequences (the first samples left
// out) is cut into 64 consecutive blocks of K = ceil(total / 64) and lane r
// walks block r, moving to the next open triangle when one ends.  A lane
// takes the triangle it moves to -- its corners' address, list mask and the
// facet that held its first sample -- from the triangle's own lane (shuffles,
// every lane active; the corners are read again from memory, which keeps 18
// registers free through the walk: 898 -> 946 M checks/s on 20 M queries),
// and a sample found outside is passed to that lane by ballot.  The
// decision -- some sample has no facet within eps -- does not depend on which
// lane tests which sample, so it is face_out_listed's.  Returns the lane's
// own triangle's decision.
__device__ bool faces_wave_listed(bool act, const double* t9, double eps2, double d, double sqrtT,
                                  LdsNode32W* lst, LdsF64c* tri, int nl, int lane)
{
    bool out = false;
    int64_t rem = 0;  // samples after the first, while the triangle is open
    uint64_t mk[2] = {0ull, 0ull};
    int hl = -1;
    if (act) {
        const V3<double> a = ld3(t9), b = ld3(t9 + 3), c = ld3(t9 + 6);
        const double l = fmax(len2(a - b), fmax(len2(b - c), len2(c - a)));
        if (!(__builtin_sqrt(l) / d <= kMaxEdgeSamples)) {
            out = true;
        } else {
            face_list_mask(lst, nl, a, b, c, sqrtT, mk);
            SampleGen g;
            g.init(a, b, c, d);
            V3<double> s;
            g.next(s);  // every sequence has a first sample
            if (!in_envelope_masked<true>(lst, tri, nl, mk, s, eps2, sqrtT, hl)) out = true;
            else rem = g.skip(INT64_MAX);
        }
    }
    const unsigned long long open = __ballot(rem > 0);
    if (open == 0ull) return out;
    int64_t incl = rem;
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    const int64_t total = __shfl(incl, 63, 64);
    const int64_t K = (total + 63) / 64;
    const int64_t first = (int64_t)lane * K;
    // the triangle holding remaining sample `first`: the first lane f with
    // incl[f] > first (a binary search, each lane its own)
    int f = 0;
    for (int step = 32; step > 0; step >>= 1) {
        const int probe = f + step - 1;
        const int64_t v = __shfl(incl, probe < 63 ? probe : 63, 64);
        if (probe < 63 && v <= first) f = probe + 1;
    }
    if (first >= total) f = 64;  // no samples for this lane
    const int fs = f < 64 ? f : lane;
    const int64_t ofs = first - (__shfl(incl, fs, 64) - __shfl(rem, fs, 64));
    // the lane's triangle: corners, mask and hint from its owner
    SampleGen g2;
    uint64_t m2[2];
    int h2;
    {
        const double* ct = (const double*)__shfl((long long)(uintptr_t)t9, fs, 64);
        m2[0] = __shfl(mk[0], fs, 64);
        m2[1] = __shfl(mk[1], fs, 64);
        h2 = __shfl(hl, fs, 64);
        if (f < 64) {
            g2.init(ld3(ct), ld3(ct + 3), ld3(ct + 6), d);
            g2.skip(1 + ofs);
        }
    }
    for (int64_t j = 0; j < K; ++j) {
        const bool have = f < 64 && first + j < total;
        V3<double> s;
        bool got = have && g2.next(s);
        const bool need = have && !got;
        if (__ballot(need)) {
            // the next open triangle after f (one exists: first + j < total)
            const unsigned long long later = f >= 63 ? 0ull : open & ~((2ull << f) - 1ull);
            const int nf = need && later ? __ffsll((long long)later) - 1 : lane;
            const double* ct = (const double*)__shfl((long long)(uintptr_t)t9, nf, 64);
            const uint64_t n0 = __shfl(mk[0], nf, 64), n1 = __shfl(mk[1], nf, 64);
            const int nh = __shfl(hl, nf, 64);
            if (need && later) {
                f = nf;
                m2[0] = n0;
                m2[1] = n1;
                h2 = nh;
                g2.init(ld3(ct), ld3(ct + 3), ld3(ct + 6), d);
                g2.skip(1);
                got = g2.next(s);
            }
        }
        const bool o2 = got && !in_envelope_masked<true>(lst, tri, nl, m2, s, eps2, sqrtT, h2);
        // the triangles found outside, to their lanes
        unsigned long long ob = __ballot(o2);
        while (ob) {
            const int l = __ffsll((long long)ob) - 1;
            ob &= ob - 1;
            if (__shfl(f, l, 64) == lane) out = true;
        }
    }
    return out;
}

// One triangle per lane, as faces_kernel, for batches walked in Morton
// order: the wave's 64 consecutive triangles lie close together, so the wave
// lists once (collect_facets, LDS) the facets whose boxes meet the box of its
// triangles padded by more than sqrt(T), with their triangles, and tests its
// samples against the list (in_envelope_list) instead of walking the tree
// per sample.  A wave whose list exceeds CAP entries walks the tree (from
// global memory: the LDS holds the lists).
template <bool F32, int CAP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void faces_list_kernel(
    BvhView B, const double* __restrict__ tri9, int64_t n, double eps2, double d, double T,
    double sqrtT, uint8_t* __restrict__ out, unsigned* n_pending, int64_t* pending,
    const int32_t* __restrict__ order, unsigned long long* lstat)
{
    constexpr int kWaveBytes = CAP * (32 + 8 + 72);
        __shared__ __align__(16) unsigned char lists[(kBlock / 64) * kWaveBytes];
    const BvhWalk W = walk(B);
    const int lane = threadIdx.x % 64;
    unsigned char* mine = lists + (threadIdx.x / 64) * kWaveBytes;
    LdsNode32W* lst = (LdsNode32W*)mine;
    LdsU32* fr = (LdsU32*)(mine + CAP * 32);
    LdsF64* tri = (LdsF64*)(mine + CAP * 40);
    int64_t i = xcd_block() * kBlock + threadIdx.x;
    bool act = false;
    const double* t = tri9;
    if (i < n) {
        if (order) i = order[i];
        t = tri9 + 9 * i;
        if (certainly_not_collinear(t, t + 3, t + 6)) {
            act = true;
        } else {
            unsigned k = atomicAdd(n_pending, 1u);
            pending[k] = i;
        }
    }
    // the box of the wave's triangles, padded
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (act)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            lo[k % 3] = t[k] < lo[k % 3] ? t[k] : lo[k % 3];
            hi[k % 3] = t[k] > hi[k % 3] ? t[k] : hi[k % 3];
        }
    double amax = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            const double ol = __shfl_xor(lo[k], off, 64), oh = __shfl_xor(hi[k], off, 64);
            lo[k] = ol < lo[k] ? ol : lo[k];
            hi[k] = oh > hi[k] ? oh : hi[k];
        }
        amax = fmax(amax, fmax(fabs(lo[k]), fabs(hi[k])));
    }
    const double pad = 1.0001 * sqrtT + 1e-9 * amax;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] -= pad;
        hi[k] += pad;
    }
    // a coordinate that is not finite can leave the lanes with different
    // boxes (min / max ignore NaN): then the whole wave walks the tree
    bool fin = true;
    if (act)
#pragma unroll
        for (int k = 0; k < 9; ++k) fin = fin && is_finite(t[k]);
    const int nl = __ballot(!fin) ? -1 : collect_facets<64>(W, lo, hi, lst, fr, CAP, lane, 0);
    if (lstat && lane == 0) atomicAdd(lstat + (nl >= 0 ? 0 : 1), 1ull);
    for (int e = lane; e < nl; e += 64) {
        const double* f = B.tri + 9 * (int64_t)lst[e].leaf;
#pragma unroll
        for (int k = 0; k < 9; ++k) tri[9 * e + k] = f[k];
    }
    list_neighbours<64>(W, lst, nl, lane);
    static_assert(CAP <= 128, "two list-mask words");
    if (nl < 0) {
        if (act) out[i] = face_out_listed<F32>(W, t, eps2, d, T, sqrtT, lst, tri, nl) ? 1 : 0;
        return;
    }
    const bool o = faces_wave_listed(act, t, eps2, d, sqrtT, lst, tri, nl, lane);
    if (act) out[i] = o ? 1 : 0;
}

// facets listed per wave in faces_group_kernel<F32, 64>
constexpr int kGroupListCap = 64;
constexpr int kGroupListBytes = kGroupListCap * (32 + 8 + 72);

// A group of G lanes per triangle, its samples spread over the lanes
// (face_out_group).  Small batches run G = 64 (a wave per triangle), so that
// few triangles with many samples each still fill the machine.  Same outputs
// and pending queue as faces_kernel.
template <bool F32, int G>
__global__ __launch_bounds__(kBlock) void faces_group_kernel(BvhView B, const double* __restrict__ tri9,
                                                             int64_t n, double eps2, double d,
                                                             double T, double sqrtT,
                                                             uint8_t* __restrict__ out,
                                                             unsigned* n_pending, int64_t* pending,
                                                             const int32_t* __restrict__ order)
{
    __shared__ __align__(16) unsigned char cache[kTreeCacheBytes];
    __shared__ __align__(16) unsigned char face_lists[G == 64 ? (kBlock / 64) * kGroupListBytes : 16];
    const BvhWalk W = cache_tree_top<F32>(B, cache);
    const int r = threadIdx.x % G;
    const int gbase = (threadIdx.x % 64) - r;
    int64_t i = (xcd_block() * (int64_t)kBlock + threadIdx.x) / G;
    if (i >= n) return;
    if (order) i = order[i];
    const double* t = tri9 + 9 * i;
    if (!certainly_not_collinear(t, t + 3, t + 6)) {
        if (r == 0) {
            unsigned k = atomicAdd(n_pending, 1u);
            pending[k] = i;
        }
        return;
    }
    int nl = -1;
    LdsNode32W* lst = nullptr;
    LdsF64* tri = nullptr;
    if (G == 64) {
        // the facets near the triangle, listed in the wave's LDS (the tree
        // walk when they do not fit or a coordinate is not finite)
        unsigned char* mine = face_lists + (threadIdx.x / 64) * kGroupListBytes;
        lst = (LdsNode32W*)mine;
        LdsU32* fr = (LdsU32*)(mine + kGroupListCap * 32);
        tri = (LdsF64*)(mine + kGroupListCap * 40);
        double lo[3], hi[3], amax = 0.0;
        bool fin = true;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            lo[j] = fmin(t[j], fmin(t[3 + j], t[6 + j]));
            hi[j] = fmax(t[j], fmax(t[3 + j], t[6 + j]));
            fin = fin && is_finite(t[j]) && is_finite(t[3 + j]) && is_finite(t[6 + j]);
            amax = fmax(amax, fmax(fabs(lo[j]), fabs(hi[j])));
        }
        const double pad = 1.0001 * sqrtT + 1e-9 * amax;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            lo[j] -= pad;
            hi[j] += pad;
        }
        // (listed from the global tree: measured ~5 % faster than through the
        // LDS copy of its top here)
        nl = fin ? collect_facets<64>(walk(B), lo, hi, lst, fr, kGroupListCap, r, 0) : -1;
        for (int e = r; e < nl; e += 64) {
            const double* f = B.tri + 9 * (int64_t)lst[e].leaf;
#pragma unroll
            for (int q = 0; q < 9; ++q) tri[9 * e + q] = f[q];
        }
        list_neighbours<64>(walk(B), lst, nl, r);
    }
    const bool o = face_out_group<F32, G, true, true>(W, ld3(t), ld3(t + 3), ld3(t + 6), eps2, d, T, sqrtT, r,
                                          gbase, lst, tri, nl);
    if (r == 0) out[i] = o ? 1 : 0;
}

template <bool F32>
__global__ __launch_bounds__(kBlock) void faces_exact_kernel(BvhView B,
                                                             const double* __restrict__ tri9,
                                                             double eps2, double d, double T,
                                                             double sqrtT, uint8_t* __restrict__ out,
                                                             const unsigned* n_pending,
                                                             const int64_t* pending)
{
    const unsigned cnt = *n_pending;
    Stats st;
    for (unsigned k = blockIdx.x * kBlock + threadIdx.x; k < cnt; k += gridDim.x * kBlock) {
        int64_t i = pending[k];
        out[i] = face_out<F32, false>(walk(B), tri9 + 9 * i, eps2, d, T, sqrtT, st) ? 1 : 0;
    }
}

__global__ void zero_u32(unsigned* p) { *p = 0u; }
__global__ void zero_u64x2(unsigned long long* p)
{
    p[0] = 0ull;
    p[1] = 0ull;
}

// batches at least this large are walked in Morton order of their queries
constexpr int64_t kSortMin = 1 << 16;
// face batches below this size run a wave per triangle (faces_group_kernel)
constexpr int64_t kWaveMax = 1 << 16;
// facets per wave list in faces_list_kernel (80: four blocks per CU; the
// kernel keeps no LDS copy of the tree top -- 48 entries with a 16 KB copy
// ran 8 % slower), and the queries per facet from which a wave's 64
// triangles (about 64 / density facets apart) usually fit it: measured on
// 1 M facets, 20 M queries ran 540 -> 720 M checks/s with the lists, 2 M
// queries 422 -> 395 (overflowing lists), so sparser batches keep faces_kernel
constexpr int kFaceListCap = 80;
constexpr int64_t kFaceListDensity = 8;
// the same for points_list_kernel (64 entries, no LDS copy of the tree top:
// 32 entries with a 16 KB copy ran 2.00 G checks/s on 20 M points, this 2.92 G)
constexpr int kPointListCap = 64;
constexpr int64_t kPointListDensity = 8;

__global__ __launch_bounds__(kBlock) void nearest_kernel(BvhView B, const double* __restrict__ p3,
                                                         int64_t n, double* __restrict__ cp3,
                                                         double* __restrict__ d2out,
                                                         int32_t* __restrict__ fout)
{
    __shared__ __align__(16) unsigned char cache[kPointCacheBytes];
    const BvhWalk W = cache_tree_top<false, kPointCacheBytes>(B, cache);
    const int64_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    V3<double> np;
    double sq;
    int best = nearest_facet(W, ld3(p3 + 3 * i), np, sq);
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
    int64_t k = 0;
    for_each_sample(ld3(t), ld3(t + 3), ld3(t + 6), d, [&](V3<double>) {
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
    int64_t k = offsets[i];
    for_each_sample(ld3(t), ld3(t + 3), ld3(t + 6), d, [&](V3<double> s) {
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
    // edge adjacency in tree order: corners identified by their coordinates'
    // bits (facets of a triangle soup share edges too); edges of one vertex
    // twice are skipped; an edge of more than two facets links them in a cycle
    std::vector<int32_t> adj(3 * nf, -1);
    {
        const int64_t nc = 3 * nf;
        auto bits = [&](int64_t c, int k) {
            uint64_t u;
            std::memcpy(&u, &tri[3 * c + k], 8);
            return u;
        };
        std::vector<int64_t> corner(nc);
        for (int64_t c = 0; c < nc; ++c) corner[c] = c;
        std::sort(corner.begin(), corner.end(), [&](int64_t a, int64_t b) {
            for (int k = 0; k < 3; ++k)
                if (bits(a, k) != bits(b, k)) return bits(a, k) < bits(b, k);
            return a < b;
        });
        std::vector<int64_t> vid(nc);
        int64_t id = -1;
        for (int64_t i = 0; i < nc; ++i) {
            const int64_t c = corner[i];
            if (i == 0 || bits(c, 0) != bits(corner[i - 1], 0) || bits(c, 1) != bits(corner[i - 1], 1) ||
                bits(c, 2) != bits(corner[i - 1], 2))
                ++id;
            vid[c] = id;
        }
        struct Edge {
            int64_t a, b;
            int64_t slot;  // 3 facet + k
        };
        std::vector<Edge> es;
        es.reserve(nc);
        for (int64_t f = 0; f < nf; ++f)
            for (int k = 0; k < 3; ++k) {
                const int64_t u = vid[3 * f + k], w = vid[3 * f + (k + 1) % 3];
                if (u != w) es.push_back({std::min(u, w), std::max(u, w), 3 * f + k});
            }
        std::sort(es.begin(), es.end(), [](const Edge& x, const Edge& y) {
            return x.a != y.a ? x.a < y.a : (x.b != y.b ? x.b < y.b : x.slot < y.slot);
        });
        for (size_t i = 0; i < es.size();) {
            size_t j = i;
            while (j < es.size() && es[j].a == es[i].a && es[j].b == es[i].b) ++j;
            if (j - i >= 2)
                for (size_t r = i; r < j; ++r) {
                    const size_t o = r + 1 < j ? r + 1 : i;
                    adj[es[r].slot] = (int32_t)(es[o].slot / 3);
                    if (j - i == 2) adj[es[o].slot] = (int32_t)(es[r].slot / 3);
                }
            i = j;
        }
    }
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
    for (int c = 0; c < 3; ++c) {
        B->lo[c] = nodes[1].lo[c];
        B->hi[c] = nodes[1].hi[c];
    }
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
    if (!r && e == hipSuccess) e = hipMalloc(&B->adj, sizeof(int32_t) * 3 * nf);
    if (!r && e == hipSuccess)
        e = hipMemcpy(B->node64, nodes.data(), sizeof(BvhNode64) * B->nnode, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess)
        e = hipMemcpy(B->node32, n32.data(), sizeof(BvhNode32) * B->nnode, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess)
        e = hipMemcpy(B->tri, tri.data(), sizeof(double) * 9 * nf, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess)
        e = hipMemcpy(B->adj, adj.data(), sizeof(int32_t) * 3 * nf, hipMemcpyHostToDevice);
    if (r || e != hipSuccess) {
        if (B->node64) (void)hipFree(B->node64);
        if (B->node32) (void)hipFree(B->node32);
        if (B->tri) (void)hipFree(B->tri);
        if (B->adj) (void)hipFree(B->adj);
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
    if (B->adj) (void)hipFree(B->adj);
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
    double T = bvh_prune_threshold(B, eps2), sT = std::sqrt(T);
    dim3 g(grid_for(n)), b(kBlock);
    BvhView v = bvh_view(B);
    const int32_t* order = nullptr;
    if (n >= kSortMin)
        if (int r = morton_order(ctx, p3, n, 3, B->lo, B->hi, &order)) return r;
    if (order && !stats && n >= kPointListDensity * B->nf) {
        if (precision == TWG_FP64)
            hipLaunchKernelGGL((points_list_kernel<false, kPointListCap>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, order, ctx->list_stats_on ? ctx->list_stats : nullptr);
        else
            hipLaunchKernelGGL((points_list_kernel<true, kPointListCap>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, order, ctx->list_stats_on ? ctx->list_stats : nullptr);
    } else if (precision == TWG_FP64) {
        if (stats) hipLaunchKernelGGL((points_kernel<false, true>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, stats, order);
        else hipLaunchKernelGGL((points_kernel<false, false>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, stats, order);
    } else {
        if (stats) hipLaunchKernelGGL((points_kernel<true, true>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, stats, order);
        else hipLaunchKernelGGL((points_kernel<true, false>), g, b, 0, ctx->cur, v, p3, n, eps2, T, sT, out, stats, order);
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
    if (int r = ensure_scratch(ctx, 256 + sizeof(int64_t) * (size_t)n)) return r;
    unsigned* n_pending = (unsigned*)ctx->scratch;
    int64_t* pending = (int64_t*)((char*)ctx->scratch + 256);
    double T = bvh_prune_threshold(B, eps2), sT = std::sqrt(T);
    dim3 g(grid_for(n)), b(kBlock);
    BvhView v = bvh_view(B);
    const int32_t* order = nullptr;
    if (n >= kSortMin)
        if (int r = morton_order(ctx, tri9, n, 9, B->lo, B->hi, &order)) return r;
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, ctx->cur, n_pending);
    TWG_LAUNCHED();
    if (n < kWaveMax && !stats) {
        dim3 gw(grid_for(n * 64));
        if (precision == TWG_FP64)
            hipLaunchKernelGGL((faces_group_kernel<false, 64>), gw, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, n_pending, pending, order);
        else
            hipLaunchKernelGGL((faces_group_kernel<true, 64>), gw, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, n_pending, pending, order);
    } else if (order && !stats && n >= kFaceListDensity * B->nf) {
        if (precision == TWG_FP64)
            hipLaunchKernelGGL((faces_list_kernel<false, kFaceListCap>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, n_pending, pending, order, ctx->list_stats_on ? ctx->list_stats : nullptr);
        else
            hipLaunchKernelGGL((faces_list_kernel<true, kFaceListCap>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, n_pending, pending, order, ctx->list_stats_on ? ctx->list_stats : nullptr);
    } else if (precision == TWG_FP64) {
        if (stats) hipLaunchKernelGGL((faces_kernel<false, true>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, stats, n_pending, pending, order);
        else hipLaunchKernelGGL((faces_kernel<false, false>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, stats, n_pending, pending, order);
    } else {
        if (stats) hipLaunchKernelGGL((faces_kernel<true, true>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, stats, n_pending, pending, order);
        else hipLaunchKernelGGL((faces_kernel<true, false>), g, b, 0, ctx->cur, v, tri9, n, eps2, sampling_dist, T, sT, out, stats, n_pending, pending, order);
    }
    TWG_LAUNCHED();
    unsigned eg = (unsigned)std::min<int64_t>(grid_for(n), 64);
    if (precision == TWG_FP64)
        hipLaunchKernelGGL((faces_exact_kernel<false>), dim3(eg), b, 0, ctx->cur, v, tri9, eps2, sampling_dist, T, sT, out, n_pending, pending);
    else
        hipLaunchKernelGGL((faces_exact_kernel<true>), dim3(eg), b, 0, ctx->cur, v, tri9, eps2, sampling_dist, T, sT, out, n_pending, pending);
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_ctx_list_stats(twg_ctx* ctx, int enable, unsigned long long* out)
{
    TWG_REQUIRE(ctx != nullptr, "twg_ctx_list_stats: ctx is NULL");
    if (int r = bind(ctx)) return r;
    if (enable) {
        if (!ctx->list_stats) TWG_HIP(hipMalloc(&ctx->list_stats, 2 * sizeof(unsigned long long)));
        hipLaunchKernelGGL(zero_u64x2, dim3(1), dim3(1), 0, ctx->cur, ctx->list_stats);
        TWG_LAUNCHED();
        ctx->list_stats_on = true;
        return TWG_OK;
    }
    unsigned long long h[2] = {0, 0};
    if (ctx->list_stats) {
        TWG_HIP(hipStreamSynchronize(ctx->cur));
        TWG_HIP(hipMemcpy(h, ctx->list_stats, sizeof h, hipMemcpyDeviceToHost));
    }
    ctx->list_stats_on = false;
    if (out) {
        out[0] = h[0];
        out[1] = h[1];
    }
    return TWG_OK;
}

int twg_nearest_points(twg_ctx* ctx, const twg_bvh* B, const double* p3, int64_t n, double* cp3,
                       double* d2, int32_t* facet)
{
    if (int r = check_env(ctx, B, p3, n, cp3, 0.0, TWG_FP64, "twg_nearest_points")) return r;
    TWG_REQUIRE(n == 0 || (d2 && facet), "twg_nearest_points: NULL output");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    hipLaunchKernelGGL(nearest_kernel, dim3(grid_for(n)), dim3(kBlock), 0, ctx->cur, bvh_view(B), p3,
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
