// smooth.hip -- (c) the VertexSmoother step on thousands of frozen one-rings
// in one persistent launch.
//
// Per problem (a vertex and the tets around it) the kernel runs the body of
// smoothSingle (src/tetwild/VertexSmoother.cpp:145-212) or smoothSurface
// (:282-455): the validity check, NewtonsMethod (:465-542) with
// NewtonsUpdate (:627-702) and getNewEnergy (:544-625), and for surface
// vertices the projection (nearest_facet), the flip and quality checks and the
// envelope test of the one-ring surface faces.
//
// Mapping: a tile of TILE lanes (16, 32 or 64) owns a problem; lanes take the
// ring's tets TILE at a time.  Sums over the ring are formed in ring order --
// the per-tet values of a chunk go through LDS and lane q adds quantity q
// sequentially -- so every result is independent of TILE and bit-identical
// to the host restatement.  Tiles pull problems from an atomic queue in
// chunks, so rings of 4..200 tets balance across the machine.  The 3x3 solve
// runs redundantly in every lane of the tile (no broadcast needed).  With
// TWG_SMOOTH_TILE_SPLIT the launch runs the problems with an envelope stage
// on 64-lane tiles first and then the others on 32-lane tiles (the widths
// each kind runs fastest on).
//
// Mixed precision: the line-search energy sums are screened in FP32 with a
// forward error bound; only sums whose accept/reject or |dE| < 1e-5 decision
// the bound cannot settle are recomputed in FP64.  Positions, decisions and
// statuses are identical to the FP64 path.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "twg_bvh.hpp"
#include "twg_internal.hpp"

namespace twg {
namespace {

constexpr int kQ = 10;  // E, J(3), H(6 unique) per tet
typedef __attribute__((address_space(3))) unsigned long long LdsU64;
// LDS copy of the surface tree's top levels (FP64 nodes: 8 levels), read by
// the projection and by the envelope test of the one-ring faces
constexpr int kSmoothTreeCacheBytes = 16384;

struct SmoothArgs {
    const double* __restrict__ x;
    const double* __restrict__ y;
    const double* __restrict__ z;
    int64_t nv;
    const int32_t* __restrict__ center;
    const int32_t* __restrict__ ring_off;
    const int4* __restrict__ ring;
    int64_t n_ring;
    int64_t n;
    const uint8_t* __restrict__ kind;
    const double* __restrict__ old_q;
    const int32_t* __restrict__ tri_off;
    const int32_t* __restrict__ tri_ab;
    int64_t n_tri;
    BvhView sf, bt;
    bool has_sf, has_bt;
    double eps2, d, T, sqrtT;
    double* __restrict__ p_out;
    uint8_t* __restrict__ status;
    double* __restrict__ new_q;
    unsigned* counter;  // the work queues: [0], and [1] for a split launch
    unsigned chunk;
};

// LDS per block, an equal slice for each tile: the tile's Newton reductions
// (kQ doubles per lane) while it solves a problem, and while it tests the
// one-ring faces its list of nearby surface facets -- per entry a 32-byte
// record, two 4-byte work-list slots (collect_facets) and the facet's FP64
// triangle (72 bytes) -- then per face being sampled (TILE / 4 at a time) its
// list mask, then the faces' sampler state: on a 64-lane tile, when the
// faces' rows fit, their segment records and headers and a row table
// (SampleGen::store_seg, seg_row), else -- and on narrower tiles -- their
// sampler records (SampleGen::store).  The reductions and the face test never
// overlap in time.  A block takes 80 KiB with the tree cache, two blocks per
// CU.
constexpr int kTilePoolBytes = 65536;
constexpr int kListEntryBytes = 32 + 8 + 72;
constexpr int kFaceRecBytes = kFaceRec * 8;
constexpr int kFaceMaskBytes = 16;  // two 64-bit words: list positions near the face
constexpr int kFaceSegBytes = kSegRec * 8 + kSegHdr * 4;
template <int TILE>
constexpr int tile_bytes() { return kTilePoolBytes / (kBlock / TILE); }
template <int TILE>
constexpr int face_chunk() { return TILE / 4; }
// row-table entries of a tile (0: the tile samples with SampleGenC)
template <int TILE>
constexpr int seg_rows() { return TILE == 64 ? 94 : 0; }
template <int TILE>
constexpr int face_area_bytes()
{
    constexpr int rec = face_chunk<TILE>() * kFaceRecBytes;
    constexpr int seg = face_chunk<TILE>() * kFaceSegBytes + seg_rows<TILE>() * kSegRow * 8;
    return seg_rows<TILE>() > 0 && seg > rec ? seg : rec;
}
template <int TILE>
constexpr int ring_cap()
{
    return (tile_bytes<TILE>() - face_chunk<TILE>() * kFaceMaskBytes - face_area_bytes<TILE>()) / kListEntryBytes;
}
static_assert(tile_bytes<16>() >= 16 * kQ * 8 && tile_bytes<64>() >= 64 * kQ * 8, "reduction rows");
static_assert(ring_cap<64>() == 92 && ring_cap<64>() <= 128 && ring_cap<16>() >= 16, "list capacity");
static_assert((ring_cap<16>() * kListEntryBytes) % 16 == 0, "record alignment");
static_assert(kSmoothTreeCacheBytes + kTilePoolBytes <= 81920, "two blocks per CU");

typedef __attribute__((address_space(3))) int LdsI32;

template <int TILE>
struct TileCtx {
    int tl;            // lane within the tile
    int tbase;         // first thread of the tile in the block
    int wbase;         // first lane of the tile in the wave
    double* red;       // the tile's slice: kQ doubles per lane of the tile
    BvhWalk sfw;       // the surface tree, its top levels in LDS
    LdsNode32W* lst;   // the same slice: facet list (ring_cap entries),
    LdsU32* fr;        // work lists (2 ring_cap entries),
    LdsF64* tri;       // the listed facets' triangles (9 doubles each)
    LdsU64* fmask;     // per face the list positions near it (2 words)
    LdsF64* frec;      // then face_chunk sampler records (kFaceRec doubles each)
                       // or the segmented layout (SegArea)
};

// The segmented layout over a tile's sampler area: face_chunk segment
// records (kSegRec doubles each), their headers (kSegHdr ints each) and
// seg_rows row-table entries (kSegRow doubles each)
template <int TILE>
struct SegArea {
    LdsF64* rec;
    LdsI32* hdr;
    LdsF64* row;
    __device__ __forceinline__ explicit SegArea(LdsF64* area)
        : rec(area),
          hdr((LdsI32*)(area + face_chunk<TILE>() * kSegRec)),
          row(area + face_chunk<TILE>() * kFaceSegBytes / 8)
    {
    }
};
static_assert(kFaceSegBytes % 8 == 0, "row table alignment");

template <int TILE>
__device__ __forceinline__ void tile_setup(TileCtx<TILE>& c, unsigned char* pool)
{
    c.tl = threadIdx.x % TILE;
    c.tbase = threadIdx.x - c.tl;
    c.wbase = (threadIdx.x % 64) - c.tl;
    constexpr int cap = ring_cap<TILE>();
    constexpr int FC = face_chunk<TILE>();
    unsigned char* slice = pool + (threadIdx.x / TILE) * tile_bytes<TILE>();
    c.red = (double*)slice;
    c.lst = (LdsNode32W*)slice;
    c.fr = (LdsU32*)(slice + cap * sizeof(BvhNode32));
    c.tri = (LdsF64*)(slice + cap * (sizeof(BvhNode32) + 2 * sizeof(uint32_t)));
    c.fmask = (LdsU64*)(slice + cap * kListEntryBytes);
    c.frec = (LdsF64*)(slice + cap * kListEntryBytes + FC * kFaceMaskBytes);
}

// acc + v[0] + v[1] + ... + v[cnt - 1] in that order (v[r] = base[r * stride]
// in LDS): the reads go out four at a time before their additions, which
// stay sequential (the same sum as a plain loop, bit for bit)
template <int TILE>
__device__ __forceinline__ double ordered_sum(double acc, const double* base, int stride, int cnt)
{
    for (int r0 = 0; r0 < cnt; r0 += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = r0 + u < cnt ? base[(r0 + u) * stride] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (r0 + u < cnt) acc += v[u];
    }
    return acc;
}

template <int TILE>
__device__ __forceinline__ bool tile_any(const TileCtx<TILE>& c, bool pred)
{
    unsigned long long b = __ballot(pred);
    unsigned long long m = TILE == 64 ? ~0ull : ((1ull << (TILE & 63)) - 1ull);
    return ((b >> c.wbase) & m) != 0ull;
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

__device__ __forceinline__ V3<double> vtx(const SmoothArgs& A, int id)
{
    return mk3(A.x[id], A.y[id], A.z[id]);
}

// tet t with vertex v at position p; rot: v first, the rest cyclic
__device__ __forceinline__ void load_ring_tet(const SmoothArgs& A, int4 t, int v, V3<double> p,
                                              bool rot, Tet& T)
{
    int id[4] = {t.x, t.y, t.z, t.w};
    int start = 0;
    if (rot) {
#pragma unroll
        for (int j = 3; j >= 0; --j)
            if (id[j] == v) start = j;  // first occurrence
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int k = id[(start + j) & 3];
        V3<double> q = k == v ? p : vtx(A, k);
        T.p[j][0] = q.x;
        T.p[j][1] = q.y;
        T.p[j][2] = q.z;
    }
}

// id slot i (0..3) of t, by selects (no run-time array index)
__device__ __forceinline__ int pick4(int4 t, int i)
{
    return i == 0 ? t.x : i == 1 ? t.y : i == 2 ? t.z : t.w;
}

// The tile's LDS copy of the ring's leading tets for the sweeps of one
// problem, after the slice's reduction rows: kRingEntry doubles per tet -- the
// three vertices other than v, from v's slot on (cyclically), and v's slot
// (-1: the tet is read from global memory; v is not in it exactly once).
// Lane l writes and reads only the entries of its own tets j = l, l + TILE, ...
// The copy lives until the face test's facet list overwrites it (a
// projection's list stays below it).
constexpr int kRingEntry = 10;
template <int TILE>
constexpr int ring_cache_cap()
{
    return (tile_bytes<TILE>() - TILE * kQ * 8) / (kRingEntry * 8);
}
static_assert(ring_cache_cap<16>() >= 32 && ring_cache_cap<64>() >= 128, "ring copy capacity");
// a projection's facet list (records and work lists, collect_facets) stays
// below the reduction rows, so the copy outlives the projection
template <int TILE>
constexpr bool proj_list_below_rows()
{
    return ring_cap<TILE>() * (int)(sizeof(BvhNode32) + 2 * sizeof(uint32_t)) <= TILE * kQ * 8;
}
static_assert(proj_list_below_rows<16>() && proj_list_below_rows<32>() && proj_list_below_rows<64>(),
              "projection list inside the reduction rows");

struct RingSrc {
    const int4* rt;  // the ring in global memory
    int cached;      // tets [0, cached) have entries in the tile's copy
};

template <int TILE>
__device__ __forceinline__ LdsF64* ring_cache(const TileCtx<TILE>& c)
{
    // the slice's start (the sampler area follows the list and the masks),
    // then the reduction rows
    return c.frec - (ring_cap<TILE>() * kListEntryBytes + face_chunk<TILE>() * kFaceMaskBytes) / 8 + TILE * kQ;
}

// checks the ids of the lane's ring tets and writes the entries of those
// among the first ring_cache_cap; true when some id is out of range
template <int TILE>
__device__ __forceinline__ bool ring_fill(const SmoothArgs& A, const TileCtx<TILE>& c, const int4* rt, int m,
                                          int v)
{
    bool bid = false;
    for (int j = c.tl; j < m; j += TILE) {
        const int4 t = rt[j];
        const bool tb = t.x < 0 || t.x >= A.nv || t.y < 0 || t.y >= A.nv || t.z < 0 || t.z >= A.nv || t.w < 0 ||
                        t.w >= A.nv;
        bid = bid || tb;
        if (j < ring_cache_cap<TILE>()) {
            const int nvv = (t.x == v) + (t.y == v) + (t.z == v) + (t.w == v);
            const int iv = tb || nvv != 1 ? -1 : t.x == v ? 0 : t.y == v ? 1 : t.z == v ? 2 : 3;
            LdsF64* e = ring_cache(c) + kRingEntry * j;
            if (iv >= 0) {
#pragma unroll
                for (int k = 1; k < 4; ++k) {
                    const V3<double> q = vtx(A, pick4(t, (iv + k) & 3));
                    e[3 * (k - 1)] = q.x;
                    e[3 * (k - 1) + 1] = q.y;
                    e[3 * (k - 1) + 2] = q.z;
                }
            }
            e[9] = (double)iv;
        }
    }
    return bid;
}

// ring tet j with v at p, as load_ring_tet, from the tile's copy when it
// holds the tet
template <int TILE>
__device__ __forceinline__ void ring_tet(const SmoothArgs& A, const TileCtx<TILE>& c, const RingSrc& R, int j,
                                         int v, V3<double> p, bool rot, Tet& T)
{
    if (j < R.cached) {
        LdsF64* e = ring_cache(c) + kRingEntry * j;
        const int iv = (int)e[9];
        if (iv >= 0) {
            double w[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) w[k] = e[k];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                // slot jj holds v (k = 0) or the k-th vertex after v
                const int k = rot ? jj : (jj - iv) & 3;
                T.p[jj][0] = k == 0 ? p.x : k == 1 ? w[0] : k == 2 ? w[3] : w[6];
                T.p[jj][1] = k == 0 ? p.y : k == 1 ? w[1] : k == 2 ? w[4] : w[7];
                T.p[jj][2] = k == 0 ? p.z : k == 1 ? w[2] : k == 2 ? w[5] : w[8];
            }
            return;
        }
    }
    load_ring_tet(A, R.rt[j], v, p, rot, T);
}

// NewtonsUpdate: sums of E, J, H over the ring (rotated tets) in ring order
template <int TILE>
__device__ bool ring_update(const SmoothArgs& A, const TileCtx<TILE>& c, const RingSrc& R, int m,
                            int v, V3<double> pos, double& E, double J[3], double H[9])
{
    double acc = 0.0;
    for (int s = 0; s < m; s += TILE) {
        int j = s + c.tl;
        double val[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) val[q] = 0.0;
        if (j < m) {
            Tet T;
            ring_tet<TILE>(A, c, R, j, v, pos, true, T);
            AmipsCore k = amips_core(T);
            double g[3], u[3], al[3], be[3], h[9];
            amips_grad(T, k, g, u, al, be);
            amips_hess(k, u, al, be, h);
            val[0] = k.E;
            val[1] = g[0]; val[2] = g[1]; val[3] = g[2];
            val[4] = h[0]; val[5] = h[1]; val[6] = h[2];
            val[7] = h[4]; val[8] = h[5]; val[9] = h[8];
        }
        double* mine = c.red + c.tl * kQ;
#pragma unroll
        for (int q = 0; q < kQ; ++q) mine[q] = val[q];
        wave_sync();
        int cnt = m - s < TILE ? m - s : TILE;
        if (c.tl < kQ) acc = ordered_sum<TILE>(acc, c.red + c.tl, kQ, cnt);
        wave_sync();
    }
    E = __shfl(acc, 0, TILE);
    J[0] = __shfl(acc, 1, TILE);
    J[1] = __shfl(acc, 2, TILE);
    J[2] = __shfl(acc, 3, TILE);
    double h00 = __shfl(acc, 4, TILE), h01 = __shfl(acc, 5, TILE), h02 = __shfl(acc, 6, TILE);
    double h11 = __shfl(acc, 7, TILE), h12 = __shfl(acc, 8, TILE), h22 = __shfl(acc, 9, TILE);
    H[0] = h00; H[1] = h01; H[2] = h02;
    H[3] = h01; H[4] = h11; H[5] = h12;
    H[6] = h02; H[7] = h12; H[8] = h22;
    if (is_inf(E)) E = kMaxEnergy;
    if (is_nan(E) || E <= 0) return false;
    for (int k = 0; k < 3; ++k)
        if (!is_finite(J[k])) return false;
    for (int k = 0; k < 9; ++k)
        if (!is_finite(H[k])) return false;
    return true;
}

// getNewEnergy with v at X: sum in ring order, clamped to MAX_ENERGY
template <int TILE>
__device__ double ring_energy(const SmoothArgs& A, const TileCtx<TILE>& c, const RingSrc& R, int m,
                              int v, V3<double> X)
{
    double acc = 0.0;
    for (int s = 0; s < m; s += TILE) {
        int j = s + c.tl;
        double e = 0.0;
        if (j < m) {
            Tet T;
            ring_tet<TILE>(A, c, R, j, v, X, false, T);
            e = amips_energy(T);
        }
        c.red[c.tl * kQ] = e;
        wave_sync();
        int cnt = m - s < TILE ? m - s : TILE;
        if (c.tl == 0) acc = ordered_sum<TILE>(acc, c.red, kQ, cnt);
        wave_sync();
    }
    double s = __shfl(acc, 0, TILE);
    if (is_inf(s) || is_nan(s) || s <= 0 || s > kMaxEnergy) s = kMaxEnergy;
    return s;
}

// FP32 screen of ring_energy: sum of FP32 energies and a bound on its distance
// to the FP64 sum (per-tet bound from amips_core32 plus the summation error of
// both sums).  ok = false when any tet's FP32 value is unusable.
template <int TILE>
__device__ void ring_energy32(const SmoothArgs& A, const TileCtx<TILE>& c, const RingSrc& R, int m,
                              int v, V3<double> X, double& sum, double& bound, bool& ok)
{
    double s = 0.0, b = 0.0;
    bool good = true;
    for (int base = 0; base < m; base += TILE) {
        int j = base + c.tl;
        if (j < m) {
            Tet T;
            ring_tet<TILE>(A, c, R, j, v, X, false, T);
            float d[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                d[0][k] = (float)(T.p[1][k] - T.p[0][k]);
                d[1][k] = (float)(T.p[2][k] - T.p[0][k]);
                d[2][k] = (float)(T.p[3][k] - T.p[0][k]);
            }
            AmipsCore32 k = amips_core32(d);
            if (!(fabsf_(k.E) < 3.0e38f) || !(k.rel < 1.0f)) good = false;
            double e = (double)k.E;
            s += e;
            b += e * (double)k.rel / (1.0 - (double)k.rel);
        }
    }
    // any summation order: the bound covers the difference to the sequential sum
    for (int off = TILE / 2; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, TILE);
        b += __shfl_xor(b, off, TILE);
    }
    good = !tile_any(c, !good);
    sum = s;
    // b bounds |E32 - E| per tet; the FP64 energies' own error is below
    // 2^-29 of that (the 2^-20 factor covers it), and both sums carry at most
    // m u sum|terms| of summation error (u = 2^-53; 2^-51 leaves a factor 2)
    bound = b * (1.0 + 0x1p-20) + (double)m * 0x1p-51 * (s + b);
    ok = good;
}

template <int TILE>
__device__ bool ring_flip(const SmoothArgs& A, const TileCtx<TILE>& c, const RingSrc& R, int m,
                          int v, V3<double> X)
{
    bool f = false;
    for (int s = 0; s < m; s += TILE) {
        int j = s + c.tl;
        if (j < m && !f) {
            Tet T;
            ring_tet<TILE>(A, c, R, j, v, X, false, T);
            f = orient3d(T.p[0], T.p[1], T.p[2], T.p[3]) != 1;
        }
    }
    return tile_any(c, f);
}

// ring_flip over the ring tets j with skip[j] == 0 (skip may be NULL)
template <int TILE>
__device__ bool ring_flip_masked(const SmoothArgs& A, const TileCtx<TILE>& c, const RingSrc& R,
                                 int m, int v, V3<double> X, const uint8_t* skip)
{
    bool f = false;
    for (int s = 0; s < m; s += TILE) {
        int j = s + c.tl;
        if (j < m && !f && !(skip && skip[j])) {
            Tet T;
            ring_tet<TILE>(A, c, R, j, v, X, false, T);
            f = orient3d(T.p[0], T.p[1], T.p[2], T.p[3]) != 1;
        }
    }
    return tile_any(c, f);
}

// 3x3 Householder QR with column pivoting (see the host restatement).
// Every array index is a compile-time constant after unrolling (pivots and
// the rank select values instead of indexing), so the arrays stay in
// registers; the arithmetic and its order are the restatement's.
__device__ __forceinline__ void solve3_qrcp(const double* Ain, const double* b, double* x)
{
    double A[3][3];
    int perm[3] = {0, 1, 2};
    double tau[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) A[i][j] = Ain[3 * i + j];
    double norm[3];
    double maxn = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        norm[j] = __builtin_sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
        if (norm[j] > maxn) maxn = norm[j];
    }
    double thr = (maxn * 0x1p-52) * (maxn * 0x1p-52) / 3.0;
    int rank = 3;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int big = k;
        double nb = norm[k];
#pragma unroll
        for (int j = k + 1; j < 3; ++j)
            if (norm[j] > nb) {
                big = j;
                nb = norm[j];
            }
        if (rank == 3 && nb * nb < thr * (double)(3 - k)) rank = k;
        // swap column k with column big (big >= k)
#pragma unroll
        for (int j = k + 1; j < 3; ++j) {
            if (big == j) {
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    double t = A[i][k];
                    A[i][k] = A[i][j];
                    A[i][j] = t;
                }
                double t = norm[k];
                norm[k] = norm[j];
                norm[j] = t;
                int p = perm[k];
                perm[k] = perm[j];
                perm[j] = p;
            }
        }
        double tail = 0.0;
#pragma unroll
        for (int i = k + 1; i < 3; ++i) tail += A[i][k] * A[i][k];
        double c0 = A[k][k];
        double beta;
        if (tail <= 0x1p-1022) {
            tau[k] = 0.0;
            beta = c0;
#pragma unroll
            for (int i = k + 1; i < 3; ++i) A[i][k] = 0.0;
        } else {
            beta = __builtin_sqrt(c0 * c0 + tail);
            if (c0 >= 0.0) beta = -beta;
#pragma unroll
            for (int i = k + 1; i < 3; ++i) A[i][k] = A[i][k] / (c0 - beta);
            tau[k] = (beta - c0) / beta;
        }
        A[k][k] = beta;
        if (tau[k] != 0.0) {
#pragma unroll
            for (int j = k + 1; j < 3; ++j) {
                double w = A[k][j];
#pragma unroll
                for (int i = k + 1; i < 3; ++i) w += A[i][k] * A[i][j];
                A[k][j] -= tau[k] * w;
#pragma unroll
                for (int i = k + 1; i < 3; ++i) A[i][j] -= tau[k] * A[i][k] * w;
            }
        }
#pragma unroll
        for (int j = k + 1; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int i = k + 1; i < 3; ++i) s += A[i][j] * A[i][j];
            norm[j] = __builtin_sqrt(s);
        }
    }
    if (rank == 0) {
        x[0] = x[1] = x[2] = 0.0;
        return;
    }
    double c[3] = {b[0], b[1], b[2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= rank || tau[k] == 0.0) continue;
        double w = c[k];
#pragma unroll
        for (int i = k + 1; i < 3; ++i) w += A[i][k] * c[i];
        c[k] -= tau[k] * w;
#pragma unroll
        for (int i = k + 1; i < 3; ++i) c[i] -= tau[k] * A[i][k] * w;
    }
#pragma unroll
    for (int i = 2; i >= 0; --i) {
        if (i >= rank) continue;
        double s = c[i];
#pragma unroll
        for (int j = i + 1; j < 3; ++j)
            if (j < rank) s -= A[i][j] * c[j];
        c[i] = s / A[i][i];
    }
    double out[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if (i < rank && perm[i] == t) out[t] = c[i];
    x[0] = out[0];
    x[1] = out[1];
    x[2] = out[2];
}

// NewtonsMethod; returns is_moved, pout = the final position
template <int TILE, bool MIXED>
__device__ bool newtons_method(const SmoothArgs& A, const TileCtx<TILE>& c, const RingSrc& R, int m,
                               int v, V3<double> p0, V3<double>& pout)
{
    V3<double> pos = p0;
    bool is_moved = false;
    for (int step = 0; step < 15; ++step) {
        double old_energy, J[3], H[9];
        if (!ring_update<TILE>(A, c, R, m, v, pos, old_energy, J, H)) break;
        V3<double> X0 = pos;
        double a = 1.0;
        bool taken = false;
        bool tiny = false;  // |new - old| < 1e-5 for the accepted step
        for (int it = 0; it < 20; ++it) {
            double rhs[3], Xs[3];
            rhs[0] = (H[0] * X0.x + H[1] * X0.y + H[2] * X0.z) - a * J[0];
            rhs[1] = (H[3] * X0.x + H[4] * X0.y + H[5] * X0.z) - a * J[1];
            rhs[2] = (H[6] * X0.x + H[7] * X0.y + H[8] * X0.z) - a * J[2];
            solve3_qrcp(H, rhs, Xs);
            if (!(is_finite(Xs[0]) && is_finite(Xs[1]) && is_finite(Xs[2]))) {
                a /= 2.0;
                continue;
            }
            V3<double> X = mk3(Xs[0], Xs[1], Xs[2]);
            if (ring_flip<TILE>(A, c, R, m, v, X)) {
                a /= 2.0;
                continue;
            }
            bool decided = false, reject = false;
            if (MIXED) {
                double s32, b32;
                bool ok;
                ring_energy32<TILE>(A, c, R, m, v, X, s32, b32, ok);
                // the FP64 sum lies in [s32 - b32, s32 + b32] and is finite
                // below MAX_ENERGY when ok; decide what the interval allows
                if (ok && s32 - b32 > 0 && s32 + b32 < kMaxEnergy) {
                    if (s32 - b32 >= old_energy) {
                        decided = true;
                        reject = true;
                    } else if (s32 + b32 < old_energy) {
                        double lo = old_energy - (s32 + b32), hi = old_energy - (s32 - b32);
                        if (lo >= 1e-5) { decided = true; reject = false; tiny = false; }
                        else if (hi < 1e-5) { decided = true; reject = false; tiny = true; }
                    }
                }
            }
            if (!decided) {
                double new_energy = ring_energy<TILE>(A, c, R, m, v, X);
                reject = new_energy >= old_energy || is_inf(new_energy) || is_nan(new_energy);
                tiny = !reject && fabs_(new_energy - old_energy) < 1e-5;
            }
            if (reject) {
                a /= 2.0;
                continue;
            }
            taken = true;
            pos = X;
            break;
        }
        if (taken && tiny) taken = false;
        if (!taken) {
            is_moved = step != 0;
            break;
        }
        is_moved = true;
    }
    pout = pos;
    return is_moved;
}

__device__ __forceinline__ double quality_at(const Tet& T)
{
    double e;
    if (orient3d(T.p[0], T.p[1], T.p[2], T.p[3]) != 1) {
        e = kMaxEnergy;
    } else {
        e = amips_energy(T);
        if (is_inf(e) || is_nan(e)) e = kMaxEnergy;
    }
    if (is_inf(e) || is_nan(e) || e <= 0) e = kMaxEnergy;
    return e;
}

// one-ring surface face k with v at pf: (pf, a, b) as the reference forms it
__device__ __forceinline__ void ring_face(const SmoothArgs& A, int64_t k, V3<double> pf,
                                          double t9[9])
{
    const V3<double> qa = vtx(A, A.tri_ab[2 * k]), qb = vtx(A, A.tri_ab[2 * k + 1]);
    t9[0] = pf.x; t9[1] = pf.y; t9[2] = pf.z;
    t9[3] = qa.x; t9[4] = qa.y; t9[5] = qa.z;
    t9[6] = qb.x; t9[7] = qb.y; t9[8] = qb.z;
}

// smoothSurface's envelope test of the one-ring surface faces
// (src/tetwild/VertexSmoother.cpp:413-429): true when some face that is not
// collinear has a sample with no facet within eps (isFaceOutEnvelop for each
// face).  The samples of all faces are spread over the tile, TILE / 4 faces
// at a time.  Lane k sets up face k's sampler (SampleGen::init); on a 64-lane
// tile whose faces' rows fit its row table, lane k stores the face's segment
// record and header, every lane then fills rows of the table (one row per
// lane, seg_row: a row needs nothing from the rows before it) and lane k sums
// its face's rows up to the first that ends; otherwise lane k stores the
// face's sampler record and counts its samples (SampleGen::store, skip).  A
// scan over the tile gives the faces' offsets, and lane r walks samples
// [r K, r K + K) of the concatenated sequence (K = ceil(total / TILE)) --
// segment by segment (SegFace), or with SampleGenC on the records -- moving
// on to the next face with samples when one ends and carrying its facet hint;
// a ballot after each step stops the tile at the first sample outside.  A
// sample is tested against the tile's LDS list of the surface facets whose
// boxes meet the faces' padded box (collect_facets; the tree walk when the
// list does not fit), starting from the facet that held the lane's previous
// sample or, at first, from `hint` (a facet near pf, or -1).  The decision
// depends neither on the list order, the hints, nor on which lane tests which
// sample, and both samplers give SampleGen's samples bit for bit.
template <int TILE>
__device__ bool ring_faces_out(const SmoothArgs& A, const TileCtx<TILE>& c, V3<double> pf,
                               int64_t t0, int64_t t1, int hint)
{
    const BvhWalk& W = c.sfw;
    Stats st;
    if (t1 <= t0) return false;
    // (the list lives in the tile's LDS slice, over its reduction rows)
    // the surface facets that can hold a sample: those whose boxes meet the
    // faces' box padded by more than sqrt(T) (the samples lie in the faces'
    // box up to their rounding)
    int nl = -1, hl = -1;
    {
        double lo[3] = {pf.x, pf.y, pf.z}, hi[3] = {pf.x, pf.y, pf.z};
        bool fin = is_finite(pf.x) && is_finite(pf.y) && is_finite(pf.z);
        for (int64_t k = t0 + c.tl; k < t1; k += TILE) {
            const V3<double> qa = vtx(A, A.tri_ab[2 * k]), qb = vtx(A, A.tri_ab[2 * k + 1]);
            const double qs[6] = {qa.x, qa.y, qa.z, qb.x, qb.y, qb.z};
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                fin = fin && is_finite(qs[j]);
                lo[j % 3] = qs[j] < lo[j % 3] ? qs[j] : lo[j % 3];
                hi[j % 3] = qs[j] > hi[j % 3] ? qs[j] : hi[j % 3];
            }
        }
        // a coordinate that is not finite can leave the lanes with different
        // boxes (min / max ignore NaN): then the tile walks the tree
        const bool all_fin = !tile_any(c, !fin);
        double amax = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            for (int off = TILE / 2; off > 0; off >>= 1) {
                const double ol = __shfl_xor(lo[k], off, TILE), oh = __shfl_xor(hi[k], off, TILE);
                lo[k] = ol < lo[k] ? ol : lo[k];
                hi[k] = oh > hi[k] ? oh : hi[k];
            }
            amax = fmax(amax, fmax(fabs(lo[k]), fabs(hi[k])));
        }
        const double pad = 1.0001 * A.sqrtT + 1e-9 * amax;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] -= pad;
            hi[k] += pad;
        }
        wave_sync();  // the slice's reduction rows are no longer read
        // (listed from the global tree: ~1 % faster than through the LDS copy)
        nl = all_fin ? collect_facets<TILE>(walk(A.sf), lo, hi, c.lst, c.fr, ring_cap<TILE>(), c.tl, c.wbase)
                     : -1;
        for (int i = c.tl; i < nl; i += TILE) {
            const double* t = W.tri + 9 * (int64_t)c.lst[i].leaf;
#pragma unroll
            for (int k = 0; k < 9; ++k) c.tri[9 * i + k] = t[k];
        }
        wave_sync();
        if (nl > 0 && hint >= 0) {  // the hint's position (a facet appears once)
            int pos = -1;
            for (int i = c.tl; i < nl; i += TILE)
                if (c.lst[i].leaf == hint) pos = i;
            const unsigned long long m = __ballot(pos >= 0) >> c.wbase;
            const unsigned long long tm = TILE == 64 ? ~0ull : ((1ull << (TILE & 63)) - 1ull);
            if (m & tm) hl = __shfl(pos, __ffsll((long long)(m & tm)) - 1, TILE);
        }
    }
    constexpr int FC = face_chunk<TILE>();
    constexpr int RC = seg_rows<TILE>();
    const SegArea<TILE> sa(c.frec);
    const unsigned long long tm = TILE == 64 ? ~0ull : ((1ull << (TILE & 63)) - 1ull);
    for (int64_t base = t0; base < t1; base += FC) {
        const int nfc = t1 - base < FC ? (int)(t1 - base) : FC;
        int64_t cnt = 0;
        bool huge = false, sampled = false;
        int rows = 0;  // the face's row count M (0 without rows)
        float flo[3] = {0.f, 0.f, 0.f}, fhi[3] = {0.f, 0.f, 0.f};  // the face's mask box
        SampleGen g;
        if (c.tl < nfc) {
            double t9[9];
            ring_face(A, base + c.tl, pf, t9);
            if (!collinear3(t9, t9 + 3, t9 + 6)) {
                const V3<double> a = ld3(t9), b = ld3(t9 + 3), cc = ld3(t9 + 6);
                const double l = fmax(len2(a - b), fmax(len2(b - cc), len2(cc - a)));
                if (!(__builtin_sqrt(l) / A.d <= kMaxEdgeSamples)) {
                    huge = true;  // reported out without sampling
                } else {
                    g.init(a, b, cc, A.d);
                    sampled = true;
                    rows = g.phase != 9 && g.M >= 1 ? g.M : 0;
                    face_mask_box(a, b, cc, A.sqrtT, flo, fhi);
                }
            }
        }
        if (tile_any(c, huge)) return true;
        // the faces' first row-table entries (an exclusive scan of the rows)
        int row0 = 0;
        bool seg = false;
        if (RC > 0) {
            int incl = rows;
            for (int off = 1; off < TILE; off <<= 1) {
                const int o = __shfl_up(incl, off, TILE);
                if (c.tl >= off) incl += o;
            }
            row0 = incl - rows;
            seg = __shfl(incl, TILE - 1, TILE) <= RC;  // the same on every lane
        }
        if (sampled) {
            if (seg) {
                g.store_seg(sa.rec + kSegRec * c.tl, sa.hdr + kSegHdr * c.tl, row0);
            } else {
                g.store(c.frec + kFaceRec * c.tl);
                cnt = g.skip(INT64_MAX);
            }
        }
        if (nl >= 0) {
            // face_list_mask for every face of the chunk at once: lane j tests
            // list entries j, j + TILE, ... against each face's box (from the
            // face's lane), and a ballot collects each face's words; the face's
            // lane keeps them (faces without samples get an unused mask)
            constexpr int per_lane = (ring_cap<TILE>() + TILE - 1) / TILE;
            static_assert(ring_cap<TILE>() <= 128, "two mask words");
            BvhNode32 er[per_lane];
#pragma unroll
            for (int i = 0; i < per_lane; ++i)
                if (c.tl + TILE * i < nl) er[i] = from_lds(c.lst + c.tl + TILE * i);
            uint64_t my0 = 0ull, my1 = 0ull;
            for (int k = 0; k < nfc; ++k) {
                float kl[3], kh[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    kl[j] = __shfl(flo[j], k, TILE);
                    kh[j] = __shfl(fhi[j], k, TILE);
                }
                uint64_t w0 = 0ull, w1 = 0ull;
#pragma unroll
                for (int i = 0; i < per_lane; ++i) {
                    const int e = c.tl + TILE * i;
                    const bool meet = e < nl && box_meets(er[i], kl, kh);
                    const uint64_t bits = (__ballot(meet) >> c.wbase) & tm;
                    const int pos = TILE * i;  // a multiple of TILE: bits stay in one word
                    if (pos < 64) w0 |= bits << pos;
                    else w1 |= bits << (pos - 64);
                }
                if (c.tl == k) {
                    my0 = w0;
                    my1 = w1;
                }
            }
            if (c.tl < nfc) {
                c.fmask[2 * c.tl] = my0;
                c.fmask[2 * c.tl + 1] = my1;
            }
        }
        wave_sync();  // the records, headers and masks, read below by every lane of the tile
        if (seg) {
            // the rows of the chunk's faces, one lane per row: entry i is row
            // i - row0 + 1 of the face with row0 <= i < row0 + rows
            // (the faces' starts are read with every lane of the tile active)
            constexpr int IPL = RC > 0 ? (RC + TILE - 1) / TILE : 1;
            const int rtot = __shfl(row0 + rows, TILE - 1, TILE);
            int fi[IPL], ri[IPL];
#pragma unroll
            for (int q = 0; q < IPL; ++q) fi[q] = ri[q] = 0;
            for (int k = 0; k < nfc; ++k) {
                const int rk = __shfl(row0, k, TILE), mk = __shfl(rows, k, TILE);
#pragma unroll
                for (int q = 0; q < IPL; ++q)
                    if (mk > 0 && rk <= c.tl + TILE * q) {
                        fi[q] = k;
                        ri[q] = rk;
                    }
            }
#pragma unroll 1
            for (int q = 0; q < IPL; ++q) {
                const int i = c.tl + TILE * q;
                if (i >= rtot) continue;
                int f = fi[0], r0 = ri[0];
#pragma unroll
                for (int u = 1; u < IPL; ++u)
                    if (q == u) {
                        f = fi[u];
                        r0 = ri[u];
                    }
                V3<double> rs = mk3(0.0, 0.0, 0.0);
                const int rc = seg_row(sa.rec + kSegRec * f, A.d, i - r0 + 1, rs);
                LdsF64* e = sa.row + kSegRow * i;
                e[0] = rs.x;
                e[1] = rs.y;
                e[2] = rs.z;
                e[3] = (double)rc;
            }
            wave_sync();  // the row table
            if (sampled) {
                // the rows before the first that ends, and their samples
                int r = 0;
                int64_t sum = 0;
                while (r < rows) {
                    const int rc = (int)sa.row[kSegRow * (row0 + r) + 3];
                    if (rc == 0) break;
                    sum += rc;
                    ++r;
                }
                sa.hdr[kSegHdr * c.tl + kShRows] = r;
                SegFace F;
                F.load(sa.hdr + kSegHdr * c.tl);
                cnt = F.total(sum);
            }
            wave_sync();  // the headers' row counts
        }
        // faces of this chunk with samples (bit k: face base + k)
        const unsigned long long has = (__ballot(cnt > 0) >> c.wbase) & tm;
        if (has == 0ull) continue;
        // inclusive scan of the counts over the tile
        int64_t incl = cnt;
        for (int off = 1; off < TILE; off <<= 1) {
            const int64_t o = __shfl_up(incl, off, TILE);
            if (c.tl >= off) incl += o;
        }
        const int64_t total = __shfl(incl, TILE - 1, TILE);
        const int64_t K = (total + TILE - 1) / TILE;
        const int64_t first = (int64_t)c.tl * K;
        // the face holding sample `first` and that face's offset
        int f = -1;
        int64_t fstart = 0;
        for (int k = 0; k < nfc; ++k) {
            const int64_t ik = __shfl(incl, k, TILE), ck = __shfl(cnt, k, TILE);
            if (f < 0 && first < ik) {
                f = k;
                fstart = ik - ck;
            }
        }
        // the lane's place: (segment s, sample k, the segment's count sc) of
        // face f, or the record sampler
        SampleGenC<LdsF64*> gc;
        SegFace F;
        int ss = 0, sk = 0, sc = 0;
        if (f >= 0) {
            if (seg) {
                F.load(sa.hdr + kSegHdr * f);
                F.seek(sa.row, first - fstart, ss, sk);
                sc = F.count(sa.row, ss);
            } else {
                gc.load(c.frec + kFaceRec * f, A.d);
                gc.skip(first - fstart);
            }
        }
        for (int64_t j = 0; j < K; ++j) {
            bool out = false;
            if (f >= 0 && first + j < total) {
                V3<double> smp;
                bool got = true;
                if (seg) {
                    // past the segments (and faces) with no samples left; one
                    // exists, since first + j < total and the counts are the
                    // segments' own
                    while (sk >= sc) {
                        sk = 0;
                        if (++ss >= F.nseg()) {
                            const unsigned long long later = f >= 63 ? 0ull : has & ~((2ull << f) - 1ull);
                            if (later == 0ull) {
                                got = false;
                                break;
                            }
                            f = __ffsll((long long)later) - 1;
                            F.load(sa.hdr + kSegHdr * f);
                            ss = 0;
                        }
                        sc = F.count(sa.row, ss);
                    }
                    if (got) smp = F.sample(sa.rec + kSegRec * f, sa.row, ss, sk++, A.d);
                } else {
                    got = gc.next(smp);
                    while (!got) {
                        // this face is done: the next face of the chunk with samples
                        const unsigned long long later = f >= 63 ? 0ull : has & ~((2ull << f) - 1ull);
                        if (later == 0ull) break;
                        f = __ffsll((long long)later) - 1;
                        gc.load(c.frec + kFaceRec * f, A.d);
                        got = gc.next(smp);
                    }
                }
                if (got)
                    out = nl >= 0 ? !in_envelope_masked<false, true>(c.lst, c.tri, nl, c.fmask + 2 * f,
                                                                     smp, A.eps2, A.sqrtT, hl)
                                  : !in_envelope<false, false>(W, smp, A.eps2, A.T, A.sqrtT, hint, st);
            }
            if (tile_any(c, out)) return true;
        }
        wave_sync();  // the records' last reads before the next chunk's
    }
    return false;
}

template <int TILE, bool MIXED>
__device__ void solve_problem(const SmoothArgs& A, const TileCtx<TILE>& c, int64_t p)
{
    const int v = A.center[p];
    const int64_t r0 = A.ring_off[p], r1 = A.ring_off[p + 1];
    const int kind = A.kind ? A.kind[p] : 0;
    int64_t t0 = 0, t1 = 0;
    if ((kind == 1 || kind == 2) && A.tri_off) {
        t0 = A.tri_off[p];
        t1 = A.tri_off[p + 1];
    }
    // input validation: every index read below stays inside its array
    bool bad = v < 0 || v >= A.nv || r0 < 0 || r1 < r0 || r1 > A.n_ring || kind > 3 ||
               ((kind == 1 || kind == 2) && (!A.has_sf || (kind == 2 && !A.has_bt) || !A.old_q || !A.new_q ||
                              t0 < 0 || t1 < t0 || t1 > A.n_tri));
    const int m = bad ? 0 : (int)(r1 - r0);
    const int4* rt = A.ring + r0;
    if (!bad) {
        bool bid = ring_fill<TILE>(A, c, rt, m, v);
        for (int64_t k = t0 + c.tl; k < t1; k += TILE) {
            int a = A.tri_ab[2 * k], b = A.tri_ab[2 * k + 1];
            bid = bid || a < 0 || a >= A.nv || b < 0 || b >= A.nv;
        }
        bad = tile_any(c, bid);
    }
    V3<double> p0 = bad ? mk3(0.0, 0.0, 0.0) : vtx(A, v);
    // the sweeps up to a surface problem's face test read the tile's copy
    const RingSrc R = {rt, bad ? 0 : (m < ring_cache_cap<TILE>() ? m : ring_cache_cap<TILE>())};
    uint8_t st = TWG_SMOOTH_MOVED;
    V3<double> pfin = p0;
    if (bad) {
        st = TWG_SMOOTH_BAD_INPUT;
    } else if (kind == 3) {
        // NewtonsMethod alone: is_moved and the loop's final position
        V3<double> pn;
        st = newtons_method<TILE, MIXED>(A, c, R, m, v, p0, pn) ? TWG_SMOOTH_MOVED
                                                                 : TWG_SMOOTH_NOT_MOVED;
        pfin = pn;
    } else if (ring_flip<TILE>(A, c, R, m, v, p0)) {
        st = TWG_SMOOTH_INVALID;
    } else {
        V3<double> pn;
        if (!newtons_method<TILE, MIXED>(A, c, R, m, v, p0, pn)) {
            st = TWG_SMOOTH_NOT_MOVED;
        } else if (kind == 0) {
            if (ring_flip<TILE>(A, c, R, m, v, pn)) st = TWG_SMOOTH_FLIP_END;
            else pfin = pn;
        } else {
            V3<double> pf;
            double sq;
            // pf lies on the facet it was projected to: when that facet is one of
            // the envelope's, it starts every lane's facet hint below
            int hint = -1;
            wave_sync();  // the slice's reduction rows are no longer read
            if (kind == 2) {
                const int f = nearest_facet_group<TILE>(walk(A.bt), pn, pf, sq, c.lst, c.fr,
                                                        ring_cap<TILE>(), c.tl, c.wbase);
                if (A.bt.n64 == A.sf.n64) hint = f;
            } else {
                hint = nearest_facet_group<TILE>(c.sfw, pn, pf, sq, c.lst, c.fr, ring_cap<TILE>(),
                                                 c.tl, c.wbase);
            }
            if (ring_flip<TILE>(A, c, R, m, v, pf)) {
                st = TWG_SMOOTH_REJECT_FLIP;
            } else {
                double om = 0.0, nm = 0.0;
                for (int j = c.tl; j < m; j += TILE) {
                    double o = A.old_q[r0 + j];
                    om = o > om ? o : om;
                    Tet T;
                    ring_tet<TILE>(A, c, R, j, v, pf, false, T);
                    double q = quality_at(T);
                    nm = q > nm ? q : nm;
                }
                om = tile_max<TILE>(om);
                nm = tile_max<TILE>(nm);
                if (!(nm < om)) {
                    st = TWG_SMOOTH_REJECT_QUALITY;
                } else {
                    if (ring_faces_out<TILE>(A, c, pf, t0, t1, hint)) {
                        st = TWG_SMOOTH_REJECT_ENVELOPE;
                    } else {
                        pfin = pf;
                        for (int j = c.tl; j < m; j += TILE) {
                            Tet T;
                            load_ring_tet(A, rt[j], v, pf, false, T);
                            A.new_q[r0 + j] = quality_at(T);
                        }
                    }
                }
            }
        }
    }
    if (c.tl == 0) {
        A.status[p] = st;
        A.p_out[3 * p + 0] = pfin.x;
        A.p_out[3 * p + 1] = pfin.y;
        A.p_out[3 * p + 2] = pfin.z;
    }
}

template <int TILE, bool MIXED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void smooth_kernel(SmoothArgs A)
{
    __shared__ __align__(16) unsigned char tcache[kSmoothTreeCacheBytes];
    __shared__ __align__(16) unsigned char pool[kTilePoolBytes];
    TileCtx<TILE> c;
    // every thread of the block, before the loop (the copy synchronises)
    c.sfw = A.has_sf ? cache_tree_top<false, kSmoothTreeCacheBytes>(A.sf, tcache) : walk(A.sf);
    tile_setup<TILE>(c, pool);
    for (;;) {
        unsigned start = 0;
        if (c.tl == 0) start = atomicAdd(A.counter, A.chunk);
        start = __shfl(start, 0, TILE);
        if ((int64_t)start >= A.n) break;
        int64_t end = (int64_t)start + A.chunk;
        if (end > A.n) end = A.n;
        for (int64_t p = start; p < end; ++p) solve_problem<TILE, MIXED>(A, c, p);
    }
}

// The tile-width split: a problem with the envelope stage (kind 1 or 2) on
// a 64-lane tile, every other problem on a 32-lane tile, in one launch.  The
// queue of counter[0] is passed first, its surface problems taken and the
// rest skipped; then the queue of counter[1], taking the others.  The long
// surface problems start first and the short ones fill in behind them.  A
// 32-lane tile's LDS slice lies inside its wave's 64-lane slice, so a wave
// that moved on never touches a slice another wave still uses.  (32 lanes
// rather than 16 for the others: the tile's copy of the ring holds 70 tets
// instead of 35, which 96 % of the benchmark's rings fit against 73 %; 5.89
// -> 5.79 ms on its 200k batch.)
__device__ __forceinline__ bool surface_kind(const SmoothArgs& A, int64_t p)
{
    const int k = A.kind ? A.kind[p] : 0;
    return k == 1 || k == 2;
}

template <int TILE, bool MIXED, int WANT>  // WANT 1: surface kinds, 0: the others
__device__ __forceinline__ void drain(const SmoothArgs& A, const TileCtx<TILE>& c, unsigned* counter)
{
    for (;;) {
        unsigned start = 0;
        if (c.tl == 0) start = atomicAdd(counter, A.chunk);
        start = __shfl(start, 0, TILE);
        if ((int64_t)start >= A.n) break;
        int64_t end = (int64_t)start + A.chunk;
        if (end > A.n) end = A.n;
        for (int64_t p = start; p < end; ++p)
            if (surface_kind(A, p) == (WANT == 1)) solve_problem<TILE, MIXED>(A, c, p);
    }
}

template <bool MIXED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void smooth_split_kernel(SmoothArgs A)
{
    __shared__ __align__(16) unsigned char tcache[kSmoothTreeCacheBytes];
    __shared__ __align__(16) unsigned char pool[kTilePoolBytes];
    static_assert(tile_bytes<32>() * 2 == tile_bytes<64>(), "32-lane slices inside the wave's");
    BvhWalk sfw = A.has_sf ? cache_tree_top<false, kSmoothTreeCacheBytes>(A.sf, tcache) : walk(A.sf);
    {
        TileCtx<64> c;
        c.sfw = sfw;
        tile_setup<64>(c, pool);
        drain<64, MIXED, 1>(A, c, A.counter);
    }
    wave_sync();  // the 64-lane slice's last reads before its 32-lane reuse
    {
        TileCtx<32> c;
        c.sfw = sfw;
        tile_setup<32>(c, pool);
        drain<32, MIXED, 0>(A, c, A.counter + 1);
    }
}

__global__ void reset_counter(unsigned* c)
{
    c[0] = 0u;
    c[1] = 0u;
}

// One NewtonsMethod step laid out for a caller that decides some flip tests
// itself: NewtonsUpdate at p0 (ring_update, the batch kernel's sums), then
// all twenty line-search candidates with their finiteness, their flip test
// over the ring tets not skipped and their ring energy (the batch kernel's
// functions, so every value is the one its loop would see).  One wave.
__global__ __launch_bounds__(64) void newton_step_kernel(SmoothArgs A, const int4* rt, int m, int v,
                                                         V3<double> p0, const uint8_t* skip,
                                                         twg_newton_step_out* out)
{
    __shared__ double red[64 * kQ];
    TileCtx<64> c;
    c.tl = threadIdx.x;
    c.tbase = 0;
    c.wbase = 0;
    c.red = red;
    bool bid = v < 0 || v >= A.nv;
    for (int j = c.tl; j < m; j += 64) {
        int4 t = rt[j];
        bid = bid || t.x < 0 || t.x >= A.nv || t.y < 0 || t.y >= A.nv || t.z < 0 ||
              t.z >= A.nv || t.w < 0 || t.w >= A.nv;
    }
    if (tile_any(c, bid)) {
        if (c.tl == 0) out->ok = -1;
        return;
    }
    double E, J[3], H[9];
    const RingSrc R = {rt, 0};
    const bool ok = ring_update<64>(A, c, R, m, v, p0, E, J, H);
    if (c.tl == 0) {
        out->ok = ok ? 1 : 0;
        out->energy = E;
        for (int k = 0; k < 3; ++k) out->J[k] = J[k];
        for (int k = 0; k < 9; ++k) out->H[k] = H[k];
    }
    if (!ok) return;
    double a = 1.0;
    for (int it = 0; it < 20; ++it) {
        double rhs[3], Xs[3];
        rhs[0] = (H[0] * p0.x + H[1] * p0.y + H[2] * p0.z) - a * J[0];
        rhs[1] = (H[3] * p0.x + H[4] * p0.y + H[5] * p0.z) - a * J[1];
        rhs[2] = (H[6] * p0.x + H[7] * p0.y + H[8] * p0.z) - a * J[2];
        solve3_qrcp(H, rhs, Xs);
        const bool fin = is_finite(Xs[0]) && is_finite(Xs[1]) && is_finite(Xs[2]);
        bool flip = false;
        double en = 0.0;
        if (fin) {
            const V3<double> X = mk3(Xs[0], Xs[1], Xs[2]);
            flip = ring_flip_masked<64>(A, c, R, m, v, X, skip);
            en = ring_energy<64>(A, c, R, m, v, X);
        }
        if (c.tl == 0) {
            for (int k = 0; k < 3; ++k) out->X[it][k] = Xs[k];
            out->finite[it] = fin ? 1 : 0;
            out->flip[it] = flip ? 1 : 0;
            out->new_energy[it] = en;
        }
        a /= 2.0;
    }
}

// twg_smooth_gather_q: old_q[r] = q[ring_tid[r]] (NaN for an id outside q)
__global__ __launch_bounds__(kBlock) void gather_q_kernel(const int32_t* __restrict__ ring_tid,
                                                          int64_t n_ring,
                                                          const double* __restrict__ q, int64_t n_q,
                                                          double* __restrict__ old_q)
{
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= n_ring) return;
    const int32_t t = ring_tid[r];
    old_q[r] = t >= 0 && t < n_q ? q[t] : __builtin_nan("");
}

// twg_smooth_apply: one thread per problem
__global__ __launch_bounds__(kBlock) void apply_kernel(SmoothArgs A, const int32_t* __restrict__ ring_tid,
                                                       double* x, double* y, double* z, double* q,
                                                       int64_t n_q)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= A.n || A.status[p] != TWG_SMOOTH_MOVED) return;
    const int kind = A.kind ? A.kind[p] : 0;
    const int v = A.center[p];
    if (kind > 2 || v < 0 || v >= A.nv) return;
    x[v] = A.p_out[3 * p];
    y[v] = A.p_out[3 * p + 1];
    z[v] = A.p_out[3 * p + 2];
    if (kind == 0 || !A.new_q || !q) return;
    const int64_t r0 = A.ring_off[p], r1 = A.ring_off[p + 1];
    for (int64_t r = r0; r < r1 && r < A.n_ring; ++r) {
        const int32_t t = ring_tid[r];
        if (t >= 0 && t < n_q) q[t] = A.new_q[r];
    }
}

using KernelFn = void (*)(SmoothArgs);

KernelFn pick(int tile, bool mixed)
{
    if (tile == TWG_SMOOTH_TILE_SPLIT) return mixed ? smooth_split_kernel<true> : smooth_split_kernel<false>;
    if (tile == 16) return mixed ? smooth_kernel<16, true> : smooth_kernel<16, false>;
    if (tile == 32) return mixed ? smooth_kernel<32, true> : smooth_kernel<32, false>;
    return mixed ? smooth_kernel<64, true> : smooth_kernel<64, false>;
}

int make_args(twg_ctx* ctx, const twg_smooth_desc* d, unsigned* counter, unsigned chunk,
              SmoothArgs& A)
{
    TWG_REQUIRE(d != nullptr, "smooth: NULL descriptor");
    TWG_REQUIRE(d->n >= 0 && d->n < (int64_t(1) << 31) - (int64_t(1) << 24),
                "smooth: n out of range");
    TWG_REQUIRE(d->n == 0 || (d->x && d->y && d->z && d->center && d->ring_off && d->ring_tets &&
                              d->p_out && d->status),
                "smooth: NULL array");
    TWG_REQUIRE(((uintptr_t)d->ring_tets & 15) == 0, "smooth: ring_tets must be 16-byte aligned");
    TWG_REQUIRE(d->nv >= 0 && d->n_ring >= 0 && d->n_tri >= 0, "smooth: negative size");
    TWG_REQUIRE(!d->sf_tree || d->sf_tree->device == ctx->device, "smooth: sf_tree on another device");
    TWG_REQUIRE(!d->b_tree || d->b_tree->device == ctx->device, "smooth: b_tree on another device");
    TWG_REQUIRE(d->sampling_dist > 0 || !d->kind, "smooth: sampling_dist must be positive");
    A.x = d->x;
    A.y = d->y;
    A.z = d->z;
    A.nv = d->nv;
    A.center = d->center;
    A.ring_off = d->ring_off;
    A.ring = reinterpret_cast<const int4*>(d->ring_tets);
    A.n_ring = d->n_ring;
    A.n = d->n;
    A.kind = d->kind;
    A.old_q = d->old_q;
    A.tri_off = d->tri_off;
    A.tri_ab = d->tri_ab;
    A.n_tri = d->tri_ab ? d->n_tri : 0;
    A.has_sf = d->sf_tree != nullptr;
    A.has_bt = d->b_tree != nullptr;
    BvhView none = no_bvh();
    A.sf = d->sf_tree ? bvh_view(d->sf_tree) : none;
    A.bt = d->b_tree ? bvh_view(d->b_tree) : none;
    A.eps2 = d->eps2;
    A.d = d->sampling_dist;
    A.T = d->sf_tree ? bvh_prune_threshold(d->sf_tree, d->eps2) : 0.0;
    A.sqrtT = std::sqrt(A.T);
    A.p_out = d->p_out;
    A.status = d->status;
    A.new_q = d->new_q;
    A.counter = counter;
    A.chunk = chunk;
    return TWG_OK;
}

int grid_blocks(twg_ctx* ctx, KernelFn k)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k),
                                                     kBlock, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return ctx->num_cus * per_cu;
}

int enqueue(twg_ctx* ctx, const SmoothArgs& A, int tile, bool mixed, hipStream_t s)
{
    KernelFn k = pick(tile, mixed);
    int blocks = grid_blocks(ctx, k);
    const int lanes = tile == TWG_SMOOTH_TILE_SPLIT ? 64 : tile;
    int64_t need = (A.n * lanes + kBlock - 1) / kBlock;  // no more tiles than problems
    if (need < blocks) blocks = (int)std::max<int64_t>(need, 1);
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, A.counter);
    TWG_LAUNCHED();
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kBlock), 0, s, A);
    TWG_LAUNCHED();
    return TWG_OK;
}

// autotuner cache: (device, mixed, size class) -> (tile, chunk)
TuneCache g_tuned;

// Scratch of twg_smooth_batch_noalloc: the persistent kernel's two
// work-queue counters (8 bytes; 256 reserved).  One copy per device, held
// until the process ends.  In host code the name is a stand-in, not a device address:
// noalloc_scratch_init looks the address up (hipGetSymbolAddress) and every
// launch takes that address.
constexpr size_t kNoallocScratchBytes = 256;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

int size_class(int64_t n) { return n < 4096 ? 0 : (n < 65536 ? 1 : 2); }

bool valid_tile(int t) { return t == 0 || t == 16 || t == 32 || t == 64 || t == TWG_SMOOTH_TILE_SPLIT; }

}  // namespace

int noalloc_scratch_init(twg_ctx* ctx)
{
    void* ptr = nullptr;
    TWG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    ctx->noalloc_scratch = ptr;
    return TWG_OK;
}

}  // namespace twg

using namespace twg;

struct twg_smooth_plan {
    twg_ctx* ctx = nullptr;
    SmoothArgs args;
    unsigned* counter = nullptr;
    int tile = 32;
    unsigned chunk = 1;
    bool mixed = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
};

extern "C" {

int twg_smooth_autotune(twg_ctx* ctx, const twg_smooth_desc* d, int precision, int* tile_out,
                        int* chunk_out, float* ms_out)
{
    TWG_REQUIRE(ctx != nullptr, "twg_smooth_autotune: ctx is NULL");
    TWG_REQUIRE(precision == TWG_FP64 || precision == TWG_MIXED, "twg_smooth_autotune: bad precision");
    if (int r = bind(ctx)) return r;
    if (int r = ensure_scratch(ctx, 256)) return r;
    SmoothArgs A;
    if (int r = make_args(ctx, d, (unsigned*)ctx->scratch, 1, A)) return r;
    const int tiles[4] = {16, 32, 64, TWG_SMOOTH_TILE_SPLIT};
    const unsigned chunks[3] = {1, 4, 16};
    EventPair ev;
    TWG_HIP(ev.create());
    float best = 1e30f;
    int bt = 32;
    unsigned bc = 1;
    for (int t : tiles)
        for (unsigned ch : chunks) {
            A.chunk = ch;
            if (int r = enqueue(ctx, A, t, precision == TWG_MIXED, ctx->cur)) return r;  // warm
            TWG_HIP(hipEventRecord(ev.e0, ctx->cur));
            for (int rep = 0; rep < 2; ++rep)
                if (int r = enqueue(ctx, A, t, precision == TWG_MIXED, ctx->cur)) return r;
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
    g_tuned.put(ctx->device, precision == TWG_MIXED, size_class(d->n), bt, bc);
    if (tile_out) *tile_out = bt;
    if (chunk_out) *chunk_out = (int)bc;
    if (ms_out) *ms_out = best;
    return TWG_OK;
}

int twg_smooth_batch(twg_ctx* ctx, const twg_smooth_desc* d, int precision, int tile, int chunk)
{
    TWG_REQUIRE(ctx != nullptr, "twg_smooth_batch: ctx is NULL");
    TWG_REQUIRE(precision == TWG_FP64 || precision == TWG_MIXED, "twg_smooth_batch: bad precision");
    TWG_REQUIRE(valid_tile(tile), "twg_smooth_batch: tile must be 0, 16, 32, 64 or TWG_SMOOTH_TILE_SPLIT");
    TWG_REQUIRE(chunk >= 0 && chunk <= 1024, "twg_smooth_batch: chunk out of range");
    TWG_REQUIRE(d != nullptr, "twg_smooth_batch: NULL descriptor");
    if (d->n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    if (int r = ensure_scratch(ctx, 256)) return r;
    int tt = 32;
    unsigned tc = 1u;
    g_tuned.get(ctx->device, precision == TWG_MIXED, size_class(d->n), tt, tc);
    int t = tile ? tile : tt;
    unsigned ch = chunk ? (unsigned)chunk : tc;
    SmoothArgs A;
    if (int r = make_args(ctx, d, (unsigned*)ctx->scratch, ch, A)) return r;
    return enqueue(ctx, A, t, precision == TWG_MIXED, ctx->cur);
}

int twg_smooth_batch_noalloc(twg_ctx* ctx, const twg_smooth_desc* d, int precision, int tile,
                             int chunk)
{
    TWG_REQUIRE(ctx != nullptr, "twg_smooth_batch_noalloc: ctx is NULL");
    TWG_REQUIRE(precision == TWG_FP64 || precision == TWG_MIXED,
                "twg_smooth_batch_noalloc: bad precision");
    TWG_REQUIRE(valid_tile(tile),
                "twg_smooth_batch_noalloc: tile must be 0, 16, 32, 64 or TWG_SMOOTH_TILE_SPLIT");
    TWG_REQUIRE(chunk >= 0 && chunk <= 1024, "twg_smooth_batch_noalloc: chunk out of range");
    TWG_REQUIRE(d != nullptr, "twg_smooth_batch_noalloc: NULL descriptor");
    // the scratch holds the two 32-bit queue counters: problems popped in
    // chunks of up to 1024 must stay below 2^31 (make_args checks n)
    TWG_REQUIRE(ctx->noalloc_scratch != nullptr, "twg_smooth_batch_noalloc: no scratch address");
    if (d->n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    int tt = 32;
    unsigned tc = 1u;
    g_tuned.get(ctx->device, precision == TWG_MIXED, size_class(d->n), tt, tc);
    int t = tile ? tile : tt;
    unsigned ch = chunk ? (unsigned)chunk : tc;
    SmoothArgs A;
    if (int r = make_args(ctx, d, (unsigned*)ctx->noalloc_scratch, ch, A)) return r;
    // enqueue zeroes the counter with a kernel (no memset node under capture)
    return enqueue(ctx, A, t, precision == TWG_MIXED, ctx->cur);
}

int twg_smooth_plan_create(twg_ctx* ctx, const twg_smooth_desc* d, int precision, int tile,
                           int chunk, twg_smooth_plan** out)
{
    TWG_REQUIRE(ctx && d && out, "twg_smooth_plan_create: NULL argument");
    TWG_REQUIRE(precision == TWG_FP64 || precision == TWG_MIXED, "twg_smooth_plan_create: bad precision");
    TWG_REQUIRE(valid_tile(tile), "twg_smooth_plan_create: bad tile");
    TWG_REQUIRE(chunk >= 0 && chunk <= 1024, "twg_smooth_plan_create: chunk out of range");
    if (int r = bind(ctx)) return r;
    twg_smooth_plan* P = new twg_smooth_plan();
    P->ctx = ctx;
    P->mixed = precision == TWG_MIXED;
    int tt = 32;
    unsigned tc = 1u;
    g_tuned.get(ctx->device, P->mixed, size_class(d->n), tt, tc);
    P->tile = tile ? tile : tt;
    P->chunk = chunk ? (unsigned)chunk : tc;
    hipError_t e = hipMalloc(&P->counter, 256);
    if (e != hipSuccess) {
        delete P;
        return hip_fail(e, "twg_smooth_plan_create", __FILE__, __LINE__);
    }
    if (int r = make_args(ctx, d, P->counter, P->chunk, P->args)) {
        (void)hipFree(P->counter);
        delete P;
        return r;
    }
    // capture [reset counter, persistent kernel] into a graph on a private
    // default-priority stream
    e = hipStreamCreateWithFlags(&P->capture_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(P->capture_stream, hipStreamCaptureModeThreadLocal);
    int r = TWG_OK;
    if (e == hipSuccess) r = enqueue(ctx, P->args, P->tile, P->mixed, P->capture_stream);
    hipError_t e2 = hipStreamEndCapture(P->capture_stream, &P->graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess && r == TWG_OK) e = hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0);
    if (e != hipSuccess || r != TWG_OK) {
        if (P->graph) (void)hipGraphDestroy(P->graph);
        if (P->capture_stream) (void)hipStreamDestroy(P->capture_stream);
        (void)hipFree(P->counter);
        delete P;
        return r ? r : hip_fail(e, "twg_smooth_plan_create capture", __FILE__, __LINE__);
    }
    *out = P;
    return TWG_OK;
}

int twg_smooth_plan_run(twg_smooth_plan* P)
{
    TWG_REQUIRE(P != nullptr, "twg_smooth_plan_run: NULL plan");
    if (int r = bind(P->ctx)) return r;
    TWG_HIP(hipGraphLaunch(P->exec, P->ctx->cur));
    return TWG_OK;
}

int twg_smooth_plan_info(const twg_smooth_plan* P, int* tile, int* chunk)
{
    TWG_REQUIRE(P != nullptr, "twg_smooth_plan_info: NULL plan");
    if (tile) *tile = P->tile;
    if (chunk) *chunk = (int)P->chunk;
    return TWG_OK;
}

int twg_smooth_gather_q(twg_ctx* ctx, const twg_smooth_desc* d, const int32_t* ring_tid,
                        const double* q, int64_t n_q)
{
    TWG_REQUIRE(ctx != nullptr && d != nullptr, "twg_smooth_gather_q: NULL argument");
    TWG_REQUIRE(d->n_ring >= 0 && n_q >= 0, "twg_smooth_gather_q: negative size");
    if (d->n_ring == 0) return TWG_OK;
    TWG_REQUIRE(ring_tid && q && d->old_q, "twg_smooth_gather_q: NULL array");
    if (int r = bind(ctx)) return r;
    hipLaunchKernelGGL(gather_q_kernel, dim3(grid_for(d->n_ring)), dim3(kBlock), 0, ctx->cur, ring_tid,
                       d->n_ring, q, n_q, const_cast<double*>(d->old_q));
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_smooth_apply(twg_ctx* ctx, const twg_smooth_desc* d, const int32_t* ring_tid, double* x,
                     double* y, double* z, double* q, int64_t n_q)
{
    TWG_REQUIRE(ctx != nullptr && d != nullptr, "twg_smooth_apply: NULL argument");
    TWG_REQUIRE(d->n >= 0 && n_q >= 0, "twg_smooth_apply: negative size");
    if (d->n == 0) return TWG_OK;
    TWG_REQUIRE(x && y && z && d->center && d->ring_off && d->p_out && d->status,
                "twg_smooth_apply: NULL array");
    TWG_REQUIRE(!q || ring_tid, "twg_smooth_apply: q needs ring_tid");
    if (int r = bind(ctx)) return r;
    SmoothArgs A;
    if (int r = make_args(ctx, d, nullptr, 1, A)) return r;
    hipLaunchKernelGGL(apply_kernel, dim3(grid_for(d->n)), dim3(kBlock), 0, ctx->cur, A, ring_tid, x, y,
                       z, q, n_q);
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_newton_step(twg_ctx* ctx, const double* x, const double* y, const double* z, int64_t nv,
                    const int32_t* ring_tets, int32_t m, int32_t v, double px, double py,
                    double pz, const uint8_t* skip, twg_newton_step_out* out)
{
    TWG_REQUIRE(ctx != nullptr, "twg_newton_step: ctx is NULL");
    TWG_REQUIRE(x && y && z && out && (m == 0 || ring_tets), "twg_newton_step: NULL array");
    TWG_REQUIRE(nv > 0 && nv < (int64_t(1) << 31) && m >= 0, "twg_newton_step: bad size");
    TWG_REQUIRE(v >= 0 && v < nv, "twg_newton_step: v out of range");
    TWG_REQUIRE(((uintptr_t)ring_tets & 15) == 0, "twg_newton_step: ring_tets must be 16-byte aligned");
    if (int r = bind(ctx)) return r;
    SmoothArgs A;
    std::memset(&A, 0, sizeof A);
    A.x = x;
    A.y = y;
    A.z = z;
    A.nv = nv;
    hipLaunchKernelGGL(newton_step_kernel, dim3(1), dim3(64), 0, ctx->cur, A,
                       reinterpret_cast<const int4*>(ring_tets), (int)m, (int)v, mk3(px, py, pz),
                       skip, out);
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_smooth_plan_destroy(twg_smooth_plan* P)
{
    if (!P) return TWG_OK;
    (void)hipSetDevice(P->ctx->device);
    (void)hipStreamSynchronize(P->ctx->cur);
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->capture_stream) (void)hipStreamDestroy(P->capture_stream);
    if (P->counter) (void)hipFree(P->counter);
    delete P;
    return TWG_OK;
}

}  // extern "C"
