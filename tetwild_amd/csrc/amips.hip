// amips.hip -- (a) per-tet AMIPS energy / gradient / Hessian, calTetQualities
// and the flip test, one thread per tet.
//
// Data path: the tet's four vertex ids come in as one 16-byte load; the twelve
// coordinates are gathered from the SoA x/y/z arrays (each vertex is shared by
// ~24 tets of a lattice, so the gathers mostly hit L2).  Results are per-tet
// rows (E[n], G[n][3], H[n][9] or, packed, H[n][6]) written through an LDS
// transpose, so each
// store instruction covers 64 consecutive elements and the kernel emits three
// sequential write streams (planar outputs -- thirteen streams -- measured
// ~35% slower on MI355X).  Vertex ids outside [0, nv) make the tet's outputs NaN (flip
// test: flipped) instead of reading out of bounds.
#include "twg_internal.hpp"
#include "twg_math.hpp"

namespace twg {
namespace {

__device__ __forceinline__ bool load_tet(const int32_t* __restrict__ tets, int64_t i,
                                         const double* __restrict__ x,
                                         const double* __restrict__ y,
                                         const double* __restrict__ z, int64_t nv, Tet& T)
{
    int4 t = reinterpret_cast<const int4*>(tets)[i];
    int id[4] = {t.x, t.y, t.z, t.w};
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 4; ++j) ok = ok && id[j] >= 0 && (int64_t)id[j] < nv;
    if (!ok) return false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        T.p[j][0] = x[id[j]];
        T.p[j][1] = y[id[j]];
        T.p[j][2] = z[id[j]];
    }
    return true;
}

// Block-cooperative row store: thread t of the block holds row (base + t) of
// K values; the rows go out through LDS so that the block writes one
// contiguous run (dst[(base + t) * K + k], rows >= n skipped), 16 bytes per
// lane and store instruction (1 KiB per wave) when the run is 16-byte
// aligned, one element per lane otherwise (a caller's offset pointer).
// Nontemporal: the outputs are streamed and never re-read by this kernel.
// Every thread of the block must call it.
template <int K, typename T>
__device__ __forceinline__ void block_store_rows(T* __restrict__ dst, int64_t base, int64_t n,
                                                 const T* v, T* lds)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) lds[t * K + k] = v[k];
    __syncthreads();
    int64_t rows = n - base < kBlock ? n - base : kBlock;
    int total = (int)rows * K;
    T* out = dst + base * K;
    constexpr int V = 16 / (int)sizeof(T);
    typedef T vec __attribute__((ext_vector_type(V)));
    if (((uintptr_t)out & 15) == 0) {
        const int nvec = total / V;
#pragma unroll
        for (int r = 0; r < (K + V - 1) / V; ++r) {
            const int j = r * kBlock + t;
            if (j < nvec) {
                vec w;
#pragma unroll
                for (int e = 0; e < V; ++e) w[e] = lds[V * j + e];
                __builtin_nontemporal_store(w, reinterpret_cast<vec*>(out) + j);
            }
        }
        const int j = nvec * V + t;  // the last total % V elements
        if (j < total) __builtin_nontemporal_store(lds[j], &out[j]);
    } else {
#pragma unroll
        for (int r = 0; r < K; ++r) {
            int j = r * kBlock + t;
            if (j < total) __builtin_nontemporal_store(lds[j], &out[j]);
        }
    }
    __syncthreads();
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void amips_eval_kernel(
    const int32_t* __restrict__ tets, int64_t n, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, int64_t nv, double* __restrict__ E,
    double* __restrict__ G, double* __restrict__ H)
{
    __shared__ double lds[kBlock * 9];
    const int64_t base = xcd_block() * kBlock;
    const int64_t i = base + threadIdx.x;
    double e = 0.0, g[3] = {0.0, 0.0, 0.0}, h[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) h[k] = 0.0;
    if (i < n) {
        Tet T;
        if (!load_tet(tets, i, x, y, z, nv, T)) {
            double q = __builtin_nan("");
            e = q;
#pragma unroll
            for (int k = 0; k < 3; ++k) g[k] = q;
#pragma unroll
            for (int k = 0; k < 9; ++k) h[k] = q;
        } else {
            AmipsCore c = amips_core(T);
            e = c.E;
            if (MODE >= 1) {
                double u[3], al[3], be[3];
                amips_grad(T, c, g, u, al, be);
                if (MODE >= 2) amips_hess(c, u, al, be, h);
            }
        }
        __builtin_nontemporal_store(e, &E[i]);
    }
    if (MODE >= 1) block_store_rows<3>(G, base, n, g, lds);
    if (MODE == 2) block_store_rows<9>(H, base, n, h, lds);
    if (MODE == 3) {
        // the upper triangle (h is symmetric: amips_hess writes both halves)
        const double u[6] = {h[0], h[1], h[2], h[4], h[5], h[8]};
        block_store_rows<6>(H, base, n, u, lds);
    }
}

// FP32 gradient / Hessian from FP32 edge vectors (mixed path)
__device__ __forceinline__ void grad_hess32(const AmipsCore32& c, const float e21[3],
                                            const float e31[3], float* g, float* h, bool want_h)
{
    float nx = e21[1] * e31[2] - e21[2] * e31[1];
    float ny = e21[2] * e31[0] - e21[0] * e31[2];
    float nz = e21[0] * e31[1] - e21[1] * e31[0];
    float rQ = 1.0f / c.Q;
    float rV = 1.0f / c.V6;
    float al[3] = {c.sx * rQ, c.sy * rQ, c.sz * rQ};
    float be[3] = {nx * rV, ny * rV, nz * rV};
    float u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u[k] = (2.0f / 3.0f) * be[k] - al[k];
        g[k] = c.E * u[k];
    }
    if (!want_h) return;
    float D = 3.0f * c.w;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            float t = u[i] * u[j] - al[i] * al[j] + (2.0f / 3.0f) * (be[i] * be[j]);
            float v = c.E * t;
            if (i == j) v += D;
            h[i * 3 + j] = v;
            h[j * 3 + i] = v;
        }
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void amips_mixed_kernel(
    const int32_t* __restrict__ tets, int64_t n, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, int64_t nv, float tol,
    double* __restrict__ E, float* __restrict__ G, float* __restrict__ H,
    uint8_t* __restrict__ fallback)
{
    __shared__ float lds[kBlock * 9];
    const int64_t base = xcd_block() * kBlock;
    const int64_t i = base + threadIdx.x;
    float g[3] = {0.f, 0.f, 0.f}, h[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) h[k] = 0.f;
    if (i < n) {
        Tet T;
        double e;
        bool fb;
        if (!load_tet(tets, i, x, y, z, nv, T)) {
            e = __builtin_nan("");
            fb = true;
#pragma unroll
            for (int k = 0; k < 3; ++k) g[k] = __builtin_nanf("");
#pragma unroll
            for (int k = 0; k < 9; ++k) h[k] = __builtin_nanf("");
        } else {
            float d[3][3], e21[3], e31[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                d[0][k] = (float)(T.p[1][k] - T.p[0][k]);
                d[1][k] = (float)(T.p[2][k] - T.p[0][k]);
                d[2][k] = (float)(T.p[3][k] - T.p[0][k]);
                e21[k] = (float)(T.p[2][k] - T.p[1][k]);
                e31[k] = (float)(T.p[3][k] - T.p[1][k]);
            }
            AmipsCore32 c = amips_core32(d);
            fb = !(c.rel <= tol) || !(fabsf_(c.E) < __builtin_inff());
            if (!fb) {
                e = (double)c.E;
                if (MODE >= 1) grad_hess32(c, e21, e31, g, h, MODE >= 2);
            } else {
                AmipsCore c64 = amips_core(T);
                e = c64.E;
                if (MODE >= 1) {
                    double gd[3], u[3], al[3], be[3], hd[9];
                    amips_grad(T, c64, gd, u, al, be);
                    for (int k = 0; k < 3; ++k) g[k] = (float)gd[k];
                    if (MODE >= 2) {
                        amips_hess(c64, u, al, be, hd);
                        for (int k = 0; k < 9; ++k) h[k] = (float)hd[k];
                    }
                }
            }
        }
        __builtin_nontemporal_store(e, &E[i]);
        if (fallback) fallback[i] = fb ? 1 : 0;
    }
    if (MODE >= 1) block_store_rows<3>(G, base, n, g, lds);
    if (MODE == 2) block_store_rows<9>(H, base, n, h, lds);
    if (MODE == 3) {
        // the upper triangle (both Hessian forms above write symmetric h)
        const float u[6] = {h[0], h[1], h[2], h[4], h[5], h[8]};
        block_store_rows<6>(H, base, n, u, lds);
    }
}

// calTetQuality_AMIPS semantics (src/tetwild/LocalOperations.cpp:862-884)
__device__ __forceinline__ double quality_from(double e)
{
    if (is_inf(e) || is_nan(e) || e <= 0) return kMaxEnergy;
    return e;
}

// ---- calTetQualities and isTetFlip ----------------------------------------
// Large batches run in two passes: the main kernel decides the orientation
// with the floating-point filter only and queues the tets it cannot decide
// (a marker goes to their output); a small kernel then finishes those with
// the exact predicate.  Keeping the exact predicate out of the main kernel
// keeps its register count (and stack) small.  Small batches use one kernel
// that calls the exact predicate inline.  Outputs are identical either way.

// 0x7ff8dead00000000: a NaN no quality result carries (bad input writes the
// default NaN)
constexpr uint64_t kPendingBits = 0x7ff8dead00000000ull;
constexpr uint8_t kPendingFlip = 2;
constexpr unsigned kPendingCap = 1u << 16;
constexpr int64_t kSplitMin = 1 << 16;

struct Pending {
    unsigned* count;
    int64_t* list;
};

// appends i (where pend) with one atomic per wave; every active lane calls it
__device__ __forceinline__ void push_pending(const Pending& P, int64_t i, bool pend)
{
    const unsigned long long m = __ballot(pend);
    if (m == 0ull) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(P.count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (pend) {
        const unsigned k = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        if (k < kPendingCap) P.list[k] = i;
    }
}

// quality of tet i given its orientation o (+1, -1 or 0); calTetQuality_AMIPS
// semantics, src/tetwild/LocalOperations.cpp:862-884
template <bool MIXED>
__device__ __forceinline__ double quality_oriented(const Tet& T, int o, float tol, uint8_t& fb)
{
    fb = 0;
    if (o != 1) return kMaxEnergy;
    if (!MIXED) return quality_from(amips_energy(T));
    float d[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d[0][k] = (float)(T.p[1][k] - T.p[0][k]);
        d[1][k] = (float)(T.p[2][k] - T.p[0][k]);
        d[2][k] = (float)(T.p[3][k] - T.p[0][k]);
    }
    AmipsCore32 c = amips_core32(d);
    if (c.rel <= tol && fabsf_(c.E) < __builtin_inff() && c.E > 0.0f) return (double)c.E;
    fb = 1;
    return quality_from(amips_energy(T));
}

// SPLIT: filter only, undecided tets queued; otherwise the exact predicate inline
template <bool MIXED, bool SPLIT>
__global__ __launch_bounds__(kBlock) void tet_quality_kernel(
    const int32_t* __restrict__ tets, int64_t n, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, int64_t nv, float tol,
    double* __restrict__ out, uint8_t* __restrict__ fallback, Pending P)
{
    const int64_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    Tet T;
    uint8_t fb = 0;
    double e;
    bool pend = false;
    if (!load_tet(tets, i, x, y, z, nv, T)) {
        e = __builtin_nan("");
        fb = 1;
    } else {
        int o = SPLIT ? orient3d_filter(T.p[0], T.p[1], T.p[2], T.p[3])
                      : orient3d(T.p[0], T.p[1], T.p[2], T.p[3]);
        if (SPLIT && o == kUndecided) {
            pend = true;
            e = __builtin_bit_cast(double, kPendingBits);
        } else {
            e = quality_oriented<MIXED>(T, o, tol, fb);
        }
    }
    out[i] = e;
    if (fallback) fallback[i] = fb;
    if (SPLIT) push_pending(P, i, pend);
}

// the queued tets with the exact predicate; when the queue overflowed, every
// tet still carrying the marker
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void tet_quality_exact_kernel(
    const int32_t* __restrict__ tets, int64_t n, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, int64_t nv, float tol,
    double* __restrict__ out, uint8_t* __restrict__ fallback, Pending P)
{
    const unsigned cnt = *P.count;
    if (cnt == 0) return;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const bool listed = cnt <= kPendingCap;
    const int64_t m = listed ? (int64_t)cnt : n;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += stride) {
        int64_t i = listed ? P.list[k] : k;
        if (!listed && __builtin_bit_cast(uint64_t, out[i]) != kPendingBits) continue;
        Tet T;
        load_tet(tets, i, x, y, z, nv, T);  // in range: only loaded tets are queued
        uint8_t fb;
        out[i] = quality_oriented<MIXED>(T, orient3d(T.p[0], T.p[1], T.p[2], T.p[3]), tol, fb);
        if (fallback) fallback[i] = fb;
    }
}

template <bool SPLIT>
__global__ __launch_bounds__(kBlock) void tet_flip_kernel(const int32_t* __restrict__ tets,
                                                          int64_t n, const double* __restrict__ x,
                                                          const double* __restrict__ y,
                                                          const double* __restrict__ z, int64_t nv,
                                                          uint8_t* __restrict__ flipped, Pending P)
{
    const int64_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    Tet T;
    bool pend = false;
    if (!load_tet(tets, i, x, y, z, nv, T)) {
        flipped[i] = 1;
    } else {
        int o = SPLIT ? orient3d_filter(T.p[0], T.p[1], T.p[2], T.p[3])
                      : orient3d(T.p[0], T.p[1], T.p[2], T.p[3]);
        pend = SPLIT && o == kUndecided;
        flipped[i] = pend ? kPendingFlip : (o != 1);
    }
    if (SPLIT) push_pending(P, i, pend);
}

__global__ __launch_bounds__(kBlock) void tet_flip_exact_kernel(
    const int32_t* __restrict__ tets, int64_t n, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, int64_t nv,
    uint8_t* __restrict__ flipped, Pending P)
{
    const unsigned cnt = *P.count;
    if (cnt == 0) return;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const bool listed = cnt <= kPendingCap;
    const int64_t m = listed ? (int64_t)cnt : n;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += stride) {
        int64_t i = listed ? P.list[k] : k;
        if (!listed && flipped[i] != kPendingFlip) continue;
        Tet T;
        load_tet(tets, i, x, y, z, nv, T);
        flipped[i] = orient3d(T.p[0], T.p[1], T.p[2], T.p[3]) != 1;
    }
}

__global__ void zero_count(unsigned* c) { *c = 0u; }

// the context slot holding the queue: count, then the list (256-byte aligned)
int pending_slot(twg_ctx* ctx, Pending& P)
{
    if (int r = ensure_slot(ctx, 0, 256 + sizeof(int64_t) * kPendingCap)) return r;
    P.count = static_cast<unsigned*>(ctx->slot[0]);
    P.list = reinterpret_cast<int64_t*>(static_cast<char*>(ctx->slot[0]) + 256);
    return TWG_OK;
}

// grid of the exact pass: enough for a full queue, bounded for the rescan
constexpr unsigned kExactBlocks = 1024;

int check_common(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x, const double* y,
                 const double* z, int64_t nv, const char* fn)
{
    TWG_REQUIRE(ctx != nullptr, "%s: ctx is NULL", fn);
    TWG_REQUIRE(n >= 0 && nv >= 0, "%s: negative size", fn);
    TWG_REQUIRE(n == 0 || (tets && x && y && z), "%s: NULL array", fn);
    TWG_REQUIRE(((uintptr_t)tets & 15) == 0, "%s: tets must be 16-byte aligned", fn);
    return TWG_OK;
}

}  // namespace
}  // namespace twg

using namespace twg;

extern "C" {

int twg_amips_eval(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                   const double* y, const double* z, int64_t nv, int mode, double* E, double* G,
                   double* H)
{
    if (int r = check_common(ctx, tets, n, x, y, z, nv, "twg_amips_eval")) return r;
    TWG_REQUIRE(mode >= TWG_AMIPS_E && mode <= TWG_AMIPS_EGH_SYM, "twg_amips_eval: bad mode %d",
                mode);
    TWG_REQUIRE(n == 0 || (E && (mode < 1 || G) && (mode < 2 || H)),
                "twg_amips_eval: NULL output");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    dim3 g(grid_for(n)), b(kBlock);
    if (mode == 0)
        hipLaunchKernelGGL(amips_eval_kernel<0>, g, b, 0, ctx->cur, tets, n, x, y, z, nv, E, G, H);
    else if (mode == 1)
        hipLaunchKernelGGL(amips_eval_kernel<1>, g, b, 0, ctx->cur, tets, n, x, y, z, nv, E, G, H);
    else if (mode == 2)
        hipLaunchKernelGGL(amips_eval_kernel<2>, g, b, 0, ctx->cur, tets, n, x, y, z, nv, E, G, H);
    else
        hipLaunchKernelGGL(amips_eval_kernel<3>, g, b, 0, ctx->cur, tets, n, x, y, z, nv, E, G, H);
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_amips_eval_mixed(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                         const double* y, const double* z, int64_t nv, int mode, double tol,
                         double* E, float* G, float* H, uint8_t* fallback)
{
    if (int r = check_common(ctx, tets, n, x, y, z, nv, "twg_amips_eval_mixed")) return r;
    TWG_REQUIRE(mode >= TWG_AMIPS_E && mode <= TWG_AMIPS_EGH_SYM, "twg_amips_eval_mixed: bad mode %d",
                mode);
    TWG_REQUIRE(n == 0 || (E && (mode < 1 || G) && (mode < 2 || H)),
                "twg_amips_eval_mixed: NULL output");
    TWG_REQUIRE(tol >= 0.0, "twg_amips_eval_mixed: negative tol");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    dim3 g(grid_for(n)), b(kBlock);
    float t = (float)tol;
    if (mode == 0)
        hipLaunchKernelGGL(amips_mixed_kernel<0>, g, b, 0, ctx->cur, tets, n, x, y, z, nv, t, E, G,
                           H, fallback);
    else if (mode == 1)
        hipLaunchKernelGGL(amips_mixed_kernel<1>, g, b, 0, ctx->cur, tets, n, x, y, z, nv, t, E, G,
                           H, fallback);
    else if (mode == 2)
        hipLaunchKernelGGL(amips_mixed_kernel<2>, g, b, 0, ctx->cur, tets, n, x, y, z, nv, t, E, G,
                           H, fallback);
    else
        hipLaunchKernelGGL(amips_mixed_kernel<3>, g, b, 0, ctx->cur, tets, n, x, y, z, nv, t, E, G,
                           H, fallback);
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_tet_qualities(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                      const double* y, const double* z, int64_t nv, int precision, double tol,
                      double* slim_energy, uint8_t* fallback)
{
    if (int r = check_common(ctx, tets, n, x, y, z, nv, "twg_tet_qualities")) return r;
    TWG_REQUIRE(precision == TWG_FP64 || precision == TWG_MIXED,
                "twg_tet_qualities: bad precision %d", precision);
    TWG_REQUIRE(n == 0 || slim_energy, "twg_tet_qualities: NULL output");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    dim3 g(grid_for(n)), b(kBlock);
    const float tf = (float)tol;
    const bool mixed = precision == TWG_MIXED;
    hipStream_t s = ctx->cur;
    Pending P = {nullptr, nullptr};
    if (n < kSplitMin) {
        if (mixed)
            hipLaunchKernelGGL((tet_quality_kernel<true, false>), g, b, 0, s, tets, n, x, y, z, nv,
                               tf, slim_energy, fallback, P);
        else
            hipLaunchKernelGGL((tet_quality_kernel<false, false>), g, b, 0, s, tets, n, x, y, z,
                               nv, tf, slim_energy, fallback, P);
        TWG_LAUNCHED();
        return TWG_OK;
    }
    if (int r = pending_slot(ctx, P)) return r;
    hipLaunchKernelGGL(zero_count, dim3(1), dim3(1), 0, s, P.count);
    TWG_LAUNCHED();
    if (mixed) {
        hipLaunchKernelGGL((tet_quality_kernel<true, true>), g, b, 0, s, tets, n, x, y, z, nv, tf,
                           slim_energy, fallback, P);
        TWG_LAUNCHED();
        hipLaunchKernelGGL(tet_quality_exact_kernel<true>, dim3(kExactBlocks), b, 0, s, tets, n,
                           x, y, z, nv, tf, slim_energy, fallback, P);
    } else {
        hipLaunchKernelGGL((tet_quality_kernel<false, true>), g, b, 0, s, tets, n, x, y, z, nv, tf,
                           slim_energy, fallback, P);
        TWG_LAUNCHED();
        hipLaunchKernelGGL(tet_quality_exact_kernel<false>, dim3(kExactBlocks), b, 0, s, tets, n,
                           x, y, z, nv, tf, slim_energy, fallback, P);
    }
    TWG_LAUNCHED();
    return TWG_OK;
}

int twg_tet_flip(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x, const double* y,
                 const double* z, int64_t nv, uint8_t* flipped)
{
    if (int r = check_common(ctx, tets, n, x, y, z, nv, "twg_tet_flip")) return r;
    TWG_REQUIRE(n == 0 || flipped, "twg_tet_flip: NULL output");
    if (n == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    dim3 g(grid_for(n)), b(kBlock);
    hipStream_t s = ctx->cur;
    Pending P = {nullptr, nullptr};
    if (n < kSplitMin) {
        hipLaunchKernelGGL(tet_flip_kernel<false>, g, b, 0, s, tets, n, x, y, z, nv, flipped, P);
        TWG_LAUNCHED();
        return TWG_OK;
    }
    if (int r = pending_slot(ctx, P)) return r;
    hipLaunchKernelGGL(zero_count, dim3(1), dim3(1), 0, s, P.count);
    TWG_LAUNCHED();
    hipLaunchKernelGGL(tet_flip_kernel<true>, g, b, 0, s, tets, n, x, y, z, nv, flipped, P);
    TWG_LAUNCHED();
    hipLaunchKernelGGL(tet_flip_exact_kernel, dim3(kExactBlocks), b, 0, s, tets, n, x, y, z, nv,
                       flipped, P);
    TWG_LAUNCHED();
    return TWG_OK;
}

}  // extern "C"
