// reduce.hip -- the energy reductions that drive pass termination in
// MeshRefinement (src/tetwild/MeshRefinement.cpp:185-206, 285-292):
// getAvgMaxEnergy, getMaxEnergy, getSecondMaxEnergy and the histogram of
// getFilterEnergy (src/tetwild/LocalOperations.cpp:605-693), in one pass over
// the per-tet slim energies resident in HBM.
//
// Each block walks a contiguous range of the tets; blocks reduce through LDS
// in a fixed tree and write partials; one block reduces the partials in the same fixed
// order.  The result is deterministic for a given device (the grid depends
// only on the CU count).  Max, second max, counts and the histogram are exact;
// the sum differs from the reference's sequential loop by rounding only.
#include <algorithm>
#include <cmath>

#include "twg_internal.hpp"
#include "twg_math.hpp"

namespace twg {
namespace {

struct Partial {
    double sum, max, second;
    unsigned long long count, bad;
    unsigned long long buckets[11];
};

constexpr int kUnroll = 4;

struct ReduceArgs {
    const double* __restrict__ e;
    int64_t n;
    const uint8_t* __restrict__ removed;
    const int4* __restrict__ tets;
    const uint8_t* __restrict__ locked;
    int64_t nv;
    double lim[11];  // filter_energy_thres - 1 + 10^j, j = 0..10
    Partial* part;
};

__device__ __forceinline__ void combine(Partial& a, const Partial& b)
{
    a.sum += b.sum;
    a.max = b.max > a.max ? b.max : a.max;
    a.second = b.second > a.second ? b.second : a.second;
    a.count += b.count;
    a.bad += b.bad;
#pragma unroll
    for (int k = 0; k < 11; ++k) a.buckets[k] += b.buckets[k];
}

// fixed-shape tree over the block: thread t combines t + s for s = 128 .. 1
__device__ void block_reduce(Partial& p, Partial* lds)
{
    const int t = threadIdx.x;
    lds[t] = p;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s) {
            Partial q = lds[t];
            combine(q, lds[t + s]);
            lds[t] = q;
        }
        __syncthreads();
    }
    p = lds[0];
}

__global__ __launch_bounds__(kBlock) void energy_partial_kernel(ReduceArgs A)
{
    __shared__ Partial lds[kBlock];
    Partial p;
    p.sum = 0.0;
    p.max = 0.0;
    p.second = 0.0;
    p.count = 0;
    p.bad = 0;
#pragma unroll
    for (int k = 0; k < 11; ++k) p.buckets[k] = 0;
    // Each block walks its own contiguous range (blocks in XCD order, so the
    // vertex flags its tets read stay in one L2), kUnroll coalesced elements
    // per thread per step with their loads issued before any is used.
    const int64_t step = (int64_t)kBlock * kUnroll;
    const int64_t per = ((A.n + gridDim.x - 1) / gridDim.x + step - 1) / step * step;
    const int64_t lo = xcd_block() * per;
    const int64_t hi = lo + per < A.n ? lo + per : A.n;
    const bool use_locks = A.tets && A.locked;
    for (int64_t base = lo + threadIdx.x; base < hi; base += step) {
        uint8_t rm[kUnroll];
        double ev[kUnroll];
        int4 tv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int64_t i = base + (int64_t)u * kBlock;
            const bool in = i < hi;
            rm[u] = in ? (A.removed ? A.removed[i] : 0) : 1;
            ev[u] = in ? A.e[i] : 0.0;
            tv[u] = in && use_locks ? A.tets[i] : make_int4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (rm[u]) continue;
            const double e = ev[u];
            // getFilterEnergy: non-removed tets, locks ignored (:656-671); the
            // bucket index first, then an unrolled update (no run-time indexing)
            int bk = -1;
            if (e > A.lim[10]) {
                bk = 10;
            } else {
#pragma unroll
                for (int j = 9; j >= 0; --j)
                    if (e > A.lim[j] && e <= A.lim[j + 1]) bk = j;
            }
#pragma unroll
            for (int k = 0; k < 11; ++k) p.buckets[k] += (bk == k) ? 1ull : 0ull;
            // isTetLocked_ui (:1377-1384)
            if (use_locks) {
                const int4 t = tv[u];
                if (t.x < 0 || t.x >= A.nv || t.y < 0 || t.y >= A.nv || t.z < 0 || t.z >= A.nv ||
                    t.w < 0 || t.w >= A.nv) {
                    p.bad++;
                    continue;
                }
                if (A.locked[t.x] || A.locked[t.y] || A.locked[t.z] || A.locked[t.w]) continue;
            }
            // getAvgMaxEnergy / getMaxEnergy (:605-635), getSecondMaxEnergy (:637-650)
            p.max = e > p.max ? e : p.max;
            if (e != kMaxEnergy) p.second = e > p.second ? e : p.second;
            p.sum += e;
            p.count++;
        }
    }
    block_reduce(p, lds);
    if (threadIdx.x == 0) A.part[blockIdx.x] = p;
}

__global__ __launch_bounds__(kBlock) void energy_final_kernel(Partial* part, int nblocks,
                                                              Partial* out)
{
    __shared__ Partial lds[kBlock];
    Partial p;
    p.sum = 0.0;
    p.max = 0.0;
    p.second = 0.0;
    p.count = 0;
    p.bad = 0;
#pragma unroll
    for (int k = 0; k < 11; ++k) p.buckets[k] = 0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) combine(p, part[b]);
    block_reduce(p, lds);
    if (threadIdx.x == 0) *out = p;
}

}  // namespace
}  // namespace twg

using namespace twg;

extern "C" int twg_energy_stats(twg_ctx* ctx, const double* slim_energy, int64_t n,
                                const uint8_t* t_is_removed, const int32_t* tets,
                                const uint8_t* v_is_locked, int64_t nv,
                                double filter_energy_thres, twg_energy_summary* out)
{
    TWG_REQUIRE(ctx != nullptr && out != nullptr, "twg_energy_stats: NULL argument");
    TWG_REQUIRE(n >= 0 && nv >= 0, "twg_energy_stats: negative size");
    TWG_REQUIRE(n == 0 || slim_energy, "twg_energy_stats: NULL slim_energy");
    TWG_REQUIRE(!tets || ((uintptr_t)tets & 15) == 0, "twg_energy_stats: tets must be 16-byte aligned");
    ReduceArgs A;
    A.e = slim_energy;
    A.n = n;
    A.removed = t_is_removed;
    A.tets = reinterpret_cast<const int4*>(tets);
    A.locked = v_is_locked;
    A.nv = nv;
    // the bucket limits as the reference forms them: thres - 1 + pow(10, j)
    for (int j = 0; j <= 10; ++j) A.lim[j] = filter_energy_thres - 1 + std::pow(10.0, (double)j);
    if (int r = bind(ctx)) return r;
    const int nblocks = (int)std::max<int64_t>(1, std::min<int64_t>(grid_for(n), (int64_t)ctx->num_cus * 4));
    if (int r = ensure_scratch(ctx, sizeof(Partial) * (size_t)(nblocks + 1))) return r;
    A.part = static_cast<Partial*>(ctx->scratch);
    Partial* dout = A.part + nblocks;
    hipLaunchKernelGGL(energy_partial_kernel, dim3(nblocks), dim3(kBlock), 0, ctx->cur, A);
    TWG_LAUNCHED();
    hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(kBlock), 0, ctx->cur, A.part, nblocks, dout);
    TWG_LAUNCHED();
    Partial h;
    TWG_HIP(hipMemcpyAsync(&h, dout, sizeof h, hipMemcpyDeviceToHost, ctx->cur));
    TWG_HIP(hipStreamSynchronize(ctx->cur));
    TWG_REQUIRE(h.bad == 0, "twg_energy_stats: %llu tets with vertex ids outside [0, nv)",
                (unsigned long long)h.bad);
    out->sum = h.sum;
    out->max = h.max;
    out->second_max = h.second;
    out->count = (int64_t)h.count;
    for (int k = 0; k < 11; ++k) out->buckets[k] = (int64_t)h.buckets[k];
    return TWG_OK;
}
