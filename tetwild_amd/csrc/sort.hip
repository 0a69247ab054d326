// sort.hip -- Morton ordering of query batches.  Envelope queries arrive in
// whatever order the caller produced them; walking them in Morton order makes
// neighbouring threads (and, through the XCD-aware block order, each XCD's
// L2) visit the same tree paths.  Results never depend on the order.
#include <hipcub/hipcub.hpp>

#include "twg_internal.hpp"
#include "twg_math.hpp"

namespace twg {
namespace {

__device__ __forceinline__ uint64_t spread3d(uint64_t v)
{
    v &= 0x1FFFFF;
    v = (v | v << 32) & 0x1F00000000FFFFull;
    v = (v | v << 16) & 0x1F0000FF0000FFull;
    v = (v | v << 8) & 0x100F00F00F00F00Full;
    v = (v | v << 4) & 0x10C30C30C30C30C3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ __launch_bounds__(kBlock) void morton_keys(const double* __restrict__ items, int64_t n,
                                                      int stride, double lox, double loy,
                                                      double loz, double sx, double sy, double sz,
                                                      uint64_t* __restrict__ keys,
                                                      int32_t* __restrict__ idx)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double* p = items + (int64_t)stride * i;
    double c[3];
    if (stride == 9) {
        for (int k = 0; k < 3; ++k) c[k] = (p[k] + p[3 + k] + p[6 + k]) * (1.0 / 3.0);
    } else {
        for (int k = 0; k < 3; ++k) c[k] = p[k];
    }
    double lo[3] = {lox, loy, loz}, s[3] = {sx, sy, sz};
    uint64_t key = 0;
    for (int k = 0; k < 3; ++k) {
        double q = (c[k] - lo[k]) * s[k];
        uint64_t qi = !(q > 0.0) ? 0 : (q >= 2097151.0 ? 2097151 : (uint64_t)q);
        key |= spread3d(qi) << k;
    }
    keys[i] = key;
    idx[i] = (int32_t)i;
}

}  // namespace

int morton_order(twg_ctx* ctx, const double* items, int64_t n, int stride, const double* lo,
                 const double* hi, const int32_t** order)
{
    TWG_REQUIRE(n < (int64_t(1) << 31), "morton_order: too many items");
    size_t nb = (size_t)n;
    // slot 1: keys in/out + values in/out
    size_t kv = nb * (8 + 8 + 4 + 4) + 256;
    if (int r = ensure_slot(ctx, 1, kv)) return r;
    char* base = (char*)ctx->slot[1];
    uint64_t* kin = (uint64_t*)base;
    uint64_t* kout = kin + nb;
    int32_t* vin = (int32_t*)(kout + nb);
    int32_t* vout = vin + nb;
    size_t temp = 0;
    TWG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, kin, kout, vin, vout, n, 0, 63,
                                               ctx->cur));
    if (int r = ensure_slot(ctx, 2, temp + 256)) return r;
    double s[3];
    for (int k = 0; k < 3; ++k) {
        double span = hi[k] - lo[k];
        s[k] = span > 0 ? 2097152.0 / span : 0.0;
    }
    hipLaunchKernelGGL(morton_keys, dim3(grid_for(n)), dim3(kBlock), 0, ctx->cur, items, n, stride,
                       lo[0], lo[1], lo[2], s[0], s[1], s[2], kin, vin);
    TWG_LAUNCHED();
    TWG_HIP(hipcub::DeviceRadixSort::SortPairs(ctx->slot[2], temp, kin, kout, vin, vout, n, 0, 63,
                                               ctx->cur));
    *order = vout;
    return TWG_OK;
}

}  // namespace twg
