// groups.hip -- isFlip + calTetQualities + getCheckQuality over many small
// groups of tets in one launch: the evaluation step of every configuration
// EdgeRemover tries (src/tetwild/EdgeRemover.cpp:138-174 for 3-2, 327-368 for
// the two rotations of 4-4, 501-575 for the tet pairs of 5-6), and of any
// other caller that checks a handful of candidate tets at a time.
//
// One thread per group (configurations hold 2..6 tets).  Orientation comes
// from the floating-point filter; a group with no tet certainly flipped but
// some tet undecided is marked and finished by a second pass with the exact
// predicate, which keeps the exact code's registers out of the main kernel.
#include <algorithm>

#include "twg_internal.hpp"
#include "twg_math.hpp"

namespace twg {
namespace {

constexpr uint8_t kGroupPending = 3;

struct GroupArgs {
    const int4* __restrict__ tets;
    const int32_t* __restrict__ off;
    int64_t n_groups, n_tets;
    const double* __restrict__ x;
    const double* __restrict__ y;
    const double* __restrict__ z;
    int64_t nv;
    uint8_t* __restrict__ flipped;
    double* __restrict__ max_q;
    double* __restrict__ q;
};

__device__ __forceinline__ bool load(const GroupArgs& A, int4 t, Tet& T)
{
    const int id[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (id[j] < 0 || id[j] >= A.nv) return false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        T.p[j][0] = A.x[id[j]];
        T.p[j][1] = A.y[id[j]];
        T.p[j][2] = A.z[id[j]];
    }
    return true;
}

// groups of at most this many tets are evaluated in one sweep
constexpr int kSweepMax = 6;

// EXACT: every orientation decided exactly; otherwise undecided groups are
// marked kGroupPending
template <bool EXACT>
__device__ void eval_group(const GroupArgs& A, int64_t g)
{
    const int64_t b = A.off[g], e = A.off[g + 1];
    if (b < 0 || e < b || e > A.n_tets) {
        A.flipped[g] = TWG_GROUP_BAD_INPUT;
        A.max_q[g] = __builtin_nan("");
        return;
    }
    if (e - b <= kSweepMax) {
        // one sweep: each tet's vertices are loaded once for its orientation
        // and its quality; the qualities are stored only when the group
        // turns out not flipped (same outputs as the two sweeps below)
        const int cnt = (int)(e - b);
        int4 tv[kSweepMax];
#pragma unroll
        for (int k = 0; k < kSweepMax; ++k) tv[k] = k < cnt ? A.tets[b + k] : make_int4(0, 0, 0, 0);
        bool flip = false, undecided = false, bad = false;
        double qv[kSweepMax];
        double m = 0.0;
#pragma unroll
        for (int k = 0; k < kSweepMax; ++k) {
            qv[k] = 0.0;
            if (k >= cnt || flip || bad) continue;
            Tet T;
            if (!load(A, tv[k], T)) {
                bad = true;
                continue;
            }
            const int o = EXACT ? orient3d(T.p[0], T.p[1], T.p[2], T.p[3])
                                : orient3d_filter(T.p[0], T.p[1], T.p[2], T.p[3]);
            if (o == kUndecided) undecided = true;
            else flip = o != 1;
            double q = amips_energy(T);
            if (is_inf(q) || is_nan(q) || q <= 0) q = kMaxEnergy;
            qv[k] = q;
            m = q > m ? q : m;
        }
        if (bad) {
            A.flipped[g] = TWG_GROUP_BAD_INPUT;
            A.max_q[g] = __builtin_nan("");
        } else if (flip) {
            A.flipped[g] = 1;
            A.max_q[g] = __builtin_nan("");
        } else if (!EXACT && undecided) {
            A.flipped[g] = kGroupPending;
        } else {
            if (A.q)
#pragma unroll
                for (int k = 0; k < kSweepMax; ++k)
                    if (k < cnt) A.q[b + k] = qv[k];
            A.flipped[g] = 0;
            A.max_q[g] = m;
        }
        return;
    }
    // isFlip (src/tetwild/LocalOperations.cpp:555-573): any tet not POSITIVE
    bool flip = false, undecided = false, bad = false;
    for (int64_t k = b; k < e && !flip && !bad; ++k) {
        Tet T;
        if (!load(A, A.tets[k], T)) {
            bad = true;
            break;
        }
        const int o = EXACT ? orient3d(T.p[0], T.p[1], T.p[2], T.p[3])
                            : orient3d_filter(T.p[0], T.p[1], T.p[2], T.p[3]);
        if (o == kUndecided) undecided = true;
        else flip = o != 1;
    }
    if (bad) {
        A.flipped[g] = TWG_GROUP_BAD_INPUT;
        A.max_q[g] = __builtin_nan("");
        return;
    }
    if (flip) {
        A.flipped[g] = 1;
        A.max_q[g] = __builtin_nan("");
        return;
    }
    if (!EXACT && undecided) {
        A.flipped[g] = kGroupPending;
        return;
    }
    // calTetQualities of positively oriented tets (src/tetwild/LocalOperations.cpp:
    // 695-773, 862-884) and getCheckQuality's max (:575-588)
    double m = 0.0;
    for (int64_t k = b; k < e; ++k) {
        Tet T;
        load(A, A.tets[k], T);
        double q = amips_energy(T);
        if (is_inf(q) || is_nan(q) || q <= 0) q = kMaxEnergy;
        if (A.q) A.q[k] = q;
        m = q > m ? q : m;
    }
    A.flipped[g] = 0;
    A.max_q[g] = m;
}

__global__ __launch_bounds__(kBlock) void groups_kernel(GroupArgs A)
{
    const int64_t g = xcd_block() * (int64_t)kBlock + threadIdx.x;
    if (g < A.n_groups) eval_group<false>(A, g);
}

__global__ __launch_bounds__(kBlock) void groups_exact_kernel(GroupArgs A)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < A.n_groups; g += stride)
        if (A.flipped[g] == kGroupPending) eval_group<true>(A, g);
}

}  // namespace
}  // namespace twg

using namespace twg;

extern "C" int twg_tet_groups(twg_ctx* ctx, const int32_t* tets, const int32_t* group_off,
                              int64_t n_groups, int64_t n_tets, const double* x, const double* y,
                              const double* z, int64_t nv, uint8_t* flipped, double* max_q,
                              double* q)
{
    TWG_REQUIRE(ctx != nullptr, "twg_tet_groups: NULL context");
    TWG_REQUIRE(n_groups >= 0 && n_tets >= 0 && nv >= 0, "twg_tet_groups: negative size");
    TWG_REQUIRE(n_groups == 0 || (tets && group_off && x && y && z && flipped && max_q),
                "twg_tet_groups: NULL array");
    TWG_REQUIRE(!tets || ((uintptr_t)tets & 15) == 0, "twg_tet_groups: tets must be 16-byte aligned");
    if (n_groups == 0) return TWG_OK;
    if (int r = bind(ctx)) return r;
    GroupArgs A;
    A.tets = reinterpret_cast<const int4*>(tets);
    A.off = group_off;
    A.n_groups = n_groups;
    A.n_tets = n_tets;
    A.x = x;
    A.y = y;
    A.z = z;
    A.nv = nv;
    A.flipped = flipped;
    A.max_q = max_q;
    A.q = q;
    hipLaunchKernelGGL(groups_kernel, dim3(grid_for(n_groups)), dim3(kBlock), 0, ctx->cur, A);
    TWG_LAUNCHED();
    const unsigned gx = std::min<unsigned>(grid_for(n_groups), 1024u);
    hipLaunchKernelGGL(groups_exact_kernel, dim3(gx), dim3(kBlock), 0, ctx->cur, A);
    TWG_LAUNCHED();
    return TWG_OK;
}
