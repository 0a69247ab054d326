// twg_internal.hpp -- context, error reporting and launch helpers shared by
// the translation units of libtetwild_amd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "../../include/tetwild_gpu.h"

struct twg_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t cur = nullptr;
    // recorded on the old stream when the caller switches streams, so work on
    // the new one starts after the old one's use of the shared scratch
    hipEvent_t order = nullptr;
    int num_cus = 256;
    // growable device scratch owned by the context
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // further growable slots: 0 the queue of tets the orientation filter left
    // undecided, 1 query ordering keys and values, 2 sort temporary storage
    void* slot[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t slot_bytes[4] = {0, 0, 0, 0};
    // device address of smooth.hip's file-scope scratch array on this
    // context's device (twg_smooth_batch_noalloc), looked up at creation
    void* noalloc_scratch = nullptr;
    // debug counters of the envelope list kernels (twg_ctx_list_stats):
    // waves that tested their LDS facet list, waves that walked the tree
    unsigned long long* list_stats = nullptr;
    bool list_stats_on = false;
};

namespace twg {

void set_error(const char* fmt, ...);

// returns nonzero (TWG_EDEVICE) and records the message on failure
int hip_fail(hipError_t e, const char* what, const char* file, int line);

int ensure_scratch(twg_ctx* ctx, size_t bytes);
int ensure_slot(twg_ctx* ctx, int i, size_t bytes);

// Morton order of n query items (points: stride 3, triangles: stride 9,
// keyed by their first three coordinates' centroid) inside the box [lo, hi]:
// writes a permutation (int32, device) into *order (a context slot).  The
// order only affects speed: every query's result is independent of it.
int morton_order(twg_ctx* ctx, const double* items, int64_t n, int stride, const double* lo,
                 const double* hi, const int32_t** order);

// Activates ctx->device for the calling thread.
int bind(twg_ctx* ctx);

// With ctx->device current: the device address of the no-allocation
// smoothing entry point's scratch array, into ctx->noalloc_scratch.
int noalloc_scratch_init(twg_ctx* ctx);

constexpr int kBlock = 256;

// Autotuner results: (device, precision flag, size class) -> (tile, chunk),
// shared by every context of the process.  Contexts may live on different
// host threads, so the map is guarded.
class TuneCache {
public:
    bool get(int device, bool flag, int size_class, int& tile, unsigned& chunk)
    {
        std::lock_guard<std::mutex> lock(m_);
        auto it = map_.find(std::make_tuple(device, flag, size_class));
        if (it == map_.end()) return false;
        tile = it->second.first;
        chunk = it->second.second;
        return true;
    }
    void put(int device, bool flag, int size_class, int tile, unsigned chunk)
    {
        std::lock_guard<std::mutex> lock(m_);
        map_[std::make_tuple(device, flag, size_class)] = {tile, chunk};
    }

private:
    std::mutex m_;
    std::map<std::tuple<int, bool, int>, std::pair<int, unsigned>> map_;
};

// The two timing events of an autotuner, destroyed on every path out of the
// scope that owns them (early error returns included).
struct EventPair {
    hipEvent_t e0 = nullptr;
    hipEvent_t e1 = nullptr;
    hipError_t create()
    {
        hipError_t e = hipEventCreate(&e0);
        return e == hipSuccess ? hipEventCreate(&e1) : e;
    }
    ~EventPair()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

inline unsigned grid_for(int64_t n, int block = kBlock)
{
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace twg

#define TWG_HIP(expr)                                                    \
    do {                                                                 \
        hipError_t _e = (expr);                                          \
        if (_e != hipSuccess) return twg::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define TWG_REQUIRE(cond, ...)                       \
    do {                                             \
        if (!(cond)) {                               \
            twg::set_error(__VA_ARGS__);             \
            return TWG_EINVAL;                       \
        }                                            \
    } while (0)

// after a kernel launch
#define TWG_LAUNCHED() TWG_HIP(hipGetLastError())
