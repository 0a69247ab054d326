// capi.hip -- context management and memory helpers of the C-ABI.
#include <cstdarg>
#include <cstring>

#include "twg_internal.hpp"

namespace {
thread_local char g_err[512] = "";
}

namespace twg {

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? TWG_ENOMEM : TWG_EDEVICE;
}

int bind(twg_ctx* ctx)
{
    TWG_HIP(hipSetDevice(ctx->device));
    return TWG_OK;
}

int ensure_scratch(twg_ctx* ctx, size_t bytes)
{
    if (bytes <= ctx->scratch_bytes) return TWG_OK;
    if (ctx->scratch) {
        TWG_HIP(hipStreamSynchronize(ctx->cur));
        TWG_HIP(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    TWG_HIP(hipMalloc(&ctx->scratch, want));
    ctx->scratch_bytes = want;
    return TWG_OK;
}

int ensure_slot(twg_ctx* ctx, int i, size_t bytes)
{
    if (bytes <= ctx->slot_bytes[i]) return TWG_OK;
    if (ctx->slot[i]) {
        TWG_HIP(hipStreamSynchronize(ctx->cur));
        TWG_HIP(hipFree(ctx->slot[i]));
        ctx->slot[i] = nullptr;
        ctx->slot_bytes[i] = 0;
    }
    size_t want = bytes + bytes / 4;
    TWG_HIP(hipMalloc(&ctx->slot[i], want));
    ctx->slot_bytes[i] = want;
    return TWG_OK;
}

}  // namespace twg

extern "C" {

const char* twg_last_error(void) { return g_err; }

int twg_ctx_create(int device, twg_ctx** out)
{
    TWG_REQUIRE(out != nullptr, "twg_ctx_create: out is NULL");
    int n = 0;
    TWG_HIP(hipGetDeviceCount(&n));
    TWG_REQUIRE(device >= 0 && device < n, "twg_ctx_create: device %d out of range (%d devices)",
                device, n);
    twg_ctx* c = new twg_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->order, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (c->own) (void)hipStreamDestroy(c->own);
        delete c;
        return twg::hip_fail(e, "twg_ctx_create", __FILE__, __LINE__);
    }
    c->cur = c->own;
    if (int r = twg::noalloc_scratch_init(c)) {
        (void)hipEventDestroy(c->order);
        (void)hipStreamDestroy(c->own);
        delete c;
        return r;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    *out = c;
    return TWG_OK;
}

int twg_ctx_destroy(twg_ctx* ctx)
{
    if (!ctx) return TWG_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->cur) (void)hipStreamSynchronize(ctx->cur);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    for (int i = 0; i < 4; ++i)
        if (ctx->slot[i]) (void)hipFree(ctx->slot[i]);
    if (ctx->list_stats) (void)hipFree(ctx->list_stats);
    if (ctx->order) (void)hipEventDestroy(ctx->order);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return TWG_OK;
}

int twg_ctx_set_stream(twg_ctx* ctx, void* s)
{
    TWG_REQUIRE(ctx != nullptr, "twg_ctx_set_stream: ctx is NULL");
    hipStream_t ns = (hipStream_t)s;
    if (ns == ctx->cur) return TWG_OK;
    if (int r = twg::bind(ctx)) return r;
    // Every call on the context shares its scratch: the new stream waits for
    // what the old one has enqueued.  A stream that is capturing a graph
    // takes no part in that (an event wait across a capture boundary is not
    // a legal capture); the capturing caller orders its own streams.
    hipStreamCaptureStatus a = hipStreamCaptureStatusNone, b = hipStreamCaptureStatusNone;
    TWG_HIP(hipStreamIsCapturing(ctx->cur, &a));
    TWG_HIP(hipStreamIsCapturing(ns, &b));
    if (a == hipStreamCaptureStatusNone && b == hipStreamCaptureStatusNone) {
        TWG_HIP(hipEventRecord(ctx->order, ctx->cur));
        TWG_HIP(hipStreamWaitEvent(ns, ctx->order, 0));
    }
    ctx->cur = ns;
    return TWG_OK;
}

void* twg_ctx_stream(twg_ctx* ctx) { return ctx ? (void*)ctx->cur : nullptr; }

int twg_ctx_sync(twg_ctx* ctx)
{
    TWG_REQUIRE(ctx != nullptr, "twg_ctx_sync: ctx is NULL");
    if (int r = twg::bind(ctx)) return r;
    TWG_HIP(hipStreamSynchronize(ctx->cur));
    return TWG_OK;
}

int twg_malloc(twg_ctx* ctx, size_t bytes, void** dptr)
{
    TWG_REQUIRE(ctx && dptr, "twg_malloc: NULL argument");
    if (int r = twg::bind(ctx)) return r;
    TWG_HIP(hipMalloc(dptr, bytes ? bytes : 1));
    return TWG_OK;
}

int twg_free(twg_ctx* ctx, void* dptr)
{
    TWG_REQUIRE(ctx != nullptr, "twg_free: ctx is NULL");
    if (!dptr) return TWG_OK;
    if (int r = twg::bind(ctx)) return r;
    TWG_HIP(hipStreamSynchronize(ctx->cur));
    TWG_HIP(hipFree(dptr));
    return TWG_OK;
}

int twg_memcpy_h2d(twg_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    TWG_REQUIRE(ctx && dst && src, "twg_memcpy_h2d: NULL argument");
    if (int r = twg::bind(ctx)) return r;
    TWG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->cur));
    TWG_HIP(hipStreamSynchronize(ctx->cur));
    return TWG_OK;
}

int twg_memcpy_d2h(twg_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    TWG_REQUIRE(ctx && dst && src, "twg_memcpy_d2h: NULL argument");
    if (int r = twg::bind(ctx)) return r;
    TWG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->cur));
    TWG_HIP(hipStreamSynchronize(ctx->cur));
    return TWG_OK;
}

}  // extern "C"
