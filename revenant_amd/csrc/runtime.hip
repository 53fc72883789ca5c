// Context, stream, buffers and errors: the part of core/gpu (context.cpp,
// buffer.cpp, kernel.cpp CommandRunner) the receive chain needs, on HIP.
//
// One context per device and process, one stream per context at the default
// priority. Buffers are plain device pointers; the three MemoryKinds of
// core/gpu/buffer.h map to hipMalloc and pinned host memory (hipHostMalloc).

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "common.h"

namespace {
thread_local std::string g_last_error;
}

namespace rvx {

void set_error(const std::string& message) { g_last_error = message; }

int32_t fail(int32_t code, const char* fmt, ...) {
    char buffer[1024];
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(buffer, sizeof buffer, fmt, args);
    va_end(args);
    g_last_error = buffer;
    return code;
}

int32_t enter(rvx_ctx* ctx, const char* who) {
    if (ctx == nullptr) return fail(RVX_EINVAL, "%s: null context", who);
    int current = -1;
    RVX_TRY_HIP(hipGetDevice(&current), who);
    if (current != ctx->device) RVX_TRY_HIP(hipSetDevice(ctx->device), who);
    return RVX_OK;
}

int32_t check_launch(const char* who) {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        return fail(RVX_EDEVICE, "%s: kernel launch failed: %s", who, hipGetErrorString(err));
    return RVX_OK;
}

}  // namespace rvx

namespace {

__global__ void zero_fill(uint32_t* __restrict__ dst, uint64_t words) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride)
        dst[i] = 0u;
}
__global__ void zero_fill_bytes(uint8_t* __restrict__ dst, uint64_t bytes) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < bytes) dst[i] = 0u;
}

}  // namespace

extern "C" {

int32_t rvx_abi_version(void) { return RVX_ABI_VERSION; }

const char* rvx_last_error(void) { return g_last_error.c_str(); }

int32_t rvx_ctx_create(int32_t device, rvx_ctx** out) {
    static const char* who = "rvx_ctx_create";
    if (out == nullptr) return rvx::fail(RVX_EINVAL, "%s: null output", who);
    *out = nullptr;
    int count = 0;
    RVX_TRY_HIP(hipGetDeviceCount(&count), who);
    if (device < 0 || device >= count)
        return rvx::fail(RVX_EINVAL, "%s: device %d out of range (%d devices)", who, device, count);
    RVX_TRY_HIP(hipSetDevice(device), who);
    auto* ctx = new rvx_ctx;
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->t0) != hipSuccess || hipEventCreate(&ctx->t1) != hipSuccess) {
        rvx_ctx_destroy(ctx);
        return rvx::fail(RVX_EDEVICE, "%s: could not create the stream or its events", who);
    }
    ctx->stream = ctx->own;
    if (int32_t rc = rvx::batch_scratch_setup(device, who)) {
        rvx_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return RVX_OK;
}

void rvx_ctx_destroy(rvx_ctx* ctx) {
    if (ctx == nullptr) return;
    if (ctx->own) {
        (void)hipStreamSynchronize(ctx->own);
        (void)hipStreamDestroy(ctx->own);
    }
    if (ctx->t0) (void)hipEventDestroy(ctx->t0);
    if (ctx->t1) (void)hipEventDestroy(ctx->t1);
    for (void* p : ctx->dft)
        if (p) (void)hipFree(p);
    delete ctx;
}

int32_t rvx_ctx_set_stream(rvx_ctx* ctx, void* hip_stream) {
    if (ctx == nullptr) return rvx::fail(RVX_EINVAL, "rvx_ctx_set_stream: null context");
    ctx->stream = static_cast<hipStream_t>(hip_stream);
    return RVX_OK;
}

int32_t rvx_ctx_use_own_stream(rvx_ctx* ctx) {
    if (ctx == nullptr) return rvx::fail(RVX_EINVAL, "rvx_ctx_use_own_stream: null context");
    ctx->stream = ctx->own;
    return RVX_OK;
}

void* rvx_ctx_stream(rvx_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int32_t rvx_stream_sync(rvx_ctx* ctx) {
    RVX_ENTER(ctx, "rvx_stream_sync");
    RVX_TRY_HIP(hipStreamSynchronize(ctx->stream), "rvx_stream_sync");
    return RVX_OK;
}

int32_t rvx_buf_alloc(rvx_ctx* ctx, uint64_t bytes, int32_t kind, void** out) {
    static const char* who = "rvx_buf_alloc";
    RVX_ENTER(ctx, who);
    if (out == nullptr) return rvx::fail(RVX_EINVAL, "%s: null output", who);
    *out = nullptr;
    if (bytes == 0) return rvx::fail(RVX_EINVAL, "%s: zero-byte buffer", who);
    hipError_t err = hipSuccess;
    if (kind == RVX_MEM_DEVICE)
        err = hipMalloc(out, bytes);
    else if (kind == RVX_MEM_UPLOAD || kind == RVX_MEM_READBACK)
        err = hipHostMalloc(out, bytes, hipHostMallocDefault);
    else
        return rvx::fail(RVX_EINVAL, "%s: unknown memory kind %d", who, kind);
    if (err != hipSuccess)
        return rvx::fail(RVX_ENOMEM, "%s: %llu bytes of kind %d: %s", who,
                         (unsigned long long)bytes, kind, hipGetErrorString(err));
    return RVX_OK;
}

int32_t rvx_buf_free(rvx_ctx* ctx, void* ptr) {
    static const char* who = "rvx_buf_free";
    RVX_ENTER(ctx, who);
    if (ptr == nullptr) return RVX_OK;
    hipPointerAttribute_t attr{};
    RVX_TRY_HIP(hipPointerGetAttributes(&attr, ptr), who);
    if (attr.type == hipMemoryTypeHost)
        RVX_TRY_HIP(hipHostFree(ptr), who);
    else
        RVX_TRY_HIP(hipFree(ptr), who);
    return RVX_OK;
}

int32_t rvx_buf_write(rvx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    static const char* who = "rvx_buf_write";
    RVX_ENTER(ctx, who);
    if (bytes == 0) return RVX_OK;
    if (dst == nullptr || src == nullptr) return rvx::fail(RVX_EINVAL, "%s: null pointer", who);
    RVX_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream), who);
    RVX_TRY_HIP(hipStreamSynchronize(ctx->stream), who);
    return RVX_OK;
}

int32_t rvx_buf_read(rvx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    static const char* who = "rvx_buf_read";
    RVX_ENTER(ctx, who);
    if (bytes == 0) return RVX_OK;
    if (dst == nullptr || src == nullptr) return rvx::fail(RVX_EINVAL, "%s: null pointer", who);
    RVX_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream), who);
    RVX_TRY_HIP(hipStreamSynchronize(ctx->stream), who);
    return RVX_OK;
}

int32_t rvx_buf_clear(rvx_ctx* ctx, void* dst, uint64_t bytes) {
    static const char* who = "rvx_buf_clear";
    RVX_ENTER(ctx, who);
    if (bytes == 0) return RVX_OK;
    if (dst == nullptr) return rvx::fail(RVX_EINVAL, "%s: null pointer", who);
    if ((reinterpret_cast<uintptr_t>(dst) & 3u) == 0 && (bytes & 3u) == 0) {
        const uint64_t words = bytes / 4;
        const uint64_t blocks = std::min<uint64_t>((words + 255) / 256, 65536);
        zero_fill<<<dim3(uint32_t(blocks)), dim3(256), 0, ctx->stream>>>(
            static_cast<uint32_t*>(dst), words);
    } else {
        if (bytes > (uint64_t(1) << 40)) return rvx::fail(RVX_EINVAL, "%s: too large", who);
        zero_fill_bytes<<<dim3(uint32_t((bytes + 255) / 256)), dim3(256), 0, ctx->stream>>>(
            static_cast<uint8_t*>(dst), bytes);
    }
    return rvx::check_launch(who);
}

int32_t rvx_timer_begin(rvx_ctx* ctx) {
    RVX_ENTER(ctx, "rvx_timer_begin");
    RVX_TRY_HIP(hipEventRecord(ctx->t0, ctx->stream), "rvx_timer_begin");
    return RVX_OK;
}

int32_t rvx_timer_end(rvx_ctx* ctx, float* elapsed_ms) {
    static const char* who = "rvx_timer_end";
    RVX_ENTER(ctx, who);
    RVX_TRY_HIP(hipEventRecord(ctx->t1, ctx->stream), who);
    RVX_TRY_HIP(hipEventSynchronize(ctx->t1), who);
    float ms = 0.0f;
    RVX_TRY_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1), who);
    if (elapsed_ms) *elapsed_ms = ms;
    return RVX_OK;
}

}  // extern "C"

// HIP graphs over the context's own stream: a step's launches are captured
// once and replayed with one call. Capture needs a stream that is not the
// null stream, so begin switches nothing: it fails unless the context runs on
// its own stream or a caller's stream.
struct rvx_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int32_t rvx_graph_begin(rvx_ctx* ctx) {
    static const char* who = "rvx_graph_begin";
    RVX_ENTER(ctx, who);
    if (ctx->stream == nullptr)
        return rvx::fail(RVX_EINVAL, "%s: the null stream cannot be captured", who);
    RVX_TRY_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), who);
    return RVX_OK;
}

int32_t rvx_graph_end(rvx_ctx* ctx, rvx_graph** out) {
    static const char* who = "rvx_graph_end";
    RVX_ENTER(ctx, who);
    if (out == nullptr) return rvx::fail(RVX_EINVAL, "%s: null output", who);
    *out = nullptr;
    hipGraph_t g = nullptr;
    RVX_TRY_HIP(hipStreamEndCapture(ctx->stream, &g), who);
    hipGraphExec_t exec = nullptr;
    const hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        (void)hipGraphDestroy(g);
        return rvx::fail(RVX_EDEVICE, "%s: hipGraphInstantiate failed: %s", who,
                         hipGetErrorString(err));
    }
    *out = new rvx_graph{g, exec};
    return RVX_OK;
}

int32_t rvx_graph_launch(rvx_ctx* ctx, rvx_graph* graph) {
    static const char* who = "rvx_graph_launch";
    RVX_ENTER(ctx, who);
    if (graph == nullptr) return rvx::fail(RVX_EINVAL, "%s: null graph", who);
    RVX_TRY_HIP(hipGraphLaunch(graph->exec, ctx->stream), who);
    return RVX_OK;
}

void rvx_graph_destroy(rvx_graph* graph) {
    if (graph == nullptr) return;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
}

}  // extern "C"
