// Internal plumbing shared by the library's translation units: the context,
// the thread-local error, and the HIP-call check.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>

#include "../../include/rvx.h"

struct rvx_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr;
    hipEvent_t t1 = nullptr;
    void* dft[2] = {nullptr, nullptr};  // matrix-core DFT operand tables, made on first use
};

namespace rvx {

void set_error(const std::string& message);
int32_t fail(int32_t code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

inline bool pow2(uint64_t v) { return v != 0 && (v & (v - 1)) == 0; }
inline uint32_t log2u(uint64_t v) {
    uint32_t s = 0;
    while ((uint64_t(1) << s) < v) ++s;
    return s;
}

// Makes the context's device current and clears any stale launch error.
int32_t enter(rvx_ctx* ctx, const char* who);
// Reports a launch failure as RVX_EDEVICE.
int32_t check_launch(const char* who);
// Resolves `device`'s address of the batch launch's static scratch (the
// device must be current); idempotent. Called by rvx_ctx_create.
int32_t batch_scratch_setup(int device, const char* who);
// Samples [first_off, first_off + count) of a code window converted (and
// corrected with `correction` unless null) into a cf32 ring at slot
// (first_abs + i) & ring_mask, as the fused chain converts and corrects them.
int32_t convert_to_ring(rvx_ctx* ctx, int32_t format, const void* codes, uint64_t code_samples,
                        uint64_t first_off, uint64_t first_abs, uint32_t count,
                        const rvx_iq_correct_params* correction, void* ring, uint32_t ring_mask,
                        const char* who);
// Ring slots of samples [first_abs, first_abs + count) into a linear cf32
// window from window_off.
int32_t ring_to_window(rvx_ctx* ctx, const void* ring, uint32_t ring_mask, uint64_t first_abs,
                       uint32_t count, void* window, uint64_t window_samples, uint64_t window_off,
                       const char* who);

}  // namespace rvx

#define RVX_TRY_HIP(call, who)                                                      \
    do {                                                                            \
        hipError_t rvx_err_ = (call);                                               \
        if (rvx_err_ != hipSuccess)                                                 \
            return rvx::fail(RVX_EDEVICE, "%s: %s failed: %s", who, #call,          \
                             hipGetErrorString(rvx_err_));                          \
    } while (0)

#define RVX_ENTER(ctx, who)                                    \
    do {                                                       \
        int32_t rvx_rc_ = rvx::enter((ctx), (who));            \
        if (rvx_rc_ != RVX_OK) return rvx_rc_;                 \
    } while (0)
