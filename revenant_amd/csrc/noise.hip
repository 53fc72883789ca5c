// The per-receiver impulse blanker, on the receiver's coarse channel before
// its fine stage.
//
//   rvx_blank_detect   noise_blank.comp pass 0 / noise_reference.cpp:158-195
//   rvx_blank_apply    pass 1 / :197-226
//   rvx_design_blanker, rvx_blanker_threshold, rvx_blanker_check   :112-156
//
// detect: flag[j] = 1 when |x[j]|^2 * W > T * (sum of |x[m]|^2 over the W
// samples ending G before j, oldest first); a zero sum never flags. apply:
// blanked[i] = 0 when any flag in [i - hang, i + lead] is set, else x[i].
// A workgroup of 256 samples stages the powers its windows read (256 + W - 1
// samples, a guard below its own) in LDS once, each power the same two
// products and one sum the twin forms, and each lane sums its window from LDS
// in the twin's order; the flags (floats, as the twin's) go to the receiver's
// short ring.

#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "common.h"

namespace {

constexpr uint32_t kMinWindow = 16, kMaxWindow = 256;  // noise_reference.h
constexpr double kWindowSeconds = 0.5e-3;              // noise_reference.cpp:81
constexpr uint32_t kHang = 2;                          // :92
constexpr uint32_t kThreads = 256;

bool is_mask(uint32_t m) { return rvx::pow2(uint64_t(m) + 1); }

uint32_t reach_below(const rvx_blanker_config& c) { return c.hang + c.guard + c.window; }

__global__ void __launch_bounds__(kThreads)
detect_kernel(rvx_blanker_config c, rvx_blanker_params p, const float2* __restrict__ chan,
              float* __restrict__ flags) {
    // powers of the window samples i0 - guard - window .. i0 + 255 - guard - 1
    __shared__ float power[kThreads + kMaxWindow];
    const uint32_t i0 = blockIdx.x * kThreads;
    const uint32_t below = c.guard + c.window;
    const uint32_t span = kThreads + c.window;
    for (uint32_t k = threadIdx.x; k < span; k += kThreads) {
        const uint32_t at = (p.chan_first + i0 + k - below) & p.chan_mask;
        const float2 r = chan[p.chan_base + at];
        power[k] = r.x * r.x + r.y * r.y;
    }
    __syncthreads();
    const uint32_t i = i0 + threadIdx.x;
    if (i >= p.count) return;
    const float2 s = chan[p.chan_base + ((p.chan_first + i) & p.chan_mask)];
    const float self = s.x * s.x + s.y * s.y;
    float sum = 0.0f;
    for (uint32_t m = 0; m < c.window; ++m) sum = sum + power[threadIdx.x + m];  // oldest first
    const float lhs = self * float(c.window);
    const float rhs = p.threshold * sum;
    flags[(p.out_first + i) & p.out_mask] = (sum > 0.0f && lhs > rhs) ? 1.0f : 0.0f;
}

__global__ void apply_kernel(rvx_blanker_config c, rvx_blanker_params p,
                             const float2* __restrict__ chan, const float* __restrict__ flags,
                             float2* __restrict__ blanked) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.count) return;
    const uint32_t ci = (p.chan_first + i) & p.chan_mask;
    const uint32_t o = (p.out_first + i) & p.out_mask;
    bool hit = false;
    for (uint32_t m = 0; m <= c.hang + c.lead; ++m)
        hit = hit || flags[(o - c.hang + m) & p.out_mask] != 0.0f;
    blanked[o] = hit ? make_float2(0.0f, 0.0f) : chan[p.chan_base + ci];
}

}  // namespace

extern "C" {

int32_t rvx_design_blanker(int64_t channel_rate, rvx_blanker_config* out) {
    if (!out) return rvx::fail(RVX_EINVAL, "rvx_design_blanker: null output");
    rvx_blanker_config c{};
    const double window = std::round(kWindowSeconds * double(channel_rate));
    c.window = uint32_t(std::clamp(window, double(kMinWindow), double(kMaxWindow)));
    c.hang = kHang;
    c.lead = kHang;
    c.guard = c.lead + 2;
    *out = c;
    return RVX_OK;
}

float rvx_blanker_threshold(double threshold_db) {
    return float(std::pow(10.0, threshold_db / 10.0));
}

int32_t rvx_blanker_check(const rvx_blanker_config* config, const rvx_blanker_params* params) {
    static const char* who = "rvx_blanker_check";
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    const rvx_blanker_config& c = *config;
    const rvx_blanker_params& p = *params;
    if (c.pass != RVX_BLANK_DETECT && c.pass != RVX_BLANK_APPLY)
        return rvx::fail(RVX_EINVAL, "%s: pass %u is neither detect nor apply", who, c.pass);
    if (c.window < kMinWindow || c.window > kMaxWindow)
        return rvx::fail(RVX_EINVAL, "%s: a %u sample window is outside [%u, %u]", who, c.window,
                         kMinWindow, kMaxWindow);
    if (!is_mask(p.chan_mask) || !is_mask(p.out_mask))
        return rvx::fail(RVX_EINVAL, "%s: masks 0x%x and 0x%x must each be a power of two minus one",
                         who, p.chan_mask, p.out_mask);
    // everything one dispatch reads is live in each ring at once
    const uint64_t reach = uint64_t(reach_below(c)) + c.lead + p.count;
    const uint64_t smaller = uint64_t(std::min(p.chan_mask, p.out_mask)) + 1;
    if (reach >= smaller)
        return rvx::fail(RVX_EINVAL, "%s: %u samples plus a reach of %u below and %u above do not "
                         "fit a %llu sample ring", who, p.count, reach_below(c), c.lead,
                         (unsigned long long)smaller);
    if (!(p.threshold > 0.0f))
        return rvx::fail(RVX_EINVAL, "%s: the threshold must be a positive power ratio", who);
    return RVX_OK;
}

int32_t rvx_blank_detect(rvx_ctx* ctx, const rvx_blanker_config* config,
                         const rvx_blanker_params* params, const void* channel_ring,
                         uint64_t ring_samples, float* flags, uint64_t n_flags) {
    static const char* who = "rvx_blank_detect";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    rvx_blanker_config c = *config;
    c.pass = RVX_BLANK_DETECT;
    if (int32_t rc = rvx_blanker_check(&c, params)) return rc;
    const rvx_blanker_params p = *params;
    if (n_flags < uint64_t(p.out_mask) + 1)
        return rvx::fail(RVX_EINVAL, "%s: %llu flags for a %u sample ring", who,
                         (unsigned long long)n_flags, p.out_mask + 1);
    if (ring_samples < uint64_t(p.chan_base) + p.chan_mask + 1)
        return rvx::fail(RVX_EINVAL, "%s: the channel ring is shorter than its base and mask", who);
    if (p.count == 0) return RVX_OK;
    if (!channel_ring || !flags) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    detect_kernel<<<dim3((p.count + kThreads - 1) / kThreads), dim3(kThreads), 0, ctx->stream>>>(
        c, p, static_cast<const float2*>(channel_ring), flags);
    return rvx::check_launch(who);
}

int32_t rvx_blank_apply(rvx_ctx* ctx, const rvx_blanker_config* config,
                        const rvx_blanker_params* params, const void* channel_ring,
                        uint64_t ring_samples, const float* flags, uint64_t n_flags,
                        void* blanked, uint64_t blanked_samples) {
    static const char* who = "rvx_blank_apply";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    rvx_blanker_config c = *config;
    c.pass = RVX_BLANK_APPLY;
    if (int32_t rc = rvx_blanker_check(&c, params)) return rc;
    const rvx_blanker_params p = *params;
    const uint64_t out_ring = uint64_t(p.out_mask) + 1;
    if (n_flags < out_ring || blanked_samples < out_ring)
        return rvx::fail(RVX_EINVAL, "%s: the flag and blanked rings must each hold out_mask + 1 "
                         "values", who);
    if (ring_samples < uint64_t(p.chan_base) + p.chan_mask + 1)
        return rvx::fail(RVX_EINVAL, "%s: the channel ring is shorter than its base and mask", who);
    if (p.count == 0) return RVX_OK;
    if (!channel_ring || !flags || !blanked) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    apply_kernel<<<dim3((p.count + 255) / 256), dim3(256), 0, ctx->stream>>>(
        c, p, static_cast<const float2*>(channel_ring), flags, static_cast<float2*>(blanked));
    return rvx::check_launch(who);
}

}  // extern "C"
