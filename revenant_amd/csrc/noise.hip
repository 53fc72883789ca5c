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
// detect: a workgroup of 1024 samples stages the powers its windows read
// (1024 + W - 1 samples, a guard below its own) in LDS once, each power the
// same two products and one sum the twin forms; a lane takes four samples and
// sums their windows in the twin's order; the flags (floats, as the twin's)
// go to the receiver's short ring. apply: the flags a workgroup's samples
// consult are staged in LDS once.

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

// Four consecutive samples per lane, 1024 per workgroup. The window powers
// are staged once; each lane runs the four samples' window sums side by side,
// reading the staged powers 16 bytes at a time and adding each power to every
// sum whose window holds it, so each sum still adds its window oldest first
// while the LDS traffic is a quarter of one read per term.
constexpr uint32_t kPer = 4;
constexpr uint32_t kDetectSpan = kThreads * kPer;

__global__ void __launch_bounds__(kThreads)
detect_kernel(rvx_blanker_config c, rvx_blanker_params p, const float2* __restrict__ chan,
              float* __restrict__ flags) {
    // powers of samples i0 - guard - window + q, q < 1024 + window - 1
    __shared__ __align__(16) float power[kDetectSpan + kMaxWindow];
    const uint32_t i0 = blockIdx.x * kDetectSpan;
    const uint32_t below = c.guard + c.window;
    const uint32_t span = kDetectSpan + c.window - 1;
    for (uint32_t k = threadIdx.x; k < span; k += kThreads) {
        const uint32_t at = (p.chan_first + i0 + k - below) & p.chan_mask;
        const float2 r = chan[p.chan_base + at];
        power[k] = r.x * r.x + r.y * r.y;
    }
    __syncthreads();
    const uint32_t i = i0 + kPer * threadIdx.x;
    if (i >= p.count) return;
    const float* w = power + kPer * threadIdx.x;  // 16-byte aligned
    const uint32_t W = c.window;                  // >= 16
    // q = 0..3: the sums start one after another
    float4 v = *reinterpret_cast<const float4*>(w);
    float s0 = 0.0f + v.x;
    s0 = s0 + v.y;
    float s1 = 0.0f + v.y;
    s0 = s0 + v.z;
    s1 = s1 + v.z;
    float s2 = 0.0f + v.z;
    s0 = s0 + v.w;
    s1 = s1 + v.w;
    s2 = s2 + v.w;
    float s3 = 0.0f + v.w;
    // q = 4.. while every sum's window holds the whole quad
    uint32_t q = 4;
    for (; q + 3 < W; q += 4) {
        v = *reinterpret_cast<const float4*>(w + q);
        s0 = s0 + v.x; s1 = s1 + v.x; s2 = s2 + v.x; s3 = s3 + v.x;
        s0 = s0 + v.y; s1 = s1 + v.y; s2 = s2 + v.y; s3 = s3 + v.y;
        s0 = s0 + v.z; s1 = s1 + v.z; s2 = s2 + v.z; s3 = s3 + v.z;
        s0 = s0 + v.w; s1 = s1 + v.w; s2 = s2 + v.w; s3 = s3 + v.w;
    }
    // the last window ends at q = W + 2: sum k takes q < W + k
    for (; q < W + 3; ++q) {
        const float pv = w[q];
        if (q < W) s0 = s0 + pv;
        if (q < W + 1) s1 = s1 + pv;
        if (q < W + 2) s2 = s2 + pv;
        s3 = s3 + pv;
    }
    const float sums[kPer] = {s0, s1, s2, s3};
    const float window = float(W);
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        if (i + k >= p.count) break;
        const float2 x = chan[p.chan_base + ((p.chan_first + i + k) & p.chan_mask)];
        const float self = x.x * x.x + x.y * x.y;
        const float lhs = self * window;
        const float rhs = p.threshold * sums[k];
        flags[(p.out_first + i + k) & p.out_mask] = (sums[k] > 0.0f && lhs > rhs) ? 1.0f : 0.0f;
    }
}

// The workgroup's flags, hang below to lead above its 256 samples, are staged
// in LDS once (up to kMaxStagedReach of hang + lead); each lane ORs its own
// run of them.
constexpr uint32_t kMaxStagedReach = 256;

template <bool kStaged>
__global__ void __launch_bounds__(kThreads)
apply_kernel(rvx_blanker_config c, rvx_blanker_params p, const float2* __restrict__ chan,
             const float* __restrict__ flags, float2* __restrict__ blanked) {
    __shared__ float staged[kStaged ? kThreads + kMaxStagedReach : 1];
    const uint32_t i0 = blockIdx.x * kThreads;
    const uint32_t i = i0 + threadIdx.x;
    const uint32_t reach = c.hang + c.lead;
    if (kStaged) {
        const uint32_t span = kThreads + reach;
        for (uint32_t k = threadIdx.x; k < span; k += kThreads)
            staged[k] = flags[(p.out_first + i0 + k - c.hang) & p.out_mask];
        __syncthreads();
    }
    if (i >= p.count) return;
    const uint32_t ci = (p.chan_first + i) & p.chan_mask;
    const uint32_t o = (p.out_first + i) & p.out_mask;
    bool hit = false;
    for (uint32_t m = 0; m <= reach; ++m) {
        const float f = kStaged ? staged[threadIdx.x + m] : flags[(o - c.hang + m) & p.out_mask];
        hit = hit || f != 0.0f;
    }
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
    detect_kernel<<<dim3((p.count + kDetectSpan - 1) / kDetectSpan), dim3(kThreads), 0,
                    ctx->stream>>>(
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
    const dim3 grid((p.count + kThreads - 1) / kThreads), block(kThreads);
    if (c.hang + c.lead <= kMaxStagedReach)
        apply_kernel<true><<<grid, block, 0, ctx->stream>>>(
            c, p, static_cast<const float2*>(channel_ring), flags, static_cast<float2*>(blanked));
    else
        apply_kernel<false><<<grid, block, 0, ctx->stream>>>(
            c, p, static_cast<const float2*>(channel_ring), flags, static_cast<float2*>(blanked));
    return rvx::check_launch(who);
}

}  // extern "C"
