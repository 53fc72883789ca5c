// Span-wide impulse excision on the sample ring, before the channelizer.
//
//   rvx_excise_segment     impulse_excise.comp pass 0 / impulse_excision_reference.cpp:250-272
//   rvx_excise_reference   pass 1 / :274-315
//   rvx_excise_flag        pass 2 / :317-340
//   rvx_excise_apply       pass 3 / :342-387
//   rvx_design_excision, rvx_excision_threshold, rvx_excision_check,
//   rvx_size_excision_rings, rvx_excision_cursor_*   the host side, :62-247
//
// The detector (core/dsp/impulse_excision.h:20-55): p[s] = sum of |x|^2 over
// segment s of S samples (aligned to the absolute index); r[s] = the median
// of p over the K segments ending one before s (a zero until a whole window
// of this run exists); a sample is flagged when |x|^2 * S > T * r; a flag
// marks [f - lead, f + hang]; a maximal marked run of at most max_width
// samples is excised (replaced by -0, its value moved to the removed ring,
// so cleaned + removed == raw bit for bit), a longer one is a transmission
// and is left alone. Each pass is one lane per segment or per sample; sums
// run in sample order from zero and the median is the twin's insertion sort,
// so every pass gives the twin's bits; the apply pass's counters are integer
// atomics, whose sums do not depend on order.

#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <new>

#include "common.h"

namespace {

constexpr uint32_t kPassSegment = 0, kPassReference = 1, kPassFlag = 2, kPassApply = 3;
constexpr uint32_t kMaxSegments = 64;     // impulse_excision.h:119
constexpr uint32_t kMaxWidth = 4096;      // :120
constexpr uint32_t kMaxSegLog2 = 12;      // :121
constexpr uint32_t kMaxGuard = 16;
constexpr uint32_t kMinGuard = 2;
constexpr uint32_t kCounters = 4;

bool is_mask(uint32_t mask) { return rvx::pow2(uint64_t(mask) + 1); }
uint64_t ring_of(uint32_t mask) { return uint64_t(mask) + 1; }

// Whether the flag `rel` samples past `first` counts (not from before the
// stage's restart, where the flag ring holds an older run).
__device__ __forceinline__ bool flag_at(const rvx_excision_params& p,
                                        const uint32_t* __restrict__ flags, int64_t rel) {
    if (rel + p.history < 0) return false;
    return flags[(p.first + uint32_t(rel)) & p.flag_mask] != 0u;
}

// A flag in [rel - hang, rel + lead] marks the sample at rel.
__device__ __forceinline__ bool marked(const rvx_excision_config& c, const rvx_excision_params& p,
                                       const uint32_t* __restrict__ flags, int64_t rel) {
    for (uint32_t m = 0; m <= c.hang + c.lead; ++m)
        if (flag_at(p, flags, rel - int64_t(c.hang) + m)) return true;
    return false;
}

__global__ void segment_kernel(rvx_excision_config c, rvx_excision_params p,
                               const float2* __restrict__ raw, float* __restrict__ seg_power) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p.count) return;
    const uint32_t base = (p.first + (j << c.seg_log2)) & p.ring_mask;
    const uint32_t samples = 1u << c.seg_log2;
    float sum = 0.0f;
    for (uint32_t m = 0; m < samples; ++m) {
        const float2 s = raw[(base + m) & p.ring_mask];
        const float q = s.x * s.x + s.y * s.y;
        sum = sum + q;
    }
    seg_power[(base >> c.seg_log2) & p.seg_mask] = sum;
}

__global__ void reference_kernel(rvx_excision_config c, rvx_excision_params p,
                                 const float* __restrict__ seg_power,
                                 float* __restrict__ seg_level) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p.count) return;
    const uint32_t base = (p.first + (j << c.seg_log2)) & p.ring_mask;
    const uint32_t slot = base >> c.seg_log2;
    const int64_t segment = int64_t(1) << c.seg_log2;
    const int64_t rel = int64_t(j) * segment;
    if (rel - (int64_t(c.segments) + 1) * segment + p.history < 0) {
        seg_level[slot & p.seg_mask] = 0.0f;  // the window is not all from this run
        return;
    }
    // Oldest first, then insertion sort: every comparison exact, so the
    // order and the median are the twin's bits.
    float v[kMaxSegments];
    for (uint32_t m = 0; m < c.segments; ++m)
        v[m] = seg_power[(slot - 1u - c.segments + m) & p.seg_mask];
    for (uint32_t a = 1; a < c.segments; ++a) {
        const float x = v[a];
        uint32_t b = a;
        while (b > 0 && v[b - 1] > x) {
            v[b] = v[b - 1];
            --b;
        }
        v[b] = x;
    }
    seg_level[slot & p.seg_mask] = v[c.segments / 2];
}

__global__ void flag_kernel(rvx_excision_config c, rvx_excision_params p,
                            const float2* __restrict__ raw, const float* __restrict__ seg_level,
                            uint32_t* __restrict__ flags) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p.count) return;
    const uint32_t slot = (p.first + j) & p.ring_mask;
    const float2 s = raw[slot];
    const float power = s.x * s.x + s.y * s.y;
    const float level = seg_level[(slot >> c.seg_log2) & p.seg_mask];
    const float lhs = power * float(1u << c.seg_log2);
    const float rhs = p.threshold * level;
    flags[(p.first + j) & p.flag_mask] = (level > 0.0f && lhs > rhs) ? 1u : 0u;
}

constexpr uint32_t kThreads = 256;

// raw and cleaned may be the same ring: a lane reads and writes only its slot.
// The workgroup first stages the flags its samples' marks read ([j0 - hang,
// j0 + 256 + lead), each already tested against the restart as flag_at
// tests it) in LDS, so the common case, an unmarked sample, reads LDS only;
// the run scans of a marked sample read the flag ring.
__global__ void __launch_bounds__(256)
apply_kernel(rvx_excision_config c, rvx_excision_params p, const float2* raw,
             const uint32_t* __restrict__ flags, float2* cleaned,
             float2* __restrict__ removed, uint32_t* __restrict__ counters) {
    __shared__ uint32_t tile[256 + 2 * kMaxGuard + 1];
    const uint32_t j0 = blockIdx.x * 256u;
    const uint32_t span = 256u + c.hang + c.lead;
    for (uint32_t k = threadIdx.x; k < span; k += 256u)
        tile[k] = flag_at(p, flags, int64_t(j0) - int64_t(c.hang) + int64_t(k)) ? 1u : 0u;
    __syncthreads();
    const uint32_t j = j0 + threadIdx.x;
    if (j >= p.count) return;
    const int64_t rel = j;
    const uint32_t slot = (p.first + j) & p.ring_mask;
    const float2 x = raw[slot];
    // any flag in [j - hang, j + lead]: a bitwise OR over the window, so the
    // reads go out together rather than one per wait as a short-circuit would
    uint32_t any = 0u;
#pragma unroll 8
    for (uint32_t m = 0; m <= c.hang + c.lead; ++m) any |= tile[threadIdx.x + m];
    const bool here = any != 0u;
    bool excise = false;
    if (here) {
        uint32_t left = 0, right = 0;
        for (uint32_t k = 1; k <= c.max_width; ++k) {
            if (!marked(c, p, flags, rel - int64_t(k))) break;
            ++left;
        }
        for (uint32_t k = 1; k <= c.max_width; ++k) {
            if (!marked(c, p, flags, rel + int64_t(k))) break;
            ++right;
        }
        excise = left + right + 1u <= c.max_width;
        if (left == 0u) atomicAdd(&counters[excise ? 1 : 2], 1u);  // events / spared
    }
    if (excise) atomicAdd(&counters[0], 1u);
    const float2 nothing = make_float2(-0.0f, -0.0f);  // adds to anything unchanged
    cleaned[slot] = excise ? nothing : x;
    removed[(p.first + j) & p.removed_mask] = excise ? x : nothing;
}

dim3 grid_for(uint32_t n) { return dim3((n + kThreads - 1) / kThreads); }

}  // namespace

struct rvx_excision_cursor {
    rvx_excision_config config{};
    uint32_t ring_mask = 0;
    rvx_excision_rings rings{};
    float threshold = 0.0f;
    bool running = false;
    uint64_t valid_from = 0, next_in = 0, seg_next = 0, flag_next = 0, apply_next = 0;

    rvx_excision_params at(uint64_t first, uint32_t count) const {
        rvx_excision_params out{};
        out.ring_mask = ring_mask;
        out.first = uint32_t(first & ring_mask);
        out.count = count;
        out.seg_mask = rings.seg_mask;
        out.flag_mask = rings.flag_mask;
        out.removed_mask = rings.removed_mask;
        // clamped well inside int32: the kernels compare it only against
        // offsets a dispatch reaches
        const int64_t clamp = int64_t(1) << 30;
        const int64_t history = int64_t(first) - int64_t(valid_from);
        out.history = int32_t(std::clamp(history, -clamp, clamp));
        out.threshold = threshold;
        return out;
    }
};

extern "C" {

int32_t rvx_design_excision(int64_t source_rate, double max_width_seconds,
                            rvx_excision_config* out) {
    if (!out) return rvx::fail(RVX_EINVAL, "rvx_design_excision: null output");
    rvx_excision_config c{};
    const double rate = double(std::max<int64_t>(source_rate, 1));
    c.pass = kPassSegment;
    c.max_width = uint32_t(std::clamp(std::round(max_width_seconds * rate), 1.0, double(kMaxWidth)));
    // 0.1 us either side, two to sixteen samples
    c.hang = uint32_t(std::clamp(std::round(0.1e-6 * rate), double(kMinGuard), double(kMaxGuard)));
    c.lead = c.hang;
    // K * S / 2 (the run a burst must cover before the median moves) is at
    // least four widest impulses: S >= W / 4 with K = 32; S >= 16.
    const uint32_t quarter = std::max(16u, (c.max_width + 3u) / 4u);
    uint32_t segment = 1;
    while (segment < quarter) segment <<= 1;
    uint32_t lg = 0;
    while ((1u << lg) < segment) ++lg;
    c.seg_log2 = std::min(lg, kMaxSegLog2);
    c.segments = 32;
    *out = c;
    return RVX_OK;
}

float rvx_excision_threshold(double threshold_db) {
    return float(std::pow(10.0, threshold_db / 10.0));
}

int32_t rvx_excision_check(const rvx_excision_config* config, const rvx_excision_params* params) {
    static const char* who = "rvx_excision_check";
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    const rvx_excision_config& c = *config;
    const rvx_excision_params& p = *params;
    if (c.pass > kPassApply) return rvx::fail(RVX_EINVAL, "%s: pass %u", who, c.pass);
    if (c.seg_log2 < 2 || c.seg_log2 > kMaxSegLog2)
        return rvx::fail(RVX_EINVAL, "%s: a segment of 2^%u samples is outside [4, %u]", who,
                         c.seg_log2, 1u << kMaxSegLog2);
    if (c.segments < 2 || c.segments > kMaxSegments)
        return rvx::fail(RVX_EINVAL, "%s: %u reference segments is outside [2, %u]", who,
                         c.segments, kMaxSegments);
    if (c.max_width < 1 || c.max_width > kMaxWidth)
        return rvx::fail(RVX_EINVAL, "%s: a %u sample maximum width is outside [1, %u]", who,
                         c.max_width, kMaxWidth);
    if (c.hang > kMaxGuard || c.lead > kMaxGuard)
        return rvx::fail(RVX_EINVAL, "%s: hang %u and lead %u must each be at most %u", who,
                         c.hang, c.lead, kMaxGuard);
    if (!is_mask(p.ring_mask) || !is_mask(p.seg_mask) || !is_mask(p.flag_mask) ||
        !is_mask(p.removed_mask))
        return rvx::fail(RVX_EINVAL, "%s: every mask must be a power of two minus one", who);
    const uint64_t ring = ring_of(p.ring_mask);
    if ((ring_of(p.seg_mask) << c.seg_log2) > ring || ring_of(p.flag_mask) > ring ||
        ring_of(p.removed_mask) > ring)
        return rvx::fail(RVX_EINVAL,
                         "%s: the segment, flag and removed rings must each be no larger than "
                         "the sample ring",
                         who);
    if (c.pass == kPassSegment || c.pass == kPassReference) {
        if ((p.first & ((1u << c.seg_log2) - 1u)) != 0u)
            return rvx::fail(RVX_EINVAL, "%s: a segment pass must start on a segment boundary",
                             who);
        if (uint64_t(p.count) + c.segments + 2u > ring_of(p.seg_mask))
            return rvx::fail(RVX_EINVAL, "%s: %u segments and a %u segment reference do not fit "
                             "a %llu segment ring", who, p.count, c.segments,
                             (unsigned long long)ring_of(p.seg_mask));
    } else {
        const uint64_t reach = uint64_t(p.count) + 2ull * c.max_width + c.hang + c.lead;
        if (reach >= ring_of(p.flag_mask) ||
            (c.pass == kPassApply && p.count > ring_of(p.removed_mask)))
            return rvx::fail(RVX_EINVAL, "%s: %u samples and their scan do not fit the flag ring "
                             "of %llu or the removed ring of %llu", who, p.count,
                             (unsigned long long)ring_of(p.flag_mask),
                             (unsigned long long)ring_of(p.removed_mask));
    }
    if (!(p.threshold > 0.0f))
        return rvx::fail(RVX_EINVAL, "%s: the threshold must be a positive power ratio", who);
    return RVX_OK;
}

int32_t rvx_size_excision_rings(const rvx_excision_config* config, uint64_t max_block,
                                uint32_t frames_in_flight, uint32_t ring_mask,
                                rvx_excision_rings* out) {
    static const char* who = "rvx_size_excision_rings";
    if (!config || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    const rvx_excision_config& c = *config;
    auto ceil2 = [](uint64_t v) {
        uint64_t r = 1;
        while (r < v) r <<= 1;
        return r;
    };
    const uint64_t frames = std::max<uint32_t>(frames_in_flight, 1u);
    const uint64_t segment = uint64_t(1) << c.seg_log2;
    const uint64_t lag = std::max<uint64_t>(uint64_t(c.max_width) + c.lead, segment);
    const uint64_t in_flight = max_block * frames;
    const uint64_t flags = ceil2(in_flight + lag + 2ull * c.max_width + c.hang + c.lead + 1);
    const uint64_t segs = ceil2(in_flight / segment + c.segments + 4);
    const uint64_t removed = ceil2((max_block + lag) * frames);
    const uint64_t ring = ring_of(ring_mask);
    if (flags > ring || removed > ring || (segs << c.seg_log2) > ring)
        return rvx::fail(RVX_EINVAL, "%s: a %llu sample ring is too small for %llu samples in "
                         "flight and their scan", who, (unsigned long long)ring,
                         (unsigned long long)in_flight);
    out->seg_mask = uint32_t(segs - 1);
    out->flag_mask = uint32_t(flags - 1);
    out->removed_mask = uint32_t(removed - 1);
    return RVX_OK;
}

int32_t rvx_excision_cursor_create(const rvx_excision_config* config, uint32_t ring_mask,
                                   const rvx_excision_rings* rings, float threshold,
                                   rvx_excision_cursor** out) {
    if (!config || !rings || !out) return rvx::fail(RVX_EINVAL, "rvx_excision_cursor_create: null");
    auto* k = new (std::nothrow) rvx_excision_cursor();
    if (!k) return rvx::fail(RVX_ENOMEM, "rvx_excision_cursor_create: out of memory");
    k->config = *config;
    k->ring_mask = ring_mask;
    k->rings = *rings;
    k->threshold = threshold;
    *out = k;
    return RVX_OK;
}

void rvx_excision_cursor_destroy(rvx_excision_cursor* k) { delete k; }

void rvx_excision_cursor_reset(rvx_excision_cursor* k) {
    if (k) k->running = false;
}

int32_t rvx_excision_cursor_plan(rvx_excision_cursor* k, uint64_t start, uint32_t count,
                                 rvx_excision_plan* out) {
    if (!k || !out) return rvx::fail(RVX_EINVAL, "rvx_excision_cursor_plan: null argument");
    rvx_excision_plan plan{};
    const uint64_t segment = uint64_t(1) << k->config.seg_log2;
    if (!k->running || start != k->next_in) {  // the first block, or a gap: restart
        k->running = true;
        plan.restarted = 1;
        k->valid_from = start;
        k->seg_next = (start + segment - 1) / segment;
        k->flag_next = start;
        k->apply_next = start;
    }
    const uint64_t end = start + count;
    k->next_in = end;
    const uint64_t seg_end = end / segment;  // every segment now whole
    if (seg_end > k->seg_next) {
        plan.segment = k->at(k->seg_next * segment, uint32_t(seg_end - k->seg_next));
        k->seg_next = seg_end;
    }
    if (end > k->flag_next) {
        const uint64_t s_lo = k->flag_next / segment, s_hi = (end - 1) / segment;
        plan.reference = k->at(s_lo * segment, uint32_t(s_hi - s_lo + 1));
        plan.flag = k->at(k->flag_next, uint32_t(end - k->flag_next));
        k->flag_next = end;
    }
    // A sample is finished when the scan max_width + lead ahead has flags and
    // its own segment is whole.
    const uint64_t reach = uint64_t(k->config.max_width) + k->config.lead;
    const uint64_t by_scan = end > reach ? end - reach : 0;
    const uint64_t by_segment = (end / segment) * segment;
    const uint64_t apply_end = std::min(by_scan, by_segment);
    if (apply_end > k->apply_next) {
        plan.apply = k->at(k->apply_next, uint32_t(apply_end - k->apply_next));
        k->apply_next = apply_end;
    }
    plan.cleaned_end = k->apply_next;
    *out = plan;
    return RVX_OK;
}

int32_t rvx_excise_segment(rvx_ctx* ctx, const rvx_excision_config* config,
                           const rvx_excision_params* params, const void* raw,
                           uint64_t raw_samples, float* seg_power, uint64_t n_seg) {
    static const char* who = "rvx_excise_segment";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    rvx_excision_config c = *config;
    c.pass = kPassSegment;
    if (int32_t rc = rvx_excision_check(&c, params)) return rc;
    const rvx_excision_params p = *params;
    if (raw_samples < ring_of(p.ring_mask) || n_seg < ring_of(p.seg_mask))
        return rvx::fail(RVX_EINVAL, "%s: a ring is shorter than its mask", who);
    if (p.count == 0) return RVX_OK;
    if (!raw || !seg_power) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    segment_kernel<<<grid_for(p.count), dim3(kThreads), 0, ctx->stream>>>(
        c, p, static_cast<const float2*>(raw), seg_power);
    return rvx::check_launch(who);
}

int32_t rvx_excise_reference(rvx_ctx* ctx, const rvx_excision_config* config,
                             const rvx_excision_params* params, const float* seg_power,
                             float* seg_level, uint64_t n_seg) {
    static const char* who = "rvx_excise_reference";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    rvx_excision_config c = *config;
    c.pass = kPassReference;
    if (int32_t rc = rvx_excision_check(&c, params)) return rc;
    const rvx_excision_params p = *params;
    if (n_seg < ring_of(p.seg_mask))
        return rvx::fail(RVX_EINVAL, "%s: a segment ring is shorter than its mask", who);
    if (p.count == 0) return RVX_OK;
    if (!seg_power || !seg_level) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    reference_kernel<<<grid_for(p.count), dim3(kThreads), 0, ctx->stream>>>(c, p, seg_power,
                                                                            seg_level);
    return rvx::check_launch(who);
}

int32_t rvx_excise_flag(rvx_ctx* ctx, const rvx_excision_config* config,
                        const rvx_excision_params* params, const void* raw, uint64_t raw_samples,
                        const float* seg_level, uint64_t n_seg, uint32_t* flags, uint64_t n_flags) {
    static const char* who = "rvx_excise_flag";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    rvx_excision_config c = *config;
    c.pass = kPassFlag;
    if (int32_t rc = rvx_excision_check(&c, params)) return rc;
    const rvx_excision_params p = *params;
    if (raw_samples < ring_of(p.ring_mask) || n_seg < ring_of(p.seg_mask) ||
        n_flags < ring_of(p.flag_mask))
        return rvx::fail(RVX_EINVAL, "%s: a ring is shorter than its mask", who);
    if (p.count == 0) return RVX_OK;
    if (!raw || !seg_level || !flags) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    flag_kernel<<<grid_for(p.count), dim3(kThreads), 0, ctx->stream>>>(
        c, p, static_cast<const float2*>(raw), seg_level, flags);
    return rvx::check_launch(who);
}

int32_t rvx_excise_apply(rvx_ctx* ctx, const rvx_excision_config* config,
                         const rvx_excision_params* params, const void* raw, uint64_t raw_samples,
                         const uint32_t* flags, uint64_t n_flags, void* cleaned,
                         uint64_t cleaned_samples, void* removed, uint64_t removed_samples,
                         uint32_t* counters) {
    static const char* who = "rvx_excise_apply";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    rvx_excision_config c = *config;
    c.pass = kPassApply;
    if (int32_t rc = rvx_excision_check(&c, params)) return rc;
    const rvx_excision_params p = *params;
    if (raw_samples < ring_of(p.ring_mask) || cleaned_samples < ring_of(p.ring_mask) ||
        n_flags < ring_of(p.flag_mask) || removed_samples < ring_of(p.removed_mask))
        return rvx::fail(RVX_EINVAL, "%s: a ring is shorter than its mask", who);
    if (p.count == 0) return RVX_OK;
    if (!raw || !flags || !cleaned || !removed || !counters)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    apply_kernel<<<grid_for(p.count), dim3(kThreads), 0, ctx->stream>>>(
        c, p, static_cast<const float2*>(raw), flags, static_cast<float2*>(cleaned),
        static_cast<float2*>(removed), counters);
    return rvx::check_launch(who);
}

}  // extern "C"
