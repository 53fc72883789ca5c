// Front-end correction on the device: the five moments a block's DC and I/Q
// imbalance estimate is read from, and the correction itself.
//
//   rvx_iq_moments        iq_moments.comp / front_end_correction_reference.cpp:46-88
//   rvx_iq_moments_codes  the same sums over native codes, each converted as
//                         convert_*_cf32.comp would (convert.cpp:126-235)
//   rvx_iq_correct        iq_correct.comp / front_end_correction_reference.cpp:90-111
//
// The moments are chunk sums in sample order from zero (64 samples a chunk,
// five sums: I, Q, II, QQ, IQ), one rounding per operation, so the order is
// fixed by the format and not by the launch. On a ring, a wave stages its 64
// chunks in LDS with coalesced loads (each chunk row padded by one sample so
// the per-lane reads down a row fall in different banks), then each lane sums
// its own chunk from LDS; on native codes each lane loads its chunk straight
// into registers. Every operation is a single IEEE add or multiply; no
// contraction, denormals flushed (the build's flags).

#pragma clang fp contract(off)

#include "common.h"

namespace {

constexpr uint32_t kChunk = 64;      // samples per chunk (front_end_correction.h:84)
constexpr uint32_t kPerChunk = 5;    // sums per chunk (:88)
constexpr uint32_t kLanes = 64;      // chunks per workgroup
constexpr uint32_t kRow = kChunk + 1;

// A cf32 ring read at (offset + k) & mask.
struct RingSource {
    const float2* ring;
    uint32_t mask, offset;
    __device__ float2 operator()(uint32_t k) const { return ring[(offset + k) & mask]; }
};

// Native codes read linearly from sample `first`, converted exactly.
template <int FORMAT>
struct CodeSource {
    const void* codes;
    uint64_t first;
    __device__ float2 operator()(uint32_t k) const {
        const uint64_t s = first + k;
        if constexpr (FORMAT == RVX_FORMAT_CS16) {
            const uint32_t w = static_cast<const uint32_t*>(codes)[s];
            const float scale = __uint_as_float(0x38000000u);  // 2^-15
            return make_float2(float(int32_t(w << 16) >> 16) * scale,
                               float(int32_t(w) >> 16) * scale);
        } else {
            const uint32_t w = static_cast<const uint16_t*>(codes)[s];
            const float bias = __uint_as_float(0x42FF0000u);   // 127.5
            const float scale = __uint_as_float(0x3C008081u);  // 1/127.5
            return make_float2((float(w & 0xFFu) - bias) * scale,
                               (float((w >> 8) & 0xFFu) - bias) * scale);
        }
    }
};

template <class Source>
__global__ void __launch_bounds__(kLanes)
iq_moments_kernel(Source src, uint32_t count, float* __restrict__ sums) {
    __shared__ float2 stage[kLanes * kRow];
    const uint32_t chunk0 = blockIdx.x * kLanes;
    const uint32_t first = chunk0 * kChunk;
    const uint32_t n = min(kLanes * kChunk, count - first);
    // coalesced: lane l loads samples l, l + 64, ... of the workgroup's span
    for (uint32_t k = threadIdx.x; k < n; k += kLanes)
        stage[(k / kChunk) * kRow + (k % kChunk)] = src(first + k);
    __syncthreads();
    const uint32_t chunk = chunk0 + threadIdx.x;
    const uint32_t c_first = chunk * kChunk;
    if (c_first >= count) return;
    const uint32_t len = min(kChunk, count - c_first);
    const float2* row = stage + threadIdx.x * kRow;
    float si = 0.0f, sq = 0.0f, sii = 0.0f, sqq = 0.0f, siq = 0.0f;
    for (uint32_t k = 0; k < len; ++k) {
        const float2 v = row[k];
        const float i = v.x, q = v.y;
        si = si + i;
        sq = sq + q;
        sii = sii + i * i;
        sqq = sqq + q * q;
        siq = siq + i * q;
    }
    float* out = sums + uint64_t(chunk) * kPerChunk;
    out[0] = si;
    out[1] = sq;
    out[2] = sii;
    out[3] = sqq;
    out[4] = siq;
}

// The native-code form reads each lane's chunk straight into registers:
// one lane per chunk, its 64 codes loaded as sixteen 16-byte vectors (cs16)
// or eight (cu8) from a linear buffer, no LDS, so occupancy is set by
// registers alone and a wave keeps a whole 16 KiB span of loads in flight.
template <int FORMAT>
__global__ void __launch_bounds__(256)
iq_moments_codes_kernel(const void* __restrict__ codes, uint64_t first, uint32_t count,
                        float* __restrict__ sums) {
    const uint32_t chunk = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c_first = chunk * kChunk;
    if (c_first >= count) return;
    const uint32_t len = min(kChunk, count - c_first);
    CodeSource<FORMAT> src{codes, first};
    float si = 0.0f, sq = 0.0f, sii = 0.0f, sqq = 0.0f, siq = 0.0f;
    auto add = [&](float2 v) {
        const float i = v.x, q = v.y;
        si = si + i;
        sq = sq + q;
        sii = sii + i * i;
        sqq = sqq + q * q;
        siq = siq + i * q;
    };
    constexpr uint32_t kBytes = FORMAT == RVX_FORMAT_CS16 ? 4u : 2u;
    const uint64_t s0 = first + c_first;
    const uint64_t at = reinterpret_cast<uintptr_t>(codes) + s0 * kBytes;
    if (len == kChunk && (at & 15u) == 0) {
        // whole, 16-byte-aligned chunk: vector loads, then the sums in order
        const uint4* p = reinterpret_cast<const uint4*>(static_cast<const uint8_t*>(codes) +
                                                        s0 * kBytes);
        constexpr uint32_t kVec = kChunk * kBytes / 16;
        uint4 w[kVec];
#pragma unroll
        for (uint32_t k = 0; k < kVec; ++k) w[k] = p[k];
#pragma unroll
        for (uint32_t k = 0; k < kVec; ++k) {
            const uint32_t words[4] = {w[k].x, w[k].y, w[k].z, w[k].w};
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                if constexpr (FORMAT == RVX_FORMAT_CS16) {
                    const float scale = __uint_as_float(0x38000000u);  // 2^-15
                    add(make_float2(float(int32_t(words[j] << 16) >> 16) * scale,
                                    float(int32_t(words[j]) >> 16) * scale));
                } else {
                    const float bias = __uint_as_float(0x42FF0000u);   // 127.5
                    const float scale = __uint_as_float(0x3C008081u);  // 1/127.5
                    for (uint32_t h = 0; h < 2; ++h) {
                        const uint32_t c = words[j] >> (16 * h);
                        add(make_float2((float(c & 0xFFu) - bias) * scale,
                                        (float((c >> 8) & 0xFFu) - bias) * scale));
                    }
                }
            }
        }
    } else {
        for (uint32_t k = 0; k < len; ++k) add(src(c_first + k));
    }
    float* out = sums + uint64_t(chunk) * kPerChunk;
    out[0] = si;
    out[1] = sq;
    out[2] = sii;
    out[3] = sqq;
    out[4] = siq;
}

__global__ void iq_correct_kernel(const float2* in, rvx_iq_correct_params p, float2* out) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= p.count) return;
    const uint32_t slot = (p.offset + index) & p.ring_mask;
    const float2 v = in[slot];
    const float i = v.x - p.dc_i;
    const float q_dc = v.y - p.dc_q;
    const float q = p.cross * i + p.scale * q_dc;
    out[slot] = make_float2(i, q);
}

int32_t ring_ok(uint64_t ring_samples, uint32_t mask, uint32_t count, const char* who) {
    const uint64_t capacity = uint64_t(mask) + 1;
    if ((capacity & (capacity - 1)) != 0)
        return rvx::fail(RVX_EINVAL, "%s: ring_mask %u is not a power of two less one", who, mask);
    if (ring_samples < capacity)
        return rvx::fail(RVX_EINVAL, "%s: the ring holds %llu samples and the mask names %llu",
                         who, (unsigned long long)ring_samples, (unsigned long long)capacity);
    if (count > capacity)
        return rvx::fail(RVX_EINVAL, "%s: a block of %u samples is longer than the %llu-sample ring",
                         who, count, (unsigned long long)capacity);
    return RVX_OK;
}

uint32_t chunks_of(uint32_t count) { return (count + kChunk - 1) / kChunk; }

}  // namespace

extern "C" {

int32_t rvx_iq_moments(rvx_ctx* ctx, const rvx_iq_moments_params* params, const void* ring,
                       uint64_t ring_samples, float* sums, uint64_t n_sums) {
    static const char* who = "rvx_iq_moments";
    RVX_ENTER(ctx, who);
    if (!params) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_iq_moments_params p = *params;
    if (int32_t rc = ring_ok(ring_samples, p.ring_mask, p.count, who)) return rc;
    const uint32_t chunks = chunks_of(p.count);
    if (n_sums < uint64_t(chunks) * kPerChunk)
        return rvx::fail(RVX_EINVAL, "%s: %u chunks need %u floats and %llu were given", who,
                         chunks, chunks * kPerChunk, (unsigned long long)n_sums);
    if (chunks == 0) return RVX_OK;
    if (!ring || !sums) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    RingSource src{static_cast<const float2*>(ring), p.ring_mask, p.src_offset};
    iq_moments_kernel<<<dim3((chunks + kLanes - 1) / kLanes), dim3(kLanes), 0, ctx->stream>>>(
        src, p.count, sums);
    return rvx::check_launch(who);
}

int32_t rvx_iq_moments_codes(rvx_ctx* ctx, int32_t format, const void* codes,
                             uint64_t code_samples, uint64_t first, uint32_t count, float* sums,
                             uint64_t n_sums) {
    static const char* who = "rvx_iq_moments_codes";
    RVX_ENTER(ctx, who);
    if (format != RVX_FORMAT_CU8 && format != RVX_FORMAT_CS16)
        return rvx::fail(RVX_EINVAL, "%s: format %d (cu8 or cs16)", who, format);
    if (first > code_samples || count > code_samples - first)
        return rvx::fail(RVX_EINVAL, "%s: samples %llu..+%u past the %llu given", who,
                         (unsigned long long)first, count, (unsigned long long)code_samples);
    const uint32_t chunks = chunks_of(count);
    if (n_sums < uint64_t(chunks) * kPerChunk)
        return rvx::fail(RVX_EINVAL, "%s: %u chunks need %u floats", who, chunks,
                         chunks * kPerChunk);
    if (chunks == 0) return RVX_OK;
    if (!codes || !sums) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    const dim3 grid((chunks + 255) / 256), block(256);
    if (format == RVX_FORMAT_CS16)
        iq_moments_codes_kernel<RVX_FORMAT_CS16><<<grid, block, 0, ctx->stream>>>(codes, first,
                                                                                  count, sums);
    else
        iq_moments_codes_kernel<RVX_FORMAT_CU8><<<grid, block, 0, ctx->stream>>>(codes, first,
                                                                                 count, sums);
    return rvx::check_launch(who);
}

int32_t rvx_iq_correct(rvx_ctx* ctx, const rvx_iq_correct_params* params, const void* in,
                       uint64_t in_samples, void* out, uint64_t out_samples) {
    static const char* who = "rvx_iq_correct";
    RVX_ENTER(ctx, who);
    if (!params) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_iq_correct_params p = *params;
    if (int32_t rc = ring_ok(in_samples, p.ring_mask, p.count, who)) return rc;
    if (out_samples < in_samples)
        return rvx::fail(RVX_EINVAL, "%s: the output ring holds %llu samples and the input %llu",
                         who, (unsigned long long)out_samples, (unsigned long long)in_samples);
    if (p.count == 0) return RVX_OK;
    if (!in || !out) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    iq_correct_kernel<<<dim3((p.count + 255) / 256), dim3(256), 0, ctx->stream>>>(
        static_cast<const float2*>(in), p, static_cast<float2*>(out));
    return rvx::check_launch(who);
}

}  // extern "C"
