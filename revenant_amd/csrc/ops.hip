// The receive chain's operators, one HIP kernel (or two) each, with host-side
// validation that refuses exactly what the reference twin refuses.
//
//   rvx_convert          convert_{cu8,cs8,cs16,cs24}_cf32.comp / convert.cpp
//   rvx_cmul             cmul.comp / complex_ops.cpp:33
//   rvx_pfb_branch       pfb_branch.comp / pfb_branch_reference.cpp:22
//   rvx_pfb_fft          pfb_fft.comp / pfb_fft_reference.cpp:143
//   rvx_fft_radix2       pfb_fft_reference.cpp:234
//   rvx_spectrum         spectrum.comp / spectrum_reference.cpp:286
//   rvx_spectrum_levels  spectrum_levels.comp / spectrum_levels_reference.cpp:38
//   rvx_ema_partial/combine  the averaged power spectrum (no reference kernel)
//
// Built with -ffp-contract=off and -fgpu-flush-denormals-to-zero: no fused
// multiply-add anywhere and FP32 denormals flushed on input and output, the
// two rules the reference's bit-exactness rests on (reference_fp.h:20-33,
// denormal_mode.h:37-72).

#pragma clang fp contract(off)

#include <algorithm>

#include "common.h"
#include "fft.h"

namespace {

using rvx::Fft;

constexpr uint32_t kMaxTransformLog2 = 14;  // 16384 points: 128 KiB of LDS

// ---------------------------------------------------------------------------
// Convert
// ---------------------------------------------------------------------------

__device__ __forceinline__ int32_t field_signed(uint32_t word, int offset, int bits) {
    return int32_t(word << (32 - offset - bits)) >> (32 - bits);
}

template <int FORMAT>
__global__ void convert_kernel(const uint32_t* __restrict__ words, rvx_convert_params p,
                               float2* __restrict__ ring) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= p.count) return;
    const uint32_t s = p.src_offset + index;
    float iv, qv;
    if constexpr (FORMAT == RVX_FORMAT_CU8) {
        const uint32_t w = words[s >> 1];
        const int shift = int((s & 1u) << 4);
        const float bias = __uint_as_float(0x42FF0000u);
        const float scale = __uint_as_float(0x3C008081u);
        iv = (float((w >> shift) & 0xFFu) - bias) * scale;
        qv = (float((w >> (shift + 8)) & 0xFFu) - bias) * scale;
    } else if constexpr (FORMAT == RVX_FORMAT_CS8) {
        const uint32_t w = words[s >> 1];
        const int shift = int((s & 1u) << 4);
        const float scale = __uint_as_float(0x3C000000u);
        iv = float(field_signed(w, shift, 8)) * scale;
        qv = float(field_signed(w, shift + 8, 8)) * scale;
    } else if constexpr (FORMAT == RVX_FORMAT_CS16) {
        const uint32_t w = words[s];
        const float scale = __uint_as_float(0x38000000u);
        iv = float(field_signed(w, 0, 16)) * scale;
        qv = float(field_signed(w, 16, 16)) * scale;
    } else {
        const uint32_t wi = s + (s >> 1);
        const uint32_t lo = words[wi];
        const uint32_t hi = words[wi + 1];
        const uint32_t shift = (s & 1u) << 4;
        const uint32_t window = (lo >> shift) | ((hi << 16) << (16 - shift));
        const uint32_t qraw = (window >> 24) | ((hi >> shift) << 8);
        const float scale = __uint_as_float(0x34000000u);
        iv = float(field_signed(window, 0, 24)) * scale;
        qv = float(field_signed(qraw, 0, 24)) * scale;
    }
    ring[(p.dst_offset + index) & p.ring_mask] = make_float2(iv, qv);
}

__global__ void cmul_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                            float2* __restrict__ out, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = rvx::cmul_exact(a[i], b[i]);
}

// ---------------------------------------------------------------------------
// Branch filter: one thread per (block, branch), as pfb_branch.comp:99-149.
// ---------------------------------------------------------------------------

__global__ void pfb_branch_kernel(const float2* __restrict__ ring, const float* __restrict__ taps,
                                  rvx_pfb_branch_params p, uint32_t M, uint32_t log2m,
                                  uint32_t taps_per_branch, uint32_t D,
                                  float2* __restrict__ out) {
    const uint64_t id = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= uint64_t(p.block_count) * M) return;
    const uint32_t block = uint32_t(id >> log2m);
    const uint32_t r = uint32_t(id) & (M - 1);
    const uint32_t base = p.base_offset + block * D;
    float acc_re = 0.0f, acc_im = 0.0f;
    for (uint32_t q = 0; q < taps_per_branch; ++q) {
        const uint32_t n = q * M + r;
        const float2 s = ring[(base - n) & p.ring_mask];
        const float c = taps[n];
        acc_re = acc_re + s.x * c;
        acc_im = acc_im + s.y * c;
    }
    const uint32_t slot = (M - r) & (M - 1);
    out[uint64_t(block) * M + slot] = make_float2(acc_re, acc_im);
}

// ---------------------------------------------------------------------------
// Transforms: BPW transforms of N points per 256-thread workgroup (or one
// transform of N/16 threads when that is larger).
// ---------------------------------------------------------------------------

template <int LOGN>
struct TransformShape {
    static constexpr int LOGV = LOGN < 4 ? LOGN : 4;
    using F = Fft<LOGN, LOGV>;
    static constexpr int THREADS = F::T >= 256 ? F::T : 256;
    static constexpr int BPW = THREADS / F::T;  // transforms per workgroup
};

// MODE 0: plain batched transform (fft_radix2). MODE 1: pfb_fft (correction +
// channel-major scatter).
template <int LOGN, int MODE>
__global__ void __launch_bounds__(TransformShape<LOGN>::THREADS)
transform_kernel(const float2* __restrict__ in, const float2* __restrict__ tw,
                 float2* __restrict__ out, uint32_t batch, rvx_pfb_fft_params p) {
    using S = TransformShape<LOGN>;
    using F = typename S::F;
    constexpr uint32_t N = F::N;
    __shared__ float2 lds[S::BPW * N];
    const uint32_t sub = threadIdx.x / F::T;
    const uint32_t t = threadIdx.x % F::T;
    const uint32_t block = blockIdx.x * S::BPW + sub;
    float2* my = lds + sub * N;
    const bool active = block < batch;
    if (active) {
        const float2* src = in + uint64_t(block) * N;
        for (uint32_t i = t; i < N; i += F::T) my[F::phys(rvx::bitrev<LOGN>(i))] = src[i];
    }
    __syncthreads();
    float2 v[F::V];
    F::template load<0>(my, v, t);
    F::template run_pass<0>(v, tw, t);
    // All subgroups run the same passes (barriers are workgroup-wide); the
    // inactive tail transforms compute on stale LDS and write nothing.
    F::template rest<1>(my, v, tw, t);
    if (!active) return;
    constexpr int LAST = F::PASSES - 1;
    if constexpr (MODE == 0) {
        float2* dst = out + uint64_t(block) * N;
#pragma unroll
        for (int r = 0; r < F::V; ++r) dst[F::template pos<LAST>(t, r)] = v[r];
    } else {
        const uint32_t m_abs = p.block_base + block;
        const uint32_t slot = m_abs & p.out_ring_mask;
        const uint32_t dm = ((m_abs & (N - 1)) * p.decimation) & (N - 1);
#pragma unroll
        for (int r = 0; r < F::V; ++r) {
            const uint32_t k = F::template pos<LAST>(t, r);
            float2 value = v[r];
            if (p.decimation != N) value = rvx::cmul_exact(value, tw[(k * dm) & (N - 1)]);
            out[uint64_t(k) * p.out_ring_blocks + slot] = value;
        }
    }
}

template <int MODE>
int32_t launch_transform(uint32_t logn, const float2* in, const float2* tw, float2* out,
                         uint32_t batch, const rvx_pfb_fft_params& p, hipStream_t stream) {
    auto go = [&](auto tag) {
        constexpr int L = decltype(tag)::value;
        using S = TransformShape<L>;
        const uint32_t groups = (batch + S::BPW - 1) / S::BPW;
        transform_kernel<L, MODE><<<dim3(groups), dim3(S::THREADS), 0, stream>>>(in, tw, out,
                                                                                 batch, p);
    };
    switch (logn) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 3: go(std::integral_constant<int, 3>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        case 5: go(std::integral_constant<int, 5>{}); break;
        case 6: go(std::integral_constant<int, 6>{}); break;
        case 7: go(std::integral_constant<int, 7>{}); break;
        case 8: go(std::integral_constant<int, 8>{}); break;
        case 9: go(std::integral_constant<int, 9>{}); break;
        case 10: go(std::integral_constant<int, 10>{}); break;
        case 11: go(std::integral_constant<int, 11>{}); break;
        case 12: go(std::integral_constant<int, 12>{}); break;
        case 13: go(std::integral_constant<int, 13>{}); break;
        case 14: go(std::integral_constant<int, 14>{}); break;
        default: return rvx::fail(RVX_EINVAL, "transform of 2^%u points is not supported", logn);
    }
    return RVX_OK;
}

// ---------------------------------------------------------------------------
// Spectrum: one transform per coarse channel (spectrum.comp:252-343).
// ---------------------------------------------------------------------------

__device__ __forceinline__ float det_recip(float v) {
    const float seed = __uint_as_float(0x7EF311C3u - __float_as_uint(v));
    const float r1 = seed * (2.0f - v * seed);
    const float r2 = r1 * (2.0f - v * r1);
    const float r3 = r2 * (2.0f - v * r2);
    const float r4 = r3 * (2.0f - v * r3);
    return r4;
}

__device__ __forceinline__ float det_log2(float v) {
    const uint32_t bits = __float_as_uint(v);
    int e = int((bits >> 23) & 0xFFu) - 127;
    float m = __uint_as_float((bits & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356237309505f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float t = (m - 1.0f) * det_recip(m + 1.0f);
    const float u = t * t;
    float p = 0.32059889797532520f;
    p = p * u + 0.41219858311113240f;
    p = p * u + 0.57707801635558536f;
    p = p * u + 0.96179669392597560f;
    p = p * u + 2.88539008177792681f;
    const float fraction = t * p;
    return float(e) + fraction;
}

template <int LOGN>
__global__ void __launch_bounds__(TransformShape<LOGN>::THREADS)
spectrum_kernel(const float2* __restrict__ ring, const float2* __restrict__ tw,
                const float* __restrict__ window, float* __restrict__ frame,
                rvx_spectrum_params p) {
    using S = TransformShape<LOGN>;
    using F = typename S::F;
    constexpr uint32_t N = F::N;
    __shared__ float2 lds[S::BPW * N];
    const uint32_t sub = threadIdx.x / F::T;
    const uint32_t t = threadIdx.x % F::T;
    const uint32_t ch = blockIdx.x * S::BPW + sub;
    float2* my = lds + sub * N;
    const bool active = ch < p.channels;
    if (active) {
        const float2* src = ring + uint64_t(ch) * p.chan_blocks;
        for (uint32_t n = t; n < N; n += F::T) {
            const float2 x = src[(p.in_offset + n) & p.chan_mask];
            const float g = window[n];
            const float re = x.x * g;
            const float im = x.y * g;
            my[F::phys(rvx::bitrev<LOGN>(n))] = make_float2(re, im);
        }
    }
    __syncthreads();
    float2 v[F::V];
    F::from_lds(my, v, tw, t);
    if (!active) return;
    constexpr int LAST = F::PASSES - 1;
    constexpr uint32_t HALF = N / 2;
    float* dst = frame + uint64_t((ch + p.channels / 2) & (p.channels - 1)) * HALF;
#pragma unroll
    for (int r = 0; r < F::V; ++r) {
        const uint32_t k = F::template pos<LAST>(t, r);
        const uint32_t bin = (k + N / 4) & (N - 1);
        if (bin >= HALF) continue;
        const float power = v[r].x * v[r].x + v[r].y * v[r].y;
        const float flat = power * window[N + bin];
        const float floored = (flat < 1.0e-20f) ? 1.0e-20f : flat;
        dst[bin] = det_log2(floored) * 3.01029995663981195f;
    }
}

// ---------------------------------------------------------------------------
// Levels: integer histogram (LDS atomics are exact, so order-free), then a
// scan for the two ranks. spectrum_levels.comp / spectrum_levels_reference.cpp.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t level_bucket(float db) {
    float t = (db - (-200.0f)) * 4.26666666666666666f;
    t = (t < 0.0f) ? 0.0f : t;
    t = (t > 1023.0f) ? 1023.0f : t;
    return uint32_t(t);
}

__global__ void __launch_bounds__(1024)
levels_kernel(const float* __restrict__ frame, rvx_spectrum_levels_params p,
              float* __restrict__ levels) {
    __shared__ uint32_t hist[1024];
    __shared__ uint32_t prefix[1024];
    const uint32_t t = threadIdx.x;
    hist[t] = 0;
    __syncthreads();
    for (uint32_t i = t; i < p.bins; i += 1024) atomicAdd(&hist[level_bucket(frame[i])], 1u);
    __syncthreads();
    // Inclusive prefix sum over the 1024 buckets (Hillis-Steele, exact ints).
    prefix[t] = hist[t];
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t add = t >= off ? prefix[t - off] : 0u;
        __syncthreads();
        prefix[t] += add;
        __syncthreads();
    }
    auto rank_of = [](uint32_t count, uint32_t permille) {
        return (count / 1000u) * permille + ((count % 1000u) * permille) / 1000u;
    };
    const uint32_t lo = min(rank_of(p.bins, p.low_permille), p.bins - 1);
    const uint32_t hi = min(rank_of(p.bins, p.high_permille), p.bins - 1);
    // The first bucket whose cumulative count exceeds the rank: exactly one
    // thread sees the crossing.
    const uint32_t before = t ? prefix[t - 1] : 0u;
    if (before <= lo && prefix[t] > lo) levels[0] = -200.0f + (float(t) + 0.5f) * 0.234375f;
    if (before <= hi && prefix[t] > hi) levels[1] = -200.0f + (float(t) + 0.5f) * 0.234375f;
}

// ---------------------------------------------------------------------------
// Averaged power spectrum (partial per frame, then the frame recurrence).
// ---------------------------------------------------------------------------

__global__ void ema_partial_kernel(const float2* __restrict__ ring, rvx_ema_params p,
                                   float* __restrict__ partial) {
    const uint64_t id = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= uint64_t(p.frames) * p.channels) return;
    const uint32_t f = uint32_t(id % p.frames);
    const uint32_t k = uint32_t(id / p.frames);
    const float2* src = ring + uint64_t(k) * p.chan_blocks;
    float s = 0.0f;
    for (uint32_t i = 0; i < p.frame_blocks; ++i) {
        const float2 v = src[(p.in_offset + f * p.frame_blocks + i) & p.chan_mask];
        const float x = v.x * v.x + v.y * v.y;
        s = s + (x - s) * p.alpha;
    }
    partial[uint64_t(f) * p.channels + k] = s;
}

// Row partials, in place into the rows buffer: S_r from zero over its frames.
__global__ void ema_rows_kernel(const float* __restrict__ partial, rvx_ema_params p,
                                float* __restrict__ rows, uint32_t nrows, bool quad) {
    const uint64_t id = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (quad) {
        // Four consecutive channels per lane, 16-byte loads and stores; each
        // channel's sum is the same chain of operations.
        const uint32_t quads = p.channels / 4;
        if (id >= uint64_t(nrows) * quads) return;
        const uint32_t r = uint32_t(id / quads), k = uint32_t(id % quads) * 4;
        float4 S = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (uint32_t g = 0; g < p.row_frames; ++g) {
            const float4 v = *reinterpret_cast<const float4*>(
                partial + uint64_t(r * p.row_frames + g) * p.channels + k);
            S.x = S.x * p.frame_decay + v.x;
            S.y = S.y * p.frame_decay + v.y;
            S.z = S.z * p.frame_decay + v.z;
            S.w = S.w * p.frame_decay + v.w;
        }
        *reinterpret_cast<float4*>(rows + uint64_t(r) * p.channels + k) = S;
        return;
    }
    if (id >= uint64_t(nrows) * p.channels) return;
    const uint32_t r = uint32_t(id / p.channels), k = uint32_t(id % p.channels);
    float S = 0.0f;
    for (uint32_t g = 0; g < p.row_frames; ++g)
        S = S * p.frame_decay + partial[uint64_t(r * p.row_frames + g) * p.channels + k];
    rows[id] = S;
}

// The serial part, one step per row: P = P * row_decay + S_r. Row values are
// fetched kBatch at a time, the next batch's loads in flight while this
// batch's steps run, so the recurrence waits on neither.
__global__ void ema_combine_kernel(rvx_ema_params p, float* __restrict__ state,
                                   float* __restrict__ rows, uint32_t nrows) {
    constexpr uint32_t kBatch = 16;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.channels) return;
    float P = state[k];
    const uint32_t whole = nrows / kBatch * kBatch;
    float S[2][kBatch];
    if (whole) {
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) S[0][i] = rows[uint64_t(i) * p.channels + k];
    }
    for (uint32_t r = 0; r < whole; r += 2 * kBatch) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t at = r + uint32_t(h) * kBatch;
            if (at >= whole) break;
            if (at + kBatch < whole) {
#pragma unroll
                for (uint32_t i = 0; i < kBatch; ++i)
                    S[h ^ 1][i] = rows[uint64_t(at + kBatch + i) * p.channels + k];
            }
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i) {
                P = P * p.row_decay + S[h][i];
                rows[uint64_t(at + i) * p.channels + k] = P;
            }
        }
    }
    for (uint32_t r = whole; r < nrows; ++r) {
        P = P * p.row_decay + rows[uint64_t(r) * p.channels + k];
        rows[uint64_t(r) * p.channels + k] = P;
    }
    state[k] = P;
}

bool fft_params_ok(const rvx_pfb_fft_params& p) {
    if (p.channels < 2 || !rvx::pow2(p.channels)) return false;
    if (p.stages >= 32 || (1u << p.stages) != p.channels) return false;
    if (p.decimation == 0 || p.decimation > p.channels || p.channels % p.decimation) return false;
    if (!rvx::pow2(p.out_ring_blocks) || p.out_ring_mask != p.out_ring_blocks - 1) return false;
    if (p.block_count > p.out_ring_blocks) return false;
    return true;
}

}  // namespace

extern "C" {

int32_t rvx_convert(rvx_ctx* ctx, int32_t format, const void* packed, uint64_t packed_bytes,
                    const rvx_convert_params* params, void* ring, uint64_t ring_samples) {
    static const char* who = "rvx_convert";
    RVX_ENTER(ctx, who);
    static const uint64_t bytes_per_sample[4] = {2, 2, 4, 6};
    if (params == nullptr || format < 0 || format > 3)
        return rvx::fail(RVX_EINVAL, "%s: unknown format %d", who, format);
    const rvx_convert_params p = *params;
    const uint64_t capacity = uint64_t(p.ring_mask) + 1;
    if ((capacity & p.ring_mask) != 0)
        return rvx::fail(RVX_EINVAL, "%s: ring_mask %u is not a power of two minus one", who,
                         p.ring_mask);
    if (ring_samples < capacity)
        return rvx::fail(RVX_EINVAL, "%s: ring holds %llu samples, mask implies %llu", who,
                         (unsigned long long)ring_samples, (unsigned long long)capacity);
    if (p.count > capacity)
        return rvx::fail(RVX_EINVAL, "%s: %u samples would alias a ring of %llu", who, p.count,
                         (unsigned long long)capacity);
    const uint64_t last = uint64_t(p.src_offset) + p.count;
    if (last > 0xFFFFFFFFull)
        return rvx::fail(RVX_EINVAL, "%s: src_offset + count overflows 32 bits", who);
    const uint64_t words = (last * bytes_per_sample[format] + 3) / 4;
    if (words > 0x100000000ull)
        return rvx::fail(RVX_EINVAL, "%s: last word index passes 2^32", who);
    if (packed_bytes < words * 4)
        return rvx::fail(RVX_EINVAL, "%s: reads %llu whole words, source holds %llu bytes", who,
                         (unsigned long long)words, (unsigned long long)packed_bytes);
    if (p.count == 0) return RVX_OK;
    if (packed == nullptr || ring == nullptr) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    const dim3 grid((p.count + 255) / 256), block(256);
    auto* w = static_cast<const uint32_t*>(packed);
    auto* r = static_cast<float2*>(ring);
    switch (format) {
        case 0: convert_kernel<0><<<grid, block, 0, ctx->stream>>>(w, p, r); break;
        case 1: convert_kernel<1><<<grid, block, 0, ctx->stream>>>(w, p, r); break;
        case 2: convert_kernel<2><<<grid, block, 0, ctx->stream>>>(w, p, r); break;
        default: convert_kernel<3><<<grid, block, 0, ctx->stream>>>(w, p, r); break;
    }
    return rvx::check_launch(who);
}

int32_t rvx_cmul(rvx_ctx* ctx, const void* lhs, const void* rhs, void* out, uint64_t count) {
    static const char* who = "rvx_cmul";
    RVX_ENTER(ctx, who);
    if (count == 0) return RVX_OK;
    if (!lhs || !rhs || !out) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    const uint64_t groups = (count + 255) / 256;
    if (groups > 0x7FFFFFFFull) return rvx::fail(RVX_EINVAL, "%s: too many samples", who);
    cmul_kernel<<<dim3(uint32_t(groups)), dim3(256), 0, ctx->stream>>>(
        static_cast<const float2*>(lhs), static_cast<const float2*>(rhs),
        static_cast<float2*>(out), count);
    return rvx::check_launch(who);
}

int32_t rvx_pfb_branch(rvx_ctx* ctx, const rvx_grid* grid, const float* taps, uint64_t n_taps,
                       const void* iq_ring, uint64_t ring_samples,
                       const rvx_pfb_branch_params* params, void* out, uint64_t out_samples) {
    static const char* who = "rvx_pfb_branch";
    RVX_ENTER(ctx, who);
    if (grid == nullptr || params == nullptr) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const uint32_t M = grid->channels, T = grid->taps_per_branch, D = grid->decimation;
    if (M < 2 || !rvx::pow2(M) || M > 65536 || T == 0 || D == 0 || M % D != 0 ||
        uint64_t(M) * T > (1u << 20))
        return rvx::fail(RVX_EINVAL, "%s: invalid grid M=%u taps=%u D=%u", who, M, T, D);
    if (n_taps != uint64_t(M) * T)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu prototype taps, got %llu", who,
                         (unsigned long long)(uint64_t(M) * T), (unsigned long long)n_taps);
    const rvx_pfb_branch_params p = *params;
    if (out_samples != uint64_t(p.block_count) * M)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu branch values, got %llu", who,
                         (unsigned long long)(uint64_t(p.block_count) * M),
                         (unsigned long long)out_samples);
    if (p.block_count == 0) return RVX_OK;
    const uint64_t capacity = uint64_t(p.ring_mask) + 1;
    if ((capacity & p.ring_mask) != 0 || ring_samples < capacity)
        return rvx::fail(RVX_EINVAL, "%s: ring mask %u / %llu samples", who, p.ring_mask,
                         (unsigned long long)ring_samples);
    if (uint64_t(M) * T + uint64_t(p.block_count - 1) * D > capacity)
        return rvx::fail(RVX_EINVAL, "%s: %u blocks need more history than the ring holds", who,
                         p.block_count);
    if (!taps || !iq_ring || !out) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    const uint64_t groups = (uint64_t(p.block_count) * M + 255) / 256;
    pfb_branch_kernel<<<dim3(uint32_t(groups)), dim3(256), 0, ctx->stream>>>(
        static_cast<const float2*>(iq_ring), taps, p, M, rvx::log2u(M), T, D,
        static_cast<float2*>(out));
    return rvx::check_launch(who);
}

int32_t rvx_pfb_fft(rvx_ctx* ctx, const rvx_pfb_fft_params* params, const void* branch,
                    uint64_t branch_samples, const void* twiddles, uint64_t n_twiddles,
                    void* channel_ring, uint64_t ring_samples) {
    static const char* who = "rvx_pfb_fft";
    RVX_ENTER(ctx, who);
    if (params == nullptr) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_pfb_fft_params p = *params;
    if (!fft_params_ok(p))
        return rvx::fail(RVX_EINVAL,
                         "%s: invalid params (M=%u D=%u stages=%u blocks=%u ring=%u mask=%u)", who,
                         p.channels, p.decimation, p.stages, p.block_count, p.out_ring_blocks,
                         p.out_ring_mask);
    if (p.stages > kMaxTransformLog2)
        return rvx::fail(RVX_EINVAL, "%s: %u channels exceeds the %u-point LDS transform", who,
                         p.channels, 1u << kMaxTransformLog2);
    if (n_twiddles != p.channels || branch_samples < uint64_t(p.block_count) * p.channels ||
        ring_samples < uint64_t(p.channels) * p.out_ring_blocks)
        return rvx::fail(RVX_EINVAL, "%s: buffer sizes do not match the params", who);
    if (p.block_count == 0) return RVX_OK;
    if (!branch || !twiddles || !channel_ring) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    if (int32_t rc = launch_transform<1>(p.stages, static_cast<const float2*>(branch),
                                         static_cast<const float2*>(twiddles),
                                         static_cast<float2*>(channel_ring), p.block_count, p,
                                         ctx->stream))
        return rc;
    return rvx::check_launch(who);
}

int32_t rvx_fft_radix2(rvx_ctx* ctx, const void* twiddles, uint64_t n_twiddles, const void* in,
                       void* out, uint32_t size, uint32_t batch) {
    static const char* who = "rvx_fft_radix2";
    RVX_ENTER(ctx, who);
    if (size < 2 || !rvx::pow2(size) || size > (1u << kMaxTransformLog2))
        return rvx::fail(RVX_EINVAL, "%s: size %u is not a power of two in [2, %u]", who, size,
                         1u << kMaxTransformLog2);
    if (n_twiddles != size) return rvx::fail(RVX_EINVAL, "%s: needs %u twiddles", who, size);
    if (batch == 0) return RVX_OK;
    if (!twiddles || !in || !out) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    if (in == out) return rvx::fail(RVX_EINVAL, "%s: input and output must not alias", who);
    rvx_pfb_fft_params unused{};
    if (int32_t rc = launch_transform<0>(rvx::log2u(size), static_cast<const float2*>(in),
                                         static_cast<const float2*>(twiddles),
                                         static_cast<float2*>(out), batch, unused, ctx->stream))
        return rc;
    return rvx::check_launch(who);
}

int32_t rvx_spectrum(rvx_ctx* ctx, const rvx_spectrum_params* params, const void* channel_ring,
                     uint64_t ring_samples, const void* twiddles, uint64_t n_twiddles,
                     const float* window, uint64_t n_window, float* frame, uint64_t n_frame) {
    static const char* who = "rvx_spectrum";
    RVX_ENTER(ctx, who);
    if (params == nullptr) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_spectrum_params p = *params;
    const uint32_t N = p.transform, M = p.channels;
    if (M < 2 || !rvx::pow2(M))
        return rvx::fail(RVX_EINVAL, "%s: channels %u not a power of two >= 2", who, M);
    if (N < 4 || !rvx::pow2(N) || N > (1u << kMaxTransformLog2))
        return rvx::fail(RVX_EINVAL, "%s: transform %u not a power of two in [4, %u]", who, N,
                         1u << kMaxTransformLog2);
    if (p.stages >= 32 || (1u << p.stages) != N)
        return rvx::fail(RVX_EINVAL, "%s: stages %u is not log2(%u)", who, p.stages, N);
    if (!rvx::pow2(p.chan_blocks) || p.chan_mask != p.chan_blocks - 1 || N > p.chan_blocks)
        return rvx::fail(RVX_EINVAL, "%s: channel ring geometry %u/%u", who, p.chan_blocks,
                         p.chan_mask);
    if (n_twiddles != N || n_window != uint64_t(N) + N / 2 ||
        ring_samples < uint64_t(M) * p.chan_blocks || n_frame < uint64_t(M) * (N / 2))
        return rvx::fail(RVX_EINVAL, "%s: buffer sizes do not match the params", who);
    if (!channel_ring || !twiddles || !window || !frame)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    auto go = [&](auto tag) {
        constexpr int L = decltype(tag)::value;
        using S = TransformShape<L>;
        spectrum_kernel<L><<<dim3((M + S::BPW - 1) / S::BPW), dim3(S::THREADS), 0, ctx->stream>>>(
            static_cast<const float2*>(channel_ring), static_cast<const float2*>(twiddles), window,
            frame, p);
    };
    switch (p.stages) {
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 3: go(std::integral_constant<int, 3>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        case 5: go(std::integral_constant<int, 5>{}); break;
        case 6: go(std::integral_constant<int, 6>{}); break;
        case 7: go(std::integral_constant<int, 7>{}); break;
        case 8: go(std::integral_constant<int, 8>{}); break;
        case 9: go(std::integral_constant<int, 9>{}); break;
        case 10: go(std::integral_constant<int, 10>{}); break;
        case 11: go(std::integral_constant<int, 11>{}); break;
        case 12: go(std::integral_constant<int, 12>{}); break;
        case 13: go(std::integral_constant<int, 13>{}); break;
        case 14: go(std::integral_constant<int, 14>{}); break;
        default: return rvx::fail(RVX_EINVAL, "%s: unsupported transform", who);
    }
    return rvx::check_launch(who);
}

int32_t rvx_spectrum_levels(rvx_ctx* ctx, const rvx_spectrum_levels_params* params,
                            const float* frame, uint64_t n_frame, float* levels) {
    static const char* who = "rvx_spectrum_levels";
    RVX_ENTER(ctx, who);
    if (params == nullptr) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_spectrum_levels_params p = *params;
    if (p.bins == 0) return rvx::fail(RVX_EINVAL, "%s: no bins", who);
    if (p.low_permille > 1000 || p.high_permille > 1000 || p.low_permille >= p.high_permille)
        return rvx::fail(RVX_EINVAL, "%s: permilles %u/%u", who, p.low_permille, p.high_permille);
    if (n_frame < p.bins) return rvx::fail(RVX_EINVAL, "%s: frame shorter than bins", who);
    if (!frame || !levels) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    levels_kernel<<<dim3(1), dim3(1024), 0, ctx->stream>>>(frame, p, levels);
    return rvx::check_launch(who);
}

static int32_t ema_valid(const rvx_ema_params& p, const char* who) {
    if (p.channels == 0) return rvx::fail(RVX_EINVAL, "%s: no channels", who);
    if (!rvx::pow2(p.chan_blocks) || p.chan_mask != p.chan_blocks - 1)
        return rvx::fail(RVX_EINVAL, "%s: channel ring geometry %u/%u", who, p.chan_blocks,
                         p.chan_mask);
    if (p.frame_blocks == 0 || uint64_t(p.frame_blocks) * p.frames > p.chan_blocks)
        return rvx::fail(RVX_EINVAL, "%s: %u frames of %u blocks do not fit %u ring blocks", who,
                         p.frames, p.frame_blocks, p.chan_blocks);
    return RVX_OK;
}

int32_t rvx_ema_partial(rvx_ctx* ctx, const rvx_ema_params* params, const void* channel_ring,
                        uint64_t ring_samples, float* partial, uint64_t n_partial) {
    static const char* who = "rvx_ema_partial";
    RVX_ENTER(ctx, who);
    if (params == nullptr) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_ema_params p = *params;
    if (int32_t rc = ema_valid(p, who)) return rc;
    if (ring_samples < uint64_t(p.channels) * p.chan_blocks ||
        n_partial < uint64_t(p.frames) * p.channels)
        return rvx::fail(RVX_EINVAL, "%s: buffer sizes do not match the params", who);
    if (p.frames == 0) return RVX_OK;
    if (!channel_ring || !partial) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    const uint64_t n = uint64_t(p.frames) * p.channels;
    ema_partial_kernel<<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, ctx->stream>>>(
        static_cast<const float2*>(channel_ring), p, partial);
    return rvx::check_launch(who);
}

int32_t rvx_ema_combine(rvx_ctx* ctx, const rvx_ema_params* params, const float* partial,
                        uint64_t n_partial, float* state, float* rows, uint64_t n_rows) {
    static const char* who = "rvx_ema_combine";
    RVX_ENTER(ctx, who);
    if (params == nullptr) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_ema_params p = *params;
    if (p.channels == 0) return rvx::fail(RVX_EINVAL, "%s: no channels", who);
    if (p.row_frames == 0 || p.frames % p.row_frames != 0)
        return rvx::fail(RVX_EINVAL, "%s: %u frames are not whole rows of %u", who, p.frames,
                         p.row_frames);
    const uint32_t nrows = p.frames / p.row_frames;
    if (n_partial < uint64_t(p.frames) * p.channels || rows == nullptr ||
        n_rows < uint64_t(nrows) * p.channels)
        return rvx::fail(RVX_EINVAL, "%s: buffer sizes do not match the params", who);
    if (!partial || !state) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    if (nrows == 0) return RVX_OK;
    const uint64_t n = uint64_t(nrows) * p.channels;
    // four channels a lane when rows of channels split into 16-byte runs
    const bool quad = p.channels % 4 == 0 && reinterpret_cast<uintptr_t>(partial) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(rows) % 16 == 0;
    const uint64_t lanes = quad ? n / 4 : n;
    ema_rows_kernel<<<dim3(uint32_t((lanes + 255) / 256)), dim3(256), 0, ctx->stream>>>(
        partial, p, rows, nrows, quad);
    ema_combine_kernel<<<dim3((p.channels + 63) / 64), dim3(64), 0, ctx->stream>>>(p, state, rows,
                                                                                  nrows);
    return rvx::check_launch(who);
}

}  // extern "C"
