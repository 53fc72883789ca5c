// The plain receive chain fused into one launch: native IQ codes -> branch
// filter -> M-point transform -> output phase correction -> channel ring, with
// the averaged power spectrum accumulated on chip.
//
// Bit-identical to convert (convert.cpp:153,235) -> pfb_branch
// (pfb_branch_reference.cpp:22) -> pfb_fft (pfb_fft_reference.cpp:143) ->
// rvx_ema_partial run one after another, which is how the reference records
// them (core/engine/graph.cpp:5120-5196): every value is computed by the same
// operations in the same order. What changes is where values live:
//   - the converted samples are never written: each tap converts its code on
//     the fly, exactly as convert would have;
//   - the branch outputs never leave the chip: each thread's branch sums are
//     the inputs of its first group of butterflies, so they go straight from
//     the tap loop into the first FFT pass in registers;
//   - the power average's state stays in registers across the frame.
//
// Work split. A workgroup owns a chunk of consecutive blocks (one averaging
// frame, or a single block) and walks it in order, so a frame's average
// accumulates in the order its definition fixes. Thread t of a transform owns
// V branches. Chunks are numbered so that the workgroups one XCD runs at the
// same time own neighbouring blocks: they share the input window in L2. A
// workgroup stages four blocks of outputs in LDS and writes each channel's
// four samples as one 32-byte run (the channel ring is channel-major, so a
// block's outputs alone would be 4096 scattered 8-byte writes).

#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "common.h"
#include "fft.h"

namespace {

using rvx::Fft;

struct ChainArgs {
    const void* input;
    const float* taps;
    const float2* tw;
    float2* ring;
    float* partial;
    uint64_t first_block;
    uint64_t input_origin;
    uint64_t input_samples;
    uint32_t decimation;
    uint32_t taps_per_branch;
    uint32_t block_count;
    uint32_t ring_mask;
    uint32_t ring_pitch;  // samples from one ring row to the next
    uint32_t chunk;   // blocks per workgroup (the averaging frame when averaging)
    uint32_t chunks;  // workgroups
    uint32_t average; // nonzero: write the frame partials
    float alpha;
    const uint32_t* rows;  // ring row of channel k (null: row k); rows >= ring_rows are not stored
    uint32_t ring_rows;
    // Front-end correction (corrected formats only): samples at window
    // offsets below corr_from take corr[0..3], the rest corr[4..7], each
    // {dc_i, dc_q, cross, scale}.
    uint32_t corr_from;
    float corr[8];
};

// Ring row of channel k, or ring_rows when the channel is not stored.
__device__ __forceinline__ uint32_t ring_row(const ChainArgs& a, uint32_t k) {
    const uint32_t row = a.rows ? a.rows[k] : k;
    return row < a.ring_rows ? row : a.ring_rows;
}

template <int LOGM, int LOGV_ = 2>
struct ChainShape {
    static constexpr int LOGV = LOGM < LOGV_ ? LOGM : LOGV_;
    using F = Fft<LOGM, LOGV>;
    static constexpr int THREADS = F::T >= 256 ? F::T : 256;
    static constexpr int BPW = THREADS / F::T;  // blocks in flight per workgroup
};

// cs16 codes read against taps that already carry the 2^-15 of the
// conversion (rvx_prescale_taps): a product code * (h * 2^-15) is the same
// real number as (code * 2^-15) * h, so it rounds to the same float, and the
// conversion's multiply leaves the tap loop.
constexpr int kCs16Scaled = 0x102;
// Codes converted and then corrected for the front end's DC and I/Q
// imbalance, exactly as rvx_iq_correct would correct the converted ring.
constexpr int kCu8Corrected = 0x200;
constexpr int kCs16Corrected = 0x202;
// Samples already converted (cf32 pairs): the engine's conditioned window,
// cleaned by impulse excision after conversion and correction.
constexpr int kCf32 = RVX_FORMAT_CF32;
__host__ __device__ constexpr bool is_cs16(int f) {
    return f == RVX_FORMAT_CS16 || f == kCs16Scaled || f == kCs16Corrected;
}
__host__ __device__ constexpr uint32_t bytes_per_sample(int f) {
    return f == kCf32 ? 8u : (is_cs16(f) ? 4u : 2u);
}
// What one sample's load returns: a native code in 32 bits, or both floats.
template <int FORMAT>
struct CodeOf {
    using type = uint32_t;
};
template <>
struct CodeOf<kCf32> {
    using type = uint64_t;
};
template <int FORMAT>
using code_t = typename CodeOf<FORMAT>::type;
__host__ __device__ constexpr bool is_corrected(int f) {
    return f == kCu8Corrected || f == kCs16Corrected;
}

// The native code of one sample (cs16: one word; cu8: two bytes).
template <int FORMAT>
__device__ __forceinline__ code_t<FORMAT> load_code(const void* input, uint32_t index) {
    if constexpr (FORMAT == kCf32)
        return static_cast<const uint64_t*>(input)[index];
    else if constexpr (is_cs16(FORMAT))
        return static_cast<const uint32_t*>(input)[index];
    else
        return static_cast<const uint16_t*>(input)[index];
}

// convert_cs16_cf32.comp / convert_cu8_cf32.comp on one code.
template <int FORMAT>
__device__ __forceinline__ float2 convert_code(code_t<FORMAT> w) {
    if constexpr (FORMAT == kCf32) {
        return make_float2(__uint_as_float(uint32_t(w)), __uint_as_float(uint32_t(w >> 32)));
    } else if constexpr (FORMAT == kCs16Corrected) {
        return convert_code<RVX_FORMAT_CS16>(w);
    } else if constexpr (FORMAT == kCu8Corrected) {
        return convert_code<RVX_FORMAT_CU8>(w);
    } else if constexpr (FORMAT == kCs16Scaled) {
        return make_float2(float(int32_t(w << 16) >> 16), float(int32_t(w) >> 16));
    } else if constexpr (FORMAT == RVX_FORMAT_CS16) {
        const float scale = __uint_as_float(0x38000000u);
        const float i = float(int32_t(w << 16) >> 16) * scale;
        const float q = float(int32_t(w) >> 16) * scale;
        return make_float2(i, q);
    } else {
        const float bias = __uint_as_float(0x42FF0000u);
        const float scale = __uint_as_float(0x3C008081u);
        const float i = (float(w & 0xFFu) - bias) * scale;
        const float q = (float((w >> 8) & 0xFFu) - bias) * scale;
        return make_float2(i, q);
    }
}

// Buffer loads: a uniform base and record count in a resource, a 32-bit
// per-lane offset and a uniform offset per tap. The tap's stride rides in
// the uniform offset, so a thread needs one offset register per branch
// instead of one address per (branch, tap), and reads past the resource
// return zero instead of faulting.
constexpr int kBufferDword3 = 0x00020000;  // raw dword access on gfx9-class parts

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint64_t bytes) {
    const uint32_t records = bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(bytes);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, records, kBufferDword3);
}

template <int FORMAT>
__device__ __forceinline__ code_t<FORMAT> buffer_code(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                      uint32_t soff) {
    if constexpr (FORMAT == kCf32) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        return uint64_t(v[0]) | (uint64_t(v[1]) << 32);
    } else if constexpr (is_cs16(FORMAT))
        return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    else
        return __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
}

template <int FORMAT>
__device__ __forceinline__ float2 load_sample(const void* input, uint32_t index) {
    return convert_code<FORMAT>(load_code<FORMAT>(input, index));
}

// i' = i - dc_i, then q' = cross * i' + scale * (q - dc_q), in that order.
__device__ __forceinline__ float2 apply_correction(float2 x, float dc_i, float dc_q, float cross,
                                                   float scale) {
    const float i = x.x - dc_i;
    const float q_dc = x.y - dc_q;
    const float q = cross * i + scale * q_dc;
    return make_float2(i, q);
}

// iq_correct.comp on one converted sample at window offset `index`
// (front_end_correction_reference.cpp:103-107): i' = i - dc_i, then
// q' = cross * i' + scale * (q - dc_q).
template <int FORMAT>
__device__ __forceinline__ float2 correct_sample(float2 x, uint32_t index, const ChainArgs& a) {
    if constexpr (!is_corrected(FORMAT)) {
        return x;
    } else {
        const bool before = index < a.corr_from;
        return apply_correction(x, before ? a.corr[0] : a.corr[4], before ? a.corr[1] : a.corr[5],
                                before ? a.corr[2] : a.corr[6], before ? a.corr[3] : a.corr[7]);
    }
}

// Taps whose sample loads are issued ahead of their use. The tap loop and
// the transform between blocks run while PREFETCH rows of codes are in
// flight, so an L2 round trip is paid once per PREFETCH taps rather than
// once per tap.
constexpr int kPrefetch = 4;

// Blocks of channel outputs staged in LDS before they are written: four
// blocks make a 32-byte run per channel, a whole memory sector.
constexpr int kStage = 4;

// Transform passes whose twiddles are read once per chunk rather than per
// block (pass 0's by scalar loads). Three of a radix-8 transform's passes fit
// beside the filter's registers without spilling and measured fastest at 4096
// points (all four: no faster; per block: 3% slower); the radix-4 shapes hold
// every pass's.
constexpr int kHoistPasses = 3;

// TAPS > 0: the tap count is a compile-time constant; REGTAPS then keeps the
// thread's V*TAPS taps in registers for the whole chunk, otherwise they are
// read (L1/L2-resident) alongside the samples, through the same prefetch.
// TAPS == 0: any tap count, read from memory tap by tap.
template <int LOGM, int LOGV_, int FORMAT, int TAPS, bool REGTAPS>
__global__ void __launch_bounds__((ChainShape<LOGM, LOGV_>::THREADS))
chain_kernel(ChainArgs a) {
    using S = ChainShape<LOGM, LOGV_>;
    using F = typename S::F;
    constexpr uint32_t M = F::N;
    constexpr int V = F::V;
    constexpr int T = F::T;
    constexpr int LOGT = F::LOGT;
    constexpr int LAST = F::PASSES - 1;
    constexpr int NT = TAPS > 0 ? TAPS : 1;

    // One transform's LDS per block in flight; with one block in flight, up
    // to kStage blocks of channel outputs are staged so that each channel's
    // outputs leave as one 32-byte run; with several, their powers are folded
    // in block order through LDS.
    constexpr int STAGE = S::BPW == 1 ? kStage : 1;
    constexpr uint32_t kLdsFloat2 =
        S::BPW * M + (STAGE > 1 ? STAGE * M : 0) + (S::BPW > 1 ? (S::BPW * M + 1) / 2 : 0);
    __shared__ float2 smem[kLdsFloat2];
    float2* lds = smem;
    float2* stage = smem + S::BPW * M;
    float* pow_lds = reinterpret_cast<float*>(smem + S::BPW * M);

    // XCD-aware chunk numbering: the hardware deals workgroups to the 8 XCDs
    // round robin, so workgroups w, w+8, w+16, ... share an XCD. Give each
    // XCD a contiguous run of chunks.
    // The grid is 8 * ceil(chunks / 8); the tail of the last XCD's run is
    // empty and those workgroups leave at once.
    const uint32_t w = blockIdx.x;
    const uint32_t per_xcd = gridDim.x / 8;
    const uint32_t chunk = (w % 8) * per_xcd + w / 8;
    if (chunk >= a.chunks) return;  // whole workgroup leaves together

    const uint32_t sub = threadIdx.x / T;
    const uint32_t t = threadIdx.x % T;
    float2* my = lds + sub * M;

    // The V branches this thread filters: slots t + T*i (i = 0..V-1), which
    // are exactly the inputs of the first-pass butterfly group that starts at
    // position bitrev(t) * V (slot s -> position bitrev_M(s)).
    uint32_t branch[V];
#pragma unroll
    for (int i = 0; i < V; ++i) branch[i] = (M - (t + uint32_t(T) * i)) & (M - 1);

    float c[V][REGTAPS ? NT : 1];
    if constexpr (TAPS > 0 && REGTAPS) {
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
            for (int q = 0; q < TAPS; ++q) c[i][q] = a.taps[q * M + branch[i]];
    }

    float ema[V];
#pragma unroll
    for (int r = 0; r < V; ++r) ema[r] = 0.0f;

    // The output correction reads w[(k * dm) mod M]; when dm is 0 or M/2 (a
    // decimation of M/2, every block) that is w[0] or w[M/2], read once here.
    const __attribute__((address_space(4))) float* tw_c =
        reinterpret_cast<const __attribute__((address_space(4))) float*>(uintptr_t(a.tw));
    const float2 tw_zero = make_float2(tw_c[0], tw_c[1]);
    const float2 tw_half = make_float2(tw_c[M], tw_c[M + 1]);

    // The transform's twiddles depend on the thread, not the block: the first
    // kHoist passes' stay in registers for the whole chunk, the rest are read
    // per block (registers do not hold them all beside the filter's).
    constexpr int kHoist =
        (F::LOG_V < 3 || kHoistPasses >= F::PASSES) ? F::PASSES : kHoistPasses;
    float2 wt[F::PASSES][F::WPP];
    F::template load_all_tw<0, kHoist>(a.tw, t, wt);

    const uint32_t first = chunk * a.chunk;
    constexpr int PF = (TAPS > 0 && TAPS < kPrefetch) ? TAPS : kPrefetch;
    auto block_base = [&](uint32_t b0) {
        const uint64_t m = a.first_block + first + b0 + sub;
        return uint32_t(m * a.decimation - a.input_origin);
    };
    auto block_active = [&](uint32_t b0) {
        return (b0 + sub < a.chunk) && (first + b0 + sub < a.block_count);
    };
    // Codes of the first PF taps of the next block, loaded while the current
    // block's transform runs.
    code_t<FORMAT> raw[PF][V];
    float ctap[REGTAPS ? 1 : PF][V];
    // The chunk's input window starts at the oldest sample its first block
    // reads; a sample's byte offset from there splits into a per-lane part
    // (block base minus branch, minus the oldest tap) and a per-tap uniform
    // part ((TAPS-1-q) rows of M samples).
    constexpr uint32_t BPS = bytes_per_sample(FORMAT);
    constexpr uint32_t NTAPS = TAPS > 0 ? TAPS : 1;
    const uint32_t window0 =
        uint32_t((a.first_block + first) * a.decimation - a.input_origin) - (NTAPS * M - 1);
    const __amdgpu_buffer_rsrc_t in_rsrc = make_rsrc(
        static_cast<const uint8_t*>(a.input) + uint64_t(window0) * BPS,
        (a.input_samples - window0) * BPS);
    const __amdgpu_buffer_rsrc_t tap_rsrc = make_rsrc(a.taps, uint64_t(NTAPS) * M * 4);
    auto issue = [&](int slot, int q, uint32_t base_) {
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const uint32_t voff = (base_ - window0 - branch[i] - (NTAPS - 1) * M) * BPS;
            raw[slot][i] = buffer_code<FORMAT>(in_rsrc, voff, uint32_t(NTAPS - 1 - q) * M * BPS);
            if constexpr (!REGTAPS)
                ctap[REGTAPS ? 0 : slot][i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                    tap_rsrc, branch[i] * 4u, uint32_t(q) * M * 4u, 0));
        }
    };
    if constexpr (TAPS > 0) {
        if (block_active(0)) {
#pragma unroll
            for (int q = 0; q < PF; ++q) issue(q, q, block_base(0));
        }
    }

    for (uint32_t b0 = 0; b0 < a.chunk; b0 += S::BPW) {
        const uint32_t rel = first + b0 + sub;           // block within the launch
        const bool active = block_active(b0);
        const uint64_t m = a.first_block + rel;           // absolute block index
        const uint32_t base = block_base(b0);

        // The twiddle loads stay inside the block loop (served by L1): hoisted
        // out of it they would take registers the prototype taps need.
        const float2* tw = a.tw;
        asm volatile("" : "+s"(tw));

        float2 v[V];
        if (active) {
            if constexpr (TAPS > 0) {
                float re[V], im[V];
#pragma unroll
                for (int i = 0; i < V; ++i) re[i] = im[i] = 0.0f;
                // Ascending q for every branch (the twin's order); the V
                // branches interleave, which no branch's sum can see.
#pragma unroll
                for (int q = 0; q < TAPS; ++q) {
                    code_t<FORMAT> cur[V];
                    float tap[V];
#pragma unroll
                    for (int i = 0; i < V; ++i) {
                        cur[i] = raw[q % PF][i];
                        if constexpr (REGTAPS)
                            tap[i] = c[i][REGTAPS ? q : 0];
                        else
                            tap[i] = ctap[REGTAPS ? 0 : q % PF][i];
                    }
                    if (q + PF < TAPS) issue(q % PF, q + PF, base);
#pragma unroll
                    for (int i = 0; i < V; ++i) {
                        const float2 x = correct_sample<FORMAT>(
                            convert_code<FORMAT>(cur[i]), base - branch[i] - uint32_t(q) * M, a);
                        re[i] = re[i] + x.x * tap[i];
                        im[i] = im[i] + x.y * tap[i];
                    }
                }
                // Register index of slot t + T*i within the group is bitrev(i).
#pragma unroll
                for (int i = 0; i < V; ++i)
                    v[rvx::bitrev<F::LOG_V>(i)] = make_float2(re[i], im[i]);
            } else {
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    float re = 0.0f, im = 0.0f;
                    const uint32_t start = base - branch[i];
                    for (uint32_t q = 0; q < a.taps_per_branch; ++q) {
                        const float2 x = correct_sample<FORMAT>(
                            load_sample<FORMAT>(a.input, start - q * M), start - q * M, a);
                        const float tap = a.taps[q * M + branch[i]];
                        re = re + x.x * tap;
                        im = im + x.y * tap;
                    }
                    v[rvx::bitrev<F::LOG_V>(i)] = make_float2(re, im);
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < V; ++i) v[i] = make_float2(0.0f, 0.0f);
        }
        // The transform's twiddles are read before the next block's codes are
        // requested: memory operations of a wave complete in issue order, so a
        // twiddle read issued after the prefetch would make the transform wait
        // for the prefetch as well. (The output correction's few reads come
        // after the transform, by which time the prefetch has landed.)
        F::template load_all_tw<kHoist>(tw, t, wt);
        if constexpr (TAPS > 0) {
            const uint32_t nb = b0 + S::BPW;
            if (nb < a.chunk && block_active(nb)) {
                const uint32_t next = block_base(nb);
#pragma unroll
                for (int q = 0; q < PF; ++q) issue(q, q, next);
            }
        }

        F::template run_pass_w<0>(v, wt[0]);
        if constexpr (F::PASSES > 1) {
            // Everyone has finished reading the previous block's last pass.
            __syncthreads();
            const uint32_t g0 = rvx::bitrev<LOGT>(t) * V;
#pragma unroll
            for (int r = 0; r < V; ++r) my[F::phys(g0 + r)] = v[r];
            __syncthreads();
            F::template load<1>(my, v, t);
            F::template run_pass_w<1>(v, wt[1]);
            F::template rest_w<2>(my, v, wt, t);
        }

        const uint32_t m_lo = uint32_t(m);
        const uint32_t slot = m_lo & a.ring_mask;
        const uint32_t dm = ((m_lo & (M - 1)) * a.decimation) & (M - 1);
        float x[V];
#pragma unroll
        for (int r = 0; r < V; ++r) {
            const uint32_t k = F::template pos<LAST>(t, r);
            float2 value = v[r];
            if (a.decimation != M) {
                float2 c;
                if ((dm & (M / 2 - 1)) == 0) {
                    const bool half = dm != 0 && (k & 1u) != 0u;
                    c.x = half ? tw_half.x : tw_zero.x;
                    c.y = half ? tw_half.y : tw_zero.y;
                } else {
                    c = tw[(k * dm) & (M - 1)];
                }
                value = rvx::cmul_exact(value, c);
            }
            if constexpr (STAGE > 1) {
                stage[(b0 % STAGE) * M + k] = value;
            } else {
                const uint32_t row = ring_row(a, k);
                if (active && row < a.ring_rows) a.ring[uint64_t(row) * a.ring_pitch + slot] = value;
            }
            x[r] = value.x * value.x + value.y * value.y;
        }
        if constexpr (STAGE > 1) {
            // Flush when the stage is full or the chunk ends. Every thread
            // flushes exactly the channels it staged, so no barrier is needed.
            const uint32_t last = min(a.chunk, a.block_count - first) - 1;
            if (b0 % STAGE == STAGE - 1 || b0 == last) {
                const uint32_t n = b0 % STAGE + 1;
                const uint32_t m0 = m_lo - (n - 1);
                const uint32_t s0 = m0 & a.ring_mask;
                const bool whole = n == STAGE && (s0 % STAGE) == 0 && STAGE == 4;
#pragma unroll
                for (int r = 0; r < V; ++r) {
                    const uint32_t k = F::template pos<LAST>(t, r);
                    const uint32_t ring_r = ring_row(a, k);
                    if (ring_r >= a.ring_rows) continue;
                    float2* row = a.ring + uint64_t(ring_r) * a.ring_pitch;
                    if (whole) {
                        const float2 v0 = stage[0 * M + k], v1 = stage[1 * M + k];
                        const float2 v2 = stage[2 * M + k], v3 = stage[3 * M + k];
                        float4* dst = reinterpret_cast<float4*>(row + s0);
                        dst[0] = make_float4(v0.x, v0.y, v1.x, v1.y);
                        dst[1] = make_float4(v2.x, v2.y, v3.x, v3.y);
                    } else {
                        for (uint32_t j = 0; j < n; ++j)
                            row[(m0 + j) & a.ring_mask] = stage[j * M + k];
                    }
                }
            }
        }
        if constexpr (S::BPW == 1) {
#pragma unroll
            for (int r = 0; r < V; ++r) ema[r] = ema[r] + (x[r] - ema[r]) * a.alpha;
        } else {
            // Blocks in flight finish together; fold their powers in block
            // order through LDS.
            __syncthreads();
#pragma unroll
            for (int r = 0; r < V; ++r) pow_lds[sub * M + F::template pos<LAST>(t, r)] = x[r];
            __syncthreads();
            if (sub == 0) {
                for (uint32_t bb = 0; bb < S::BPW; ++bb) {
                    if (b0 + bb >= a.chunk || first + b0 + bb >= a.block_count) break;
#pragma unroll
                    for (int r = 0; r < V; ++r) {
                        const float xr = pow_lds[bb * M + F::template pos<LAST>(t, r)];
                        ema[r] = ema[r] + (xr - ema[r]) * a.alpha;
                    }
                }
            }
        }
    }
    if (a.average && sub == 0) {
#pragma unroll
        for (int r = 0; r < V; ++r)
            a.partial[uint64_t(chunk) * M + F::template pos<LAST>(t, r)] = ema[r];
    }
}

// ---------------------------------------------------------------------------
// The two-times-oversampled grid (D = M/2) with the sample stream shared
// between a branch pair.
//
// At D = M/2 the sample branch r reads for tap q of block m is
// x[(m - 2q) D - r] and branch r + D reads x[(m - 2q - 1) D - r]: both walk the
// one sequence s_r[j] = x[j D - r], one step apart. A thread owns its V
// branches as V/2 such pairs (slots t + T i and t + T (i + V/2) are branches D
// apart), so for a chunk of C blocks it needs C + 2 TAPS - 1 samples per pair
// instead of 2 C TAPS: it loads and converts each once, newest first, and
// adds it into every (branch, block) sum whose tap reaches it. Streaming the
// samples newest first visits each sum's terms in ascending q, the twin's
// order, so every sum is the same chain of roundings as before.
//
// All C blocks' branch sums stay in registers; each block's transform then
// runs in place, so after the chunk a thread holds its channels' outputs of
// all C blocks and writes each channel's run straight from registers (no LDS
// staging). LDS holds one transform's exchange only.
// ---------------------------------------------------------------------------
template <int LOGM, int LOGV_, int FORMAT, int TAPS, int C, int SUBS>
__global__ void __launch_bounds__((ChainShape<LOGM, LOGV_>::THREADS))
chain_pair_kernel(ChainArgs a) {
    using S = ChainShape<LOGM, LOGV_>;
    using F = typename S::F;
    constexpr uint32_t M = F::N;
    constexpr uint32_t D = M / 2;
    constexpr int V = F::V;
    constexpr int P = V / 2;  // branch pairs per thread
    constexpr int T = F::T;
    constexpr int LOGT = F::LOGT;
    constexpr int LAST = F::PASSES - 1;
    constexpr int NS = C + 2 * TAPS - 1;  // samples per pair per sub-chunk
    constexpr uint32_t BPS = bytes_per_sample(FORMAT);

    __shared__ float2 smem[S::BPW * M];

    const uint32_t w = blockIdx.x;
    const uint32_t per_xcd = gridDim.x / 8;
    const uint32_t group = (w % 8) * per_xcd + w / 8;  // XCD-major numbering
    const uint32_t sub = threadIdx.x / T;
    const uint32_t t = threadIdx.x % T;
    const uint32_t chunk = group * S::BPW + sub;
    // A sub-transform past the last chunk still runs every barrier.
    const bool live = chunk < a.chunks;
    float2* my = smem + sub * M;

    uint32_t branch[V];
#pragma unroll
    for (int i = 0; i < V; ++i) branch[i] = (M - (t + uint32_t(T) * i)) & (M - 1);
    // Pair i is branches lo and lo + D. Slot t + T i holds lo only at t = 0,
    // i = 0 (branch 0); every other pair's first slot holds lo + D.
    uint32_t lo[P], voff[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        lo[i] = branch[i] < D ? branch[i] : branch[i + P];
        voff[i] = (D - 1u - lo[i]) * BPS;
    }
    const bool lo_first = t == 0;
    const __amdgpu_buffer_rsrc_t tap_rsrc = make_rsrc(a.taps, uint64_t(TAPS) * M * 4);

    float2 wt[F::PASSES][F::WPP];
    const __attribute__((address_space(4))) float* tw_c =
        reinterpret_cast<const __attribute__((address_space(4))) float*>(uintptr_t(a.tw));
    const float2 tw_zero = make_float2(tw_c[0], tw_c[1]);
    const float2 tw_half = make_float2(tw_c[M], tw_c[M + 1]);

    float ema[V];
#pragma unroll
    for (int r = 0; r < V; ++r) ema[r] = 0.0f;

    for (int sc = 0; sc < SUBS; ++sc) {
        const uint32_t first = chunk * (C * SUBS) + uint32_t(sc * C);  // block within the launch
        // Oldest sample of the sub-chunk: j_min D - (D - 1), j_min = m0 - 2 TAPS + 1.
        const uint32_t base0 = uint32_t((a.first_block + first) * a.decimation - a.input_origin);
        const uint32_t window0 = base0 - uint32_t(TAPS) * M + 1u;
        const __amdgpu_buffer_rsrc_t in_rsrc = make_rsrc(
            static_cast<const uint8_t*>(a.input) + uint64_t(window0) * BPS,
            live ? (a.input_samples - window0) * BPS : 0);

        // acc[i][b]: pair i's sums for block b, (x, y) of lo then of lo + D
        float4 acc[P][C];
#pragma unroll
        for (int i = 0; i < P; ++i)
#pragma unroll
            for (int b = 0; b < C; ++b) acc[i][b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

        // Each pair's samples in segments of SEG, a segment's codes and the taps
        // it first reaches loaded one segment ahead of its arithmetic. The
        // scheduling barriers stop later loads from being hoisted further, so
        // two segments of loads are in registers beside the sums.
        constexpr int SEG = 8;
        constexpr int NSEG = (NS + SEG - 1) / SEG;
        code_t<FORMAT> code[P][NS];
        float h_lo[P][TAPS], h_hi[P][TAPS];
        auto load_seg = [&](int i, int sg) {
#pragma unroll
            for (int n = sg * SEG; n < (sg + 1) * SEG && n < NS; ++n)
                code[i][n] = buffer_code<FORMAT>(in_rsrc, voff[i], uint32_t(NS - 1 - n) * D * BPS);
#pragma unroll
            for (int q = sg * (SEG / 2); q < (sg + 1) * (SEG / 2) && q < TAPS; ++q) {
                h_lo[i][q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                    tap_rsrc, lo[i] * 4u, uint32_t(q) * M * 4u, 0));
                h_hi[i][q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                    tap_rsrc, (lo[i] + D) * 4u, uint32_t(q) * M * 4u, 0));
            }
        };
        // Sample n of a segment is j = m0 + C - 1 - n; it meets block m0 + b of
        // branch lo at tap q = (b - C + 1 + n) / 2 and of lo + D at
        // q = (b - C + n) / 2 (whole, in range), first at n = 2q, so a tap is
        // loaded in segment q / (SEG/2) at the latest.
        auto run_seg = [&](int i, int sg) {
#pragma unroll
            for (int n = sg * SEG; n < (sg + 1) * SEG && n < NS; ++n) {
                const uint32_t idx = base0 + uint32_t(C - 1 - n) * D - lo[i];
                const float2 x =
                    correct_sample<FORMAT>(convert_code<FORMAT>(code[i][n]), idx, a);
#pragma unroll
                for (int b = 0; b < C; ++b) {
                    const int dl = b - C + 1 + n;
                    if (dl >= 0 && (dl & 1) == 0 && dl / 2 < TAPS) {
                        const float h = h_lo[i][dl / 2 < TAPS ? dl / 2 : 0];
                        acc[i][b].x = acc[i][b].x + x.x * h;
                        acc[i][b].y = acc[i][b].y + x.y * h;
                    }
                    const int dh = b - C + n;
                    if (dh >= 0 && (dh & 1) == 0 && dh / 2 < TAPS) {
                        const float h = h_hi[i][dh / 2 < TAPS ? dh / 2 : 0];
                        acc[i][b].z = acc[i][b].z + x.x * h;
                        acc[i][b].w = acc[i][b].w + x.y * h;
                    }
                }
            }
        };
        load_seg(0, 0);
#pragma unroll
        for (int g = 0; g < P * NSEG; ++g) {
            __builtin_amdgcn_sched_barrier(0);
            if (g + 1 < P * NSEG) load_seg((g + 1) / NSEG, (g + 1) % NSEG);
            run_seg(g / NSEG, g % NSEG);
        }
        __builtin_amdgcn_sched_barrier(0);

        // The transform's twiddles (pass 0's by scalar loads), read after the
        // filter so that they do not hold registers through it.
        F::template load_all_tw<0, F::PASSES>(a.tw, t, wt);
        float2 out[C][V];
#pragma unroll
        for (int b = 0; b < C; ++b) {
            float2 v[V];
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const float2 s_lo = make_float2(acc[i][b].x, acc[i][b].y);
                const float2 s_hi = make_float2(acc[i][b].z, acc[i][b].w);
                // slot t + T i is the pair's lo + D branch except at t = 0, i = 0
                v[rvx::bitrev<F::LOG_V>(i)] = (i == 0 && lo_first) ? s_lo : s_hi;
                v[rvx::bitrev<F::LOG_V>(i + P)] = (i == 0 && lo_first) ? s_hi : s_lo;
            }
            F::template run_pass_w<0>(v, wt[0]);
            if constexpr (F::PASSES > 1) {
                __syncthreads();  // the previous block's last pass has been read
                const uint32_t g0 = rvx::bitrev<LOGT>(t) * V;
#pragma unroll
                for (int r = 0; r < V; ++r) my[F::phys(g0 + r)] = v[r];
                __syncthreads();
                F::template load<1>(my, v, t);
                F::template run_pass_w<1>(v, wt[1]);
                F::template rest_w<2>(my, v, wt, t);
            }
            // Output correction: at D = M/2 block m turns channel k by
            // w[(k m D) mod M], which is w[M/2] for odd k of odd m and w[0] otherwise.
            const bool odd_block = (uint32_t(a.first_block + first + b) & 1u) != 0u;
#pragma unroll
            for (int r = 0; r < V; ++r) {
                const uint32_t k = F::template pos<LAST>(t, r);
                const bool half = odd_block && (k & 1u) != 0u;
                float2 c;
                c.x = half ? tw_half.x : tw_zero.x;
                c.y = half ? tw_half.y : tw_zero.y;
                out[b][r] = rvx::cmul_exact(v[r], c);
            }
        }

        // Powers into the frame's average, and each channel's run of C outputs.
        const uint32_t active =
            live && first < a.block_count ? min(uint32_t(C), a.block_count - first) : 0u;
        const uint32_t m0 = uint32_t(a.first_block + first);
        const uint32_t s0 = m0 & a.ring_mask;
        const bool whole = active == uint32_t(C) && (s0 % C) == 0 && (C % 2) == 0;
#pragma unroll
        for (int r = 0; r < V; ++r) {
            const uint32_t k = F::template pos<LAST>(t, r);
            if (a.average) {
#pragma unroll
                for (int b = 0; b < C; ++b) {
                    const float x = out[b][r].x * out[b][r].x + out[b][r].y * out[b][r].y;
                    ema[r] = ema[r] + (x - ema[r]) * a.alpha;
                }
            }
            const uint32_t ring_r = ring_row(a, k);
            if (ring_r >= a.ring_rows || active == 0) continue;
            float2* row = a.ring + uint64_t(ring_r) * a.ring_pitch;
            if (whole) {
#pragma unroll
                for (int b = 0; b < C; b += 2) {
                    float4* dst = reinterpret_cast<float4*>(row + s0 + b);
                    *dst = make_float4(out[b][r].x, out[b][r].y, out[b + 1][r].x,
                                       out[b + 1][r].y);
                }
            } else {
#pragma unroll
                for (int b = 0; b < C; ++b)
                    if (uint32_t(b) < active) row[(m0 + b) & a.ring_mask] = out[b][r];
            }
        }
    }
    if (a.average && live) {
#pragma unroll
        for (int r = 0; r < V; ++r)
            a.partial[uint64_t(chunk) * M + F::template pos<LAST>(t, r)] = ema[r];
    }
}

// Eight values per thread: radix-8 register passes (four for 4096 points).
// The 512-point grid runs radix-4 passes: two transforms per workgroup at
// half the registers per thread, 1.6x faster than radix-8 there.
constexpr int kChainLogV = 3;

// Blocks per chunk of the pair-streaming kernel (and its averaging frame),
// filtered kPairSub blocks at a time.
constexpr uint32_t kPairChunk = 4;
constexpr int kPairSub = 4;

// Whether a grid runs on chain_pair_kernel: D = M/2 with a compile-time tap
// count, and an averaging frame (if any) of one chunk.
bool pair_eligible(uint32_t M, uint32_t taps, uint32_t D, uint32_t frame) {
    return M >= 4 && 2 * D == M && (taps == 16 || taps == 17) &&
           (frame == 0 || frame == kPairChunk);
}

template <int LOGM, int FORMAT>
void launch_chain_pair(const ChainArgs& a, uint32_t taps, hipStream_t stream) {
    constexpr int kLogV = LOGM == 9 ? 2 : kChainLogV;
    using S = ChainShape<LOGM, kLogV>;
    const uint32_t groups = (a.chunks + S::BPW - 1) / S::BPW;
    const dim3 grid(8 * ((groups + 7) / 8)), block(S::THREADS);
    constexpr int C = kPairSub, SUBS = kPairChunk / kPairSub;
    if (taps == 16)
        chain_pair_kernel<LOGM, kLogV, FORMAT, 16, C, SUBS><<<grid, block, 0, stream>>>(a);
    else
        chain_pair_kernel<LOGM, kLogV, FORMAT, 17, C, SUBS><<<grid, block, 0, stream>>>(a);
}

template <int LOGM, int FORMAT>
void launch_chain_m(const ChainArgs& a, uint32_t taps, hipStream_t stream) {
    if constexpr (LOGM >= 2) {
        if (pair_eligible(1u << LOGM, taps, a.decimation, a.average ? a.chunk : 0)) {
            launch_chain_pair<LOGM, FORMAT>(a, taps, stream);
            return;
        }
    }
    // Radix-8 grids from 1024 channels up keep their taps in memory: in
    // registers they take half the register file and starve the prefetch. The
    // 512-channel grid's radix-4 threads filter four branches, whose 64 taps
    // fit at the same occupancy and save re-reading the prototype every block
    // (10% faster there).
    constexpr bool kRegTaps = LOGM <= 9;
    constexpr int kLogV = LOGM == 9 ? 2 : kChainLogV;
    using S = ChainShape<LOGM, kLogV>;
    const dim3 grid(8 * ((a.chunks + 7) / 8)), block(S::THREADS);
    if (taps == 16)
        chain_kernel<LOGM, kLogV, FORMAT, 16, kRegTaps><<<grid, block, 0, stream>>>(a);
    else if (taps == 17)
        chain_kernel<LOGM, kLogV, FORMAT, 17, kRegTaps><<<grid, block, 0, stream>>>(a);
    else
        chain_kernel<LOGM, kLogV, FORMAT, 0, false><<<grid, block, 0, stream>>>(a);
}

template <int FORMAT>
int32_t launch_chain(uint32_t logm, const ChainArgs& a, uint32_t taps, hipStream_t stream) {
    switch (logm) {
        case 1: launch_chain_m<1, FORMAT>(a, taps, stream); break;
        case 2: launch_chain_m<2, FORMAT>(a, taps, stream); break;
        case 3: launch_chain_m<3, FORMAT>(a, taps, stream); break;
        case 4: launch_chain_m<4, FORMAT>(a, taps, stream); break;
        case 5: launch_chain_m<5, FORMAT>(a, taps, stream); break;
        case 6: launch_chain_m<6, FORMAT>(a, taps, stream); break;
        case 7: launch_chain_m<7, FORMAT>(a, taps, stream); break;
        case 8: launch_chain_m<8, FORMAT>(a, taps, stream); break;
        case 9: launch_chain_m<9, FORMAT>(a, taps, stream); break;
        case 10: launch_chain_m<10, FORMAT>(a, taps, stream); break;
        case 11: launch_chain_m<11, FORMAT>(a, taps, stream); break;
        case 12: launch_chain_m<12, FORMAT>(a, taps, stream); break;
        default: return rvx::fail(RVX_EINVAL, "rvx_chain: 2^%u channels is not supported", logm);
    }
    return RVX_OK;
}

// Conditioning: samples [first, first + count) of a code window converted
// (and corrected, when `correct`) into a cf32 ring at slot (absolute & mask),
// by the same operations the chain applies to each code it reads.
template <int FORMAT>
__global__ void __launch_bounds__(256)
convert_ring_kernel(const void* __restrict__ codes, uint64_t first_off, uint64_t first_abs,
                    uint32_t count, float2* __restrict__ ring, uint32_t mask, float4 corr,
                    int correct) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float2 x = convert_code<FORMAT>(load_code<FORMAT>(codes, uint32_t(first_off + i)));
    if (correct) x = apply_correction(x, corr.x, corr.y, corr.z, corr.w);
    ring[uint32_t(first_abs + i) & mask] = x;
}

// Ring slots of samples [first_abs, first_abs + count) copied to a linear
// window at window_off.
__global__ void __launch_bounds__(256)
ring_to_window_kernel(const float2* __restrict__ ring, uint32_t mask, uint64_t first_abs,
                      uint32_t count, float2* __restrict__ window, uint64_t window_off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    window[window_off + i] = ring[uint32_t(first_abs + i) & mask];
}

}  // namespace

extern "C" int32_t rvx_chain(rvx_ctx* ctx, const rvx_chain_desc* desc, const void* input,
                             const float* taps, uint64_t n_taps, const void* twiddles,
                             uint64_t n_twiddles, void* channel_ring, uint64_t ring_samples,
                             float* partial, uint64_t n_partial) {
    static const char* who = "rvx_chain";
    RVX_ENTER(ctx, who);
    if (desc == nullptr) return rvx::fail(RVX_EINVAL, "%s: null descriptor", who);
    const rvx_chain_desc d = *desc;
    const uint32_t M = d.channels, T = d.taps_per_branch, D = d.decimation;
    if (d.taps_scaled > 1 || (d.taps_scaled && d.format != RVX_FORMAT_CS16))
        return rvx::fail(RVX_EINVAL, "%s: prescaled taps belong to cs16 input", who);
    if (d.format == RVX_FORMAT_CF32 && d.correction)
        return rvx::fail(RVX_EINVAL, "%s: cf32 input arrives corrected (the correction applies "
                         "where samples are converted)", who);
    if (d.correction && d.taps_scaled)
        return rvx::fail(RVX_EINVAL,
                         "%s: a front-end correction needs the converted samples, not prescaled "
                         "taps", who);
    if (d.format != RVX_FORMAT_CU8 && d.format != RVX_FORMAT_CS16 && d.format != RVX_FORMAT_CF32)
        return rvx::fail(RVX_EINVAL, "%s: format %d (the fused chain reads cu8, cs16 or cf32)",
                         who, d.format);
    if (M < 2 || !rvx::pow2(M) || M > 4096 || T == 0 || D == 0 || M % D != 0 ||
        uint64_t(M) * T > (1u << 20))
        return rvx::fail(RVX_EINVAL, "%s: invalid grid M=%u taps=%u D=%u (M <= 4096)", who, M, T,
                         D);
    if (n_taps != uint64_t(M) * T || n_twiddles != M)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu taps and %u twiddles", who,
                         (unsigned long long)(uint64_t(M) * T), M);
    if (d.block_count == 0) return RVX_OK;
    if (!rvx::pow2(d.out_ring_blocks) || d.block_count > d.out_ring_blocks)
        return rvx::fail(RVX_EINVAL,
                         "%s: %u blocks into a %u-block ring (power of two, >= blocks)", who,
                         d.block_count, d.out_ring_blocks);
    const uint32_t rows = d.channel_rows ? d.ring_rows : M;
    if (d.channel_rows && (d.ring_rows == 0 || d.ring_rows > M))
        return rvx::fail(RVX_EINVAL, "%s: %u ring rows for %u channels", who, d.ring_rows, M);
    const uint32_t pitch = d.ring_pitch ? d.ring_pitch : d.out_ring_blocks;
    if (pitch < d.out_ring_blocks || (d.ring_pitch && d.ring_pitch % 4))
        return rvx::fail(RVX_EINVAL, "%s: ring pitch %u (at least %u blocks, a multiple of 4)", who,
                         d.ring_pitch, d.out_ring_blocks);
    const uint64_t ring_need = uint64_t(rows - 1) * pitch + d.out_ring_blocks;
    if (ring_samples < ring_need)
        return rvx::fail(RVX_EINVAL, "%s: channel ring holds %llu samples, needs %llu", who,
                         (unsigned long long)ring_samples, (unsigned long long)ring_need);
    const uint64_t N = uint64_t(M) * T;
    // byte offsets into the input are 32-bit
    const uint64_t max_in = d.format == RVX_FORMAT_CF32 ? (uint64_t(1) << 29) : (uint64_t(1) << 31);
    if (d.input_samples >= max_in)
        return rvx::fail(RVX_EINVAL, "%s: %llu input samples exceed %llu", who,
                         (unsigned long long)d.input_samples, (unsigned long long)max_in);
    const uint64_t first_base = d.first_block * D;
    const uint64_t last_base = (d.first_block + d.block_count - 1) * D;
    if (first_base < d.input_origin + (N - 1))
        return rvx::fail(RVX_EINVAL, "%s: block %llu needs samples before the input starts", who,
                         (unsigned long long)d.first_block);
    if (last_base >= d.input_origin + d.input_samples)
        return rvx::fail(RVX_EINVAL, "%s: the last block needs samples past the input's end", who);
    // Four blocks per workgroup unless averaging (then one frame): the XCD
    // numbering hands an XCD's concurrent workgroups consecutive chunks, and
    // each writes a 32-byte run per channel, so the channel ring's 128-byte
    // lines (16 blocks of one channel) fill from a few workgroups while they
    // are still in that XCD's L2.
    uint32_t chunk = pair_eligible(M, T, D, d.frame_blocks) ? kPairChunk : 4;
    if (d.frame_blocks != 0) {
        if (d.block_count % d.frame_blocks)
            return rvx::fail(RVX_EINVAL, "%s: frame of %u blocks does not divide %u blocks", who,
                             d.frame_blocks, d.block_count);
        if (n_partial < uint64_t(d.block_count / d.frame_blocks) * M || partial == nullptr)
            return rvx::fail(RVX_EINVAL, "%s: partial output too small", who);
        chunk = d.frame_blocks;
    }
    if (!input || !taps || !twiddles || !channel_ring)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);

    ChainArgs a{};
    a.input = input;
    a.taps = taps;
    a.tw = static_cast<const float2*>(twiddles);
    a.ring = static_cast<float2*>(channel_ring);
    a.partial = partial;
    a.first_block = d.first_block;
    a.input_origin = d.input_origin;
    a.input_samples = d.input_samples;
    a.decimation = D;
    a.taps_per_branch = T;
    a.block_count = d.block_count;
    a.ring_mask = d.out_ring_blocks - 1;
    a.ring_pitch = pitch;
    a.chunk = chunk;
    a.chunks = (d.block_count + chunk - 1) / chunk;
    a.average = d.frame_blocks != 0;
    a.alpha = d.alpha;
    a.rows = d.channel_rows;
    a.ring_rows = rows;
    if (d.correction) {
        const rvx_iq_correct_params& now = *d.correction;
        const rvx_iq_correct_params& before = d.correction_before ? *d.correction_before : now;
        const float vals[8] = {before.dc_i, before.dc_q, before.cross, before.scale,
                               now.dc_i,    now.dc_q,    now.cross,    now.scale};
        for (int k = 0; k < 8; ++k) {
            if (!std::isfinite(vals[k]))
                return rvx::fail(RVX_EINVAL, "%s: a correction coefficient is not finite", who);
            a.corr[k] = vals[k];
        }
        // window offset of the first sample that takes `correction`
        const uint64_t from = d.correction_from <= d.input_origin ? 0
                              : d.correction_from - d.input_origin;
        a.corr_from = uint32_t(std::min<uint64_t>(from, d.input_samples));
    }
    const uint32_t logm = rvx::log2u(M);
    int32_t rc;
    if (d.correction)
        rc = d.format == RVX_FORMAT_CU8 ? launch_chain<kCu8Corrected>(logm, a, T, ctx->stream)
                                        : launch_chain<kCs16Corrected>(logm, a, T, ctx->stream);
    else if (d.format == RVX_FORMAT_CF32)
        rc = launch_chain<kCf32>(logm, a, T, ctx->stream);
    else
        rc = d.format == RVX_FORMAT_CU8 ? launch_chain<RVX_FORMAT_CU8>(logm, a, T, ctx->stream)
             : d.taps_scaled            ? launch_chain<kCs16Scaled>(logm, a, T, ctx->stream)
                                        : launch_chain<RVX_FORMAT_CS16>(logm, a, T, ctx->stream);
    if (rc) return rc;
    return rvx::check_launch(who);
}

namespace rvx {

int32_t convert_to_ring(rvx_ctx* ctx, int32_t format, const void* codes, uint64_t code_samples,
                        uint64_t first_off, uint64_t first_abs, uint32_t count,
                        const rvx_iq_correct_params* correction, void* ring, uint32_t ring_mask,
                        const char* who) {
    if (count == 0) return RVX_OK;
    if (first_off + count > code_samples || first_off + count > (uint64_t(1) << 32))
        return rvx::fail(RVX_EINVAL, "%s: conversion past the window", who);
    if (count > uint64_t(ring_mask) + 1)
        return rvx::fail(RVX_EINVAL, "%s: %u samples into a %llu-sample ring", who, count,
                         (unsigned long long)ring_mask + 1);
    const float4 c = correction ? make_float4(correction->dc_i, correction->dc_q,
                                              correction->cross, correction->scale)
                                : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    const dim3 grid((count + 255) / 256), block(256);
    float2* r = static_cast<float2*>(ring);
    if (format == RVX_FORMAT_CS16)
        convert_ring_kernel<RVX_FORMAT_CS16><<<grid, block, 0, ctx->stream>>>(
            codes, first_off, first_abs, count, r, ring_mask, c, correction ? 1 : 0);
    else if (format == RVX_FORMAT_CU8)
        convert_ring_kernel<RVX_FORMAT_CU8><<<grid, block, 0, ctx->stream>>>(
            codes, first_off, first_abs, count, r, ring_mask, c, correction ? 1 : 0);
    else
        return rvx::fail(RVX_EINVAL, "%s: format %d", who, format);
    return rvx::check_launch(who);
}

int32_t ring_to_window(rvx_ctx* ctx, const void* ring, uint32_t ring_mask, uint64_t first_abs,
                       uint32_t count, void* window, uint64_t window_samples, uint64_t window_off,
                       const char* who) {
    if (count == 0) return RVX_OK;
    if (window_off + count > window_samples || count > uint64_t(ring_mask) + 1)
        return rvx::fail(RVX_EINVAL, "%s: copy past the window or the ring", who);
    ring_to_window_kernel<<<dim3((count + 255) / 256), dim3(256), 0, ctx->stream>>>(
        static_cast<const float2*>(ring), ring_mask, first_abs, count,
        static_cast<float2*>(window), window_off);
    return rvx::check_launch(who);
}

}  // namespace rvx
