// The 4096-channel chain with its transform on the matrix cores.
//
// Block m: branch sums u_r = sum_q h[qM + r] * x[mD - qM - r] (fp32, VALU:
// every branch has its own taps, so the FIR has no operand a GEMM could
// share), y at slot (M - r) mod M, then the 4096-point DFT factored 64 x 64
// (n = 64 n1 + n2, k = k1 + 64 k2):
//     Z[k1][n2] = sum_n1 y[64 n1 + n2] e^{-j2pi k1 n1 / 64}       (GEMM 1)
//     Z'        = Z * e^{-j2pi k1 n2 / 4096}                         (twiddle)
//     X[k1 + 64 k2] = sum_n2 Z'[k1][n2] e^{-j2pi k2 n2 / 64}         (GEMM 2)
// and the output correction W^{k (mD mod M)} of pfb_fft.comp. Each complex
// GEMM is one real 128 x 128 x 64 product ([[C, S], [-S, C]] times [re; im]).
// PREC 0: split bf16 -- a = a_hi + a_lo, each product a_hi b_hi + a_hi b_lo +
// a_lo b_hi on v_mfma_f32_16x16x16_bf16; PREC 1: fp16 operands (inputs scaled
// by 2^8 to keep weak signals normal) on v_mfma_f32_16x16x16_f16. Both
// accumulate in fp32. This form is not bit-exact with the twins (its
// precision is measured against FP64); the plain chain is the exact one.
//
// A workgroup (4 waves) owns a frame of blocks. Wave w computes output row
// tiles w (real parts of k1 or k2 in 16w..16w+15) and w + 4 (their imaginary
// parts) for all 4 column tiles, so every complex value it needs for the
// twiddle and the correction is in its own registers.

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kM = 4096;
constexpr int kThreads = 256;
constexpr int kKS = 128 + 8;  // padded K stride of the transposed B operand (uint16 entries)
constexpr float kHalfScale = 256.0f;  // fp16: inputs times 2^8, outputs times 2^-8

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ uint16_t f2h(float f) {
    const _Float16 h = _Float16(f);
    return *reinterpret_cast<const uint16_t*>(&h);
}

// One B operand entry: split bf16 (hi, lo), or fp16 of v * scale.
template <int PREC>
__device__ __forceinline__ void put(uint16_t* hi, uint16_t* lo, int idx, float v, float scale) {
    if constexpr (PREC == 0) {
        const uint16_t h = f2bf(v);
        hi[idx] = h;
        lo[idx] = f2bf(v - bf2f(h));
    } else {
        hi[idx] = f2h(v * scale);
    }
}

struct MfmaArgs {
    const void* input;
    const float* taps;
    const float2* tw;       // W_4096^j
    const uint16_t* dft;    // A = [[C, S], [-S, C]] 128 x 128: hi then lo (bf16) or fp16
    float2* ring;
    float* partial;
    uint64_t first_block;
    uint64_t input_origin;
    uint64_t input_samples;
    uint32_t decimation;
    uint32_t taps_per_branch;
    uint32_t block_count;
    uint32_t ring_mask;
    uint32_t ring_blocks;
    uint32_t chunk;
    uint32_t chunks;
    uint32_t average;
    float alpha;
};

template <int PREC>
__device__ __forceinline__ float4_t mma(const uint16_t* a_row, const uint16_t* b, float4_t acc,
                                        int grp, int row, int tn, int ks) {
    // a_row: this lane's A row (its M index); b: the transposed B operand,
    // b[n * kKS + k]. Both reads are four consecutive 16-bit entries.
    const int k0 = ks * 16 + grp * 4;
    const uint2 araw = *reinterpret_cast<const uint2*>(a_row + k0);
    const uint2 braw = *reinterpret_cast<const uint2*>(b + (tn * 16 + row) * kKS + k0);
    if constexpr (PREC == 0) {
        const short4_t av = __builtin_bit_cast(short4_t, araw);
        const short4_t bv = __builtin_bit_cast(short4_t, braw);
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    } else {
        const half4_t av = __builtin_bit_cast(half4_t, araw);
        const half4_t bv = __builtin_bit_cast(half4_t, braw);
        return __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv, acc, 0, 0, 0);
    }
}

// C (8 row tiles x 4 col tiles, this wave's rows w and w + 4) = A * B.
template <int PREC>
__device__ __forceinline__ void gemm(const uint16_t* ahi, const uint16_t* alo, const uint16_t* bhi,
                                     const uint16_t* blo, float4_t (&acc)[2][4], int wave,
                                     int lane) {
    const int row = lane % 16, grp = lane / 16;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int tn = 0; tn < 4; ++tn) acc[h][tn] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int m = (wave + 4 * h) * 16 + row;  // A row of this lane
        const uint16_t* ah = ahi + m * 128;
        const uint16_t* al = alo + m * 128;
#pragma unroll 2
        for (int ks = 0; ks < 8; ++ks) {
#pragma unroll
            for (int tn = 0; tn < 4; ++tn) {
                acc[h][tn] = mma<PREC>(ah, bhi, acc[h][tn], grp, row, tn, ks);
                if constexpr (PREC == 0) {
                    acc[h][tn] = mma<PREC>(ah, blo, acc[h][tn], grp, row, tn, ks);
                    acc[h][tn] = mma<PREC>(al, bhi, acc[h][tn], grp, row, tn, ks);
                }
            }
        }
    }
}

template <int PREC, int TAPS>
__global__ void __launch_bounds__(kThreads) chain_mfma_kernel(MfmaArgs a) {
    constexpr int NA = PREC == 0 ? 2 : 1;
    __shared__ __align__(16) uint16_t amat[NA][128 * 128];
    __shared__ __align__(16) uint16_t bhi[64 * kKS];
    __shared__ __align__(16) uint16_t blo[PREC == 0 ? 64 * kKS : 4];
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int row = lane % 16, grp = lane / 16;

    const uint32_t w = blockIdx.x, per_xcd = gridDim.x / 8;
    const uint32_t chunk = (w % 8) * per_xcd + w / 8;  // XCD-aware, as the plain chain
    if (chunk >= a.chunks) return;

    for (int e = tid; e < NA * 128 * 128; e += kThreads) amat[e / (128 * 128)][e % (128 * 128)] = a.dft[e];
    const uint16_t* ahi = amat[0];
    const uint16_t* alo = amat[NA - 1];

    const uint32_t first = chunk * a.chunk;
    const uint32_t M = kM, D = a.decimation, L = a.taps_per_branch;
    float ema[4][4];  // this lane's 16 channels
#pragma unroll
    for (int tn = 0; tn < 4; ++tn)
#pragma unroll
        for (int i = 0; i < 4; ++i) ema[tn][i] = 0.f;
    const float in_scale = PREC == 1 ? kHalfScale : 1.0f;
    const float out_scale = PREC == 1 ? 1.0f / kHalfScale : 1.0f;

    for (uint32_t b0 = 0; b0 < a.chunk; ++b0) {
        if (first + b0 >= a.block_count) break;
        const uint64_t m = a.first_block + first + b0;
        const uint32_t base = uint32_t(m * D - a.input_origin);
        __syncthreads();  // the previous block's B reads are done
        // FIR: branches r = tid + 256 j. With a compile-time tap count every
        // load of two branches is issued before their sums start.
        const uint32_t* in = static_cast<const uint32_t*>(a.input);
        auto emit = [&](uint32_t r, float ur, float ui) {
            const uint32_t s = (M - r) & (M - 1);
            const int n1 = int(s >> 6), n2 = int(s & 63);
            // B1[k = n1 | 64 + n1][n = n2], stored transposed
            put<PREC>(bhi, blo, n2 * kKS + n1, ur, in_scale);
            put<PREC>(bhi, blo, n2 * kKS + 64 + n1, ui, in_scale);
        };
        if constexpr (TAPS > 0) {
            for (int j = 0; j < 16; j += 2) {
                uint32_t code[2][TAPS];
                float h[2][TAPS];
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const uint32_t r = uint32_t(tid) + 256u * (j + jj);
#pragma unroll
                    for (int q = 0; q < TAPS; ++q) {
                        code[jj][q] = in[base - uint32_t(q) * M - r];
                        h[jj][q] = a.taps[uint32_t(q) * M + r];
                    }
                }
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    float ur = 0.f, ui = 0.f;
#pragma unroll
                    for (int q = 0; q < TAPS; ++q) {
                        const float xr = float(int32_t(code[jj][q] << 16) >> 16);
                        const float xi = float(int32_t(code[jj][q]) >> 16);
                        const float pr = xr * h[jj][q], pi = xi * h[jj][q];
                        ur = ur + pr;
                        ui = ui + pi;
                    }
                    // 2^-15 once per branch (exact: a power of two)
                    emit(uint32_t(tid) + 256u * (j + jj), ur * __uint_as_float(0x38000000u),
                         ui * __uint_as_float(0x38000000u));
                }
            }
        } else {
            for (int j = 0; j < 16; ++j) {
                const uint32_t r = uint32_t(tid) + 256u * j;
                float ur = 0.f, ui = 0.f;
                for (uint32_t q = 0; q < L; ++q) {
                    const uint32_t code = in[base - q * M - r];
                    const float xr = float(int32_t(code << 16) >> 16);
                    const float xi = float(int32_t(code) >> 16);
                    const float h = a.taps[q * M + r];
                    const float pr = xr * h, pi = xi * h;
                    ur = ur + pr;
                    ui = ui + pi;
                }
                emit(r, ur * __uint_as_float(0x38000000u), ui * __uint_as_float(0x38000000u));
            }
        }
        __syncthreads();
        float4_t acc[2][4];
        gemm<PREC>(ahi, alo, bhi, blo, acc, wave, lane);
        __syncthreads();  // every wave has read B
        // twiddle Z[k1][n2] by W_4096^(k1 n2); lane holds k1 = 16 wave + 4 grp + i,
        // n2 = 16 tn + row, re in acc[0], im in acc[1]
#pragma unroll
        for (int tn = 0; tn < 4; ++tn) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k1 = wave * 16 + grp * 4 + i, n2 = tn * 16 + row;
                const float2 t = a.tw[(k1 * n2) & (kM - 1)];
                const float zr = acc[0][tn][i], zi = acc[1][tn][i];  // fp16: already x 2^8
                const float vr = zr * t.x - zi * t.y;
                const float vi = zr * t.y + zi * t.x;
                // B2[k = n2 | 64 + n2][n = k1], stored transposed
                put<PREC>(bhi, blo, k1 * kKS + n2, vr, 1.0f);
                put<PREC>(bhi, blo, k1 * kKS + 64 + n2, vi, 1.0f);
            }
        }
        __syncthreads();
        gemm<PREC>(ahi, alo, bhi, blo, acc, wave, lane);
        // X[k2][k1]: lane holds k2 = 16 wave + 4 grp + i, k1 = 16 tn + row
        const uint32_t dm = uint32_t((m & (M - 1)) * D) & (M - 1);
        const uint32_t slot = uint32_t(m) & a.ring_mask;
        const bool active = true;
#pragma unroll
        for (int tn = 0; tn < 4; ++tn) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t k2 = uint32_t(wave * 16 + grp * 4 + i), k1 = uint32_t(tn * 16 + row);
                const uint32_t k = k1 + 64u * k2;
                const float xr = acc[0][tn][i] * out_scale, xi = acc[1][tn][i] * out_scale;
                float2 v = make_float2(xr, xi);
                if (D != M) {
                    const float2 c = a.tw[(k * dm) & (M - 1)];
                    v = make_float2(xr * c.x - xi * c.y, xr * c.y + xi * c.x);
                }
                if (active) a.ring[uint64_t(k) * a.ring_blocks + slot] = v;
                const float p = v.x * v.x + v.y * v.y;
                ema[tn][i] = ema[tn][i] + (p - ema[tn][i]) * a.alpha;
            }
        }
    }
    if (a.average) {
#pragma unroll
        for (int tn = 0; tn < 4; ++tn)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t k2 = uint32_t(wave * 16 + grp * 4 + i), k1 = uint32_t(tn * 16 + row);
                a.partial[uint64_t(chunk) * kM + k1 + 64u * k2] = ema[tn][i];
            }
    }
}

// The DFT64 operand table: A[i][j], i < 64 real rows, j < 64 real inputs.
std::vector<uint16_t> dft_table(int prec) {
    const double two_pi = 6.283185307179586476925286766559;
    std::vector<uint16_t> out(size_t(prec == 0 ? 2 : 1) * 128 * 128);
    auto bf = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return uint16_t(u >> 16);
    };
    auto bff = [](uint16_t b) {
        uint32_t u = uint32_t(b) << 16;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    };
    for (int i = 0; i < 128; ++i)
        for (int j = 0; j < 128; ++j) {
            const int k = i % 64, n = j % 64;
            const double ang = two_pi * double((k * n) % 64) / 64.0;
            const double c = std::cos(ang), s = std::sin(ang);
            // [[C, S], [-S, C]]
            const double v = i < 64 ? (j < 64 ? c : s) : (j < 64 ? -s : c);
            const float f = float(v);
            if (prec == 0) {
                const uint16_t h = bf(f);
                out[size_t(i) * 128 + j] = h;
                out[size_t(128 * 128) + size_t(i) * 128 + j] = bf(float(v - double(bff(h))));
            } else {
                const _Float16 hv = _Float16(f);
                std::memcpy(&out[size_t(i) * 128 + j], &hv, 2);
            }
        }
    return out;
}

}  // namespace

extern "C" int32_t rvx_chain_mfma(rvx_ctx* ctx, const rvx_chain_desc* desc, int32_t precision,
                                  const void* input, const float* taps, uint64_t n_taps,
                                  const void* twiddles, uint64_t n_twiddles, void* channel_ring,
                                  uint64_t ring_samples, float* partial, uint64_t n_partial) {
    static const char* who = "rvx_chain_mfma";
    RVX_ENTER(ctx, who);
    if (!desc) return rvx::fail(RVX_EINVAL, "%s: null descriptor", who);
    const rvx_chain_desc d = *desc;
    if (precision != RVX_PREC_BF16X3 && precision != RVX_PREC_FP16)
        return rvx::fail(RVX_EINVAL, "%s: precision %d", who, precision);
    if (d.format != RVX_FORMAT_CS16)
        return rvx::fail(RVX_EINVAL, "%s: the matrix-core chain reads cs16", who);
    if (d.correction)
        return rvx::fail(RVX_EINVAL, "%s: no front-end correction in the matrix-core chain", who);
    if (d.channels != uint32_t(kM) || d.decimation == 0 || kM % d.decimation != 0 ||
        d.taps_per_branch == 0 || d.taps_per_branch > 64)
        return rvx::fail(RVX_EINVAL, "%s: grid M=%u L=%u D=%u (M = 4096)", who, d.channels,
                         d.taps_per_branch, d.decimation);
    if (d.channel_rows || d.taps_scaled)
        return rvx::fail(RVX_EINVAL, "%s: channel rows and prescaled taps are not supported", who);
    if (n_taps != uint64_t(kM) * d.taps_per_branch || n_twiddles != uint64_t(kM))
        return rvx::fail(RVX_EINVAL, "%s: wants %llu taps and 4096 twiddles", who,
                         (unsigned long long)(uint64_t(kM) * d.taps_per_branch));
    if (d.block_count == 0) return RVX_OK;
    if (!rvx::pow2(d.out_ring_blocks) || d.block_count > d.out_ring_blocks ||
        ring_samples < uint64_t(kM) * d.out_ring_blocks)
        return rvx::fail(RVX_EINVAL, "%s: ring of %u blocks for %u blocks", who, d.out_ring_blocks,
                         d.block_count);
    if (d.input_samples >= (uint64_t(1) << 31))
        return rvx::fail(RVX_EINVAL, "%s: input exceeds 2^31 samples", who);
    const uint64_t N = uint64_t(kM) * d.taps_per_branch;
    if (d.first_block * d.decimation < d.input_origin + (N - 1) ||
        (d.first_block + d.block_count - 1) * d.decimation >= d.input_origin + d.input_samples)
        return rvx::fail(RVX_EINVAL, "%s: the blocks reach outside the input", who);
    uint32_t chunk = 4;
    if (d.frame_blocks) {
        if (d.block_count % d.frame_blocks || !partial ||
            n_partial < uint64_t(d.block_count / d.frame_blocks) * kM)
            return rvx::fail(RVX_EINVAL, "%s: frames do not fit", who);
        chunk = d.frame_blocks;
    }
    if (!input || !taps || !twiddles || !channel_ring)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    const int prec = precision == RVX_PREC_BF16X3 ? 0 : 1;
    if (!ctx->dft[prec]) {
        const std::vector<uint16_t> t = dft_table(prec);
        RVX_TRY_HIP(hipMalloc(&ctx->dft[prec], t.size() * 2), who);
        RVX_TRY_HIP(hipMemcpy(ctx->dft[prec], t.data(), t.size() * 2, hipMemcpyHostToDevice), who);
    }
    MfmaArgs a{};
    a.input = input;
    a.taps = taps;
    a.tw = static_cast<const float2*>(twiddles);
    a.dft = static_cast<const uint16_t*>(ctx->dft[prec]);
    a.ring = static_cast<float2*>(channel_ring);
    a.partial = partial;
    a.first_block = d.first_block;
    a.input_origin = d.input_origin;
    a.input_samples = d.input_samples;
    a.decimation = d.decimation;
    a.taps_per_branch = d.taps_per_branch;
    a.block_count = d.block_count;
    a.ring_mask = d.out_ring_blocks - 1;
    a.ring_blocks = d.out_ring_blocks;
    a.chunk = chunk;
    a.chunks = (d.block_count + chunk - 1) / chunk;
    a.average = d.frame_blocks != 0;
    a.alpha = d.alpha;
    const dim3 grid(8 * ((a.chunks + 7) / 8)), block(kThreads);
    const bool t16 = d.taps_per_branch == 16;
    if (prec == 0 && t16)
        chain_mfma_kernel<0, 16><<<grid, block, 0, ctx->stream>>>(a);
    else if (prec == 0)
        chain_mfma_kernel<0, 0><<<grid, block, 0, ctx->stream>>>(a);
    else if (t16)
        chain_mfma_kernel<1, 16><<<grid, block, 0, ctx->stream>>>(a);
    else
        chain_mfma_kernel<1, 0><<<grid, block, 0, ctx->stream>>>(a);
    return rvx::check_launch(who);
}

namespace {

// float -> OCP e4m3 (bias 7, max 448, no infinities), round to nearest even,
// saturating; NaN -> 0x7F.
__device__ __forceinline__ uint8_t to_e4m3(float v) {
    const uint32_t bits = __float_as_uint(v);
    const uint32_t sign = (bits >> 24) & 0x80u;
    const uint32_t u = bits & 0x7FFFFFFFu;
    if (u > 0x7F800000u) return uint8_t(sign | 0x7Fu);
    if (u >= 0x43E00000u) return uint8_t(sign | 0x7Eu);  // >= 448
    if (u < 0x3C800000u) {                                // below 2^-6: subnormal steps of 2^-9
        const float q = __uint_as_float(u) * 512.0f;      // exact
        return uint8_t(sign | uint32_t(__builtin_rintf(q)));
    }
    uint32_t e = (u >> 23) - 120u;
    uint32_t m = (u >> 20) & 7u;
    const uint32_t rem = u & 0xFFFFFu;
    if (rem > 0x80000u || (rem == 0x80000u && (m & 1u))) ++m;
    if (m == 8u) {
        m = 0;
        ++e;
    }
    if (e > 15u || (e == 15u && m == 7u)) return uint8_t(sign | 0x7Eu);
    return uint8_t(sign | (e << 3) | m);
}

__global__ void spectrum_fp8_kernel(const float* __restrict__ p, uint32_t n,
                                    uint8_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float db = 10.0f * log10f(fmaxf(p[i], 1e-20f));
    out[i] = to_e4m3(db * 0.25f);
}

}  // namespace

extern "C" int32_t rvx_spectrum_fp8(rvx_ctx* ctx, const float* power, uint64_t n, uint8_t* out) {
    static const char* who = "rvx_spectrum_fp8";
    RVX_ENTER(ctx, who);
    if (n == 0) return RVX_OK;
    if (!power || !out || n > 0xFFFFFFFFull) return rvx::fail(RVX_EINVAL, "%s: bad buffers", who);
    spectrum_fp8_kernel<<<dim3(uint32_t((n + 255) / 256)), dim3(256), 0, ctx->stream>>>(
        power, uint32_t(n), out);
    return rvx::check_launch(who);
}
