// The 4096-channel chain of the twice-oversampled grid (D = 2048, 16 taps per
// branch, cs16 input) with its transform on the matrix cores: the fast form of
// rvx_chain, not bit-exact with the twins (its error is measured against FP64).
//
// Block m: branch sums u_r = sum_q h[qM + r] x[mD - qM - r] on VALU in fp32 with
// fused multiply-adds (every branch has its own taps, so the filter has no
// operand a GEMM could share), u_r at slot s = (M - r) mod M, then the
// 4096-point DFT as two 64-point DFT-GEMMs (n = 64 n1 + n2, k = k1 + 64 k2):
//     Z[k1][n2]     = sum_n1 y[64 n1 + n2] W64^(k1 n1)      (GEMM 1)
//     Z'[k1][n2]    = Z[k1][n2] W4096^(k1 n2)               (twiddle, VALU)
//     X[k1 + 64 k2] = sum_n2 Z'[k1][n2] W64^(k2 n2)         (GEMM 2)
// and the output correction W^(k (mD mod M)) of pfb_fft.comp, which at D = M/2
// negates the odd channels of odd blocks. Each complex GEMM is one real
// product: rows are the output index's real then imaginary parts, K the 64
// inputs' real and imaginary parts, [[C, S], [-S, C]] times [re; im]. Operands
// are split bf16 (a = a_hi + a_lo, each product a_hi b_hi + a_hi b_lo +
// a_lo b_hi) on v_mfma_f32_16x16x32_bf16 with fp32 accumulation.
//
// Work split. A workgroup (8 waves) walks chunks of 4 consecutive blocks.
// Filter: thread t = 64 a + n2 owns the four branch pairs (r, r + D) whose
// slots are 64 n1 + n2 and 64 (n1 + 32) + n2 for n1 in {2a, 2a+1, 2a+16,
// 2a+17}. At D = M/2 a pair walks one sample sequence x[jD - r], one step
// apart, and blocks m and m + 2 share 15 of a branch's 16 samples, so the
// thread loads each of a pair's 35 samples of the chunk once and adds it into
// every (branch, block) sum its taps reach. The sums of the chunk's four
// blocks go to LDS as GEMM 1's B operand (columns: input index n2 and block),
// split into hi and lo as they are stored. GEMMs: wave w takes output rows
// 16 (w & 3) .. + 15 (real and imaginary) and half of the columns of all four
// blocks; its A fragments (the DFT matrix) are the same for both GEMMs, read
// once per chunk. GEMM 1's outputs are twiddled in registers and stored as
// GEMM 2's B operand; GEMM 2 leaves each lane eight channels' outputs for all
// four blocks, so the averaged power runs in block order in registers and
// each channel's four outputs leave as one 32-byte store.
//
// K order. K is a summation index, so both GEMMs may order it as they like
// provided A and B agree; both use kappa -> (input index 32 h + 16 e + c,
// component p) with p = kappa & 1, e = (kappa >> 1) & 1, c = (kappa >> 2) & 15,
// h = kappa >> 6. Under it a filter thread's 16 values of a block fill two
// 16-byte runs of its column, and a GEMM 1 lane's twiddled values for one
// (k1, block) fill one 8-byte run of GEMM 2's column, and one A table serves
// both GEMMs.

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kM = 4096;
constexpr uint32_t kD = 2048;
constexpr int kTaps = 16;
constexpr int kC = 4;                     // blocks per chunk (the averaging frame)
constexpr int kThreads = 512;             // 8 waves
constexpr int kNS = kC + 2 * kTaps - 1;   // samples per branch pair per chunk
constexpr int kCols = 64 * kC;            // B operand columns: (index, block)
constexpr int kRS = 128 + 8;              // LDS stride of a column, bf16 (padded off bank conflicts)
constexpr int kFrags = 16;                // A fragments per lane: 2 row tiles x 4 K steps x hi/lo
constexpr int kBufferDword3 = 0x00020000; // raw dword access on gfx9-class parts

struct FastArgs {
    const void* input;
    const float* taps;
    const float2* tw;     // W_4096^j
    const uint4* dft;     // A fragments: [w & 3][16][64 lanes] x 8 bf16
    float2* ring;
    float* partial;
    uint64_t first_block;
    uint64_t input_origin;
    uint64_t input_samples;
    uint32_t ring_mask;
    uint32_t ring_pitch;  // samples from one ring row to the next
    uint32_t chunks;
    uint32_t chunks_per_xcd;
    uint32_t average;
    float alpha;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint64_t bytes) {
    const uint32_t records = bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(bytes);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, records, kBufferDword3);
}

// bf16 hi and lo of two floats, packed two to a word each.
__device__ __forceinline__ void split2(float x0, float x1, uint32_t& hi, uint32_t& lo) {
    const __bf16 h0 = __bf16(x0), h1 = __bf16(x1);
    const __bf16 l0 = __bf16(x0 - float(h0)), l1 = __bf16(x1 - float(h1));
    hi = uint32_t(__builtin_bit_cast(uint16_t, h0)) | (uint32_t(__builtin_bit_cast(uint16_t, h1)) << 16);
    lo = uint32_t(__builtin_bit_cast(uint16_t, l0)) | (uint32_t(__builtin_bit_cast(uint16_t, l1)) << 16);
}

__device__ __forceinline__ f32x4 mma(uint4 a, uint4 b, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
}

// out[blk][e][mt] (rows 16 (w & 3) + 64 mt .. + 15, columns 64 blk + 16 (2 nh + e) .. + 15)
// = A x B, B read from the operand columns in LDS, three products per K step.
__device__ __forceinline__ void gemm(const uint16_t* bhi, const uint16_t* blo,
                                     const uint4 (&af)[2][4][2], f32x4 (&out)[kC][2][2], int nh,
                                     int g, int c) {
#pragma unroll
    for (int blk = 0; blk < kC; ++blk) {
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) out[blk][e][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            uint4 bh[2], bl[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int col = 64 * blk + 16 * (2 * nh + e) + c;
                const int at = col * kRS + 32 * ks + 8 * g;
                bh[e] = *reinterpret_cast<const uint4*>(bhi + at);
                bl[e] = *reinterpret_cast<const uint4*>(blo + at);
            }
            // the small terms first, into four independent accumulators
#pragma unroll
            for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) out[blk][e][mt] = mma(af[mt][ks][1], bh[e], out[blk][e][mt]);
#pragma unroll
            for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) out[blk][e][mt] = mma(af[mt][ks][0], bl[e], out[blk][e][mt]);
#pragma unroll
            for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) out[blk][e][mt] = mma(af[mt][ks][0], bh[e], out[blk][e][mt]);
        }
    }
}

__global__ void __launch_bounds__(kThreads) chain_fast_kernel(FastArgs a) {
    __shared__ __align__(16) uint16_t ops_hi[kCols * kRS];
    __shared__ __align__(16) uint16_t ops_lo[kCols * kRS];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15, mp = wave & 3, nh = wave >> 2;

    // Filter roles: pairs p = 0..3 at n1 = 2 wave + (p & 1) + 16 (p >> 1) and
    // n1 + 32, input index n2 = lane. Branch b < D of a pair sits at the
    // n1 + 32 slot, b + D at the n1 slot, except b = 0 (slot 0).
    uint32_t br[4], voff[4];
    bool zero_slot[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t n1 = 2u * wave + (p & 1) + 16u * (p >> 1);
        const uint32_t s = 64u * n1 + uint32_t(lane);
        br[p] = (kD - s) & (kD - 1);
        zero_slot[p] = s == 0;
        voff[p] = (kD - 1 - br[p]) * 4u;
    }
    const __amdgpu_buffer_rsrc_t tap_rsrc = make_rsrc(a.taps, uint64_t(kTaps) * kM * 4);

    // A chunk's outputs are stored during the next chunk's GEMMs, which wait
    // on LDS only: loads and stores share one memory counter, so stores issued
    // at the end of a chunk would hold up the next chunk's sample loads.
    float4 held[2][4][2];
    float held_ema[2][4];
    uint32_t held_chunk = 0;
    bool have_held = false;
    auto store_held = [&]() {
        const uint64_t hm0 = a.first_block + uint64_t(held_chunk) * kC;
        const uint32_t s0 = uint32_t(hm0) & a.ring_mask;
        const bool whole = (s0 % kC) == 0 && a.ring_mask >= uint32_t(kC - 1);
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t k = uint32_t(32 * nh + 16 * e + c) + 64u * uint32_t(16 * mp + 4 * g + i);
                float2* row = a.ring + uint64_t(k) * a.ring_pitch;
                const float4 lo = held[e][i][0], hi = held[e][i][1];
                if (whole) {
                    *reinterpret_cast<float4*>(row + s0) = lo;
                    *reinterpret_cast<float4*>(row + s0 + 2) = hi;
                } else {
                    row[(uint32_t(hm0) + 0) & a.ring_mask] = make_float2(lo.x, lo.y);
                    row[(uint32_t(hm0) + 1) & a.ring_mask] = make_float2(lo.z, lo.w);
                    row[(uint32_t(hm0) + 2) & a.ring_mask] = make_float2(hi.x, hi.y);
                    row[(uint32_t(hm0) + 3) & a.ring_mask] = make_float2(hi.z, hi.w);
                }
                if (a.average) a.partial[uint64_t(held_chunk) * kM + k] = held_ema[e][i];
            }
    };

    const uint32_t xcd = blockIdx.x % 8, per_xcd = gridDim.x / 8;
    for (uint32_t ci = blockIdx.x / 8; ci < a.chunks_per_xcd; ci += per_xcd) {
        const uint32_t chunk = xcd * a.chunks_per_xcd + ci;
        if (chunk >= a.chunks) break;  // uniform across the workgroup
        const uint64_t m0 = a.first_block + uint64_t(chunk) * kC;
        // Oldest sample of the chunk: (m0 - 2 kTaps + 1) D - (D - 1).
        const uint64_t window0 = m0 * kD - a.input_origin - (uint64_t(kTaps) * kM - 1);
        const __amdgpu_buffer_rsrc_t in_rsrc =
            make_rsrc(static_cast<const uint8_t*>(a.input) + window0 * 4,
                      (a.input_samples - window0) * 4);

        // acc[p][blk]: branch b's (re, im), then branch b + D's.
        float4 acc[4][kC];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int blk = 0; blk < kC; ++blk) acc[p][blk] = make_float4(0.f, 0.f, 0.f, 0.f);

        // Sample n (newest first) of pair p is x[(m0 + kC - 1 - n) D - b]; it
        // meets block m0 + blk of branch b at tap (blk - kC + 1 + n) / 2 and of
        // b + D at (blk - kC + n) / 2 (whole, in range). Samples stream in
        // segments of 8 with the next segment's loads in flight; a tap is
        // loaded in the segment that first reaches it.
        constexpr int SEG = 8, NSEG = (kNS + SEG - 1) / SEG;
        uint32_t code[4][kNS];
        float hb[4][kTaps], hd[4][kTaps];
        auto load_seg = [&](int p, int sg) {
#pragma unroll
            for (int n = sg * SEG; n < (sg + 1) * SEG && n < kNS; ++n)
                code[p][n] = __builtin_amdgcn_raw_buffer_load_b32(in_rsrc, voff[p],
                                                                  uint32_t(kNS - 1 - n) * kD * 4u, 0);
#pragma unroll
            for (int q = sg * (SEG / 2); q < (sg + 1) * (SEG / 2) && q < kTaps; ++q) {
                hb[p][q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                    tap_rsrc, br[p] * 4u, uint32_t(q) * kM * 4u, 0));
                hd[p][q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                    tap_rsrc, (br[p] + kD) * 4u, uint32_t(q) * kM * 4u, 0));
            }
        };
        auto run_seg = [&](int p, int sg) {
#pragma unroll
            for (int n = sg * SEG; n < (sg + 1) * SEG && n < kNS; ++n) {
                const float xr = float(int32_t(code[p][n] << 16) >> 16);
                const float xi = float(int32_t(code[p][n]) >> 16);
#pragma unroll
                for (int blk = 0; blk < kC; ++blk) {
                    const int dl = blk - kC + 1 + n;
                    if (dl >= 0 && (dl & 1) == 0 && dl / 2 < kTaps) {
                        const float h = hb[p][dl / 2 < kTaps ? dl / 2 : 0];
                        acc[p][blk].x = __builtin_fmaf(xr, h, acc[p][blk].x);
                        acc[p][blk].y = __builtin_fmaf(xi, h, acc[p][blk].y);
                    }
                    const int dh = blk - kC + n;
                    if (dh >= 0 && (dh & 1) == 0 && dh / 2 < kTaps) {
                        const float h = hd[p][dh / 2 < kTaps ? dh / 2 : 0];
                        acc[p][blk].z = __builtin_fmaf(xr, h, acc[p][blk].z);
                        acc[p][blk].w = __builtin_fmaf(xi, h, acc[p][blk].w);
                    }
                }
            }
        };
        load_seg(0, 0);
#pragma unroll
        for (int s = 0; s < 4 * NSEG; ++s) {
            __builtin_amdgcn_sched_barrier(0);
            if (s + 1 < 4 * NSEG) load_seg((s + 1) / NSEG, (s + 1) % NSEG);
            run_seg(s / NSEG, s % NSEG);
        }
        __builtin_amdgcn_sched_barrier(0);

        __syncthreads();  // the previous chunk's GEMM 2 has read the operands
        // GEMM 1's B operand: column 64 blk + n2, kappa = 8 wave + 4 (p & 1) +
        // 2 (p >> 1) + component (+ 64 for the n1 + 32 slot). The codes'
        // 2^-15 is applied here (exact).
        const float scale = __uint_as_float(0x38000000u);
#pragma unroll
        for (int blk = 0; blk < kC; ++blk) {
            float v0[8], v1[8];  // the n1 slots, then the n1 + 32 slots
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const float4 s = acc[p][blk];
                const int j = 4 * (p & 1) + 2 * (p >> 1);
                const float br_x = s.x * scale, br_y = s.y * scale;  // branch b
                const float bd_x = s.z * scale, bd_y = s.w * scale;  // branch b + D
                v0[j] = zero_slot[p] ? br_x : bd_x;
                v0[j + 1] = zero_slot[p] ? br_y : bd_y;
                v1[j] = zero_slot[p] ? bd_x : br_x;
                v1[j + 1] = zero_slot[p] ? bd_y : br_y;
            }
            uint4 h0, l0, h1, l1;
            split2(v0[0], v0[1], h0.x, l0.x);
            split2(v0[2], v0[3], h0.y, l0.y);
            split2(v0[4], v0[5], h0.z, l0.z);
            split2(v0[6], v0[7], h0.w, l0.w);
            split2(v1[0], v1[1], h1.x, l1.x);
            split2(v1[2], v1[3], h1.y, l1.y);
            split2(v1[4], v1[5], h1.z, l1.z);
            split2(v1[6], v1[7], h1.w, l1.w);
            const int at = (64 * blk + lane) * kRS + 8 * wave;
            *reinterpret_cast<uint4*>(ops_hi + at) = h0;
            *reinterpret_cast<uint4*>(ops_lo + at) = l0;
            *reinterpret_cast<uint4*>(ops_hi + at + 64) = h1;
            *reinterpret_cast<uint4*>(ops_lo + at + 64) = l1;
        }
        // This wave's A fragments and GEMM 1's twiddles, in flight across the barrier.
        uint4 af[2][4][2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                for (int hl = 0; hl < 2; ++hl)
                    af[mt][ks][hl] = a.dft[(mp * kFrags + mt * 8 + ks * 2 + hl) * 64 + lane];
        float2 tw[2][4];  // W4096^(k1 n2), k1 = 16 mp + 4 g + i, n2 = 32 nh + 16 e + c
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                tw[e][i] = a.tw[(uint32_t(16 * mp + 4 * g + i) * uint32_t(32 * nh + 16 * e + c)) &
                                (kM - 1)];
        if (have_held) store_held();
        __syncthreads();

        f32x4 z[kC][2][2];
        gemm(ops_hi, ops_lo, af, z, nh, g, c);
        __syncthreads();  // every wave has read GEMM 1's operand
        // Z' = Z W4096^(k1 n2) into GEMM 2's B operand: column 64 blk + k1,
        // kappa = 4 (c + 16 nh) + 2 e + component.
#pragma unroll
        for (int blk = 0; blk < kC; ++blk)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v[4];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const float zr = z[blk][e][0][i], zi = z[blk][e][1][i];
                    const float2 t = tw[e][i];
                    v[2 * e] = __builtin_fmaf(zr, t.x, -(zi * t.y));
                    v[2 * e + 1] = __builtin_fmaf(zr, t.y, zi * t.x);
                }
                uint2 h, l;
                split2(v[0], v[1], h.x, l.x);
                split2(v[2], v[3], h.y, l.y);
                const int at = (64 * blk + 16 * mp + 4 * g + i) * kRS + 4 * (c + 16 * nh);
                *reinterpret_cast<uint2*>(ops_hi + at) = h;
                *reinterpret_cast<uint2*>(ops_lo + at) = l;
            }
        __syncthreads();
        f32x4 x[kC][2][2];
        gemm(ops_hi, ops_lo, af, x, nh, g, c);

        // Lane holds X[k1 + 64 k2][blk] for k2 = 16 mp + 4 g + i, k1 = 32 nh + 16 e + c:
        // corrected, averaged in block order, and held for the next pass.
        const bool odd_m0 = (m0 & 1u) != 0u;
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float ema = 0.f;
                float2 o[kC];
#pragma unroll
                for (int blk = 0; blk < kC; ++blk) {
                    // W^(k (m D mod M)) = (-1)^k for odd m, 1 for even m
                    const bool neg = ((odd_m0 != ((blk & 1) != 0)) && (c & 1));
                    const float xr = x[blk][e][0][i], xi = x[blk][e][1][i];
                    o[blk] = neg ? make_float2(-xr, -xi) : make_float2(xr, xi);
                    const float pw = __builtin_fmaf(o[blk].x, o[blk].x, o[blk].y * o[blk].y);
                    ema = __builtin_fmaf(pw - ema, a.alpha, ema);
                }
                held[e][i][0] = make_float4(o[0].x, o[0].y, o[1].x, o[1].y);
                held[e][i][1] = make_float4(o[2].x, o[2].y, o[3].x, o[3].y);
                held_ema[e][i] = ema;
            }
        held_chunk = chunk;
        have_held = true;
    }
    if (have_held) store_held();
}

// A fragments of the DFT64 real form [[C, S], [-S, C]] under the K order
// above: entry (w & 3, mt, ks, hl, lane, j) is row 16 (w & 3) + 64 mt +
// lane % 16, kappa = 32 ks + 8 (lane / 16) + j, hi or lo bf16 part.
std::vector<uint16_t> dft_fragments() {
    const double two_pi = 6.283185307179586476925286766559;
    std::vector<uint16_t> out(size_t(4) * kFrags * 64 * 8);
    auto bf = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return uint16_t(u >> 16);
    };
    auto bff = [](uint16_t b) {
        const uint32_t u = uint32_t(b) << 16;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    };
    for (int mp = 0; mp < 4; ++mp)
        for (int mt = 0; mt < 2; ++mt)
            for (int ks = 0; ks < 4; ++ks)
                for (int hl = 0; hl < 2; ++hl)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int j = 0; j < 8; ++j) {
                            const int kout = 16 * mp + lane % 16;  // k1 or k2
                            const int kappa = 32 * ks + 8 * (lane / 16) + j;
                            const int comp = kappa & 1, e = (kappa >> 1) & 1;
                            const int cc = (kappa >> 2) & 15, h = kappa >> 6;
                            const int idx = 32 * h + 16 * e + cc;
                            const double ang = two_pi * double((kout * idx) % 64) / 64.0;
                            const double cs = std::cos(ang), sn = std::sin(ang);
                            // real row: C re + S im; imaginary row: -S re + C im
                            const double v = mt == 0 ? (comp == 0 ? cs : sn) : (comp == 0 ? -sn : cs);
                            const uint16_t hi = bf(float(v));
                            const uint16_t val = hl == 0 ? hi : bf(float(v - double(bff(hi))));
                            out[((size_t(mp * kFrags + mt * 8 + ks * 2 + hl) * 64 + lane) * 8) + j] = val;
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
    if (precision == RVX_PREC_FP16)
        return rvx::fail(RVX_EINVAL,
                         "%s: the fp16 form is withdrawn (its error, -71 dBc against FP64, misses "
                         "-80 dBc); use RVX_PREC_BF16X3",
                         who);
    if (precision != RVX_PREC_BF16X3) return rvx::fail(RVX_EINVAL, "%s: precision %d", who, precision);
    if (d.format != RVX_FORMAT_CS16)
        return rvx::fail(RVX_EINVAL, "%s: the matrix-core chain reads cs16", who);
    if (d.correction)
        return rvx::fail(RVX_EINVAL, "%s: no front-end correction in the matrix-core chain", who);
    if (d.channels != kM || d.decimation != kD || d.taps_per_branch != uint32_t(kTaps))
        return rvx::fail(RVX_EINVAL, "%s: grid M=%u L=%u D=%u (takes M = 4096, D = 2048, 16 taps)",
                         who, d.channels, d.taps_per_branch, d.decimation);
    if (d.channel_rows || d.taps_scaled)
        return rvx::fail(RVX_EINVAL, "%s: channel rows and prescaled taps are not supported", who);
    if (n_taps != uint64_t(kM) * kTaps || n_twiddles != uint64_t(kM))
        return rvx::fail(RVX_EINVAL, "%s: wants %llu taps and 4096 twiddles", who,
                         (unsigned long long)(uint64_t(kM) * kTaps));
    if (d.block_count == 0) return RVX_OK;
    if (d.block_count % kC)
        return rvx::fail(RVX_EINVAL, "%s: %u blocks is not a whole number of %d-block chunks", who,
                         d.block_count, kC);
    if (!rvx::pow2(d.out_ring_blocks) || d.block_count > d.out_ring_blocks)
        return rvx::fail(RVX_EINVAL, "%s: ring of %u blocks for %u blocks", who, d.out_ring_blocks,
                         d.block_count);
    const uint32_t pitch = d.ring_pitch ? d.ring_pitch : d.out_ring_blocks;
    if (pitch < d.out_ring_blocks || (d.ring_pitch && d.ring_pitch % 4) ||
        ring_samples < uint64_t(kM - 1) * pitch + d.out_ring_blocks)
        return rvx::fail(RVX_EINVAL, "%s: ring pitch %u or ring of %llu samples for %u-block rows",
                         who, d.ring_pitch, (unsigned long long)ring_samples, d.out_ring_blocks);
    if (d.input_samples >= (uint64_t(1) << 31))
        return rvx::fail(RVX_EINVAL, "%s: input exceeds 2^31 samples", who);
    const uint64_t N = uint64_t(kM) * kTaps;
    if (d.first_block * kD < d.input_origin + (N - 1) ||
        (d.first_block + d.block_count - 1) * kD >= d.input_origin + d.input_samples)
        return rvx::fail(RVX_EINVAL, "%s: the blocks reach outside the input", who);
    if (d.frame_blocks && d.frame_blocks != uint32_t(kC))
        return rvx::fail(RVX_EINVAL, "%s: averaging frame of %u blocks (takes %d)", who,
                         d.frame_blocks, kC);
    if (d.frame_blocks && (!partial || n_partial < uint64_t(d.block_count / kC) * kM))
        return rvx::fail(RVX_EINVAL, "%s: frames do not fit", who);
    if (!input || !taps || !twiddles || !channel_ring)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    if (!ctx->dft[0]) {
        const std::vector<uint16_t> t = dft_fragments();
        RVX_TRY_HIP(hipMalloc(&ctx->dft[0], t.size() * 2), who);
        RVX_TRY_HIP(hipMemcpy(ctx->dft[0], t.data(), t.size() * 2, hipMemcpyHostToDevice), who);
    }
    FastArgs a{};
    a.input = input;
    a.taps = taps;
    a.tw = static_cast<const float2*>(twiddles);
    a.dft = static_cast<const uint4*>(ctx->dft[0]);
    a.ring = static_cast<float2*>(channel_ring);
    a.partial = partial;
    a.first_block = d.first_block;
    a.input_origin = d.input_origin;
    a.input_samples = d.input_samples;
    a.ring_mask = d.out_ring_blocks - 1;
    a.ring_pitch = pitch;
    a.chunks = d.block_count / kC;
    a.chunks_per_xcd = (a.chunks + 7) / 8;
    a.average = d.frame_blocks != 0;
    a.alpha = d.alpha;
    // One workgroup per CU fits (the operands take 136 KiB of LDS); up to 128
    // per XCD, each walking its XCD's chunks with a stride of the XCD's
    // workgroups so that neighbours run together and share the input in L2.
    const uint32_t per_xcd = a.chunks_per_xcd < 128u ? a.chunks_per_xcd : 128u;
    chain_fast_kernel<<<dim3(8 * per_xcd), dim3(kThreads), 0, ctx->stream>>>(a);
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
