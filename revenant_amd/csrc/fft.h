// Radix-2 decimation-in-time transform, register-blocked for CDNA4.
//
// The butterfly graph is exactly the one core/dsp/pfb_fft_reference.cpp:54-104
// fixes (pfb_fft.comp:291-332): bit-reversed load, log2 N stages, stage s
// pairing (ia, ia + 2^s) with twiddle W[j * (N >> (s + 1))], j = ia mod 2^s,
// and the six-statement butterfly
//     vr = b.x*w.x - b.y*w.y;  vi = b.x*w.y + b.y*w.x;
//     a' = a + v;              b' = a - v.
// Each butterfly reads only its own two values, so its result depends on
// (N, stage, index) and not on the order butterflies are visited. That is what
// lets a thread hold V = 2^LOGV values and run LOGV consecutive stages in
// registers (one "pass") between LDS exchanges: every butterfly is the
// reference's, with the reference's twiddle, computed the same way.
//
// Pass p covers stages [p*LOGV, p*LOGV + LS) with LS = LOGV except possibly the
// last. A pass partitions the N positions into groups of G = 2^LS:
//     group g: j = g mod 2^S0, base = (g >> S0) << (S0 + LS),
//     element i at base + j + (i << S0)
// and a thread owns V/G whole groups. Stage S0 + l pairs elements i and
// i + 2^l (bit l of i clear) with twiddle index
//     (j + ((i mod 2^l) << S0)) * (N >> (S0 + l + 1)).
//
// LDS holds the N values between passes under a bank swizzle that XORs low
// position bits with higher ones; it is a bijection on [0, N), so it moves no
// value anywhere it is not read back from.
#pragma once

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rvx {

__device__ __forceinline__ void butterfly(float2& a, float2& b, const float2 w) {
    const float vr = b.x * w.x - b.y * w.y;
    const float vi = b.x * w.y + b.y * w.x;
    const float sr = a.x + vr;
    const float si = a.y + vi;
    const float dr = a.x - vr;
    const float di = a.y - vi;
    a = make_float2(sr, si);
    b = make_float2(dr, di);
}

__device__ __forceinline__ float2 cmul_exact(const float2 v, const float2 c) {
    const float pr = v.x * c.x - v.y * c.y;
    const float pi = v.x * c.y + v.y * c.x;
    return make_float2(pr, pi);
}

template <int BITS>
__host__ __device__ constexpr uint32_t bitrev(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < BITS; ++b) r |= ((v >> b) & 1u) << (BITS - 1 - b);
    return r;
}

template <int LOGN, int LOGV>
struct Fft {
    static_assert(LOGV >= 1 && LOGV <= LOGN, "need 2 <= V <= N");
    static constexpr int N = 1 << LOGN;
    static constexpr int LOG_V = LOGV;
    static constexpr int V = 1 << LOGV;
    static constexpr int LOGT = LOGN - LOGV;
    static constexpr int T = 1 << LOGT;  // threads per transform
    static constexpr int PASSES = (LOGN + LOGV - 1) / LOGV;

    static constexpr int stages_in(int p) {
        return (p + 1) * LOGV <= LOGN ? LOGV : LOGN - p * LOGV;
    }

    // LDS position swizzle: XOR of higher position bits into the low bits, a
    // bijection on [0, N). For radix-8 passes over 2048+ points five bits from
    // position bit LOGV+4 up spread a wave's 64 accesses over the banks in
    // every exchange (a bank model of the exchange patterns puts it at about
    // half the conflicted cycles of the V-bit form, which the smaller shapes
    // keep).
    __device__ __forceinline__ static uint32_t phys(uint32_t p) {
        if constexpr (LOGV == 3 && LOGN >= 11)
            return p ^ ((p >> (LOGV + 4)) & 31u);
        else if constexpr (LOGV == 2 && LOGN == 9)  // radix-4 over 512 points
            return p ^ ((p >> 4) & 15u) ^ (((p >> 5) & 3u) << 3);
        else
            return p ^ ((p >> (LOGV + 2)) & (V - 1));
    }

    // Position of register r (= u * G + i) of thread t in pass P.
    template <int P>
    __device__ __forceinline__ static uint32_t pos(uint32_t t, int r) {
        constexpr int S0 = P * LOGV;
        constexpr int LS = stages_in(P);
        constexpr int G = 1 << LS;
        const uint32_t u = uint32_t(r) / G, i = uint32_t(r) % G;
        const uint32_t g = t * (V / G) + u;
        const uint32_t j = g & ((1u << S0) - 1u);
        const uint32_t base = (g >> S0) << (S0 + LS);
        return base + j + (i << S0);
    }

    // The butterflies of pass P on the thread's registers.
    template <int P>
    __device__ __forceinline__ static void run_pass(float2 (&v)[V], const float2* __restrict__ tw,
                                                    uint32_t t) {
        constexpr int S0 = P * LOGV;
        constexpr int LS = stages_in(P);
        constexpr int G = 1 << LS;
#pragma unroll
        for (int u = 0; u < V / G; ++u) {
            const uint32_t g = t * (V / G) + u;
            const uint32_t j = g & ((1u << S0) - 1u);
#pragma unroll
            for (int l = 0; l < LS; ++l) {
                const int s = S0 + l;
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    if (i & (1 << l)) continue;
                    const uint32_t widx = (j + (uint32_t(i & ((1 << l) - 1)) << S0)) *
                                          uint32_t(N >> (s + 1));
                    butterfly(v[u * G + i], v[u * G + i + (1 << l)], tw[widx]);
                }
            }
        }
    }

    // Twiddles of pass P for thread t held in registers: for each group u and
    // stage l, the 2^l distinct values w[(j + (m << S0)) * (N >> (s + 1))],
    // m = 0 .. 2^l - 1, the ones run_pass reads at i & (2^l - 1). A pass has
    // at most V - 1 of them. Loading them apart from the butterflies lets a
    // caller issue the reads before other memory traffic of its own.
    //
    // The upper half of a stage's values sits N/4 entries past the lower half
    // (below N/4), and the table is exactly quadrant-symmetric: entry k + N/4
    // is entry k turned by -i, (x, y) -> (y, -x), bit for bit (the reference
    // builds it from the first octant with exact negations). So only the lower
    // half is read and the upper half is formed in registers: the same values
    // the table holds, in 4 reads instead of 7 for a radix-8 pass.
    static constexpr int WPP = V - 1;
    template <int P>
    __device__ __forceinline__ static void load_tw(const float2* __restrict__ tw, uint32_t t,
                                                   float2 (&w)[WPP]) {
        constexpr int S0 = P * LOGV;
        constexpr int LS = stages_in(P);
        constexpr int G = 1 << LS;
#pragma unroll
        for (int u = 0; u < V / G; ++u) {
            const uint32_t g = t * (V / G) + u;
            const uint32_t j = g & ((1u << S0) - 1u);
#pragma unroll
            for (int l = 0; l < LS; ++l) {
                const int s = S0 + l;
#pragma unroll
                for (int m = 0; m < (1 << l); ++m) {
                    const uint32_t idx = (j + (uint32_t(m) << S0)) * uint32_t(N >> (s + 1));
                    if (l >= 1 && m >= (1 << (l - 1))) {
                        const float2 lo = w[u * (G - 1) + (1 << l) - 1 + m - (1 << (l - 1))];
                        w[u * (G - 1) + (1 << l) - 1 + m] = make_float2(lo.y, -lo.x);
                    } else if constexpr (P == 0) {
                        // pass 0 runs at t = 0: every lane reads the same entries,
                        // which scalar loads bring into scalar registers
                        const __attribute__((address_space(4))) float* c =
                            reinterpret_cast<const __attribute__((address_space(4))) float*>(
                                uintptr_t(tw));
                        w[u * (G - 1) + (1 << l) - 1 + m] = make_float2(c[2 * idx], c[2 * idx + 1]);
                    } else {
                        w[u * (G - 1) + (1 << l) - 1 + m] = tw[idx];
                    }
                }
            }
        }
    }
    template <int P>
    __device__ __forceinline__ static void run_pass_w(float2 (&v)[V], const float2 (&w)[WPP]) {
        constexpr int LS = stages_in(P);
        constexpr int G = 1 << LS;
#pragma unroll
        for (int u = 0; u < V / G; ++u) {
#pragma unroll
            for (int l = 0; l < LS; ++l) {
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    if (i & (1 << l)) continue;
                    butterfly(v[u * G + i], v[u * G + i + (1 << l)],
                              w[u * (G - 1) + (1 << l) - 1 + (i & ((1 << l) - 1))]);
                }
            }
        }
    }
    // The twiddles of passes P .. END - 1 for thread t (pass 0 at t = 0, as
    // the callers run it).
    template <int P = 0, int END = PASSES>
    __device__ __forceinline__ static void load_all_tw(const float2* __restrict__ tw, uint32_t t,
                                                       float2 (&w)[PASSES][WPP]) {
        if constexpr (P < END && P < PASSES) {
            load_tw<P>(tw, P == 0 ? 0u : t, w[P]);
            load_all_tw<P + 1, END>(tw, t, w);
        }
    }

    template <int P>
    __device__ __forceinline__ static void store(float2* lds, const float2 (&v)[V], uint32_t t) {
#pragma unroll
        for (int r = 0; r < V; ++r) lds[phys(pos<P>(t, r))] = v[r];
    }
    template <int P>
    __device__ __forceinline__ static void load(const float2* lds, float2 (&v)[V], uint32_t t) {
#pragma unroll
        for (int r = 0; r < V; ++r) v[r] = lds[phys(pos<P>(t, r))];
    }

    // Passes 1 .. PASSES-1, entering with pass 0 already done in registers and
    // leaving the last pass's results in registers (natural order positions
    // given by pos<PASSES-1>). Ends with the LDS free for reuse only after a
    // barrier the caller places.
    template <int P = 1>
    __device__ __forceinline__ static void rest(float2* lds, float2 (&v)[V],
                                                const float2* __restrict__ tw, uint32_t t) {
        if constexpr (P < PASSES) {
            store<P - 1>(lds, v, t);
            __syncthreads();
            load<P>(lds, v, t);
            run_pass<P>(v, tw, t);
            rest<P + 1>(lds, v, tw, t);
        }
    }

    // rest() with the twiddles already in registers.
    template <int P = 1>
    __device__ __forceinline__ static void rest_w(float2* lds, float2 (&v)[V],
                                                  const float2 (&w)[PASSES][WPP], uint32_t t) {
        if constexpr (P < PASSES) {
            store<P - 1>(lds, v, t);
            __syncthreads();
            load<P>(lds, v, t);
            run_pass_w<P>(v, w[P]);
            rest_w<P + 1>(lds, v, w, t);
        }
    }

    // Whole transform of data already in LDS at bit-reversed positions
    // (lds[phys(i)] = x[bitrev(i)]); results left in registers.
    __device__ __forceinline__ static void from_lds(float2* lds, float2 (&v)[V],
                                                    const float2* __restrict__ tw, uint32_t t) {
        load<0>(lds, v, t);
        run_pass<0>(v, tw, t);
        rest<1>(lds, v, tw, t);
    }
};

}  // namespace rvx
