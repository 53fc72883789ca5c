// The carrier loop: a phase-locked loop on SAM's transmitted carrier, a
// Costas loop on DSB's suppressed one.
//
//   rvx_vrx_carrier        vrx_carrier.comp / vrx_reference.cpp:1052-1185
//   rvx_vrx_carrier_batch  many receivers' loops in one launch
//   rvx_vrx_carrier_check  validate(VrxCarrierConfig, VrxCarrierParams),
//                          vrx_reference.cpp:986-1050
//
// The loop is a recurrence: each sample's rotation depends on the phase the
// previous sample left, so one receiver's samples run in order on one lane.
// The reference runs one single-invocation dispatch per receiver; here a
// receiver is a lane and a launch carries up to 64 receivers per wave, so a
// set of SAM/DSB receivers costs one loop's latency, not the sum of them.
//
// Per sample: rotate the fine sample by the NCO (table lookup on the top
// nco_log2 bits of a 0.32 phase) and write it back; a level one-pole (the
// magnitude for the PLL, the power for Costas), primed by the first sample;
// a lock-test one-pole on the sample (PLL) or its square (Costas); the lock
// test real > 2|imag| with hold-in / hold-out counters choosing the acquire or
// track gains; the phase detector (imag, or real*imag for Costas) over the
// level, clamped to +-1; a proportional-integral update with the integrator
// clamped to the passband's reach; the phase advanced by the float advance
// times 2^32/(2 pi) converted to int32. Every operation is the twin's in its
// order, no contraction, denormals flushed (the build's flags), and the
// clamps and maxima written with the comparisons std::min/std::max use, so a
// NaN takes the same path as on the CPU.

#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "common.h"
#include "plan.h"
#include "vrx_math.h"

namespace {

// The kernel's constants, rounded to float from the same decimals
// (vrx_reference.cpp:909-911).
constexpr float kTurnsPerRadianF = static_cast<float>(683565275.576431632);
constexpr float kLevelFloorF = static_cast<float>(1.0e-24);
constexpr float kLockRatioF = static_cast<float>(2.0);
constexpr double kClampCeiling = 0.5 * 3.141592653589793238462643383279502884;
constexpr uint32_t kMaxNcoLog2 = 24;
constexpr uint32_t kLanes = 256;          // receivers per workgroup (four waves)
constexpr uint32_t kChunk = 8;            // fine samples loaded one chunk ahead
constexpr uint32_t kLdsNcoLog2 = 16;      // largest table whose quarter fits in LDS
constexpr uint32_t kLdsEntries = 1u << (kLdsNcoLog2 - 2);

// The NCO table (rvx_build_nco_table) is exactly quadrant-symmetric: built
// from the first octant, reflected and rotated with exact negations, zeros
// stored as +0. Entry q*N/4 + r is entry r turned by q quarter turns: (x, y),
// (y, -x), (-x, -y), (-y, x), with a zero written back as +0. So a workgroup
// keeps the first quarter in LDS (128 KiB at 2^16) and every lookup is an LDS
// read. The identity is checked for every table size on the CPU.
__device__ __forceinline__ float2 quarter_lookup(const float2* tab, uint32_t index,
                                                 uint32_t quarter_log2) {
    const float2 e = tab[index & ((1u << quarter_log2) - 1u)];
    const uint32_t q = (index >> quarter_log2) & 3u;
    // the sign flips depend on the index alone and are ready before the read
    const uint32_t flip_x = (q & 2u) << 30;         // quadrants 2, 3
    const uint32_t flip_y = ((q + 1u) & 2u) << 30;  // quadrants 1, 2
    const bool swap = (q & 1u) != 0u;
    const float x = __uint_as_float(__float_as_uint(swap ? e.y : e.x) ^ flip_x);
    const float y = __uint_as_float(__float_as_uint(swap ? e.x : e.y) ^ flip_y);
    // + 0 turns a negated zero into the +0 the table stores and leaves every
    // other value as it is (no table entry is denormal)
    return make_float2(x + 0.0f, y + 0.0f);
}

// std::max(v, lo) then std::min(., hi): GLSL's clamp as the twin spells it.
__device__ __forceinline__ float clamp_to(float v, float lo, float hi) {
    const float low = (v < lo) ? lo : v;
    return (hi < low) ? hi : low;
}

__device__ __forceinline__ bool item_fits(const rvx_vrx_carrier_item& it, uint32_t nco_log2) {
    const rvx_vrx_carrier_config& c = it.config;
    const rvx_vrx_carrier_params& p = it.params;
    if (c.power != RVX_CARRIER_PLL && c.power != RVX_CARRIER_COSTAS) return false;
    if (c.nco_log2 != nco_log2) return false;
    const uint64_t capacity = uint64_t(p.in_mask) + 1;
    if ((capacity & uint64_t(p.in_mask)) != 0 || p.count > capacity) return false;
    if (it.fine_samples < capacity) return false;
    return it.fine_ring != nullptr && it.state != nullptr;
}

// The fine ring and the state in the global address space, so their loads
// and stores are global (vector-memory counter only) rather than flat ones,
// which the LDS table lookups would otherwise wait behind.
typedef __attribute__((address_space(1))) float global_sample;
typedef __attribute__((address_space(1))) uint32_t global_u32;

__device__ __forceinline__ float2 load_sample(const global_sample* ring, uint32_t slot) {
    return make_float2(ring[2 * uint64_t(slot)], ring[2 * uint64_t(slot) + 1]);
}
__device__ __forceinline__ void store_sample(global_sample* ring, uint32_t slot, float2 v) {
    ring[2 * uint64_t(slot)] = v.x;
    ring[2 * uint64_t(slot) + 1] = v.y;
}

// det_sqrt with its zero case as a select rather than a branch: the same
// operations for v > 0, and 0 otherwise.
__device__ __forceinline__ float sqrt_or_zero(float v) {
    const float s = v * rvx::det_rsqrt(v);
    return (v > 0.0f) ? s : 0.0f;
}

// The loop's carried values.
struct Loop {
    uint32_t phase;
    float integrator, level, lock_real, lock_imag;
    uint32_t locked, held, primed;
};

// One sample: the twin's operations in its order. Written with selects in
// place of the twin's branches (each select keeps exactly the value the
// taken branch computes), so lanes of different loops and lock states run
// one instruction stream.
template <bool kLds>
__device__ __forceinline__ float2 loop_step(Loop& s, const rvx_vrx_carrier_params& p, bool pll,
                                            float2 z, const float2* __restrict__ nco,
                                            uint32_t shift_down, uint32_t quarter_log2,
                                            float bound_lo, float bound_hi) {
    const uint32_t index = s.phase >> shift_down;
    const float2 r = kLds ? quarter_lookup(nco, index, quarter_log2) : nco[index];
    const float y_real = z.x * r.x - z.y * r.y;
    const float y_imag = z.x * r.y + z.y * r.x;

    const float real_square = y_real * y_real;
    const float imag_square = y_imag * y_imag;
    const float power = real_square + imag_square;
    const float cross = y_real * y_imag;

    const float magnitude = pll ? sqrt_or_zero(power) : power;
    const float w_real = pll ? y_real : real_square - imag_square;
    const float w_imag = pll ? y_imag : cross + cross;

    {
        const float step = p.level_alpha * (magnitude - s.level);
        const float moved = s.level + step;
        s.level = (s.primed == 0u) ? magnitude : moved;
        s.primed = (s.primed == 0u) ? 1u : s.primed;
    }
    {
        const float step_real = p.lock_alpha * (w_real - s.lock_real);
        const float step_imag = p.lock_alpha * (w_imag - s.lock_imag);
        s.lock_real = s.lock_real + step_real;
        s.lock_imag = s.lock_imag + step_imag;
    }

    const float bound = kLockRatioF * fabsf(s.lock_imag);
    const bool inside = s.lock_real > bound;
    const bool was_locked = s.locked != 0u;
    const uint32_t held_locked = inside ? ((s.held > 0u) ? s.held - 1u : 0u) : s.held + 1u;
    const uint32_t held_open = inside ? s.held + 1u : 0u;
    const uint32_t held = was_locked ? held_locked : held_open;
    const bool flip = held >= (was_locked ? p.hold_out : p.hold_in);
    s.locked = flip ? (was_locked ? 0u : 1u) : s.locked;
    s.held = flip ? 0u : held;

    const float floored = (s.level < kLevelFloorF) ? kLevelFloorF : s.level;
    const float inverse = rvx::det_recip(floored);
    const float detected = (pll ? y_imag : cross) * inverse;
    const float error = clamp_to(detected, -1.0f, 1.0f);

    const float k1 = (s.locked != 0u) ? p.k1_track : p.k1_acquire;
    const float k2 = (s.locked != 0u) ? p.k2_track : p.k2_acquire;

    const float integrated = s.integrator + k2 * error;
    s.integrator = clamp_to(integrated, bound_lo, bound_hi);
    const float advance = k1 * error + s.integrator;

    const float turns = advance * kTurnsPerRadianF;
    s.phase = s.phase + static_cast<uint32_t>(static_cast<int32_t>(turns));
    return make_float2(y_real, y_imag);
}

template <bool kLds>
__device__ void run_loop(const rvx_vrx_carrier_item& it, const float2* __restrict__ nco,
                         uint32_t nco_log2) {
    const rvx_vrx_carrier_params& p = it.params;
    const bool pll = it.config.power == RVX_CARRIER_PLL;
    global_sample* ring = (global_sample*)(it.fine_ring);
    global_u32* st = (global_u32*)(it.state);

    const bool reset = (p.flags & RVX_CARRIER_FLAG_RESET) != 0u;
    Loop s;
    s.phase = reset ? 0u : st[0];
    s.integrator = reset ? 0.0f : __uint_as_float(st[1]);
    s.level = reset ? 0.0f : __uint_as_float(st[2]);
    s.lock_real = reset ? 0.0f : __uint_as_float(st[3]);
    s.lock_imag = reset ? 0.0f : __uint_as_float(st[4]);
    s.locked = reset ? 0u : st[5];
    s.held = reset ? 0u : st[6];
    s.primed = reset ? 0u : st[7];

    const float bound_hi = p.clamp_radians, bound_lo = -p.clamp_radians;
    {
        const float moved = s.integrator + p.shift;
        s.integrator = clamp_to(moved, bound_lo, bound_hi);
    }

    const uint32_t shift_down = 32u - nco_log2;
    const uint32_t quarter_log2 = nco_log2 - 2;
    const uint32_t count = p.count, off = p.in_offset, mask = p.in_mask;
    // Whole chunks: each chunk's fine samples are loaded while the previous
    // chunk runs (they do not depend on the loop), so their latency hides
    // behind its arithmetic. Then the tail, one sample at a time.
    const uint32_t whole = count - count % kChunk;
    float2 cur[kChunk];
    if (whole > 0) {
#pragma unroll
        for (uint32_t j = 0; j < kChunk; ++j) cur[j] = load_sample(ring, (off + j) & mask);
    }
    for (uint32_t base = 0; base < whole; base += kChunk) {
        float2 nxt[kChunk];
        const bool more = base + kChunk < whole;
#pragma unroll
        for (uint32_t j = 0; j < kChunk; ++j)
            nxt[j] = more ? load_sample(ring, (off + base + kChunk + j) & mask)
                          : make_float2(0.0f, 0.0f);
#pragma unroll
        for (uint32_t j = 0; j < kChunk; ++j) {
            const float2 y = loop_step<kLds>(s, p, pll, cur[j], nco, shift_down, quarter_log2,
                                             bound_lo, bound_hi);
            store_sample(ring, (off + base + j) & mask, y);
        }
#pragma unroll
        for (uint32_t j = 0; j < kChunk; ++j) cur[j] = nxt[j];
    }
    for (uint32_t i = whole; i < count; ++i) {
        const uint32_t slot = (off + i) & mask;
        const float2 y = loop_step<kLds>(s, p, pll, load_sample(ring, slot), nco, shift_down,
                                         quarter_log2, bound_lo, bound_hi);
        store_sample(ring, slot, y);
    }

    st[0] = s.phase;
    st[1] = __float_as_uint(s.integrator);
    st[2] = __float_as_uint(s.level);
    st[3] = __float_as_uint(s.lock_real);
    st[4] = __float_as_uint(s.lock_imag);
    st[5] = s.locked;
    st[6] = s.held;
    st[7] = s.primed;
}

// One lane per receiver, every receiver on the one NCO table. `items` null:
// the single receiver `one`. kLds: the table's first quarter is staged in
// LDS (nco_log2 in [2, 16]); otherwise lookups read the table in memory.
template <bool kLds>
__global__ void __launch_bounds__(kLanes)
carrier_kernel(const rvx_vrx_carrier_item* __restrict__ items, uint32_t n_items,
               rvx_vrx_carrier_item one, const float2* __restrict__ nco, uint32_t nco_log2) {
    __shared__ float2 tab[kLds ? kLdsEntries : 1];
    if (kLds) {
        const uint32_t quarter = 1u << (nco_log2 - 2);
        for (uint32_t e = threadIdx.x; e < quarter; e += kLanes) tab[e] = nco[e];
        __syncthreads();
    }
    const uint32_t k = blockIdx.x * kLanes + threadIdx.x;
    if (k >= n_items) return;
    const rvx_vrx_carrier_item it = items ? items[k] : one;
    if (!item_fits(it, nco_log2)) return;
    run_loop<kLds>(it, kLds ? tab : nco, nco_log2);
}

int32_t launch(rvx_ctx* ctx, const rvx_vrx_carrier_item* items, uint32_t n_items,
               const rvx_vrx_carrier_item& one, const void* nco, uint32_t nco_log2,
               const char* who) {
    const uint32_t blocks = (n_items + kLanes - 1) / kLanes;
    const float2* table = static_cast<const float2*>(nco);
    if (nco_log2 >= 2 && nco_log2 <= kLdsNcoLog2)
        carrier_kernel<true><<<dim3(blocks), dim3(kLanes), 0, ctx->stream>>>(items, n_items, one,
                                                                             table, nco_log2);
    else
        carrier_kernel<false><<<dim3(blocks), dim3(kLanes), 0, ctx->stream>>>(items, n_items, one,
                                                                              table, nco_log2);
    return rvx::check_launch(who);
}

}  // namespace

namespace rvx {

int32_t vrx_carrier_valid(const rvx_vrx_carrier_config& c, const rvx_vrx_carrier_params& p,
                          const char* who) {
    if (c.power != RVX_CARRIER_PLL && c.power != RVX_CARRIER_COSTAS)
        return fail(RVX_EINVAL,
                    "%s: power %u is neither the carrier loop (1) nor the Costas loop (2)", who,
                    c.power);
    if (c.nco_log2 == 0 || c.nco_log2 > kMaxNcoLog2)
        return fail(RVX_EINVAL, "%s: NCO table log2 %u is outside [1, %u]", who, c.nco_log2,
                    kMaxNcoLog2);
    const uint64_t capacity = uint64_t(p.in_mask) + 1;
    if ((capacity & uint64_t(p.in_mask)) != 0)
        return fail(RVX_EINVAL, "%s: ring mask %u is not a power of two minus one", who,
                    p.in_mask);
    if (p.count > capacity)
        return fail(RVX_EINVAL, "%s: %u samples in a ring of %llu would rotate some of them twice",
                    who, p.count, (unsigned long long)capacity);
    if ((p.flags & ~RVX_CARRIER_FLAG_RESET) != 0)
        return fail(RVX_EINVAL, "%s: flags 0x%x carry bits nothing reads", who, p.flags);
    if (!std::isfinite(p.shift))
        return fail(RVX_EINVAL, "%s: the retune shift is not a finite number", who);
    if (!(p.clamp_radians > 0.0f) || !(double(p.clamp_radians) <= kClampCeiling))
        return fail(RVX_EINVAL,
                    "%s: an integrator bound of %g radians per sample is outside (0, pi/2]", who,
                    double(p.clamp_radians));
    for (const float k : {p.k1_acquire, p.k1_track})
        if (!(k > 0.0f) || !(k < 1.0f))
            return fail(RVX_EINVAL,
                        "%s: a proportional gain of %g is outside (0, 1), which the phase "
                        "conversion's range rests on",
                        who, double(k));
    for (const float k : {p.k2_acquire, p.k2_track})
        if (!(k > 0.0f) || !(k < 1.0f))
            return fail(RVX_EINVAL, "%s: an integral gain of %g is outside (0, 1)", who,
                        double(k));
    for (const float a : {p.level_alpha, p.lock_alpha})
        if (!(a > 0.0f) || !(a <= 1.0f))
            return fail(RVX_EINVAL, "%s: a one-pole coefficient of %g is outside (0, 1]", who,
                        double(a));
    if (p.hold_in == 0 || p.hold_out == 0)
        return fail(RVX_EINVAL, "%s: a lock hold of zero samples switches on noise", who);
    return RVX_OK;
}

}  // namespace rvx

extern "C" {

int32_t rvx_vrx_carrier_check(const rvx_vrx_carrier_config* config,
                              const rvx_vrx_carrier_params* params) {
    static const char* who = "rvx_vrx_carrier_check";
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    return rvx::vrx_carrier_valid(*config, *params, who);
}

int32_t rvx_vrx_carrier_item_check(const rvx_vrx_carrier_item* item, uint64_t n_nco) {
    static const char* who = "rvx_vrx_carrier_item_check";
    if (!item) return rvx::fail(RVX_EINVAL, "%s: null item", who);
    const rvx_vrx_carrier_item& it = *item;
    if (int32_t rc = rvx::vrx_carrier_valid(it.config, it.params, who)) return rc;
    if (n_nco != (uint64_t(1) << it.config.nco_log2))
        return rvx::fail(RVX_EINVAL, "%s: the NCO table holds %llu entries and log2 %u implies %llu",
                         who, (unsigned long long)n_nco, it.config.nco_log2,
                         (unsigned long long)(uint64_t(1) << it.config.nco_log2));
    const uint64_t capacity = uint64_t(it.params.in_mask) + 1;
    if (it.fine_samples < capacity)
        return rvx::fail(RVX_EINVAL, "%s: fine ring holds %llu samples, mask %u implies %llu", who,
                         (unsigned long long)it.fine_samples, it.params.in_mask,
                         (unsigned long long)capacity);
    if (!it.fine_ring || !it.state) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    return RVX_OK;
}

int32_t rvx_vrx_carrier(rvx_ctx* ctx, const rvx_vrx_carrier_config* config,
                        const rvx_vrx_carrier_params* params, const void* nco, uint64_t n_nco,
                        void* fine_ring, uint64_t fine_samples, uint32_t* state) {
    static const char* who = "rvx_vrx_carrier";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (!nco) return rvx::fail(RVX_EINVAL, "%s: null NCO table", who);
    rvx_vrx_carrier_item it{};
    it.config = *config;
    it.params = *params;
    it.fine_ring = fine_ring;
    it.fine_samples = fine_samples;
    it.state = state;
    if (int32_t rc = rvx_vrx_carrier_item_check(&it, n_nco)) return rc;
    return launch(ctx, nullptr, 1, it, nco, config->nco_log2, who);
}

int32_t rvx_vrx_carrier_batch(rvx_ctx* ctx, const rvx_vrx_carrier_item* items, uint32_t n_items,
                              const void* nco, uint64_t n_nco) {
    static const char* who = "rvx_vrx_carrier_batch";
    RVX_ENTER(ctx, who);
    if (n_items == 0) return RVX_OK;
    if (!items || !nco) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    if (!rvx::pow2(n_nco) || n_nco < 2 || n_nco > (uint64_t(1) << kMaxNcoLog2))
        return rvx::fail(RVX_EINVAL, "%s: NCO table of %llu entries", who,
                         (unsigned long long)n_nco);
    return launch(ctx, items, n_items, rvx_vrx_carrier_item{}, nco, rvx::log2u(n_nco), who);
}

}  // extern "C"
