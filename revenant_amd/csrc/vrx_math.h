// Receiver arithmetic shared by the per-receiver and the batched kernels.
//
// det_*: core/dsp/vrx_reference.cpp:227-306 (vrx_demod.comp:329-428), fixed
// Newton steps and polynomials so any IEEE float implementation agrees.
// fine_output: one output of the stateless resampler + complex FIR + 0.64
// NCO rotation, vrx_reference.cpp:489-620 (vrx_fine.comp:271-369); every
// value is a function of the output index alone.
#pragma once

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/rvx.h"

namespace rvx {

__device__ __forceinline__ float det_recip(float v) {
    const float seed = __uint_as_float(0x7EF311C3u - __float_as_uint(v));
    const float r1 = seed * (2.0f - v * seed);
    const float r2 = r1 * (2.0f - v * r1);
    const float r3 = r2 * (2.0f - v * r2);
    const float r4 = r3 * (2.0f - v * r3);
    return r4;
}

__device__ __forceinline__ float det_rsqrt(float v) {
    const float seed = __uint_as_float(0x5F375A86u - (__float_as_uint(v) >> 1));
    const float h = 0.5f * v;
    const float r1 = seed * (1.5f - h * (seed * seed));
    const float r2 = r1 * (1.5f - h * (r1 * r1));
    const float r3 = r2 * (1.5f - h * (r2 * r2));
    const float r4 = r3 * (1.5f - h * (r3 * r3));
    return r4;
}

__device__ __forceinline__ float det_sqrt(float v) {
    if (!(v > 0.0f)) return 0.0f;
    return v * det_rsqrt(v);
}

__device__ __forceinline__ float det_atan2(float y, float x) {
    const float ay = fabsf(y), ax = fabsf(x);
    if (ay == 0.0f && ax == 0.0f) return 0.0f;
    const float num = (ax < ay) ? ax : ay;
    const float den = (ay < ax) ? ax : ay;
    const float t = num * det_recip(den);
    const float u = t * t;
    float p = float(-0.0021090222);
    p = p * u + float(0.0122847765);
    p = p * u + float(-0.0336503503);
    p = p * u + float(0.0603537785);
    p = p * u + float(-0.0853800324);
    p = p * u + float(0.1099615976);
    p = p * u + float(-0.1427217912);
    p = p * u + float(0.1999923872);
    p = p * u + float(-0.3333331929);
    p = p * u + float(0.9999999998);
    float angle = t * p;
    if (ay > ax) angle = float(1.57079632679489661923) - angle;
    if (x < 0.0f) angle = float(3.14159265358979323846) - angle;
    return (y < 0.0f) ? -angle : angle;
}

// Fine-stage output i of a dispatch with params p; `ch` is the receiver's
// channel (channel ring + chan_base).
__device__ __forceinline__ float2 fine_output(const rvx_vrx_fine_config& c,
                                              const rvx_vrx_fine_params& p,
                                              const float2* __restrict__ ch,
                                              const float2* __restrict__ taps,
                                              const float2* __restrict__ nco, uint32_t i) {
    const uint32_t acc_idx = p.frac0 + i * p.step_rem;
    const uint32_t carry = acc_idx / p.out_rate;
    const uint32_t frac = acc_idx - carry * p.out_rate;
    const uint32_t n_rel = i * p.step_whole + carry;
    const uint32_t scaled = frac * c.phases;
    const uint32_t phase = scaled / p.out_rate;
    const uint32_t blend_num = scaled - phase * p.out_rate;
    const float blend = float(blend_num) * p.inv_out_rate;
    const float comp = 1.0f - blend;
    const uint32_t base = p.in_offset + n_rel;
    float ar = 0.0f, ai = 0.0f;
    for (uint32_t k = 0; k < c.taps; ++k) {
        const uint32_t slot = phase + k * c.phases;
        const float2 g0 = taps[slot];
        const float2 g1 = taps[slot + 1];
        const float tr = g0.x * comp + g1.x * blend;
        const float ti = g0.y * comp + g1.y * blend;
        const float2 s = ch[(base - k) & p.chan_mask];
        ar = ar + (s.x * tr - s.y * ti);
        ai = ai + (s.x * ti + s.y * tr);
    }
    const uint32_t prod_lo = i * p.nco_delta_low;
    const uint32_t prod_hi = __umulhi(i, p.nco_delta_low);
    const uint32_t ph_lo = p.nco_phase_low + prod_lo;
    const uint32_t carry_bit = ph_lo < p.nco_phase_low ? 1u : 0u;
    const uint32_t ph_hi = p.nco_phase_high + i * p.nco_delta_high + prod_hi + carry_bit;
    const uint32_t ni = c.nco_log2 == 0 ? 0u : (ph_hi >> (32 - c.nco_log2));
    const float2 r = nco[ni];
    const float orr = ar * r.x - ai * r.y;
    const float oi = ar * r.y + ai * r.x;
    return make_float2(orr, oi);
}

// Resampler phase period: fine indices i and i + period share (phase, blend),
// since period * step_rem is a multiple of out_rate.
__host__ __device__ __forceinline__ uint32_t phase_period(const rvx_vrx_fine_params& p) {
    uint32_t a = p.step_rem, b = p.out_rate;
    while (a != 0) {
        const uint32_t t = b % a;
        b = a;
        a = t;
    }
    return p.out_rate / b;  // b = gcd(step_rem, out_rate); step_rem = 0 gives 1
}

// The interpolated taps of phase class cls (fine indices i = cls mod period):
// exactly the (tr, ti) fine_output forms for such an i, written to row[0..taps).
__device__ __forceinline__ void fine_taps_row(const rvx_vrx_fine_config& c,
                                              const rvx_vrx_fine_params& p,
                                              const float2* __restrict__ taps, uint32_t cls,
                                              uint32_t k, float2* row) {
    const uint32_t frac = uint32_t((uint64_t(p.frac0) + uint64_t(cls) * p.step_rem) % p.out_rate);
    const uint32_t scaled = frac * c.phases;
    const uint32_t phase = scaled / p.out_rate;
    const uint32_t blend_num = scaled - phase * p.out_rate;
    const float blend = float(blend_num) * p.inv_out_rate;
    const float comp = 1.0f - blend;
    const uint32_t slot = phase + k * c.phases;
    const float2 g0 = taps[slot];
    const float2 g1 = taps[slot + 1];
    const float tr = g0.x * comp + g1.x * blend;
    const float ti = g0.y * comp + g1.y * blend;
    row[k] = make_float2(tr, ti);
}

// fine_output with the interpolated taps of i's phase class already in `row`.
__device__ __forceinline__ float2 fine_output_row(const rvx_vrx_fine_config& c,
                                                  const rvx_vrx_fine_params& p,
                                                  const float2* __restrict__ ch,
                                                  const float2* row,
                                                  const float2* __restrict__ nco, uint32_t i) {
    const uint32_t acc_idx = p.frac0 + i * p.step_rem;
    const uint32_t carry = acc_idx / p.out_rate;
    const uint32_t n_rel = i * p.step_whole + carry;
    const uint32_t base = p.in_offset + n_rel;
    // The ring through the global address space: its reads then wait on the
    // vector-memory counter alone, not on the LDS reads of the taps as well.
    typedef float ring_f2 __attribute__((ext_vector_type(2)));
    typedef const __attribute__((address_space(1))) ring_f2* ring_ptr;
    const ring_ptr gch = reinterpret_cast<ring_ptr>(uintptr_t(ch));
    float ar = 0.0f, ai = 0.0f;
    uint32_t k = 0;
    const uint32_t b0 = base & p.chan_mask;
    if (b0 >= c.taps + 7u) {
        // The window does not wrap: tap k reads q[-k], one pointer with
        // immediate offsets, and each group of eight taps' reads is issued
        // while the group before it is summed (the last reads run up to 15
        // samples past the window, still inside the ring, and go unused). The
        // sums run in the same order.
        const ring_ptr q = gch + b0;
        auto group = [&](const ring_f2 (&s)[8], uint32_t k0) {
            float2 t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) t[u] = row[k0 + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                ar = ar + (s[u].x * t[u].x - s[u].y * t[u].y);
                ai = ai + (s[u].x * t[u].y + s[u].y * t[u].x);
            }
        };
        auto fetch = [&](ring_f2 (&s)[8], uint32_t k0) {
            const ring_ptr qk = q - k0;
#pragma unroll
            for (int u = 0; u < 8; ++u) s[u] = qk[-u];
        };
        // two buffers in turn: group k's samples are summed while group
        // k + 8's are read
        ring_f2 sa[8], sb[8];
        fetch(sa, 0);
        for (; k + 16 <= c.taps; k += 16) {
            fetch(sb, k + 8);
            group(sa, k);
            fetch(sa, k + 16);
            group(sb, k + 8);
        }
        if (k + 8 <= c.taps) {
            fetch(sb, k + 8);
            group(sa, k);
            k += 8;
#pragma unroll
            for (int u = 0; u < 8; ++u) sa[u] = sb[u];
        }
        // sa holds taps k .. k + 7; the tail uses those below c.taps
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            if (k + uint32_t(u) < c.taps) {
                const float2 t = row[k + u];
                ar = ar + (sa[u].x * t.x - sa[u].y * t.y);
                ai = ai + (sa[u].x * t.y + sa[u].y * t.x);
            }
        }
    } else {
        for (; k + 8 <= c.taps; k += 8) {
            float2 s[8], t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const ring_f2 v = gch[(base - (k + u)) & p.chan_mask];
                s[u] = make_float2(v.x, v.y);
                t[u] = row[k + u];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                ar = ar + (s[u].x * t[u].x - s[u].y * t[u].y);
                ai = ai + (s[u].x * t[u].y + s[u].y * t[u].x);
            }
        }
        for (; k < c.taps; ++k) {
            const float2 s = ch[(base - k) & p.chan_mask];
            const float2 t = row[k];
            ar = ar + (s.x * t.x - s.y * t.y);
            ai = ai + (s.x * t.y + s.y * t.x);
        }
    }
    const uint32_t prod_lo = i * p.nco_delta_low;
    const uint32_t prod_hi = __umulhi(i, p.nco_delta_low);
    const uint32_t ph_lo = p.nco_phase_low + prod_lo;
    const uint32_t carry_bit = ph_lo < p.nco_phase_low ? 1u : 0u;
    const uint32_t ph_hi = p.nco_phase_high + i * p.nco_delta_high + prod_hi + carry_bit;
    const uint32_t ni = c.nco_log2 == 0 ? 0u : (ph_hi >> (32 - c.nco_log2));
    const float2 r = nco[ni];
    const float orr = ar * r.x - ai * r.y;
    const float oi = ar * r.y + ai * r.x;
    return make_float2(orr, oi);
}

}  // namespace rvx
