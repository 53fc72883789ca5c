// Per-receiver fine stage and demodulator.
//
//   rvx_vrx_fine   vrx_fine.comp / vrx_reference.cpp:489-620
//   rvx_vrx_demod  vrx_demod.comp / vrx_reference.cpp:714-905
//
// The fine stage is one thread per output exactly as the reference, every
// index 32-bit and wrapping, the 0.64 NCO phase carried in two words.
//
// The demodulator is tiled. The reference evaluates detect(idx - t) inside the
// audio filter's tap loop, so an AM output recomputes its DC window audio_taps
// times over. detect() is a pure function of the fine index, so a workgroup
// here evaluates it once per fine sample of its tile into LDS (magnitudes
// first for AM, then the detector) and every output's filter reads it from
// there: the same values, summed in the same order, computed once.

#pragma clang fp contract(off)

#include <algorithm>

#include "common.h"
#include "plan.h"
#include "vrx_math.h"

namespace {

using rvx::det_atan2;
using rvx::det_recip;
using rvx::det_sqrt;

__global__ void vrx_fine_kernel(rvx_vrx_fine_config c, rvx_vrx_fine_params p,
                                const float2* __restrict__ chan, const float2* __restrict__ taps,
                                const float2* __restrict__ nco, float2* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.count) return;
    out[(p.out_offset + i) & p.out_mask] = rvx::fine_output(c, p, chan + p.chan_base, taps, nco, i);
}

bool complex_tap(uint32_t mode) { return mode == 0 || (mode >= 8 && mode <= 11); }

struct DemodTile {
    uint32_t outputs;     // audio outputs per workgroup
    uint32_t det_reach;   // how far below an output's index detect() is evaluated
    uint32_t det_hist;    // samples below j that detect(j) reads (dc_taps-1 / 1 / 0)
};

constexpr uint32_t kDemodLdsFloats = 12288;  // 48 KiB of detector values + magnitudes

// Tiled demodulator for the real-output modes (AM, SAM, NFM, WFM mono/stereo,
// USB, LSB, DSB, CW).
__global__ void __launch_bounds__(256)
vrx_demod_kernel(rvx_vrx_demod_config c, rvx_vrx_demod_params p, DemodTile tile,
                 const float2* __restrict__ fine, const float* __restrict__ w,
                 float* __restrict__ audio) {
    extern __shared__ float lds[];
    const uint32_t o0 = blockIdx.x * tile.outputs;
    const uint32_t n = min(tile.outputs, p.count - o0);
    const uint32_t first_idx = p.in_offset + o0 * c.decimation;
    const uint32_t det_lo = first_idx - tile.det_reach;          // lowest detect() index
    const uint32_t det_span = (n - 1) * c.decimation + tile.det_reach + 1;
    float* det = lds;                                           // det_span values
    float* mag = lds + det_span;                                // det_span + det_hist
    const uint32_t mag_span = det_span + tile.det_hist;
    const uint32_t mag_lo = det_lo - tile.det_hist;
    auto sample = [&](uint32_t j) { return fine[j & p.in_mask]; };

    if (c.mode == RVX_DEMOD_AM || c.mode == 12) {
        for (uint32_t s = threadIdx.x; s < mag_span; s += blockDim.x) {
            const float2 z = sample(mag_lo + s);
            if (c.mode == RVX_DEMOD_AM) {
                const float sq = z.x * z.x + z.y * z.y;
                mag[s] = det_sqrt(sq);
            } else {
                mag[s] = z.x;
            }
        }
        __syncthreads();
        const float* dcw = w + c.audio_taps;
        for (uint32_t s = threadIdx.x; s < det_span; s += blockDim.x) {
            // detect(j) with j = det_lo + s; mag index of j - k is s + det_hist - k
            float dc = 0.0f;
            for (uint32_t k = 0; k < c.dc_taps; ++k) {
                const float term = mag[s + tile.det_hist - k] * dcw[k];
                dc = dc + term;
            }
            det[s] = mag[s + tile.det_hist] - dc;
        }
    } else if (c.mode == RVX_DEMOD_NFM || c.mode == RVX_DEMOD_WFM) {
        for (uint32_t s = threadIdx.x; s < det_span; s += blockDim.x) {
            const uint32_t j = det_lo + s;
            const float2 z = sample(j);
            const float2 q = sample(j - 1);
            const float pr = z.x * q.x + z.y * q.y;
            const float pi = z.y * q.x - z.x * q.y;
            det[s] = det_atan2(pi, pr);
        }
    } else {
        for (uint32_t s = threadIdx.x; s < det_span; s += blockDim.x) det[s] = sample(det_lo + s).x;
    }
    __syncthreads();

    const uint32_t pilot_base = c.audio_taps + c.dc_taps;
    const uint32_t rot_base = pilot_base + 2 * c.pilot_taps;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t o = o0 + i;
        const uint32_t rel = i * c.decimation + tile.det_reach;  // det index of idx
        if (c.channels == 2) {
            const uint32_t at = rel - c.audio_taps / 2;
            float pr = 0.0f, pi = 0.0f;
            for (uint32_t k = 0; k < c.pilot_taps; ++k) {
                const float comp = det[at - k];
                const float tr = comp * w[pilot_base + 2 * k];
                const float ti = comp * w[pilot_base + 2 * k + 1];
                pr = pr + tr;
                pi = pi + ti;
            }
            const float sr = pr * p.gain, si = pi * p.gain;
            const float rsq = sr * sr, isq = si * si;
            const float power = rsq + isq;
            const float inv = det_recip(fmaxf(power, 2.0e-5f));
            const float cross = sr * si;
            const float cos2 = (isq - rsq) * inv;
            const float sin2 = (-2.0f * cross) * inv;
            const float locked = power >= 2.0e-5f ? 1.0f : 0.0f;
            float sum = 0.0f, diff = 0.0f;
            for (uint32_t t = 0; t < c.audio_taps; ++t) {
                const float weighted = det[rel - t] * w[t];
                sum = sum + weighted;
                const float ref = sin2 * w[rot_base + 2 * t] - cos2 * w[rot_base + 2 * t + 1];
                const float term = weighted * ref;
                diff = diff + term;
            }
            const float gated = diff * locked;
            audio[2 * uint64_t(o)] = (sum + gated) * p.gain;
            audio[2 * uint64_t(o) + 1] = (sum - gated) * p.gain;
        } else {
            float acc = 0.0f;
            for (uint32_t t = 0; t < c.audio_taps; ++t) {
                const float term = det[rel - t] * w[t];
                acc = acc + term;
            }
            audio[o] = acc * p.gain;
        }
    }
}

// Complex taps (raw and the digital-voice taps): the baseband times the gain.
__global__ void vrx_demod_complex_kernel(rvx_vrx_demod_params p, const float2* __restrict__ fine,
                                         float* __restrict__ audio) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.count) return;
    const float2 z = fine[(p.in_offset + i) & p.in_mask];
    audio[2 * uint64_t(i)] = z.x * p.gain;
    audio[2 * uint64_t(i) + 1] = z.y * p.gain;
}

}  // namespace

namespace rvx {

// Validation shared with the batched launch (vrx_reference.cpp:429-486).
int32_t vrx_fine_valid(const rvx_vrx_fine_config& c, const rvx_vrx_fine_params& p,
                       const char* who) {
    if (c.taps < 2 || c.taps > 256)
        return fail(RVX_EINVAL, "%s: %u taps outside [2, 256]", who, c.taps);
    if (c.phases < 1 || c.phases > 4096)
        return fail(RVX_EINVAL, "%s: %u phases outside [1, 4096]", who, c.phases);
    if (c.nco_log2 > 24) return fail(RVX_EINVAL, "%s: NCO log2 %u above 24", who, c.nco_log2);
    if (p.out_rate == 0) return fail(RVX_EINVAL, "%s: output rate is zero", who);
    if (p.frac0 >= p.out_rate)
        return fail(RVX_EINVAL, "%s: frac0 %u not below the output rate %u", who, p.frac0,
                    p.out_rate);
    const uint64_t cap = uint64_t(p.chan_mask) + 1, ocap = uint64_t(p.out_mask) + 1;
    if ((cap & p.chan_mask) != 0 || (ocap & p.out_mask) != 0)
        return fail(RVX_EINVAL, "%s: ring masks must be powers of two minus one", who);
    if (uint64_t(p.frac0) + uint64_t(p.count) * p.step_rem > 0xFFFFFFFFull)
        return fail(RVX_EINVAL, "%s: the resampler recurrence overflows 32 bits", who);
    if (uint64_t(p.out_rate - 1) * c.phases > 0xFFFFFFFFull)
        return fail(RVX_EINVAL, "%s: phase selection overflows 32 bits", who);
    return RVX_OK;
}

uint32_t demod_fine_history(const rvx_vrx_demod_config& c) {
    const uint32_t det = (c.mode == RVX_DEMOD_AM || c.mode == 12) ? c.dc_taps - 1
                         : (c.mode == RVX_DEMOD_NFM || c.mode == RVX_DEMOD_WFM) ? 1u
                                                                                : 0u;
    const uint32_t audio_reach = c.audio_taps - 1;
    const uint32_t pilot_reach = c.pilot_taps == 0 ? 0u : c.audio_taps / 2 + c.pilot_taps - 1;
    return std::max(audio_reach, pilot_reach) + det;
}

int32_t vrx_demod_valid(const rvx_vrx_demod_config& c, const rvx_vrx_demod_params& p,
                        const char* who) {
    if (c.mode > 12) return fail(RVX_EINVAL, "%s: mode %u unknown", who, c.mode);
    const bool cx = complex_tap(c.mode);
    if (c.decimation == 0) return fail(RVX_EINVAL, "%s: zero decimation", who);
    if (c.audio_taps == 0 || c.audio_taps > 4096)
        return fail(RVX_EINVAL, "%s: %u audio taps outside [1, 4096]", who, c.audio_taps);
    if (c.dc_taps == 0 || c.dc_taps > 4096)
        return fail(RVX_EINVAL, "%s: %u DC taps outside [1, 4096]", who, c.dc_taps);
    if (cx && (c.decimation != 1 || c.audio_taps != 1))
        return fail(RVX_EINVAL, "%s: a complex tap carries no audio filter", who);
    if (c.channels != 1 && c.channels != 2)
        return fail(RVX_EINVAL, "%s: %u channels", who, c.channels);
    if (cx && c.channels != 2) return fail(RVX_EINVAL, "%s: a complex tap is two channels", who);
    if (c.pilot_taps > 2048) return fail(RVX_EINVAL, "%s: %u pilot taps", who, c.pilot_taps);
    if (c.pilot_taps != 0 && (c.mode != RVX_DEMOD_WFM || c.channels != 2))
        return fail(RVX_EINVAL, "%s: a pilot filter belongs to two-channel WFM", who);
    if (c.channels == 2 && !cx && c.pilot_taps == 0)
        return fail(RVX_EINVAL, "%s: two channels without a pilot filter", who);
    const uint64_t cap = uint64_t(p.in_mask) + 1;
    if ((cap & p.in_mask) != 0)
        return fail(RVX_EINVAL, "%s: input ring mask %u", who, p.in_mask);
    const uint64_t span =
        p.count == 0 ? 0
                     : uint64_t(p.count - 1) * c.decimation + demod_fine_history(c) + 1;
    if (span > cap)
        return fail(RVX_EINVAL, "%s: %u outputs need %llu fine samples, ring holds %llu", who,
                    p.count, (unsigned long long)span, (unsigned long long)cap);
    return RVX_OK;
}

}  // namespace rvx

extern "C" {

int32_t rvx_vrx_fine(rvx_ctx* ctx, const rvx_vrx_fine_config* config,
                     const rvx_vrx_fine_params* params, const void* channel_ring,
                     uint64_t ring_samples, const void* taps, uint64_t n_taps, const void* nco,
                     uint64_t n_nco, void* fine_ring, uint64_t fine_samples) {
    static const char* who = "rvx_vrx_fine";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_vrx_fine_config c = *config;
    const rvx_vrx_fine_params p = *params;
    if (int32_t rc = rvx::vrx_fine_valid(c, p, who)) return rc;
    if (n_taps != uint64_t(c.phases) * c.taps + 1)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu tap entries", who,
                         (unsigned long long)(uint64_t(c.phases) * c.taps + 1));
    if (n_nco != (uint64_t(1) << c.nco_log2))
        return rvx::fail(RVX_EINVAL, "%s: wants %llu NCO entries", who,
                         (unsigned long long)(uint64_t(1) << c.nco_log2));
    if (ring_samples < uint64_t(p.chan_base) + uint64_t(p.chan_mask) + 1)
        return rvx::fail(RVX_EINVAL, "%s: channel ring too small for base %u mask %u", who,
                         p.chan_base, p.chan_mask);
    if (fine_samples < uint64_t(p.out_mask) + 1 || p.count > uint64_t(p.out_mask) + 1)
        return rvx::fail(RVX_EINVAL, "%s: output ring too small", who);
    if (p.count == 0) return RVX_OK;
    if (!channel_ring || !taps || !nco || !fine_ring)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    vrx_fine_kernel<<<dim3((p.count + 255) / 256), dim3(256), 0, ctx->stream>>>(
        c, p, static_cast<const float2*>(channel_ring), static_cast<const float2*>(taps),
        static_cast<const float2*>(nco), static_cast<float2*>(fine_ring));
    return rvx::check_launch(who);
}

int32_t rvx_vrx_demod(rvx_ctx* ctx, const rvx_vrx_demod_config* config,
                      const rvx_vrx_demod_params* params, const void* fine_ring,
                      uint64_t fine_samples, const float* weights, uint64_t n_weights,
                      float* audio, uint64_t n_audio) {
    static const char* who = "rvx_vrx_demod";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_vrx_demod_config c = *config;
    const rvx_vrx_demod_params p = *params;
    if (int32_t rc = rvx::vrx_demod_valid(c, p, who)) return rc;
    const uint64_t pilot_base = uint64_t(c.audio_taps) + c.dc_taps;
    const uint64_t want =
        c.pilot_taps == 0 ? pilot_base : pilot_base + 2ull * c.pilot_taps + 2ull * c.audio_taps;
    if (n_weights != want)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu weights, got %llu", who,
                         (unsigned long long)want, (unsigned long long)n_weights);
    if (fine_samples < uint64_t(p.in_mask) + 1)
        return rvx::fail(RVX_EINVAL, "%s: fine ring smaller than its mask", who);
    if (n_audio != uint64_t(p.count) * c.channels)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu audio values", who,
                         (unsigned long long)(uint64_t(p.count) * c.channels));
    if (p.count == 0) return RVX_OK;
    if (!fine_ring || !weights || !audio) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    if (complex_tap(c.mode)) {
        vrx_demod_complex_kernel<<<dim3((p.count + 255) / 256), dim3(256), 0, ctx->stream>>>(
            p, static_cast<const float2*>(fine_ring), audio);
        return rvx::check_launch(who);
    }
    DemodTile tile{};
    const uint32_t audio_reach = c.audio_taps - 1;
    const uint32_t pilot_reach = c.pilot_taps == 0 ? 0u : c.audio_taps / 2 + c.pilot_taps - 1;
    tile.det_reach = std::max(audio_reach, pilot_reach);
    tile.det_hist = (c.mode == RVX_DEMOD_AM || c.mode == 12) ? c.dc_taps - 1 : 0u;
    // det_span + mag_span = 2*((n-1)*R + reach + 1) + hist floats must fit.
    const uint64_t fixed = 2ull * (tile.det_reach + 1) + tile.det_hist;
    if (fixed > kDemodLdsFloats)
        return rvx::fail(RVX_EINVAL, "%s: filter reach %u does not fit the LDS tile", who,
                         tile.det_reach);
    const uint64_t fit = (kDemodLdsFloats - fixed) / (2ull * c.decimation) + 1;
    tile.outputs = uint32_t(std::min<uint64_t>({fit, 1024, p.count}));
    const uint32_t groups = (p.count + tile.outputs - 1) / tile.outputs;
    const uint64_t det_span = uint64_t(tile.outputs - 1) * c.decimation + tile.det_reach + 1;
    const size_t lds_bytes = size_t(2 * det_span + tile.det_hist) * sizeof(float);
    vrx_demod_kernel<<<dim3(groups), dim3(256), lds_bytes, ctx->stream>>>(
        c, p, tile, static_cast<const float2*>(fine_ring), weights, audio);
    return rvx::check_launch(who);
}

}  // extern "C"
