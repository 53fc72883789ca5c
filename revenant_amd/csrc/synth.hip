// Synthetic IQ source: a scene of carriers over Gaussian noise, rendered on
// the device as a pure function of the absolute sample index.
//
// Modelled on core/dsp/synth (wideband.h, channel.h:105-118): every sample is
// computed from its index alone, so any block of a stream can be rendered on
// its own and spot-checked on the CPU. Carrier and tone phases are reduced
// exactly in integers, (f * n) mod rate, before they become angles, so the
// phase at index 2^27 is as accurate as at index 0. Noise is Box-Muller over a
// counter-based hash (splitmix64) of (seed, index). Arithmetic is FP64; the
// sum is rounded once to float and then quantised to the native format.

#include <cmath>

#include "common.h"

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double unit_open(uint64_t bits) {
    // (0, 1]: 53 random bits, never zero, so log() is finite.
    return (double((bits >> 11) + 1)) * (1.0 / 9007199254740992.0);
}

// Turns of a frequency at an index: ((f mod rate) * (n mod rate)) mod rate,
// over rate. Exact as long as rate^2 < 2^63.
__device__ __forceinline__ double turns(int64_t f, uint64_t n, int64_t rate) {
    int64_t fr = f % rate;
    if (fr < 0) fr += rate;
    const uint64_t nr = n % uint64_t(rate);
    const uint64_t prod = (uint64_t(fr) * nr) % uint64_t(rate);
    return double(prod) / double(rate);
}

__global__ void synth_kernel(const rvx_carrier* __restrict__ carriers, uint32_t ncarriers,
                             rvx_synth_params p, void* __restrict__ out) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.count) return;
    const uint64_t n = p.start + i;
    double re = 0.0, im = 0.0;
    for (uint32_t c = 0; c < ncarriers; ++c) {
        const rvx_carrier k = carriers[c];
        double ph = turns(k.freq_hz, n, p.rate);
        double env = k.amplitude;
        if (k.kind == RVX_CARRIER_FM) {
            ph += k.index * sinpi(2.0 * turns(k.tone_hz, n, p.rate)) / (2.0 * M_PI);
        } else if (k.kind == RVX_CARRIER_AM) {
            env = k.amplitude * (1.0 + k.depth * cospi(2.0 * turns(k.tone_hz, n, p.rate))) /
                  (1.0 + k.depth);
        }
        double s, co;
        sincospi(2.0 * ph, &s, &co);
        re += env * co;
        im += env * s;
    }
    if (p.noise_sigma > 0.0) {
        const uint64_t h = splitmix64(p.seed ^ splitmix64(n));
        const uint64_t h2 = splitmix64(h);
        const double r = sqrt(-2.0 * log(unit_open(h)));
        double s, co;
        sincospi(2.0 * unit_open(h2), &s, &co);
        re += p.noise_sigma * r * co;
        im += p.noise_sigma * r * s;
    }
    const float fr = float(re), fi = float(im);
    if (p.format == RVX_FORMAT_CS16) {
        const float qr = fminf(fmaxf(rintf(fr * 32768.0f), -32768.0f), 32767.0f);
        const float qi = fminf(fmaxf(rintf(fi * 32768.0f), -32768.0f), 32767.0f);
        const uint32_t w = (uint32_t(int32_t(qr)) & 0xFFFFu) | (uint32_t(int32_t(qi)) << 16);
        static_cast<uint32_t*>(out)[i] = w;
    } else if (p.format == RVX_FORMAT_CU8) {
        const float qr = fminf(fmaxf(rintf(fr * 127.5f + 127.5f), 0.0f), 255.0f);
        const float qi = fminf(fmaxf(rintf(fi * 127.5f + 127.5f), 0.0f), 255.0f);
        static_cast<uint16_t*>(out)[i] = uint16_t(uint32_t(qr) | (uint32_t(qi) << 8));
    } else {
        static_cast<float2*>(out)[i] = make_float2(fr, fi);
    }
}

}  // namespace

extern "C" int32_t rvx_synth(rvx_ctx* ctx, const rvx_carrier* carriers, uint32_t ncarriers,
                             const rvx_synth_params* params, void* out, uint64_t out_bytes) {
    static const char* who = "rvx_synth";
    RVX_ENTER(ctx, who);
    if (params == nullptr) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_synth_params p = *params;
    if (p.rate <= 0 || p.rate > (int64_t(1) << 31))
        return rvx::fail(RVX_EINVAL, "%s: rate %lld outside (0, 2^31]", who, (long long)p.rate);
    const uint64_t bps = p.format == RVX_FORMAT_CS16 ? 4 : p.format == RVX_FORMAT_CU8 ? 2 : 8;
    if (p.format != RVX_FORMAT_CS16 && p.format != RVX_FORMAT_CU8 && p.format != -1)
        return rvx::fail(RVX_EINVAL, "%s: format %d (cs16, cu8 or -1 for cf32)", who, p.format);
    if (out_bytes < p.count * bps) return rvx::fail(RVX_EINVAL, "%s: output too small", who);
    if (p.count == 0) return RVX_OK;
    if (!out || (ncarriers && !carriers)) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    // The carrier table goes to the device on the context's stream.
    rvx_carrier* dev = nullptr;
    if (ncarriers) {
        RVX_TRY_HIP(hipMallocAsync(reinterpret_cast<void**>(&dev), ncarriers * sizeof(rvx_carrier),
                                   ctx->stream),
                    who);
        RVX_TRY_HIP(hipMemcpyAsync(dev, carriers, ncarriers * sizeof(rvx_carrier),
                                   hipMemcpyHostToDevice, ctx->stream),
                    who);
    }
    const uint64_t groups = (p.count + 255) / 256;
    synth_kernel<<<dim3(uint32_t(groups)), dim3(256), 0, ctx->stream>>>(dev, ncarriers, p, out);
    const int32_t rc = rvx::check_launch(who);
    if (dev) (void)hipFreeAsync(dev, ctx->stream);
    RVX_TRY_HIP(hipStreamSynchronize(ctx->stream), who);
    return rc;
}
