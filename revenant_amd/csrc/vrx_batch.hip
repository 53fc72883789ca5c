// Batched receivers: every VRX of a step in one persistent launch.
//
// The reference records two or three dispatches per receiver per block, each
// reading and writing device memory (core/engine/vrx_stage.cpp:1306-1399).
// Here a persistent workgroup takes (receiver, audio tile) items from a
// queue and, for one tile, computes on chip:
//   1. the fine samples the tile's audio filter reaches, with the receiver's
//      fine-block params (fine_output, exactly vrx_fine.comp's arithmetic);
//   2. detect() once per fine sample (AM magnitudes first), as vrx.hip's
//      tiled demodulator does;
//   3. the decimating audio filter, mono, FM stereo or complex tap.
// Tiles overlap only by the filter's history, recomputed: every value is a
// function of its index, so the bits equal rvx_vrx_fine + rvx_vrx_demod.
// Receivers of any mode and rate mix freely in one launch; the host orders
// the work largest-first and the queue balances the rest.

#pragma clang fp contract(off)

#include <algorithm>
#include <mutex>

#include "common.h"
#include "plan.h"
#include "vrx_math.h"

// The no-allocation entry point's scratch: the persistent launch's queue
// counter in word 0 of a 256-byte array. Each device holds its own copy until
// the process ends; every rvx_vrx_batch_noalloc call on a device shares it.
static __device__ __align__(256) uint32_t scratch[64];

namespace {

constexpr uint32_t kBatchThreads = 256;
// Fine samples per tile, 8 B each. The detector's values (and AM's
// magnitudes) reuse the same bytes once the fine samples are read: the
// magnitudes in the lower half as floats, the detector output in the upper.
constexpr uint32_t kMaxSpan = RVX_VRX_TILE_SPAN;
constexpr uint32_t kTapEntries = 2040;  // interpolated taps of every phase class (8 B each)
constexpr uint32_t kPerThread = kMaxSpan / kBatchThreads;
static_assert(kMaxSpan % kBatchThreads == 0, "whole detector rounds per thread");
// Global memory seen through the constant address space: the caller's
// promise that nothing writes it while the kernel runs.
typedef const __attribute__((address_space(4))) float kconst_float;
// 40,900 bytes of LDS: four workgroups fit a CU's 160 KiB.
static_assert(kTapEntries * 8 + kMaxSpan * 8 + 4 == 40900, "LDS budget of four workgroups a CU");
static_assert(sizeof(rvx_vrx_desc) == 160, "descriptor layout is part of the ABI");

__host__ __device__ bool complex_tap(uint32_t mode) { return mode == 0 || (mode >= 8 && mode <= 11); }

__host__ __device__ uint32_t detector_history(const rvx_vrx_demod_config& c) {
    return (c.mode == RVX_DEMOD_AM || c.mode == 12) ? c.dc_taps - 1
           : (c.mode == RVX_DEMOD_NFM || c.mode == RVX_DEMOD_WFM) ? 1u
                                                                  : 0u;
}

__global__ void zero_word(uint32_t* p) { *p = 0u; }

__global__ void __launch_bounds__(kBatchThreads)
vrx_batch_kernel(const rvx_vrx_desc* __restrict__ descs, uint32_t n_desc,
                 const rvx_vrx_work* __restrict__ work, uint32_t n_work, uint32_t* __restrict__ counter,
                 const float2* __restrict__ nco, uint32_t nco_log2_max) {
    __shared__ float2 tap_lds[kTapEntries];
    __shared__ float2 fine[kMaxSpan];
    __shared__ uint32_t item_sh;
    float* mag = reinterpret_cast<float*>(fine);  // AM magnitudes replace the fine samples
    float* det = mag + kMaxSpan;                  // and the detector output follows them
    const uint32_t tid = threadIdx.x;

    for (;;) {
        if (tid == 0) item_sh = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(item_sh);
        if (item >= n_work) break;  // uniform: every thread read the same item
        const rvx_vrx_work w = work[item];
        if (w.vrx >= n_desc) {  // a malformed work list computes nothing
            __syncthreads();
            continue;
        }
        const rvx_vrx_desc& d = descs[w.vrx];
        const rvx_vrx_fine_config fc = d.fine_config;
        const rvx_vrx_fine_params fp = d.fine;
        const rvx_vrx_demod_config dc = d.demod_config;
        const rvx_vrx_demod_params dp = d.demod;
        const uint32_t R = dc.decimation;
        const uint32_t a0 = w.tile * d.tile;
        // Descriptors pass rvx_vrx_desc_check on the host; these guards keep an
        // unchecked one inside the NCO table and the LDS tile.
        const uint32_t n = a0 < dp.count ? min(d.tile, dp.count - a0) : 0u;
        const uint64_t want = n ? uint64_t(n - 1) * R + d.hist + 1 : 0u;
        const uint32_t pilot_reach = dc.pilot_taps ? dc.audio_taps / 2 + dc.pilot_taps - 1 : 0u;
        const uint32_t filter_reach = max(dc.audio_taps - 1, pilot_reach);
        const bool valid = fc.nco_log2 <= nco_log2_max && want <= kMaxSpan &&
                           d.hist >= detector_history(dc) + filter_reach;
        const uint32_t span = valid ? uint32_t(want) : 0u;
        const uint32_t dh = detector_history(dc);
        const float2* ch = static_cast<const float2*>(d.channel_ring) + fp.chan_base;
        const float2* taps = static_cast<const float2*>(d.taps);
        // The demod tables are read-only for the launch and every lane of a
        // workgroup reads the same tap at once: through the constant address
        // space those reads are scalar loads, issued ahead of the taps that
        // use them, instead of a vector load and a full wait per tap.
        const kconst_float* wts = reinterpret_cast<const kconst_float*>(uintptr_t(d.weights));
        float* audio = d.audio;

        // 1. fine samples a0*R .. a0*R + span - 1 of the receiver's fine block.
        // Fine indices one phase period apart share their interpolated taps,
        // so when the period's taps fit they are formed once per tile.
        // Rows sit taps + 1 entries apart: lanes of different phase classes
        // read the same tap index, and an odd row stride keeps them on
        // different banks.
        const uint32_t period = rvx::phase_period(fp);
        const uint32_t stride = fc.taps + 1;
        if (uint64_t(period) * stride <= kTapEntries) {
            const uint32_t entries = period * fc.taps;
            for (uint32_t e = tid; span && e < entries; e += kBatchThreads)
                rvx::fine_taps_row(fc, fp, taps, e / fc.taps, e % fc.taps,
                                   tap_lds + (e / fc.taps) * stride);
            __syncthreads();
            for (uint32_t s = tid; s < span; s += kBatchThreads) {
                const uint32_t i = a0 * R + s;
                fine[s] = rvx::fine_output_row(fc, fp, ch, tap_lds + (i % period) * stride, nco, i);
            }
        } else {
            for (uint32_t s = tid; s < span; s += kBatchThreads)
                fine[s] = rvx::fine_output(fc, fp, ch, taps, nco, a0 * R + s);
        }
        __syncthreads();

        if (complex_tap(dc.mode)) {
            for (uint32_t i = tid; span && i < n; i += kBatchThreads) {
                const float2 z = fine[i * R + d.hist];
                audio[2 * uint64_t(a0 + i)] = z.x * dp.gain;
                audio[2 * uint64_t(a0 + i) + 1] = z.y * dp.gain;
            }
            __syncthreads();
            continue;
        }

        // 2. detector: det[s] is detect() at fine index s + dh of the tile
        const uint32_t det_span = span ? span - dh : 0u;
        if (dc.mode == RVX_DEMOD_AM || dc.mode == 12) {
            // every fine sample is read before any magnitude overwrites it
            float m[kPerThread];
#pragma unroll
            for (uint32_t q = 0; q < kPerThread; ++q) {
                const uint32_t s = tid + q * kBatchThreads;
                m[q] = 0.0f;
                if (s < span) {
                    const float2 z = fine[s];
                    if (dc.mode == RVX_DEMOD_AM) {
                        const float sq = z.x * z.x + z.y * z.y;
                        m[q] = rvx::det_sqrt(sq);
                    } else {
                        m[q] = z.x;
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t q = 0; q < kPerThread; ++q) {
                const uint32_t s = tid + q * kBatchThreads;
                if (s < span) mag[s] = m[q];
            }
            __syncthreads();
            const kconst_float* dcw = wts + dc.audio_taps;
            for (uint32_t s = tid; s < det_span; s += kBatchThreads) {
                float acc = 0.0f;
                const kconst_float* dw = dcw;
#pragma unroll 8
                for (uint32_t k = 0; k < dc.dc_taps; ++k, ++dw) {
                    const float term = mag[s + dh - k] * dw[0];
                    acc = acc + term;
                }
                det[s] = mag[s + dh] - acc;
            }
        } else {
            // every fine sample is read before the detector output overwrites it
            const bool fm = dc.mode == RVX_DEMOD_NFM || dc.mode == RVX_DEMOD_WFM;
            float v[kPerThread];
#pragma unroll
            for (uint32_t q = 0; q < kPerThread; ++q) {
                const uint32_t s = tid + q * kBatchThreads;
                v[q] = 0.0f;
                if (s < det_span) {
                    if (fm) {
                        const float2 z = fine[s + dh];
                        const float2 p = fine[s + dh - 1];
                        const float pr = z.x * p.x + z.y * p.y;
                        const float pi = z.y * p.x - z.x * p.y;
                        v[q] = rvx::det_atan2(pi, pr);
                    } else {
                        v[q] = fine[s].x;
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t q = 0; q < kPerThread; ++q) {
                const uint32_t s = tid + q * kBatchThreads;
                if (s < det_span) det[s] = v[q];
            }
        }
        __syncthreads();

        // 3. audio filter; output i sits at detector index i*R + (hist - dh)
        const uint32_t reach = d.hist - dh;
        const uint32_t pilot_base = dc.audio_taps + dc.dc_taps;
        const uint32_t rot_base = pilot_base + 2 * dc.pilot_taps;
        for (uint32_t i = tid; span && i < n; i += kBatchThreads) {
            const uint32_t rel = i * R + reach;
            const uint32_t o = a0 + i;
            if (dc.channels == 2) {
                const uint32_t at = rel - dc.audio_taps / 2;
                float pr = 0.0f, pi = 0.0f;
                const kconst_float* pw = wts + pilot_base;
#pragma unroll 8
                for (uint32_t k = 0; k < dc.pilot_taps; ++k, pw += 2) {
                    const float comp = det[at - k];
                    const float tr = comp * pw[0];
                    const float ti = comp * pw[1];
                    pr = pr + tr;
                    pi = pi + ti;
                }
                const float sr = pr * dp.gain, si = pi * dp.gain;
                const float rsq = sr * sr, isq = si * si;
                const float power = rsq + isq;
                const float inv = rvx::det_recip(fmaxf(power, 2.0e-5f));
                const float cross = sr * si;
                const float cos2 = (isq - rsq) * inv;
                const float sin2 = (-2.0f * cross) * inv;
                const float locked = power >= 2.0e-5f ? 1.0f : 0.0f;
                float sum = 0.0f, diff = 0.0f;
                const kconst_float* aw = wts;
                const kconst_float* rw = wts + rot_base;
#pragma unroll 8
                for (uint32_t t = 0; t < dc.audio_taps; ++t, ++aw, rw += 2) {
                    const float weighted = det[rel - t] * aw[0];
                    sum = sum + weighted;
                    const float ref = sin2 * rw[0] - cos2 * rw[1];
                    const float term = weighted * ref;
                    diff = diff + term;
                }
                const float gated = diff * locked;
                audio[2 * uint64_t(o)] = (sum + gated) * dp.gain;
                audio[2 * uint64_t(o) + 1] = (sum - gated) * dp.gain;
            } else {
                float acc = 0.0f;
                const kconst_float* aw = wts;
#pragma unroll 8
                for (uint32_t t = 0; t < dc.audio_taps; ++t, ++aw) {
                    const float term = det[rel - t] * aw[0];
                    acc = acc + term;
                }
                audio[o] = acc * dp.gain;
            }
        }
        __syncthreads();  // the tile's LDS is free and item_sh may be rewritten
    }
}

}  // namespace

extern "C" {

int32_t rvx_vrx_desc_check(const rvx_vrx_desc* desc) {
    static const char* who = "rvx_vrx_desc_check";
    if (desc == nullptr) return rvx::fail(RVX_EINVAL, "%s: null descriptor", who);
    const rvx_vrx_desc& d = *desc;
    if (int32_t rc = rvx::vrx_fine_valid(d.fine_config, d.fine, who)) return rc;
    rvx_vrx_demod_params dp = d.demod;
    const uint32_t hist = rvx::demod_fine_history(d.demod_config);
    if (d.hist != hist)
        return rvx::fail(RVX_EINVAL, "%s: hist %u, the demodulator reaches %u", who, d.hist, hist);
    // The fine block must hold every fine sample the audio outputs reach.
    const uint64_t need =
        d.demod.count == 0 ? 0 : uint64_t(d.demod.count - 1) * d.demod_config.decimation + hist + 1;
    if (d.fine.count < need)
        return rvx::fail(RVX_EINVAL, "%s: fine block of %u outputs, the audio needs %llu", who,
                         d.fine.count, (unsigned long long)need);
    dp.in_mask = 0xFFFFFFFFu;  // the fine block is on chip: no ring to alias
    if (int32_t rc = rvx::vrx_demod_valid(d.demod_config, dp, who)) return rc;
    if (d.tile == 0) return rvx::fail(RVX_EINVAL, "%s: zero tile", who);
    const uint64_t span = uint64_t(d.tile - 1) * d.demod_config.decimation + hist + 1;
    if (span > kMaxSpan)
        return rvx::fail(RVX_EINVAL, "%s: a tile of %u outputs needs %llu fine samples; at most %u fit",
                         who, d.tile, (unsigned long long)span, kMaxSpan);
    if (!d.channel_ring || !d.taps || !d.weights || !d.audio)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    return RVX_OK;
}

}  // extern "C"

namespace {

// The queue counter runs from 0 to n_work + workgroups - 1 in one uint32.
constexpr uint32_t kMaxWorkItems = 0x7FFFFFFFu;

int32_t launch_batch(rvx_ctx* ctx, const rvx_vrx_desc* descs, uint32_t n_desc,
                     const rvx_vrx_work* work, uint32_t n_work, uint32_t* counter,
                     const void* nco, uint64_t n_nco, uint32_t workgroups, const char* who) {
    if (!descs || !work || !counter || !nco || n_desc == 0)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    if (n_work > kMaxWorkItems)
        return rvx::fail(RVX_EINVAL, "%s: %u work items exceed the queue's %u", who, n_work,
                         kMaxWorkItems);
    if (!rvx::pow2(n_nco) || n_nco > (1u << 24))
        return rvx::fail(RVX_EINVAL, "%s: NCO table of %llu entries", who, (unsigned long long)n_nco);
    if (workgroups == 0) {
        int cus = 0;
        RVX_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
                    who);
        workgroups = uint32_t(std::max(1, cus)) * 4;  // four workgroups fit a CU
    }
    workgroups = std::min(workgroups, n_work);
    zero_word<<<dim3(1), dim3(1), 0, ctx->stream>>>(counter);
    vrx_batch_kernel<<<dim3(workgroups), dim3(kBatchThreads), 0, ctx->stream>>>(
        descs, n_desc, work, n_work, counter, static_cast<const float2*>(nco), rvx::log2u(n_nco));
    return rvx::check_launch(who);
}

// Each device's address of `scratch`, resolved once with that device current.
constexpr int kMaxDevices = 64;
std::mutex scratch_lock;
uint32_t* scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace rvx {

int32_t batch_scratch_setup(int device, const char* who) {
    if (device < 0 || device >= kMaxDevices)
        return fail(RVX_EINVAL, "%s: device %d beyond the %d the library tracks", who, device,
                    kMaxDevices);
    std::lock_guard<std::mutex> hold(scratch_lock);
    if (scratch_ptr[device]) return RVX_OK;
    void* ptr = nullptr;
    // `scratch` names a host-side stand-in; the device's copy is at the
    // address the runtime reports for the current device.
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || ptr == nullptr)
        return fail(RVX_EDEVICE, "%s: hipGetSymbolAddress(scratch) failed: %s", who,
                    hipGetErrorString(err));
    scratch_ptr[device] = static_cast<uint32_t*>(ptr);
    return RVX_OK;
}

}  // namespace rvx

extern "C" {

int32_t rvx_vrx_batch(rvx_ctx* ctx, const rvx_vrx_desc* descs, uint32_t n_desc,
                      const rvx_vrx_work* work, uint32_t n_work, uint32_t* counter,
                      const void* nco, uint64_t n_nco, uint32_t workgroups) {
    static const char* who = "rvx_vrx_batch";
    RVX_ENTER(ctx, who);
    if (n_work == 0) return RVX_OK;
    return launch_batch(ctx, descs, n_desc, work, n_work, counter, nco, n_nco, workgroups, who);
}

int32_t rvx_vrx_batch_noalloc(rvx_ctx* ctx, const rvx_vrx_desc* descs, uint32_t n_desc,
                              const rvx_vrx_work* work, uint32_t n_work, const void* nco,
                              uint64_t n_nco, uint32_t workgroups) {
    static const char* who = "rvx_vrx_batch_noalloc";
    RVX_ENTER(ctx, who);
    if (n_work == 0) return RVX_OK;
    // set up by rvx_ctx_create; resolved here if this device's context came
    // from elsewhere (RVX_ENTER made the device current)
    if (int32_t rc = rvx::batch_scratch_setup(ctx->device, who)) return rc;
    uint32_t* counter;
    {
        std::lock_guard<std::mutex> hold(scratch_lock);
        counter = scratch_ptr[ctx->device];
    }
    return launch_batch(ctx, descs, n_desc, work, n_work, counter, nco, n_nco, workgroups, who);
}

}  // extern "C"
