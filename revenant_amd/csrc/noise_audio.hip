// The per-receiver audio noise stages: the line kernel (manual notch cascade
// and adaptive line enhancer) and spectral noise reduction.
//
//   rvx_noise_line       noise_line.comp / noise_reference.cpp:231-421
//   rvx_noise_spectral   noise_spectral.comp / :556-700
//   rvx_design_notch, rvx_notch_section_db      :424-478
//   rvx_spectral_frame_for, rvx_design_spectral_tables, rvx_spectral_params_for  :493-553
//
// Both are recursions over one receiver's audio, so a receiver is one
// workgroup and a batch launch runs many receivers side by side.
//
// Line: one wavefront. The history ring and the weights are staged in LDS
// for the dispatch; the notch sections (and the guard) run on every lane as
// uniform values; lane l owns taps l, l + lanes, ... and sums its products in
// ascending tap order, then every lane folds the lanes' partial sums in
// ascending lane order through readlane, the twin's order, without a trip
// through LDS. Audio is staged through LDS 256 samples at a time so the
// per-sample loop reads no global memory.
//
// Spectral: frame N threads. Streaming a hop in and the queued output out is
// one sample per thread; at a full hop the analysis window, the forward DFT
// (one bin per thread, a direct sum in ascending time order over the LDS
// tables), the smoothed power, Doblinger's minimum and the subtraction gain
// (per bin), the inverse DFT and synthesis window (one sample per thread) and
// the overlap-add each run between barriers, every element the twin's
// operations in its order.

#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "common.h"
#include "vrx_math.h"

namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kChunk = 256;  // audio samples staged per step
// taps + history, floats: with the audio chunk, 64 KiB of LDS
constexpr uint32_t kMaxLineLds = 16384 - kChunk;
constexpr uint32_t kMoreSections = RVX_MAX_NOTCHES - 1;
constexpr double kPi = 3.141592653589793238462643383279502884;

// noise_reference.cpp:67-80, the spectral stage's fixed figures
constexpr double kSpectralSmoothing = 0.7;
constexpr double kSpectralGamma = 0.998;
constexpr double kSpectralBeta = 0.96;
constexpr double kSpectralBias = 2.4;
constexpr double kSpectralEps = 1.0e-12;

__host__ __device__ inline uint32_t notch_sections(uint32_t flags) {
    return (flags & RVX_LINE_NOTCH) == 0
               ? 0
               : 1 + ((flags & RVX_LINE_NOTCH_COUNT_MASK) >> RVX_LINE_NOTCH_COUNT_SHIFT);
}

__device__ __forceinline__ float section(float x, float b0, float b1, float b2, float a1, float a2,
                                         float& x1, float& x2, float& y1, float& y2) {
    const float t0 = b0 * x;
    const float t1 = b1 * x1;
    const float t2 = b2 * x2;
    const float t3 = a1 * y1;
    const float t4 = a2 * y2;
    float acc = t0 + t1;
    acc = acc + t2;
    acc = acc - t3;
    acc = acc - t4;
    x2 = x1;
    x1 = x;
    y2 = y1;
    y1 = acc;
    return acc;
}

__device__ __forceinline__ float lane_value(float v, uint32_t lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), int(lane)));
}

// kLanes / kPerLane nonzero: the reference's default shape compiled in, so
// the per-tap loops and the lane fold unroll; zero: read from the config.
template <uint32_t kLanes, uint32_t kPerLane>
__global__ void __launch_bounds__(kWave)
line_kernel(rvx_line_config c, const rvx_noise_line_item* __restrict__ items,
            rvx_noise_line_item one) {
    extern __shared__ float lds[];
    const rvx_noise_line_item it = items ? items[blockIdx.x] : one;
    const rvx_line_params& p = it.params;
    float* __restrict__ state = it.state;
    float* __restrict__ audio = it.audio;
    float* const weights = lds;
    float* const history = lds + c.taps;
    float* const chunk = history + c.history;
    const uint32_t lane = threadIdx.x;
    const bool ale = (p.flags & RVX_LINE_ALE) != 0;
    const bool guard = ale && (p.flags & RVX_LINE_GUARD) != 0;
    const uint32_t sections = notch_sections(p.flags);
    const uint32_t lanes = kLanes ? kLanes : c.lanes;
    const uint32_t per_lane = kPerLane ? kPerLane : c.taps / c.lanes;
    const uint32_t hmask = c.history - 1;

    if (ale) {
        for (uint32_t i = lane; i < c.taps + c.history; i += kWave)
            lds[i] = state[RVX_LINE_STATE_WEIGHTS + i];
    }
    float x1 = state[0], x2 = state[1], y1 = state[2], y2 = state[3];
    float more[kMoreSections][4];  // registers: every index below is a constant
#pragma unroll
    for (uint32_t s = 0; s < kMoreSections; ++s)
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
            more[s][k] = s + 1 < sections ? state[RVX_LINE_STATE_MORE + 4 * s + k] : 0.0f;
    uint32_t position = uint32_t(state[RVX_LINE_STATE_POSITION]);
    float gx1 = state[RVX_LINE_STATE_GUARD + 0], gx2 = state[RVX_LINE_STATE_GUARD + 1];
    float gy1 = state[RVX_LINE_STATE_GUARD + 2], gy2 = state[RVX_LINE_STATE_GUARD + 3];

    for (uint32_t n0 = 0; n0 < p.count; n0 += kChunk) {
        const uint32_t here = min(kChunk, p.count - n0);
        for (uint32_t i = lane; i < here; i += kWave) chunk[i] = audio[n0 + i];
        __syncthreads();
        for (uint32_t i = 0; i < here; ++i) {
            const float x = chunk[i];
            float v = x;
            if (sections != 0) {
                v = section(x, p.b0, p.b1, p.b2, p.a1, p.a2, x1, x2, y1, y2);
#pragma unroll
                for (uint32_t s = 0; s < kMoreSections; ++s) {
                    if (s + 1 < sections) {
                        const rvx_line_section& q = p.more[s];
                        v = section(v, q.b0, q.b1, q.b2, q.a1, q.a2, more[s][0], more[s][1],
                                    more[s][2], more[s][3]);
                    }
                }
            }
            float out = v;
            if (ale) {
                float py = 0.0f, pp = 0.0f;
                if (lane < lanes) {
#pragma unroll
                    for (uint32_t j = 0; j < per_lane; ++j) {
                        const uint32_t k = lane + j * lanes;
                        const float r = history[(position - p.delay - k * p.stride) & hmask];
                        const float wy = weights[k] * r;
                        py = py + wy;
                        const float rr = r * r;
                        pp = pp + rr;
                    }
                }
                float y = 0.0f, pw = 0.0f;
#pragma unroll
                for (uint32_t l = 0; l < lanes; ++l) {
                    y = y + lane_value(py, l);
                    pw = pw + lane_value(pp, l);
                }
                const float e = v - y;
                const float g = p.mu * rvx::det_recip(pw + p.eps);
                const float ge = g * e;
                if (lane < lanes) {
#pragma unroll
                    for (uint32_t j = 0; j < per_lane; ++j) {
                        const uint32_t k = lane + j * lanes;
                        const float r = history[(position - p.delay - k * p.stride) & hmask];
                        const float kept = p.leak * weights[k];
                        const float step = ge * r;
                        weights[k] = kept + step;
                    }
                }
                __syncthreads();  // every lane's reads of the ring precede the write
                if (lane == 0) history[position & hmask] = v;
                out = e;
                if (guard) {
                    const float acc =
                        section(y, p.gb0, p.gb1, p.gb2, p.ga1, p.ga2, gx1, gx2, gy1, gy2);
                    out = v - acc;
                }
                __syncthreads();  // the write is seen by the next sample's reads
            }
            if (lane == 0) chunk[i] = out;
            position = position + 1u;
        }
        __syncthreads();
        for (uint32_t i = lane; i < here; i += kWave) audio[n0 + i] = chunk[i];
        __syncthreads();  // the chunk is refilled next
    }

    if (lane == 0) {
        state[0] = x1;
        state[1] = x2;
        state[2] = y1;
        state[3] = y2;
#pragma unroll
        for (uint32_t s = 0; s < kMoreSections; ++s)
            if (s + 1 < sections)
#pragma unroll
                for (uint32_t k = 0; k < 4; ++k)
                    state[RVX_LINE_STATE_MORE + 4 * s + k] = more[s][k];
        state[RVX_LINE_STATE_POSITION] = float(position & hmask);
        state[RVX_LINE_STATE_GUARD + 0] = gx1;
        state[RVX_LINE_STATE_GUARD + 1] = gx2;
        state[RVX_LINE_STATE_GUARD + 2] = gy1;
        state[RVX_LINE_STATE_GUARD + 3] = gy2;
    }
    if (ale) {
        for (uint32_t i = lane; i < c.taps + c.history; i += kWave)
            state[RVX_LINE_STATE_WEIGHTS + i] = lds[i];
    }
}

// ---------------------------------------------------------------------------

struct SpectralLds {
    float cosine[RVX_MAX_SPECTRAL_FRAME], sine[RVX_MAX_SPECTRAL_FRAME];
    float analysis[RVX_MAX_SPECTRAL_FRAME], synthesis[RVX_MAX_SPECTRAL_FRAME];
    float hist[RVX_MAX_SPECTRAL_FRAME], work[RVX_MAX_SPECTRAL_FRAME];
    float tail[RVX_MAX_SPECTRAL_FRAME / 2], queue[RVX_MAX_SPECTRAL_FRAME / 2];
    float smoothed[RVX_MAX_SPECTRAL_FRAME / 2 + 1], minimum[RVX_MAX_SPECTRAL_FRAME / 2 + 1];
    float y_real[RVX_MAX_SPECTRAL_FRAME / 2 + 1], y_imag[RVX_MAX_SPECTRAL_FRAME / 2 + 1];
};

__global__ void __launch_bounds__(RVX_MAX_SPECTRAL_FRAME)
spectral_kernel(uint32_t n, const float* __restrict__ tables,
                const rvx_noise_spectral_item* __restrict__ items, rvx_noise_spectral_item one) {
    __shared__ SpectralLds m;
    const rvx_noise_spectral_item it = items ? items[blockIdx.x] : one;
    const rvx_spectral_params& p = it.params;
    float* __restrict__ audio = it.audio;
    float* __restrict__ state = it.state;
    const uint32_t h = n / 2, bins = h + 1, mask = n - 1;
    const uint32_t t = threadIdx.x;  // blockDim.x == n
    // state: [phase, seen, hist(n), tail(h), queue(h), smoothed(bins), minimum(bins)]
    const float* st_hist = state + 2;
    const float* st_tail = st_hist + n;
    const float* st_queue = st_tail + h;
    const float* st_smoothed = st_queue + h;
    const float* st_minimum = st_smoothed + bins;
    m.cosine[t] = tables[t];
    m.sine[t] = tables[n + t];
    m.analysis[t] = tables[2 * n + t];
    m.synthesis[t] = tables[3 * n + t];
    m.hist[t] = st_hist[t];
    if (t < h) {
        m.tail[t] = st_tail[t];
        m.queue[t] = st_queue[t];
    }
    if (t < bins) {
        m.smoothed[t] = st_smoothed[t];
        m.minimum[t] = st_minimum[t];
    }
    uint32_t phase = uint32_t(state[0]);
    float seen = state[1];
    __syncthreads();

    uint32_t s = 0;
    while (s < p.count) {
        const uint32_t length = min(h - phase, p.count - s);
        if (t < length) {
            const float x = audio[s + t];
            m.hist[h + phase + t] = x;
            audio[s + t] = m.queue[phase + t];
        }
        phase += length;
        s += length;
        if (phase != h) continue;
        __syncthreads();

        m.work[t] = m.hist[t] * m.analysis[t];
        __syncthreads();

        if (t < bins) {
            const uint32_t k = t;
            float re = 0.0f, im = 0.0f;
            for (uint32_t u = 0; u < n; ++u) {
                const uint32_t at = (k * u) & mask;
                const float xc = m.work[u] * m.cosine[at];
                re = re + xc;
                const float xs = m.work[u] * m.sine[at];
                im = im - xs;
            }
            const float rr = re * re;
            const float ii = im * im;
            const float power = rr + ii;
            float level = power;
            float floor_power = power;
            if (seen != 0.0f) {
                const float previous = m.smoothed[k];
                const float kept = p.smoothing * previous;
                const float fresh = p.smoothing_complement * power;
                level = kept + fresh;
                const float held = m.minimum[k];
                if (held < level) {
                    const float decayed = p.gamma * held;
                    const float lagged = p.beta * previous;
                    const float lift = level - lagged;
                    const float risen = p.rise * lift;
                    floor_power = decayed + risen;
                } else {
                    floor_power = level;
                }
            }
            m.smoothed[k] = level;
            m.minimum[k] = floor_power;
            const float noise = p.bias * floor_power;
            const float ratio = noise * rvx::det_recip(level + p.eps);
            const float cut = p.alpha * ratio;
            float gain = 1.0f - cut;
            gain = (gain < p.floor_gain) ? p.floor_gain : gain;  // std::max(gain, floor)
            m.y_real[k] = gain * re;
            m.y_imag[k] = gain * im;
        }
        __syncthreads();

        {
            float acc = 0.0f;
            for (uint32_t k = 1; k < h; ++k) {
                const uint32_t at = (k * t) & mask;
                const float a = m.y_real[k] * m.cosine[at];
                const float b = m.y_imag[k] * m.sine[at];
                const float term = a - b;
                acc = acc + term;
            }
            const float nyquist = (t & 1u) ? -m.y_real[h] : m.y_real[h];
            const float doubled = 2.0f * acc;
            float sum = m.y_real[0] + nyquist;
            sum = sum + doubled;
            m.work[t] = sum * m.synthesis[t];
        }
        __syncthreads();

        if (t < h) {
            m.queue[t] = m.tail[t] + m.work[t];
            m.tail[t] = m.work[h + t];
            m.hist[t] = m.hist[h + t];
        }
        __syncthreads();
        phase = 0;
        const float next = seen + 1.0f;
        seen = (2.0f < next) ? 2.0f : next;  // std::min(seen + 1, 2)
    }
    __syncthreads();

    float* out_hist = state + 2;
    float* out_tail = out_hist + n;
    float* out_queue = out_tail + h;
    float* out_smoothed = out_queue + h;
    float* out_minimum = out_smoothed + bins;
    out_hist[t] = m.hist[t];
    if (t < h) {
        out_tail[t] = m.tail[t];
        out_queue[t] = m.queue[t];
    }
    if (t < bins) {
        out_smoothed[t] = m.smoothed[t];
        out_minimum[t] = m.minimum[t];
    }
    if (t == 0) {
        state[0] = float(phase);
        state[1] = seen;
    }
}

int32_t line_valid(const rvx_line_config& c, const rvx_line_params& p, const char* who) {
    using rvx::fail;
    if (c.lanes == 0 || c.taps == 0 || c.taps % c.lanes != 0)
        return fail(RVX_EINVAL, "%s: %u taps do not divide over %u lanes", who, c.taps, c.lanes);
    if (!rvx::pow2(c.history))
        return fail(RVX_EINVAL, "%s: a history of %u is not a power of two", who, c.history);
    if (p.stride == 0 || p.stride > RVX_MAX_LINE_STRIDE)
        return fail(RVX_EINVAL, "%s: stride %u is outside [1, %u]", who, p.stride,
                    RVX_MAX_LINE_STRIDE);
    if (p.delay == 0)
        return fail(RVX_EINVAL, "%s: the predictor's delay must be at least one sample, or it "
                    "predicts the sample it is given", who);
    const uint64_t reach = uint64_t(p.delay) + uint64_t(c.taps - 1) * p.stride;
    if (reach >= c.history)
        return fail(RVX_EINVAL, "%s: a predictor reaching %llu samples back does not fit a %u "
                    "sample history", who, (unsigned long long)reach, c.history);
    if (!(p.eps > 0.0f))
        return fail(RVX_EINVAL, "%s: the normalisation's eps must be positive, or silence "
                    "divides by zero", who);
    if ((p.flags & ~(RVX_LINE_NOTCH | RVX_LINE_ALE | RVX_LINE_GUARD | RVX_LINE_NOTCH_COUNT_MASK)) !=
        0)
        return fail(RVX_EINVAL, "%s: flags 0x%x set a bit the kernel does not read", who, p.flags);
    // this kernel's own limits: one wavefront of lanes, its LDS copy of the state
    if (c.lanes > kWave)
        return fail(RVX_EINVAL, "%s: %u lanes is more than the one wavefront (%u) this kernel "
                    "sums over", who, c.lanes, kWave);
    if (uint64_t(c.taps) + c.history > kMaxLineLds)
        return fail(RVX_EINVAL, "%s: %u taps and a %u sample history exceed the %u floats the "
                    "kernel keeps in LDS", who, c.taps, c.history, kMaxLineLds);
    return RVX_OK;
}

size_t line_lds_bytes(const rvx_line_config& c) {
    return (size_t(c.taps) + c.history + kChunk) * sizeof(float);
}

void launch_line(const rvx_line_config& c, uint32_t groups, const rvx_noise_line_item* items,
                 const rvx_noise_line_item& one, hipStream_t stream) {
    const size_t lds = line_lds_bytes(c);
    if (c.lanes == 32 && c.taps == 128)
        line_kernel<32, 4><<<dim3(groups), dim3(kWave), lds, stream>>>(c, items, one);
    else
        line_kernel<0, 0><<<dim3(groups), dim3(kWave), lds, stream>>>(c, items, one);
}

int32_t spectral_valid(uint32_t frame, const rvx_spectral_params& p, const char* who) {
    if (!rvx::pow2(frame) || frame < RVX_MIN_SPECTRAL_FRAME || frame > RVX_MAX_SPECTRAL_FRAME)
        return rvx::fail(RVX_EINVAL, "%s: a frame of %u is not a power of two in [%u, %u]", who,
                         frame, RVX_MIN_SPECTRAL_FRAME, RVX_MAX_SPECTRAL_FRAME);
    if (!(p.eps > 0.0f))
        return rvx::fail(RVX_EINVAL, "%s: eps must be positive, or a silent bin divides by zero",
                         who);
    return RVX_OK;
}

}  // namespace

extern "C" {

// ---- line ------------------------------------------------------------------

int32_t rvx_design_notch(double audio_hz, double depth_db, double width_hz, int64_t audio_rate,
                         double* out5) {
    static const char* who = "rvx_design_notch";
    if (!out5) return rvx::fail(RVX_EINVAL, "%s: null output", who);
    if (audio_rate <= 0) return rvx::fail(RVX_EINVAL, "%s: no audio rate to design at", who);
    const double fa = double(audio_rate);
    if (!(audio_hz > 0.0) || !(audio_hz < 0.5 * fa))
        return rvx::fail(RVX_EINVAL, "%s: %g Hz is not inside the audio band of a %lld S/s stream",
                         who, audio_hz, (long long)audio_rate);
    if (!(width_hz > 0.0) || !(width_hz < 0.25 * fa))
        return rvx::fail(RVX_EINVAL, "%s: a %g Hz width does not fit a %lld S/s stream", who,
                         width_hz, (long long)audio_rate);
    // zero pair and pole pair on the notch angle; poles set the width, zeros the depth
    const double omega = 2.0 * kPi * audio_hz / fa;
    const double rp = 1.0 - kPi * width_hz / fa;
    const double rz = 1.0 - (1.0 - rp) * std::pow(10.0, -depth_db / 20.0);
    const double c = std::cos(omega);
    double b0 = 1.0, b1 = -2.0 * rz * c, b2 = rz * rz;
    const double a1 = -2.0 * rp * c, a2 = rp * rp;
    // unit gain at the band edge (z = -1 Nyquist, z = 1 DC) further from the notch
    const double z = (omega < 0.5 * kPi) ? -1.0 : 1.0;
    const double numerator = b0 + b1 * z + b2;
    const double denominator = 1.0 + a1 * z + a2;
    const double scale = denominator / numerator;
    b0 *= scale;
    b1 *= scale;
    b2 *= scale;
    out5[0] = b0;
    out5[1] = b1;
    out5[2] = b2;
    out5[3] = a1;
    out5[4] = a2;
    return RVX_OK;
}

double rvx_notch_section_db(const double* d, double audio_hz, int64_t audio_rate) {
    if (!d) return std::nan("");
    auto r = [](double v) { return double(float(v)); };
    const double b0 = r(d[0]), b1 = r(d[1]), b2 = r(d[2]), a1 = r(d[3]), a2 = r(d[4]);
    const double omega = 2.0 * kPi * audio_hz / double(std::max<int64_t>(audio_rate, 1));
    // H(e^jw) with z^-1 = e^-jw
    const double c1 = std::cos(-omega), s1 = std::sin(-omega);
    const double c2 = c1 * c1 - s1 * s1, s2 = 2.0 * c1 * s1;
    const double nr = b0 + b1 * c1 + b2 * c2, ni = b1 * s1 + b2 * s2;
    const double dr = 1.0 + a1 * c1 + a2 * c2, di = a1 * s1 + a2 * s2;
    const double power = (nr * nr + ni * ni) / std::max(dr * dr + di * di, 1.0e-300);
    return 10.0 * std::log10(std::max(power, 1.0e-30));
}

uint64_t rvx_line_state_size(const rvx_line_config* config) {
    if (!config) return 0;
    return uint64_t(RVX_LINE_STATE_WEIGHTS) + config->taps + config->history;
}

int32_t rvx_line_check(const rvx_line_config* config, const rvx_line_params* params) {
    if (!config || !params) return rvx::fail(RVX_EINVAL, "rvx_line_check: null argument");
    return line_valid(*config, *params, "rvx_line_check");
}

int32_t rvx_noise_line_item_check(const rvx_line_config* config, const rvx_noise_line_item* item,
                                  uint64_t n_audio, uint64_t n_state) {
    static const char* who = "rvx_noise_line_item_check";
    if (!config || !item) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (int32_t rc = line_valid(*config, item->params, who)) return rc;
    if (n_audio < item->params.count)
        return rvx::fail(RVX_EINVAL, "%s: %u samples asked of a %llu sample buffer", who,
                         item->params.count, (unsigned long long)n_audio);
    if (n_state < rvx_line_state_size(config))
        return rvx::fail(RVX_EINVAL, "%s: a %llu float state where %llu are needed", who,
                         (unsigned long long)n_state,
                         (unsigned long long)rvx_line_state_size(config));
    if (!item->state || (item->params.count && !item->audio))
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    return RVX_OK;
}

int32_t rvx_noise_line(rvx_ctx* ctx, const rvx_line_config* config, const rvx_line_params* params,
                       float* audio, uint64_t n_audio, float* state, uint64_t n_state) {
    static const char* who = "rvx_noise_line";
    RVX_ENTER(ctx, who);
    if (!config || !params) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    rvx_noise_line_item one{};
    one.params = *params;
    one.audio = audio;
    one.state = state;
    if (int32_t rc = rvx_noise_line_item_check(config, &one, n_audio, n_state)) return rc;
    launch_line(*config, 1, nullptr, one, ctx->stream);
    return rvx::check_launch(who);
}

int32_t rvx_noise_line_batch(rvx_ctx* ctx, const rvx_line_config* config,
                             const rvx_noise_line_item* items, uint32_t n_items) {
    static const char* who = "rvx_noise_line_batch";
    RVX_ENTER(ctx, who);
    if (!config) return rvx::fail(RVX_EINVAL, "%s: null config", who);
    if (n_items == 0) return RVX_OK;
    if (!items) return rvx::fail(RVX_EINVAL, "%s: null items", who);
    rvx_line_params probe{};
    probe.stride = 1;
    probe.delay = 1;
    probe.eps = 1.0f;
    if (int32_t rc = line_valid(*config, probe, who)) return rc;  // the shared config
    launch_line(*config, n_items, items, rvx_noise_line_item{}, ctx->stream);
    return rvx::check_launch(who);
}

// ---- spectral ----------------------------------------------------------------

uint32_t rvx_spectral_frame_for(int64_t audio_rate) {
    const uint64_t wanted = audio_rate > 0 ? uint64_t(audio_rate) / 90u : RVX_MIN_SPECTRAL_FRAME;
    uint64_t floor = 1;
    while (floor * 2 <= std::max<uint64_t>(wanted, 1)) floor *= 2;
    return uint32_t(std::clamp<uint64_t>(floor, RVX_MIN_SPECTRAL_FRAME, RVX_MAX_SPECTRAL_FRAME));
}

uint64_t rvx_spectral_state_size(uint32_t frame) {
    const uint64_t n = frame;
    return 2 + n + n / 2 + n / 2 + 2 * (n / 2 + 1);
}

int32_t rvx_design_spectral_tables(uint32_t frame, float* out, uint64_t n_floats) {
    static const char* who = "rvx_design_spectral_tables";
    if (!rvx::pow2(frame) || frame < RVX_MIN_SPECTRAL_FRAME || frame > RVX_MAX_SPECTRAL_FRAME)
        return rvx::fail(RVX_EINVAL, "%s: a frame of %u is not a power of two in [%u, %u]", who,
                         frame, RVX_MIN_SPECTRAL_FRAME, RVX_MAX_SPECTRAL_FRAME);
    if (!out || n_floats != 4ull * frame)
        return rvx::fail(RVX_EINVAL, "%s: wants %u floats", who, 4 * frame);
    const uint32_t n = frame;
    const double step = 2.0 * kPi / double(n);
    for (uint32_t k = 0; k < n; ++k) {
        double c = std::cos(step * double(k)), s = std::sin(step * double(k));
        // the quarter and half points exactly, so the table is exactly odd about N/2
        if (k == 0) {
            c = 1.0;
            s = 0.0;
        } else if (4 * k == n) {
            c = 0.0;
            s = 1.0;
        } else if (2 * k == n) {
            c = -1.0;
            s = 0.0;
        } else if (4 * k == 3 * n) {
            c = 0.0;
            s = -1.0;
        }
        out[k] = float(c);
        out[n + k] = float(s);
        // square-root Hann offset half a sample: w^2(k) + w^2(k + N/2) = 1
        const double w = std::sin(kPi * (double(k) + 0.5) / double(n));
        out[2 * n + k] = float(w);
        out[3 * n + k] = float(w / double(n));
    }
    return RVX_OK;
}

int32_t rvx_spectral_params_for(double strength, rvx_spectral_params* out) {
    if (!out) return rvx::fail(RVX_EINVAL, "rvx_spectral_params_for: null output");
    const double s = std::clamp(strength, 0.0, 1.0);
    rvx_spectral_params p{};
    p.smoothing = float(kSpectralSmoothing);
    p.smoothing_complement = float(1.0 - kSpectralSmoothing);
    p.gamma = float(kSpectralGamma);
    p.beta = float(kSpectralBeta);
    p.rise = float((1.0 - kSpectralGamma) / (1.0 - kSpectralBeta));
    p.bias = float(kSpectralBias);
    p.alpha = float(1.0 + s);
    p.floor_gain = float(std::pow(10.0, -(6.0 + 12.0 * s) / 20.0));
    p.eps = float(kSpectralEps);
    *out = p;
    return RVX_OK;
}

int32_t rvx_spectral_check(uint32_t frame, const rvx_spectral_params* params) {
    if (!params) return rvx::fail(RVX_EINVAL, "rvx_spectral_check: null params");
    return spectral_valid(frame, *params, "rvx_spectral_check");
}

int32_t rvx_noise_spectral_item_check(uint32_t frame, const rvx_noise_spectral_item* item,
                                      uint64_t n_audio, uint64_t n_state) {
    static const char* who = "rvx_noise_spectral_item_check";
    if (!item) return rvx::fail(RVX_EINVAL, "%s: null item", who);
    if (int32_t rc = spectral_valid(frame, item->params, who)) return rc;
    if (n_audio < item->params.count)
        return rvx::fail(RVX_EINVAL, "%s: %u samples asked of a %llu sample buffer", who,
                         item->params.count, (unsigned long long)n_audio);
    if (n_state < rvx_spectral_state_size(frame))
        return rvx::fail(RVX_EINVAL, "%s: a %llu float state where %llu are needed", who,
                         (unsigned long long)n_state,
                         (unsigned long long)rvx_spectral_state_size(frame));
    if (!item->state || (item->params.count && !item->audio))
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    return RVX_OK;
}

int32_t rvx_noise_spectral(rvx_ctx* ctx, uint32_t frame, const rvx_spectral_params* params,
                           const float* tables, uint64_t n_tables, float* audio, uint64_t n_audio,
                           float* state, uint64_t n_state) {
    static const char* who = "rvx_noise_spectral";
    RVX_ENTER(ctx, who);
    if (!params) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    rvx_noise_spectral_item one{};
    one.params = *params;
    one.audio = audio;
    one.state = state;
    if (int32_t rc = rvx_noise_spectral_item_check(frame, &one, n_audio, n_state)) return rc;
    if (!tables || n_tables < 4ull * frame)
        return rvx::fail(RVX_EINVAL, "%s: %llu table entries where %u are needed", who,
                         (unsigned long long)n_tables, 4 * frame);
    spectral_kernel<<<dim3(1), dim3(frame), 0, ctx->stream>>>(frame, tables, nullptr, one);
    return rvx::check_launch(who);
}

int32_t rvx_noise_spectral_batch(rvx_ctx* ctx, uint32_t frame, const float* tables,
                                 uint64_t n_tables, const rvx_noise_spectral_item* items,
                                 uint32_t n_items) {
    static const char* who = "rvx_noise_spectral_batch";
    RVX_ENTER(ctx, who);
    rvx_spectral_params probe{};
    probe.eps = 1.0f;
    if (int32_t rc = spectral_valid(frame, probe, who)) return rc;
    if (n_items == 0) return RVX_OK;
    if (!items || !tables || n_tables < 4ull * frame)
        return rvx::fail(RVX_EINVAL, "%s: null items or a short table", who);
    spectral_kernel<<<dim3(n_items), dim3(frame), 0, ctx->stream>>>(frame, tables, items,
                                                                    rvx_noise_spectral_item{});
    return rvx::check_launch(who);
}

}  // extern "C"
