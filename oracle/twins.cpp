// CPU oracle for the receive chain: a restatement of Revenant's CPU twins.
//
// TEST INFRASTRUCTURE. Only the test suite, the smoke check and the CPU
// baseline leg of the benchmark load this library, and only to check or time
// the GPU path. Nothing in the product calls it.
//
// Every function here fixes the same floating-point operation order as the
// reference twin it cites: no contraction (-ffp-contract=off),
// denormals flushed on input and output (MXCSR FTZ|DAZ, 0x8040, exactly the
// policy of core/dsp/denormal_mode.h and lw-radio-dsp primitives.cpp:319-348),
// and 32-bit index arithmetic that wraps where the kernels' does.
//
// Pinning (what checks each function against something other than itself) is
// listed beside each function. The tests named there live in tests/.

// Contraction is switched off by the build (-ffp-contract=off in oracle/Makefile;
// g++ ignores the standard FP_CONTRACT pragma).

#include <xmmintrin.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct FlushGuard {
    unsigned saved;
    FlushGuard() : saved(_mm_getcsr()) { _mm_setcsr(saved | 0x8040U); }
    ~FlushGuard() { _mm_setcsr(saved); }
};

inline float bits_to_float(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
inline uint32_t float_to_bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}
inline bool pow2(uint64_t v) { return v != 0 && (v & (v - 1)) == 0; }

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Parameter blocks, byte-identical to the reference's POD push-constant
// structs (sizes static_asserted there and here).
// ---------------------------------------------------------------------------

struct OConvertParams {  // core/dsp/convert.h:116-141
    uint32_t ring_mask, dst_offset, src_offset, count;
};
struct OGrid {  // core/dsp/pfb.h:32-74
    uint32_t channels, taps_per_branch, decimation;
};
struct OBranchParams {  // core/dsp/pfb_branch_reference.h:337-360
    uint32_t ring_mask, base_offset, block_count;
};
struct OFftParams {  // core/dsp/pfb_fft_reference.h:42-82
    uint32_t channels, decimation, stages, block_base, block_count, out_ring_blocks,
        out_ring_mask;
};
struct OSpectrumParams {  // core/dsp/spectrum_reference.h:55-79
    uint32_t channels, transform, stages, chan_blocks, chan_mask, in_offset;
};
struct OLevelsParams {  // core/dsp/spectrum_levels_reference.h
    uint32_t bins, low_permille, high_permille;
};
// The exponentially averaged power spectrum. The reference has no GPU
// operator for it (its average is the host detector's, detector.cpp:621-627),
// so this block and its operation order are this project's definition.
struct OEmaParams {
    uint32_t channels, chan_blocks, chan_mask, in_offset, frame_blocks, frames;
    float alpha, frame_decay;
    uint32_t row_frames;
    float row_decay;
};
struct OFineConfig {  // core/dsp/vrx_reference.h:178-195
    uint32_t taps, phases, nco_log2;
};
struct OFineParams {  // core/dsp/vrx_reference.h:232-271
    uint32_t chan_base, chan_mask, in_offset, out_offset, out_mask, count, step_whole, step_rem,
        out_rate, frac0, nco_phase_high, nco_phase_low, nco_delta_high, nco_delta_low;
    float inv_out_rate;
};
struct ODemodConfig {  // core/dsp/vrx_reference.h:506-544
    uint32_t mode, decimation, audio_taps, dc_taps, channels, pilot_taps;
};
struct ODemodParams {  // core/dsp/vrx_reference.h:569-584
    uint32_t in_mask, in_offset, count;
    float gain;
};

static_assert(sizeof(OConvertParams) == 16);
static_assert(sizeof(OBranchParams) == 12);
static_assert(sizeof(OFineParams) == 60);
static_assert(sizeof(ODemodParams) == 16);

// ---------------------------------------------------------------------------
// Format conversion. core/dsp/convert.cpp:126-290 (cu8_to_float .. cs24) and
// convert_*_cf32.comp. Words are assembled little-endian and fields extracted
// as the kernels do.
// Pinned by: every cu8/cs8 code and the cs16/cs24 ranges against the exact
// closed forms (tests/test_oracle_convert.py), the constant words
// 0x42FF0000 / 0x3C008081 and the +-1 endpoints (convert.h:82-105).
// ---------------------------------------------------------------------------

static uint32_t load_word(const uint8_t* p, uint64_t word) {
    const uint8_t* b = p + word * 4;
    return uint32_t(b[0]) | (uint32_t(b[1]) << 8) | (uint32_t(b[2]) << 16) | (uint32_t(b[3]) << 24);
}
static int32_t sext(uint32_t word, int offset, int bits) {
    const int spare = 32 - offset - bits;
    return int32_t(word << spare) >> (32 - bits);
}

int oracle_convert(int format, const uint8_t* packed, uint64_t packed_bytes,
                   const OConvertParams* p, float* ring, uint64_t ring_samples) {
    static const uint64_t bytes_per_sample[4] = {2, 2, 4, 6};
    if (format < 0 || format > 3) return -1;
    const uint64_t capacity = uint64_t(p->ring_mask) + 1;
    if ((capacity & p->ring_mask) != 0 || ring_samples < capacity || p->count > capacity) return -1;
    const uint64_t last = uint64_t(p->src_offset) + p->count;
    if (last > 0xFFFFFFFFull) return -1;
    const uint64_t words = (last * bytes_per_sample[format] + 3) / 4;
    if (packed_bytes < words * 4) return -1;
    if (p->count == 0) return 0;
    FlushGuard guard;
    const float kCu8Bias = bits_to_float(0x42FF0000u);   // 127.5
    const float kCu8Scale = bits_to_float(0x3C008081u);  // nearest float to 2/255
    const float kCs8Scale = bits_to_float(0x3C000000u);  // 2^-7
    const float kCs16Scale = bits_to_float(0x38000000u); // 2^-15
    const float kCs24Scale = bits_to_float(0x34000000u); // 2^-23
    for (uint32_t index = 0; index < p->count; ++index) {
        const uint32_t s = p->src_offset + index;
        float iv = 0.0f, qv = 0.0f;
        if (format == 0 || format == 1) {
            const uint32_t w = load_word(packed, s >> 1);
            const int shift = int((s & 1u) << 4);
            if (format == 0) {
                const uint32_t ic = (w >> shift) & 0xFFu;
                const uint32_t qc = (w >> (shift + 8)) & 0xFFu;
                iv = (float(ic) - kCu8Bias) * kCu8Scale;
                qv = (float(qc) - kCu8Bias) * kCu8Scale;
            } else {
                iv = float(sext(w, shift, 8)) * kCs8Scale;
                qv = float(sext(w, shift + 8, 8)) * kCs8Scale;
            }
        } else if (format == 2) {
            const uint32_t w = load_word(packed, s);
            iv = float(sext(w, 0, 16)) * kCs16Scale;
            qv = float(sext(w, 16, 16)) * kCs16Scale;
        } else {
            const uint32_t wi = s + (s >> 1);
            const uint32_t lo = load_word(packed, wi);
            const uint32_t hi = load_word(packed, uint64_t(wi) + 1);
            const uint32_t shift = (s & 1u) << 4;
            const uint32_t window = (lo >> shift) | ((hi << 16) << (16 - shift));
            const uint32_t qraw = (window >> 24) | ((hi >> shift) << 8);
            iv = float(sext(window, 0, 24)) * kCs24Scale;
            qv = float(sext(qraw, 0, 24)) * kCs24Scale;
        }
        const uint64_t slot = (p->dst_offset + index) & p->ring_mask;
        ring[2 * slot] = iv;
        ring[2 * slot + 1] = qv;
    }
    return 0;
}

// Complex multiply, core/dsp/complex_ops.cpp:33-60 (cmul.comp). Used as the
// device denormal probe and the plumbing check.
int oracle_cmul(const float* a, const float* b, float* out, uint64_t n) {
    FlushGuard guard;
    for (uint64_t i = 0; i < n; ++i) {
        const float ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        const float re = ar * br - ai * bi;
        const float im = ar * bi + ai * br;
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Grid validation, core/dsp/pfb_design.cpp:301-347, with the 2048-channel
// ceiling lifted: that ceiling was Vulkan's 16 KiB shared-memory floor, and
// the butterfly graph is the same at any power of two.
// ---------------------------------------------------------------------------

static bool grid_ok(const OGrid* g) {
    if (g->channels < 2 || !pow2(g->channels) || g->channels > 65536) return false;
    if (g->taps_per_branch == 0 || g->decimation == 0) return false;
    if (g->channels % g->decimation != 0) return false;
    if (uint64_t(g->channels) * g->taps_per_branch > (1u << 20)) return false;
    return true;
}

// ---------------------------------------------------------------------------
// Branch filter, core/dsp/pfb_branch_reference.cpp:22-147 (pfb_branch.comp).
// u_r[m] = sum over ascending q of h[qM+r] * x[(base - qM - r) & mask], each
// step acc = acc + s*c, written to slot (M - r) & (M - 1).
// Pinned by: the FP64 direct channelizer definition through the FFT
// (tests/test_oracle_pfb.py), the tone-in-its-channel test, and a brute-force
// single-branch sum on tiny grids.
// ---------------------------------------------------------------------------

int oracle_pfb_branch(const OGrid* g, const float* taps, uint64_t ntaps, const float* ring,
                      uint64_t ring_samples, const OBranchParams* p, float* out,
                      uint64_t out_samples) {
    if (!grid_ok(g)) return -1;
    const uint32_t M = g->channels, T = g->taps_per_branch, D = g->decimation;
    if (ntaps != uint64_t(M) * T) return -1;
    if (out_samples != uint64_t(p->block_count) * M) return -1;
    if (p->block_count == 0) return 0;
    const uint64_t capacity = uint64_t(p->ring_mask) + 1;
    if ((capacity & p->ring_mask) != 0 || ring_samples < capacity) return -1;
    if (uint64_t(M) * T + uint64_t(p->block_count - 1) * D > capacity) return -1;
    FlushGuard guard;
    for (uint32_t block = 0; block < p->block_count; ++block) {
        const uint32_t base = p->base_offset + block * D;
        for (uint32_t r = 0; r < M; ++r) {
            float acc_re = 0.0f, acc_im = 0.0f;
            for (uint32_t q = 0; q < T; ++q) {
                const uint32_t n = q * M + r;
                const uint64_t idx = (base - n) & p->ring_mask;
                const float c = taps[n];
                acc_re = acc_re + ring[2 * idx] * c;
                acc_im = acc_im + ring[2 * idx + 1] * c;
            }
            const uint32_t slot = (M - r) & (M - 1);
            const uint64_t o = uint64_t(block) * M + slot;
            out[2 * o] = acc_re;
            out[2 * o + 1] = acc_im;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Radix-2 DIT transform, core/dsp/pfb_fft_reference.cpp:54-104 (pfb_fft.comp
// stage loop). Bit-reversed load, log2 M stages, the six-statement butterfly.
// Pinned by: an FP64 brute-force DFT on sizes 2..4096 (tests), Parseval, and a
// unit impulse giving an exact all-ones spectrum.
// ---------------------------------------------------------------------------

static uint32_t reverse_bits(uint32_t v, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t b = 0; b < bits; ++b) r |= ((v >> b) & 1u) << (bits - 1 - b);
    return r;
}

static void transform(const float* in, const float* tw, float* s, uint32_t n, uint32_t stages) {
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t src = reverse_bits(i, stages);
        s[2 * i] = in[2 * src];
        s[2 * i + 1] = in[2 * src + 1];
    }
    for (uint32_t stage = 0; stage < stages; ++stage) {
        const uint32_t half = 1u << stage;
        const uint32_t step = n >> (stage + 1);
        for (uint32_t b = 0; b < n / 2; ++b) {
            const uint32_t j = b & (half - 1);
            const uint32_t ia = ((b >> stage) << (stage + 1)) + j;
            const uint32_t ib = ia + half;
            const float ar = s[2 * ia], ai = s[2 * ia + 1];
            const float br = s[2 * ib], bi = s[2 * ib + 1];
            const float wr = tw[2 * (j * step)], wi = tw[2 * (j * step) + 1];
            const float vr = br * wr - bi * wi;
            const float vi = br * wi + bi * wr;
            const float sr = ar + vr, si = ai + vi;
            const float dr = ar - vr, di = ai - vi;
            s[2 * ia] = sr;
            s[2 * ia + 1] = si;
            s[2 * ib] = dr;
            s[2 * ib + 1] = di;
        }
    }
}

static uint32_t log2u(uint32_t v) {
    uint32_t s = 0;
    while ((1u << s) < v) ++s;
    return s;
}

int oracle_fft_radix2(const float* tw, uint64_t ntw, const float* in, float* out, uint32_t size,
                      uint32_t batch) {
    if (size < 2 || !pow2(size) || ntw != size) return -1;
    FlushGuard guard;
    const uint32_t stages = log2u(size);
    for (uint32_t b = 0; b < batch; ++b)
        transform(in + 2 * uint64_t(b) * size, tw, out + 2 * uint64_t(b) * size, size, stages);
    return 0;
}

// core/dsp/pfb_fft_reference.cpp:106-141 (validate)
static bool fft_params_ok(const OFftParams* p) {
    if (p->channels < 2 || !pow2(p->channels)) return false;
    if (p->stages >= 32 || (1u << p->stages) != p->channels) return false;
    if (p->decimation == 0 || p->decimation > p->channels || p->channels % p->decimation) return false;
    if (!pow2(p->out_ring_blocks) || p->out_ring_mask != p->out_ring_blocks - 1) return false;
    if (p->block_count > p->out_ring_blocks) return false;
    return true;
}

// core/dsp/pfb_fft_reference.cpp:143-232: transform, then the output phase
// correction x W[(k * dm) & (M - 1)] with dm = ((m & (M-1)) * D) & (M-1)
// whenever D != M, written channel-major at k * out_ring_blocks + (m & mask).
int oracle_pfb_fft(const OFftParams* p, const float* branch, uint64_t nbranch, const float* tw,
                   uint64_t ntw, float* ring, uint64_t ring_samples) {
    if (!fft_params_ok(p)) return -1;
    const uint32_t M = p->channels;
    if (ntw != M || nbranch < uint64_t(p->block_count) * M) return -1;
    if (ring_samples < uint64_t(M) * p->out_ring_blocks) return -1;
    FlushGuard guard;
    std::vector<float> s(2 * size_t(M));
    for (uint32_t block = 0; block < p->block_count; ++block) {
        transform(branch + 2 * uint64_t(block) * M, tw, s.data(), M, p->stages);
        const uint32_t m_abs = p->block_base + block;
        const uint32_t slot = m_abs & p->out_ring_mask;
        const uint32_t dm = ((m_abs & (M - 1)) * p->decimation) & (M - 1);
        for (uint32_t k = 0; k < M; ++k) {
            float vr = s[2 * k], vi = s[2 * k + 1];
            if (p->decimation != M) {
                const uint32_t c = (k * dm) & (M - 1);
                const float cr = tw[2 * c], ci = tw[2 * c + 1];
                const float pr = vr * cr - vi * ci;
                const float pi = vr * ci + vi * cr;
                vr = pr;
                vi = pi;
            }
            const uint64_t o = uint64_t(k) * p->out_ring_blocks + slot;
            ring[2 * o] = vr;
            ring[2 * o + 1] = vi;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Deterministic transcendentals, core/dsp/vrx_reference.cpp:227-306 and
// spectrum_reference.cpp:123 (det_log2). Fixed Newton iterations and fixed
// polynomials, no library calls, so any IEEE float implementation agrees.
// Pinned by: libm log2/atan2/sqrt/1/x within stated ulp bounds on dense
// sweeps (tests/test_oracle_det.py).
// ---------------------------------------------------------------------------

static float det_recip(float v) {
    const float seed = bits_to_float(0x7EF311C3u - float_to_bits(v));
    const float r1 = seed * (2.0f - v * seed);
    const float r2 = r1 * (2.0f - v * r1);
    const float r3 = r2 * (2.0f - v * r2);
    const float r4 = r3 * (2.0f - v * r3);
    return r4;
}
static float det_rsqrt(float v) {
    const float seed = bits_to_float(0x5F375A86u - (float_to_bits(v) >> 1));
    const float h = 0.5f * v;
    const float r1 = seed * (1.5f - h * (seed * seed));
    const float r2 = r1 * (1.5f - h * (r1 * r1));
    const float r3 = r2 * (1.5f - h * (r2 * r2));
    const float r4 = r3 * (1.5f - h * (r3 * r3));
    return r4;
}
static float det_sqrt(float v) {
    if (!(v > 0.0f)) return 0.0f;
    return v * det_rsqrt(v);
}
static float det_atan_unit(float t) {
    static const float c[10] = {
        float(0.9999999998),  float(-0.3333331929), float(0.1999923872), float(-0.1427217912),
        float(0.1099615976),  float(-0.0853800324), float(0.0603537785), float(-0.0336503503),
        float(0.0122847765),  float(-0.0021090222)};
    const float u = t * t;
    float p = c[9];
    for (int i = 8; i >= 0; --i) p = p * u + c[i];
    return t * p;
}
static float det_atan2(float y, float x) {
    const float kHalfPi = float(1.57079632679489661923);
    const float kPi = float(3.14159265358979323846);
    const float ay = std::fabs(y), ax = std::fabs(x);
    if (ay == 0.0f && ax == 0.0f) return 0.0f;
    const float num = (ax < ay) ? ax : ay;
    const float den = (ay < ax) ? ax : ay;
    const float ratio = num * det_recip(den);
    float angle = det_atan_unit(ratio);
    if (ay > ax) angle = kHalfPi - angle;
    if (x < 0.0f) angle = kPi - angle;
    return (y < 0.0f) ? -angle : angle;
}
static float det_log2(float v) {
    const float kSqrt2 = 1.41421356237309505f;
    const uint32_t bits = float_to_bits(v);
    int e = int((bits >> 23) & 0xFFu) - 127;
    float m = bits_to_float((bits & 0x007FFFFFu) | 0x3F800000u);
    if (m > kSqrt2) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float t = (m - 1.0f) * det_recip(m + 1.0f);
    const float u = t * t;
    float p = 0.32059889797532520f;
    p = p * u + 0.41219858311113240f;
    p = p * u + 0.57707801635558536f;
    p = p * u + 0.96179669392597560f;
    p = p * u + 2.88539008177792681f;
    const float fraction = t * p;
    return float(e) + fraction;
}

int oracle_det(int which, const float* a, const float* b, float* out, uint64_t n) {
    FlushGuard guard;
    for (uint64_t i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = det_log2(a[i]); break;
            case 1: out[i] = det_recip(a[i]); break;
            case 2: out[i] = det_rsqrt(a[i]); break;
            case 3: out[i] = det_sqrt(a[i]); break;
            case 4: out[i] = det_atan2(a[i], b[i]); break;
            case 5: out[i] = det_atan_unit(a[i]); break;
            default: return -1;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Full-span spectrum, core/dsp/spectrum_reference.cpp:286-394 (spectrum.comp).
// Per channel: window x N samples from in_offset, N-point DIT, keep the central
// half (bin + N - N/4) & (N-1), p = re^2 + im^2, x correction, floor 1e-20,
// det_log2 * 3.0103 dB, written at slot (k + M/2) & (M-1).
// Pinned by: a tone landing in its bin, the unit coherent gain of the window
// and the det_log2 checks (tests/test_oracle_spectrum.py).
// ---------------------------------------------------------------------------

int oracle_spectrum(const OSpectrumParams* p, const float* ring, uint64_t ring_samples,
                    const float* tw, uint64_t ntw, const float* window, uint64_t nwindow,
                    float* frame, uint64_t nframe) {
    const uint32_t N = p->transform, M = p->channels;
    if (M < 2 || !pow2(M) || N < 4 || !pow2(N) || N > 65536) return -1;
    if (p->stages >= 32 || (1u << p->stages) != N) return -1;
    if (!pow2(p->chan_blocks) || p->chan_mask != p->chan_blocks - 1 || N > p->chan_blocks) return -1;
    if (ntw != N || nwindow != uint64_t(N) + N / 2) return -1;
    if (ring_samples < uint64_t(M) * p->chan_blocks) return -1;
    const uint32_t half = N / 2;
    if (nframe < uint64_t(M) * half) return -1;
    FlushGuard guard;
    const float kDbPerOctave = 3.01029995663981195f;
    std::vector<float> win(2 * size_t(N)), out(2 * size_t(N));
    for (uint32_t ch = 0; ch < M; ++ch) {
        const uint64_t origin = uint64_t(ch) * p->chan_blocks;
        for (uint32_t n = 0; n < N; ++n) {
            const uint64_t slot = (p->in_offset + n) & p->chan_mask;
            const float g = window[n];
            const float re = ring[2 * (origin + slot)] * g;
            const float im = ring[2 * (origin + slot) + 1] * g;
            win[2 * n] = re;
            win[2 * n + 1] = im;
        }
        transform(win.data(), tw, out.data(), N, p->stages);
        const uint64_t out_origin = uint64_t((ch + M / 2) & (M - 1)) * half;
        for (uint32_t bin = 0; bin < half; ++bin) {
            const uint32_t src = (bin + N - N / 4) & (N - 1);
            const float vr = out[2 * src], vi = out[2 * src + 1];
            const float power = vr * vr + vi * vi;
            const float flat = power * window[N + bin];
            const float floored = (flat < 1.0e-20f) ? 1.0e-20f : flat;
            frame[out_origin + bin] = det_log2(floored) * kDbPerOctave;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Spectrum levels, core/dsp/spectrum_levels_reference.cpp:38-120: a 1024
// bucket histogram over [-200, 40) dB and the buckets holding two permille
// ranks. Pinned by: bucket boundaries and an exact sort of the frame agreeing
// to within half a bucket (tests/test_oracle_spectrum.py).
// ---------------------------------------------------------------------------

static uint32_t level_bucket(float db) {
    float t = (db - (-200.0f)) * 4.26666666666666666f;
    t = (t < 0.0f) ? 0.0f : t;
    t = (t > 1023.0f) ? 1023.0f : t;
    return uint32_t(t);
}
static float level_value(uint32_t b) { return -200.0f + (float(b) + 0.5f) * 0.234375f; }
static uint32_t rank_of(uint32_t count, uint32_t permille) {
    return (count / 1000u) * permille + ((count % 1000u) * permille) / 1000u;
}

int oracle_spectrum_levels(const OLevelsParams* p, const float* frame, uint64_t nframe,
                           float* levels) {
    if (p->bins == 0 || p->low_permille > 1000 || p->high_permille > 1000 ||
        p->low_permille >= p->high_permille || nframe < p->bins)
        return -1;
    FlushGuard guard;
    std::vector<uint32_t> hist(1024, 0);
    for (uint32_t i = 0; i < p->bins; ++i) ++hist[level_bucket(frame[i])];
    const uint32_t lo_rank = std::min(rank_of(p->bins, p->low_permille), p->bins - 1);
    const uint32_t hi_rank = std::min(rank_of(p->bins, p->high_permille), p->bins - 1);
    uint32_t lo = 0, hi = 0, cum = 0;
    bool lo_found = false, hi_found = false;
    for (uint32_t b = 0; b < 1024; ++b) {
        cum += hist[b];
        if (!lo_found && cum > lo_rank) { lo = b; lo_found = true; }
        if (!hi_found && cum > hi_rank) { hi = b; hi_found = true; }
    }
    levels[0] = level_value(lo);
    levels[1] = level_value(hi);
    return 0;
}

// ---------------------------------------------------------------------------
// Exponentially averaged power spectrum (this project's operator; the
// reference's average lives in the host detector, detector.cpp:621-627).
//
// Partial: for each frame f of frame_blocks consecutive channel samples and
// each channel k, s starts at 0 and for ascending blocks
//     x = re*re + im*im;  s = s + (x - s) * alpha
// Row: S = 0 and S = S * frame_decay + s over a row's row_frames frames in
// ascending order. Combine: P = P * row_decay + S over ascending rows; P after
// row r is row r of the averaged spectrum. frame_decay = (1-alpha)^F and
// row_decay = (1-alpha)^(F*G), each rounded once to float.
// Pinned by: the FP64 exponential average of the same powers (relative error
// bound), a constant input converging to its value, and alpha = 1 reducing to
// the last power (tests/test_oracle_ema.py).
// ---------------------------------------------------------------------------

int oracle_ema_partial(const OEmaParams* p, const float* ring, uint64_t ring_samples,
                       float* partial, uint64_t npartial) {
    if (p->channels == 0 || !pow2(p->chan_blocks) || p->chan_mask != p->chan_blocks - 1) return -1;
    if (p->frame_blocks == 0 || uint64_t(p->frame_blocks) * p->frames > p->chan_blocks) return -1;
    if (ring_samples < uint64_t(p->channels) * p->chan_blocks) return -1;
    if (npartial < uint64_t(p->frames) * p->channels) return -1;
    FlushGuard guard;
    for (uint32_t f = 0; f < p->frames; ++f) {
        for (uint32_t k = 0; k < p->channels; ++k) {
            float s = 0.0f;
            for (uint32_t i = 0; i < p->frame_blocks; ++i) {
                const uint64_t slot = (p->in_offset + f * p->frame_blocks + i) & p->chan_mask;
                const uint64_t o = uint64_t(k) * p->chan_blocks + slot;
                const float re = ring[2 * o], im = ring[2 * o + 1];
                const float x = re * re + im * im;
                s = s + (x - s) * p->alpha;
            }
            partial[uint64_t(f) * p->channels + k] = s;
        }
    }
    return 0;
}

int oracle_ema_combine(const OEmaParams* p, const float* partial, uint64_t npartial, float* state,
                       float* rows, uint64_t nrows) {
    const uint32_t G = p->row_frames;
    if (G == 0 || p->frames % G != 0) return -1;
    const uint32_t R = p->frames / G;
    if (npartial < uint64_t(p->frames) * p->channels) return -1;
    if (rows == nullptr || nrows < uint64_t(R) * p->channels) return -1;
    FlushGuard guard;
    for (uint32_t k = 0; k < p->channels; ++k) {
        float P = state[k];
        for (uint32_t r = 0; r < R; ++r) {
            float S = 0.0f;
            for (uint32_t g = 0; g < G; ++g)
                S = S * p->frame_decay + partial[uint64_t(r * G + g) * p->channels + k];
            P = P * p->row_decay + S;
            rows[uint64_t(r) * p->channels + k] = P;
        }
        state[k] = P;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Receiver fine stage, core/dsp/vrx_reference.cpp:429-620 (vrx_fine.comp).
// Stateless rational resampler with a T-tap complex polyphase FIR whose taps
// are lerped between adjacent phases, then a rotation by the NCO table entry
// the 0.64 phase's top nco_log2 bits select. Every index is 32-bit and wraps.
// Pinned by: blocking invariance (any split of a run gives the same bits),
// a unit tone at the centre coming out at unit level, and the FP64 resampler
// definition (tests/test_oracle_vrx.py).
// ---------------------------------------------------------------------------

static int fine_valid(const OFineConfig* c, const OFineParams* p) {
    if (c->taps < 2 || c->taps > 256 || c->phases < 1 || c->phases > 4096 || c->nco_log2 > 24)
        return -1;
    if (p->out_rate == 0 || p->frac0 >= p->out_rate) return -1;
    const uint64_t cap = uint64_t(p->chan_mask) + 1, ocap = uint64_t(p->out_mask) + 1;
    if ((cap & p->chan_mask) != 0 || (ocap & p->out_mask) != 0) return -1;
    if (uint64_t(p->frac0) + uint64_t(p->count) * p->step_rem > 0xFFFFFFFFull) return -1;
    if (uint64_t(p->out_rate - 1) * c->phases > 0xFFFFFFFFull) return -1;
    return 0;
}

int oracle_vrx_fine(const OFineConfig* c, const OFineParams* p, const float* chan,
                    uint64_t chan_samples, const float* taps, uint64_t ntaps, const float* nco,
                    uint64_t nnco, float* out, uint64_t out_samples) {
    if (fine_valid(c, p)) return -1;
    if (ntaps != uint64_t(c->phases) * c->taps + 1) return -1;
    if (nnco != (uint64_t(1) << c->nco_log2)) return -1;
    if (chan_samples < uint64_t(p->chan_base) + uint64_t(p->chan_mask) + 1) return -1;
    if (out_samples < uint64_t(p->out_mask) + 1 || p->count > uint64_t(p->out_mask) + 1) return -1;
    FlushGuard guard;
    for (uint32_t i = 0; i < p->count; ++i) {
        const uint32_t acc_idx = p->frac0 + i * p->step_rem;
        const uint32_t carry = acc_idx / p->out_rate;
        const uint32_t frac = acc_idx - carry * p->out_rate;
        const uint32_t n_rel = i * p->step_whole + carry;
        const uint32_t scaled = frac * c->phases;
        const uint32_t phase = scaled / p->out_rate;
        const uint32_t blend_num = scaled - phase * p->out_rate;
        const float blend = float(blend_num) * p->inv_out_rate;
        const float comp = 1.0f - blend;
        const uint32_t base = p->in_offset + n_rel;
        float ar = 0.0f, ai = 0.0f;
        for (uint32_t k = 0; k < c->taps; ++k) {
            const uint64_t slot = phase + uint64_t(k) * c->phases;
            const float g0r = taps[2 * slot], g0i = taps[2 * slot + 1];
            const float g1r = taps[2 * slot + 2], g1i = taps[2 * slot + 3];
            const float tr = g0r * comp + g1r * blend;
            const float ti = g0i * comp + g1i * blend;
            const uint64_t idx = uint64_t(p->chan_base) + ((base - k) & p->chan_mask);
            const float sr = chan[2 * idx], si = chan[2 * idx + 1];
            ar = ar + (sr * tr - si * ti);
            ai = ai + (sr * ti + si * tr);
        }
        const uint64_t prod = uint64_t(i) * p->nco_delta_low;
        const uint32_t ph_lo = p->nco_phase_low + uint32_t(prod);
        const uint32_t carry_bit = ph_lo < p->nco_phase_low ? 1u : 0u;
        const uint32_t ph_hi =
            p->nco_phase_high + i * p->nco_delta_high + uint32_t(prod >> 32) + carry_bit;
        const uint32_t ni = c->nco_log2 == 0 ? 0u : (ph_hi >> (32 - c->nco_log2));
        const float rr = nco[2 * ni], ri = nco[2 * ni + 1];
        const float orr = ar * rr - ai * ri;
        const float oi = ar * ri + ai * rr;
        const uint64_t o = (p->out_offset + i) & p->out_mask;
        out[2 * o] = orr;
        out[2 * o + 1] = oi;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Receiver demodulator, core/dsp/vrx_reference.cpp:621-905 (vrx_demod.comp):
// detect() per mode, then the decimating audio FIR sum_t detect(idx - t) * w[t]
// in ascending t, times the gain. FM stereo (channels 2 with pilot taps)
// follows the reference's pilot / rotation / matrix order.
// Pinned by: tone-modulated carriers recovered at the stated level for AM,
// NFM/WFM, USB/LSB/CW, and detect() composed from the det_* functions pinned
// against libm (tests/test_oracle_vrx.py).
// ---------------------------------------------------------------------------

static bool complex_tap(uint32_t mode) { return mode == 0 || (mode >= 8 && mode <= 11); }

static int demod_valid(const ODemodConfig* c, const ODemodParams* p) {
    if (c->mode > 12) return -1;
    const bool cx = complex_tap(c->mode);
    if (c->decimation == 0 || c->audio_taps == 0 || c->audio_taps > 4096) return -1;
    if (c->dc_taps == 0 || c->dc_taps > 4096) return -1;
    if (cx && (c->decimation != 1 || c->audio_taps != 1)) return -1;
    if (c->channels != 1 && c->channels != 2) return -1;
    if (cx && c->channels != 2) return -1;
    if (c->pilot_taps > 2048) return -1;
    if (c->pilot_taps != 0 && (c->mode != 3 || c->channels != 2)) return -1;
    if (c->channels == 2 && !cx && c->pilot_taps == 0) return -1;
    const uint64_t cap = uint64_t(p->in_mask) + 1;
    if ((cap & p->in_mask) != 0) return -1;
    const uint32_t det = (c->mode == 1 || c->mode == 12) ? c->dc_taps - 1
                         : (c->mode == 2 || c->mode == 3) ? 1u : 0u;
    const uint32_t audio_reach = c->audio_taps - 1;
    const uint32_t pilot_reach = c->pilot_taps == 0 ? 0u : c->audio_taps / 2 + c->pilot_taps - 1;
    const uint64_t reach = uint64_t(std::max(audio_reach, pilot_reach)) + det;
    const uint64_t span = p->count == 0 ? 0 : uint64_t(p->count - 1) * c->decimation + reach + 1;
    if (span > cap) return -1;
    return 0;
}

int oracle_vrx_demod(const ODemodConfig* c, const ODemodParams* p, const float* fine,
                     uint64_t fine_samples, const float* w, uint64_t nw, float* audio,
                     uint64_t naudio) {
    if (demod_valid(c, p)) return -1;
    const uint64_t pilot_base = uint64_t(c->audio_taps) + c->dc_taps;
    const uint64_t rot_base = pilot_base + 2ull * c->pilot_taps;
    const uint64_t want = c->pilot_taps == 0 ? pilot_base : rot_base + 2ull * c->audio_taps;
    if (nw != want) return -1;
    if (fine_samples < uint64_t(p->in_mask) + 1) return -1;
    if (naudio != uint64_t(p->count) * c->channels) return -1;
    if (p->count == 0) return 0;
    FlushGuard guard;
    auto re = [&](uint32_t i) { return fine[2 * uint64_t(i & p->in_mask)]; };
    auto im = [&](uint32_t i) { return fine[2 * uint64_t(i & p->in_mask) + 1]; };
    auto mag = [&](uint32_t i) {
        const float a = re(i), b = im(i);
        const float sq = a * a + b * b;
        return det_sqrt(sq);
    };
    auto detect = [&](uint32_t i) -> float {
        if (c->mode == 1) {
            float dc = 0.0f;
            for (uint32_t k = 0; k < c->dc_taps; ++k) {
                const float term = mag(i - k) * w[c->audio_taps + k];
                dc = dc + term;
            }
            return mag(i) - dc;
        }
        if (c->mode == 12) {
            float dc = 0.0f;
            for (uint32_t k = 0; k < c->dc_taps; ++k) {
                const float term = re(i - k) * w[c->audio_taps + k];
                dc = dc + term;
            }
            return re(i) - dc;
        }
        if (c->mode == 2 || c->mode == 3) {
            const float zr = re(i), zi = im(i), pr = re(i - 1), pi = im(i - 1);
            const float prod_r = zr * pr + zi * pi;
            const float prod_i = zi * pr - zr * pi;
            return det_atan2(prod_i, prod_r);
        }
        return re(i);
    };
    const float kPilotLock = 2.0e-5f;
    for (uint32_t i = 0; i < p->count; ++i) {
        const uint32_t idx = p->in_offset + i * c->decimation;
        if (complex_tap(c->mode)) {
            audio[2 * uint64_t(i)] = re(idx) * p->gain;
            audio[2 * uint64_t(i) + 1] = im(idx) * p->gain;
            continue;
        }
        if (c->channels == 2) {
            const uint32_t at = idx - c->audio_taps / 2;
            float pr = 0.0f, pi = 0.0f;
            for (uint32_t k = 0; k < c->pilot_taps; ++k) {
                const float comp = detect(at - k);
                const float tr = comp * w[pilot_base + 2 * k];
                const float ti = comp * w[pilot_base + 2 * k + 1];
                pr = pr + tr;
                pi = pi + ti;
            }
            const float sr = pr * p->gain, si = pi * p->gain;
            const float rsq = sr * sr, isq = si * si;
            const float power = rsq + isq;
            const float inv = det_recip(std::max(power, kPilotLock));
            const float cross = sr * si;
            const float cos2 = (isq - rsq) * inv;
            const float sin2 = (-2.0f * cross) * inv;
            const float locked = power >= kPilotLock ? 1.0f : 0.0f;
            float sum = 0.0f, diff = 0.0f;
            for (uint32_t t = 0; t < c->audio_taps; ++t) {
                const float weighted = detect(idx - t) * w[t];
                sum = sum + weighted;
                const float ref = sin2 * w[rot_base + 2 * t] - cos2 * w[rot_base + 2 * t + 1];
                const float term = weighted * ref;
                diff = diff + term;
            }
            const float gated = diff * locked;
            audio[2 * uint64_t(i)] = (sum + gated) * p->gain;
            audio[2 * uint64_t(i) + 1] = (sum - gated) * p->gain;
            continue;
        }
        float acc = 0.0f;
        for (uint32_t t = 0; t < c->audio_taps; ++t) {
            const float term = detect(idx - t) * w[t];
            acc = acc + term;
        }
        audio[i] = acc * p->gain;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// The plain chain as a composition of the twins above: convert the input
// codes into a ring, branch filter, transform + correction into a channel
// ring, exponentially averaged power per frame. Input index 0 is absolute
// sample 0; block m (absolute) is taken at base m*D and needs samples
// m*D - N + 1 .. m*D. This is the CPU baseline of the benchmark and the
// referee of the fused GPU chain.
// ---------------------------------------------------------------------------

struct OChain {
    int format;                // 0 cu8, 2 cs16
    uint32_t channels, taps_per_branch, decimation;
    uint64_t first_block;      // absolute block index of block 0
    uint32_t block_count;
    uint32_t out_ring_blocks;  // channel ring capacity in blocks (power of two)
    uint32_t frame_blocks;     // EMA frame; 0 skips the average
    float alpha, frame_decay;
};

static int chain_range(const OChain* c, const uint8_t* input, uint64_t input_bytes,
                       const float* taps, const float* tw, uint64_t b0, uint32_t nb,
                       float* ring_out, float* partial) {
    const uint32_t M = c->channels, D = c->decimation, N = M * c->taps_per_branch;
    const uint64_t bps = (c->format == 0) ? 2 : 4;
    OGrid g{M, c->taps_per_branch, D};
    // Work in sub-batches whose window fits a local ring.
    const uint32_t sub = std::max<uint32_t>(1, std::min<uint32_t>(nb, 64));
    uint64_t cap = 1;
    while (cap < uint64_t(N) + uint64_t(sub - 1) * D) cap <<= 1;
    std::vector<float> ring(2 * cap), branch(2 * size_t(sub) * M);
    for (uint32_t done = 0; done < nb; done += sub) {
        const uint32_t n = std::min(sub, nb - done);
        const uint64_t m0 = c->first_block + b0 + done;
        const uint64_t first_sample = m0 * D - (N - 1);
        const uint64_t count = uint64_t(N) + uint64_t(n - 1) * D;
        if ((first_sample + count) * bps > input_bytes) return -1;
        OConvertParams cp{uint32_t(cap - 1), uint32_t(first_sample), 0, uint32_t(count)};
        const uint64_t whole = (count * bps + 3) & ~uint64_t(3);
        const uint8_t* src = input + first_sample * bps;
        std::vector<uint8_t> padded;
        if (first_sample * bps + whole > input_bytes) {
            // The conversion loads whole words; give the odd tail its padding.
            padded.assign(whole, 0);
            std::memcpy(padded.data(), src, count * bps);
            src = padded.data();
        }
        if (oracle_convert(c->format, src, whole, &cp, ring.data(), cap)) return -1;
        OBranchParams bp{uint32_t(cap - 1), uint32_t(m0 * D), n};
        if (oracle_pfb_branch(&g, taps, uint64_t(N), ring.data(), cap, &bp, branch.data(),
                              uint64_t(n) * M))
            return -1;
        OFftParams fp{M, D, log2u(M), uint32_t(m0), n, c->out_ring_blocks, c->out_ring_blocks - 1};
        if (oracle_pfb_fft(&fp, branch.data(), uint64_t(n) * M, tw, M, ring_out,
                           uint64_t(M) * c->out_ring_blocks))
            return -1;
    }
    if (c->frame_blocks != 0 && partial != nullptr) {
        const uint32_t frames = nb / c->frame_blocks;
        OEmaParams ep{M, c->out_ring_blocks, c->out_ring_blocks - 1,
                      uint32_t(c->first_block + b0), c->frame_blocks, frames, c->alpha,
                      c->frame_decay, 1, c->frame_decay};
        if (frames && oracle_ema_partial(&ep, ring_out, uint64_t(M) * c->out_ring_blocks,
                                         partial + (b0 / c->frame_blocks) * M,
                                         uint64_t(frames) * M))
            return -1;
    }
    return 0;
}

// input: packed codes from absolute sample 0. ring_out: M x out_ring_blocks
// cf32. partial: (block_count / frame_blocks) x M floats, or null.
int oracle_chain(const OChain* c, const uint8_t* input, uint64_t input_bytes, const float* taps,
                 uint64_t ntaps, const float* tw, uint64_t ntw, float* ring_out,
                 uint64_t ring_samples, float* partial, uint64_t npartial, int threads) {
    const uint32_t M = c->channels;
    OGrid g{M, c->taps_per_branch, c->decimation};
    if (!grid_ok(&g) || (c->format != 0 && c->format != 2)) return -1;
    if (ntaps != uint64_t(M) * c->taps_per_branch || ntw != M) return -1;
    if (!pow2(c->out_ring_blocks) || c->block_count > c->out_ring_blocks) return -1;
    if (ring_samples < uint64_t(M) * c->out_ring_blocks) return -1;
    const uint64_t N = uint64_t(M) * c->taps_per_branch;
    if (c->first_block * c->decimation < N - 1) return -1;
    if (c->frame_blocks != 0) {
        if (c->block_count % c->frame_blocks) return -1;
        if (partial && npartial < uint64_t(c->block_count / c->frame_blocks) * M) return -1;
    }
    if (threads < 1) threads = 1;
    // Split on frame boundaries so every frame's average is one thread's.
    const uint32_t unit = c->frame_blocks ? c->frame_blocks : 1;
    const uint32_t units = c->block_count / unit;
    const uint32_t per = (units + threads - 1) / threads;
    std::vector<int> rc(threads, 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        const uint64_t u0 = uint64_t(t) * per;
        if (u0 >= units) break;
        const uint32_t nu = uint32_t(std::min<uint64_t>(per, units - u0));
        pool.emplace_back([&, t, u0, nu] {
            rc[t] = chain_range(c, input, input_bytes, taps, tw, u0 * unit, nu * unit, ring_out,
                                partial);
        });
    }
    for (auto& th : pool) th.join();
    for (int r : rc)
        if (r) return r;
    return 0;
}


// ---------------------------------------------------------------------------
// Front-end correction. core/dsp/front_end_correction_reference.cpp:46-88
// (reference_iq_moments: five sums per 64-sample chunk, in sample order from
// zero) and :90-111 (reference_iq_correct: i' = i - dc_i, q' = cross * i' +
// scale * (q - dc_q) on a ring block).
// Pinned by (tests/test_oracle_front_end.py): the moments against numpy
// float64 sums of the same chunks within float rounding and exactly on
// integer-valued inputs (where every partial sum is exact), the correction
// undoing the imbalance model I_r = I, Q_r = g (Q cos phi + I sin phi) to
// 1e-6 (test_iq_correction.cpp:166-199), and the identity parameters leaving
// every sample's bits alone.
// ---------------------------------------------------------------------------

struct OIqMomentsParams {  // core/dsp/front_end_correction.h:96-107
    uint32_t ring_mask, src_offset, count;
};
struct OIqCorrectParams {  // core/dsp/front_end_correction.h:116-130
    uint32_t ring_mask, offset, count;
    float dc_i, dc_q, cross, scale;
};

static bool fe_ring_ok(uint64_t ring_samples, uint32_t mask, uint32_t count) {
    const uint64_t capacity = uint64_t(mask) + 1;
    return pow2(capacity) && ring_samples >= capacity && count <= capacity;
}

int oracle_iq_moments(const OIqMomentsParams* p, const float* ring, uint64_t ring_samples,
                      float* sums, uint64_t n_sums) {
    if (!fe_ring_ok(ring_samples, p->ring_mask, p->count)) return -1;
    const uint32_t chunks = (p->count + 63) / 64;
    if (n_sums < uint64_t(chunks) * 5) return -1;
    FlushGuard flush;
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t first = c * 64, last = std::min(first + 64, p->count);
        float si = 0.0f, sq = 0.0f, sii = 0.0f, sqq = 0.0f, siq = 0.0f;
        for (uint32_t k = first; k < last; ++k) {
            const uint64_t slot = (p->src_offset + k) & p->ring_mask;
            const float i = ring[2 * slot], q = ring[2 * slot + 1];
            si = si + i;
            sq = sq + q;
            sii = sii + i * i;
            sqq = sqq + q * q;
            siq = siq + i * q;
        }
        float* o = sums + uint64_t(c) * 5;
        o[0] = si;
        o[1] = sq;
        o[2] = sii;
        o[3] = sqq;
        o[4] = siq;
    }
    return 0;
}

int oracle_iq_correct(const OIqCorrectParams* p, const float* in, uint64_t in_samples, float* out,
                      uint64_t out_samples) {
    if (!fe_ring_ok(in_samples, p->ring_mask, p->count) || out_samples < in_samples) return -1;
    FlushGuard flush;
    for (uint32_t k = 0; k < p->count; ++k) {
        const uint64_t slot = (p->offset + k) & p->ring_mask;
        const float i = in[2 * slot] - p->dc_i;
        const float q_dc = in[2 * slot + 1] - p->dc_q;
        const float q = p->cross * i + p->scale * q_dc;
        out[2 * slot] = i;
        out[2 * slot + 1] = q;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Impulse excision. core/dsp/impulse_excision_reference.cpp:250-387, the four
// passes: segment powers (sample order from zero), the median reference
// (insertion sort of the K segments ending one before, zero until a whole
// window of this run), flags (|x|^2 * S > T * r with r > 0), and apply (a
// flag marks [f - lead, f + hang]; a maximal marked run of at most max_width
// is replaced by -0 and moved to the removed ring; counters of excised
// samples, events and spared runs).
// Pinned by (tests/test_oracle_excision.py), with the reference's measurements
// (test_impulse_excision.cpp:344-477): forty impulses each found and removed,
// a tone kept to 0.1 dB, cleaned + removed == raw to the bit, no sample moved,
// a strong clean carrier left alone, and bursts longer than the maximum width
// spared; the median against numpy's on the same windows.
// ---------------------------------------------------------------------------

struct OExcisionConfig {  // core/dsp/impulse_excision.h:155-168
    uint32_t pass, seg_log2, segments, max_width, hang, lead;
};
struct OExcisionParams {  // :185-201
    uint32_t ring_mask, first, count, seg_mask, flag_mask, removed_mask;
    int32_t history;
    float threshold;
};
static_assert(sizeof(OExcisionParams) == 32);

static bool ex_flag_at(const OExcisionParams* p, const uint32_t* flags, int64_t rel) {
    if (rel + p->history < 0) return false;
    return flags[(p->first + uint32_t(rel)) & p->flag_mask] != 0u;
}
static bool ex_marked(const OExcisionConfig* c, const OExcisionParams* p, const uint32_t* flags,
                      int64_t rel) {
    for (uint32_t m = 0; m <= c->hang + c->lead; ++m)
        if (ex_flag_at(p, flags, rel - int64_t(c->hang) + m)) return true;
    return false;
}

int oracle_excise_segment(const OExcisionConfig* c, const OExcisionParams* p, const float* raw,
                          float* seg_power) {
    FlushGuard flush;
    const uint32_t samples = 1u << c->seg_log2;
    for (uint32_t j = 0; j < p->count; ++j) {
        const uint32_t base = (p->first + (j << c->seg_log2)) & p->ring_mask;
        float sum = 0.0f;
        for (uint32_t m = 0; m < samples; ++m) {
            const uint64_t k = (base + m) & p->ring_mask;
            const float q = raw[2 * k] * raw[2 * k] + raw[2 * k + 1] * raw[2 * k + 1];
            sum = sum + q;
        }
        seg_power[(base >> c->seg_log2) & p->seg_mask] = sum;
    }
    return 0;
}

int oracle_excise_reference(const OExcisionConfig* c, const OExcisionParams* p,
                            const float* seg_power, float* seg_level) {
    const int64_t segment = int64_t(1) << c->seg_log2;
    for (uint32_t j = 0; j < p->count; ++j) {
        const uint32_t base = (p->first + (j << c->seg_log2)) & p->ring_mask;
        const uint32_t slot = base >> c->seg_log2;
        const int64_t rel = int64_t(j) * segment;
        if (rel - (int64_t(c->segments) + 1) * segment + p->history < 0) {
            seg_level[slot & p->seg_mask] = 0.0f;
            continue;
        }
        float v[64];
        for (uint32_t m = 0; m < c->segments; ++m)
            v[m] = seg_power[(slot - 1u - c->segments + m) & p->seg_mask];
        for (uint32_t a = 1; a < c->segments; ++a) {
            const float x = v[a];
            uint32_t b = a;
            while (b > 0 && v[b - 1] > x) {
                v[b] = v[b - 1];
                --b;
            }
            v[b] = x;
        }
        seg_level[slot & p->seg_mask] = v[c->segments / 2];
    }
    return 0;
}

int oracle_excise_flag(const OExcisionConfig* c, const OExcisionParams* p, const float* raw,
                       const float* seg_level, uint32_t* flags) {
    FlushGuard flush;
    const float samples = float(1u << c->seg_log2);
    for (uint32_t j = 0; j < p->count; ++j) {
        const uint32_t slot = (p->first + j) & p->ring_mask;
        const float pw = raw[2 * uint64_t(slot)] * raw[2 * uint64_t(slot)] +
                         raw[2 * uint64_t(slot) + 1] * raw[2 * uint64_t(slot) + 1];
        const float level = seg_level[(slot >> c->seg_log2) & p->seg_mask];
        const float lhs = pw * samples;
        const float rhs = p->threshold * level;
        flags[(p->first + j) & p->flag_mask] = (level > 0.0f && lhs > rhs) ? 1u : 0u;
    }
    return 0;
}

int oracle_excise_apply(const OExcisionConfig* c, const OExcisionParams* p, const float* raw,
                        const uint32_t* flags, float* cleaned, float* removed,
                        uint32_t* counters) {
    for (uint32_t j = 0; j < p->count; ++j) {
        const int64_t rel = j;
        const uint64_t slot = (p->first + j) & p->ring_mask;
        const float xr = raw[2 * slot], xi = raw[2 * slot + 1];
        bool excise = false;
        if (ex_marked(c, p, flags, rel)) {
            uint32_t left = 0, right = 0;
            for (uint32_t k = 1; k <= c->max_width; ++k) {
                if (!ex_marked(c, p, flags, rel - int64_t(k))) break;
                ++left;
            }
            for (uint32_t k = 1; k <= c->max_width; ++k) {
                if (!ex_marked(c, p, flags, rel + int64_t(k))) break;
                ++right;
            }
            excise = left + right + 1u <= c->max_width;
            if (left == 0u) counters[excise ? 1 : 2] += 1u;
        }
        if (excise) counters[0] += 1u;
        const uint64_t r = (p->first + j) & p->removed_mask;
        cleaned[2 * slot] = excise ? -0.0f : xr;
        cleaned[2 * slot + 1] = excise ? -0.0f : xi;
        removed[2 * r] = excise ? xr : -0.0f;
        removed[2 * r + 1] = excise ? xi : -0.0f;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Comb cancellation, the device half. core/dsp/comb_kernels_reference.cpp:
// 100-141 (project: per line and sub-block, x * conj(tone) summed in chunks of
// 64 from zero, the chunk sums from zero in chunk order) and :143-180
// (subtract: the lines' a * tone summed from -0 in line order, subtracted).
// tone(n) = table[(phase0 + n * step) >> 20], 32-bit wrapping phase.
// Pinned by (tests/test_oracle_comb.py): a pure tone at a line's step
// projects to amplitude times the sub-block length (comb_kernels :237-268);
// subtracting a comb with its true amplitudes leaves the table's quantisation
// floor (:270-326); no lines copies the ring and removes -0 (:328-346); a
// subtract across the ring's end equals the same subtract unwrapped
// (:348-381); projections against float64 sums to float rounding.
// ---------------------------------------------------------------------------

struct OCombLine {  // core/dsp/comb_cancel.h:248-254
    uint32_t phase0, step, pad0, pad1;
    float amp_re, amp_im;
};
struct OCombParams {  // :276-285
    uint32_t ring_mask, first, count, lines, sub_blocks, removed_mask, pad0, pad1;
};
static_assert(sizeof(OCombLine) == 24 && sizeof(OCombParams) == 32);

int oracle_comb_project(const OCombParams* p, const float* raw, const OCombLine* lines,
                        const float* table, float* proj) {
    if (p->sub_blocks == 0 || p->count % p->sub_blocks) return -1;
    FlushGuard flush;
    const uint32_t length = p->count / p->sub_blocks;
    for (uint32_t l = 0; l < p->lines; ++l) {
        for (uint32_t s = 0; s < p->sub_blocks; ++s) {
            const uint32_t begin = s * length;
            float tr = 0.0f, ti = 0.0f;
            for (uint32_t c = 0; c < length; c += 64) {
                const uint32_t end = std::min(c + 64, length);
                float cr = 0.0f, ci = 0.0f;
                for (uint32_t m = c; m < end; ++m) {
                    const uint32_t n = begin + m;
                    const uint64_t k = (p->first + n) & p->ring_mask;
                    const uint32_t t = (lines[l].phase0 + n * lines[l].step) >> 20;
                    const float xr = raw[2 * k], xi = raw[2 * k + 1];
                    const float c0 = table[2 * t], s0 = table[2 * t + 1];
                    const float re = xr * c0 + xi * s0;
                    const float im = xi * c0 - xr * s0;
                    cr = cr + re;
                    ci = ci + im;
                }
                tr = tr + cr;
                ti = ti + ci;
            }
            proj[2 * (uint64_t(l) * p->sub_blocks + s)] = tr;
            proj[2 * (uint64_t(l) * p->sub_blocks + s) + 1] = ti;
        }
    }
    return 0;
}

int oracle_comb_subtract(const OCombParams* p, const float* raw, const OCombLine* lines,
                         const float* table, float* cleaned, float* removed) {
    FlushGuard flush;
    for (uint32_t n = 0; n < p->count; ++n) {
        const uint64_t slot = (p->first + n) & p->ring_mask;
        const float xr = raw[2 * slot], xi = raw[2 * slot + 1];
        float sr = -0.0f, si = -0.0f;
        for (uint32_t l = 0; l < p->lines; ++l) {
            const uint32_t t = (lines[l].phase0 + n * lines[l].step) >> 20;
            const float c0 = table[2 * t], s0 = table[2 * t + 1];
            const float re = lines[l].amp_re * c0 - lines[l].amp_im * s0;
            const float im = lines[l].amp_re * s0 + lines[l].amp_im * c0;
            sr = sr + re;
            si = si + im;
        }
        if (p->lines == 0) {
            cleaned[2 * slot] = xr;
            cleaned[2 * slot + 1] = xi;
        } else {
            cleaned[2 * slot] = xr - sr;
            cleaned[2 * slot + 1] = xi - si;
        }
        const uint64_t r = (p->first + n) & p->removed_mask;
        removed[2 * r] = sr;
        removed[2 * r + 1] = si;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// The per-receiver impulse blanker. core/dsp/noise_reference.cpp:158-226:
// detect (flag when |x|^2 * W > T * the window's power sum, oldest first, a
// zero sum never flags) and apply (zero when a flag lies in [i - hang,
// i + lead]). Flags are floats, as the twin's.
// Pinned by (tests/test_oracle_noise.py): the window sums against float64
// sums and the flag inequality against numpy away from ties; one five-sample
// impulse 30 dB over noise blanked with its hang and lead and nothing else
// (tests/reference/test_noise.cpp:270-322).
// ---------------------------------------------------------------------------

struct OBlankerConfig {  // core/dsp/noise_reference.h:117
    uint32_t pass, window, guard, hang, lead;
};
struct OBlankerParams {  // :152
    uint32_t chan_base, chan_mask, chan_first, out_mask, out_first, count;
    float threshold;
};

int oracle_blank_detect(const OBlankerConfig* c, const OBlankerParams* p, const float* chan,
                        float* flags) {
    FlushGuard flush;
    const float window = float(c->window);
    for (uint32_t i = 0; i < p->count; ++i) {
        const uint32_t ci = (p->chan_first + i) & p->chan_mask;
        const uint32_t o = (p->out_first + i) & p->out_mask;
        const uint64_t k = uint64_t(p->chan_base) + ci;
        const float pw = chan[2 * k] * chan[2 * k] + chan[2 * k + 1] * chan[2 * k + 1];
        float sum = 0.0f;
        for (uint32_t m = 0; m < c->window; ++m) {
            const uint32_t at = (ci - c->guard - c->window + m) & p->chan_mask;
            const uint64_t r = uint64_t(p->chan_base) + at;
            const float q = chan[2 * r] * chan[2 * r] + chan[2 * r + 1] * chan[2 * r + 1];
            sum = sum + q;
        }
        const float lhs = pw * window;
        const float rhs = p->threshold * sum;
        flags[o] = (sum > 0.0f && lhs > rhs) ? 1.0f : 0.0f;
    }
    return 0;
}

int oracle_blank_apply(const OBlankerConfig* c, const OBlankerParams* p, const float* chan,
                       const float* flags, float* blanked) {
    for (uint32_t i = 0; i < p->count; ++i) {
        const uint32_t ci = (p->chan_first + i) & p->chan_mask;
        const uint32_t o = (p->out_first + i) & p->out_mask;
        bool hit = false;
        for (uint32_t m = 0; m <= c->hang + c->lead; ++m)
            if (flags[(o - c->hang + m) & p->out_mask] != 0.0f) hit = true;
        const uint64_t k = uint64_t(p->chan_base) + ci;
        blanked[2 * uint64_t(o)] = hit ? 0.0f : chan[2 * k];
        blanked[2 * uint64_t(o) + 1] = hit ? 0.0f : chan[2 * k + 1];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// The line kernel. core/dsp/noise_reference.cpp:266-421: a cascade of notch
// sections (y = b0 x + b1 x1 + b2 x2 - a1 y1 - a2 y2, summed in that order),
// then a normalized leaky LMS predictor over the history ring, its taps split
// over `lanes` partial sums (lane l: taps l, l + lanes, ..., ascending) added
// in lane order; the error is the output, or with the guard, v less the
// guard section run on the prediction. State: notch [0,4), position [4],
// guard [8,12), sections 1..3 [12,24), weights, history.
// Pinned by (tests/test_oracle_noise.py): a tone through a notch loses what
// the float-rounded section's analytic response says, at the centre and at
// the half-width edges (tests/reference/test_noise.cpp:517-567); a steady
// heterodyne is learned down past 10 dB while a gliding tone keeps within
// 1 dB (:618-676); on CW the guard keeps the pitch and the predictor still
// takes a heterodyne, and without the guard takes the pitch (:782-857); the
// cascade state equals sections run one after another; lanes = taps (one tap
// per lane) and lanes = 1 give the same output within float rounding.
// ---------------------------------------------------------------------------

struct OLineConfig {
    uint32_t lanes, taps, history;
};
struct OLineSection {
    float b0, b1, b2, a1, a2;
};
struct OLineParams {
    uint32_t count, flags, stride, delay;
    float b0, b1, b2, a1, a2;
    float mu, leak, eps;
    float gb0, gb1, gb2, ga1, ga2;
    OLineSection more[3];
};

static float o_section(float x, float b0, float b1, float b2, float a1, float a2, float* st) {
    // st: x1, x2, y1, y2
    const float t0 = b0 * x;
    const float t1 = b1 * st[0];
    const float t2 = b2 * st[1];
    const float t3 = a1 * st[2];
    const float t4 = a2 * st[3];
    float acc = t0 + t1;
    acc = acc + t2;
    acc = acc - t3;
    acc = acc - t4;
    st[1] = st[0];
    st[0] = x;
    st[3] = st[2];
    st[2] = acc;
    return acc;
}

int oracle_noise_line(const OLineConfig* c, const OLineParams* p, float* audio, float* state) {
    FlushGuard flush;
    const bool notch = (p->flags & 1u) != 0;
    const bool ale = (p->flags & 2u) != 0;
    const bool guard = ale && (p->flags & 4u) != 0;
    const uint32_t sections = notch ? 1 + ((p->flags >> 8) & 3u) : 0;
    const uint32_t lanes = c->lanes, per_lane = c->taps / c->lanes, hmask = c->history - 1;
    float notch_st[4] = {state[0], state[1], state[2], state[3]};
    float more_st[3][4] = {};
    for (uint32_t s = 0; s + 1 < sections; ++s)
        for (int k = 0; k < 4; ++k) more_st[s][k] = state[12 + 4 * s + k];
    float guard_st[4] = {state[8], state[9], state[10], state[11]};
    uint32_t position = static_cast<uint32_t>(state[4]);
    float* weights = state + 24;
    float* history = weights + c->taps;
    std::vector<float> part_y(lanes), part_p(lanes), ref(c->taps);
    for (uint32_t n = 0; n < p->count; ++n) {
        float v = audio[n];
        if (notch) {
            v = o_section(v, p->b0, p->b1, p->b2, p->a1, p->a2, notch_st);
            for (uint32_t s = 0; s + 1 < sections; ++s) {
                const OLineSection& q = p->more[s];
                v = o_section(v, q.b0, q.b1, q.b2, q.a1, q.a2, more_st[s]);
            }
        }
        float out = v;
        if (ale) {
            for (uint32_t l = 0; l < lanes; ++l) {
                float py = 0.0f, pp = 0.0f;
                for (uint32_t j = 0; j < per_lane; ++j) {
                    const uint32_t k = l + j * lanes;
                    const float r = history[(position - p->delay - k * p->stride) & hmask];
                    ref[k] = r;
                    const float wy = weights[k] * r;
                    py = py + wy;
                    const float rr = r * r;
                    pp = pp + rr;
                }
                part_y[l] = py;
                part_p[l] = pp;
            }
            float y = 0.0f, pw = 0.0f;
            for (uint32_t l = 0; l < lanes; ++l) {
                y = y + part_y[l];
                pw = pw + part_p[l];
            }
            const float e = v - y;
            const float g = p->mu * det_recip(pw + p->eps);
            const float ge = g * e;
            for (uint32_t k = 0; k < c->taps; ++k) {
                const float kept = p->leak * weights[k];
                const float step = ge * ref[k];
                weights[k] = kept + step;
            }
            history[position & hmask] = v;
            out = e;
            if (guard) {
                const float acc = o_section(y, p->gb0, p->gb1, p->gb2, p->ga1, p->ga2, guard_st);
                out = v - acc;
            }
        }
        audio[n] = out;
        position = position + 1u;
    }
    for (int k = 0; k < 4; ++k) state[k] = notch_st[k];
    for (uint32_t s = 0; s + 1 < sections; ++s)
        for (int k = 0; k < 4; ++k) state[12 + 4 * s + k] = more_st[s][k];
    state[4] = static_cast<float>(position & hmask);
    for (int k = 0; k < 4; ++k) state[8 + k] = guard_st[k];
    return 0;
}

// ---------------------------------------------------------------------------
// Spectral noise reduction. core/dsp/noise_reference.cpp:568-700: hop H = N/2
// of input into the history while the queued output goes out; at a full hop
// the analysis window, a direct DFT per bin (ascending time, tables indexed
// (k t) mod N, im accumulated by subtraction), power smoothing and Doblinger's
// minimum (the first frame seeds both), gain 1 - alpha bias minimum / (level
// + eps) floored, the inverse sum over bins 1..H-1 in ascending order plus DC
// and the alternating Nyquist, the synthesis window, and the overlap-add.
// State: phase, frames seen (saturating at 2), history N, tail H, queue H,
// smoothed N/2+1, minimum N/2+1.
// Pinned by (tests/test_oracle_noise.py): at unit gain the output is the input
// one frame late within 2e-5 and the first hop is zero (the Allen overlap-add
// identity; test_noise.cpp:911-938); on four seconds of white noise the bias
// times the tracked minimum lands within 1.5 dB of sigma^2 N / 2 (:940-976);
// the tables against numpy's cos/sin/Hann; a stationary tone is passed at a
// gain within 0.5 dB of unity while noise-only bins are floored.
// ---------------------------------------------------------------------------

struct OSpectralParams {
    uint32_t count;
    float smoothing, smoothing_complement, gamma, beta, rise, bias, alpha, floor_gain, eps;
};

int oracle_noise_spectral(uint32_t n, const OSpectralParams* p, const float* tables, float* audio,
                          float* state) {
    FlushGuard flush;
    const uint32_t h = n / 2, bins = h + 1, mask = n - 1;
    const float* cosine = tables;
    const float* sine = tables + n;
    const float* analysis = tables + 2 * n;
    const float* synthesis = tables + 3 * n;
    float* hist = state + 2;
    float* tail = hist + n;
    float* queue = tail + h;
    float* smoothed = queue + h;
    float* minimum = smoothed + bins;
    uint32_t phase = static_cast<uint32_t>(state[0]);
    float seen = state[1];
    std::vector<float> work(n), yr(bins), yi(bins);
    uint32_t s = 0;
    while (s < p->count) {
        const uint32_t length = std::min(h - phase, p->count - s);
        for (uint32_t i = 0; i < length; ++i) {
            const float x = audio[s + i];
            hist[h + phase + i] = x;
            audio[s + i] = queue[phase + i];
        }
        phase += length;
        s += length;
        if (phase != h) continue;
        for (uint32_t t = 0; t < n; ++t) work[t] = hist[t] * analysis[t];
        for (uint32_t k = 0; k < bins; ++k) {
            float re = 0.0f, im = 0.0f;
            for (uint32_t t = 0; t < n; ++t) {
                const uint32_t at = (k * t) & mask;
                const float xc = work[t] * cosine[at];
                re = re + xc;
                const float xs = work[t] * sine[at];
                im = im - xs;
            }
            const float rr = re * re;
            const float ii = im * im;
            const float power = rr + ii;
            float level = power, floor_power = power;
            if (seen != 0.0f) {
                const float previous = smoothed[k];
                const float kept = p->smoothing * previous;
                const float fresh = p->smoothing_complement * power;
                level = kept + fresh;
                const float held = minimum[k];
                if (held < level) {
                    const float decayed = p->gamma * held;
                    const float lagged = p->beta * previous;
                    const float lift = level - lagged;
                    const float risen = p->rise * lift;
                    floor_power = decayed + risen;
                } else {
                    floor_power = level;
                }
            }
            smoothed[k] = level;
            minimum[k] = floor_power;
            const float noise = p->bias * floor_power;
            const float ratio = noise * det_recip(level + p->eps);
            const float cut = p->alpha * ratio;
            const float gain = std::max(1.0f - cut, p->floor_gain);
            yr[k] = gain * re;
            yi[k] = gain * im;
        }
        for (uint32_t t = 0; t < n; ++t) {
            float acc = 0.0f;
            for (uint32_t k = 1; k < h; ++k) {
                const uint32_t at = (k * t) & mask;
                const float a = yr[k] * cosine[at];
                const float b = yi[k] * sine[at];
                const float term = a - b;
                acc = acc + term;
            }
            const float nyquist = (t & 1u) ? -yr[h] : yr[h];
            const float doubled = 2.0f * acc;
            float sum = yr[0] + nyquist;
            sum = sum + doubled;
            work[t] = sum * synthesis[t];
        }
        for (uint32_t t = 0; t < h; ++t) {
            queue[t] = tail[t] + work[t];
            tail[t] = work[h + t];
            hist[t] = hist[h + t];
        }
        phase = 0;
        seen = std::min(seen + 1.0f, 2.0f);
    }
    state[0] = static_cast<float>(phase);
    state[1] = seen;
    return 0;
}

// ---------------------------------------------------------------------------
// The carrier loop. core/dsp/vrx_reference.cpp:1052-1185
// (reference_vrx_carrier) and its validate(), :986-1050: PLL (power 1) for
// SAM, Costas (power 2) for DSB, one serial pass over the fine ring rotating
// it in place, eight words of state carried between calls.
// Pinned by (tests/test_oracle_carrier.py), with the reference's own
// measurements from tests/reference/test_vrx_carrier.cpp:684-900: both loops
// lock on an offset carrier within the times the reference states and settle
// within 1 Hz of it (a sign slip in the detector, the NCO or the advance
// fails to lock or locks on the mirror), a DSB tone comes back at unit level
// at sixteen carrier phases (0.995..1.005; a wrong Costas detector or level
// fails), a drifting carrier is followed to the passband's reach and never
// past it (the integrator clamp), a restart equals a fresh loop, and the
// result does not depend on where dispatches fall.
// ---------------------------------------------------------------------------

struct OCarrierConfig {  // core/dsp/vrx_reference.h:686-690
    uint32_t power, nco_log2;
};
struct OCarrierParams {  // core/dsp/vrx_reference.h:699-714
    uint32_t in_mask, in_offset, count, flags;
    float shift, k1_acquire, k2_acquire, k1_track, k2_track, level_alpha, lock_alpha,
        clamp_radians;
    uint32_t hold_in, hold_out;
};
static_assert(sizeof(OCarrierParams) == 56);

static bool carrier_valid(const OCarrierConfig* c, const OCarrierParams* p) {
    if (c->power != 1 && c->power != 2) return false;
    if (c->nco_log2 == 0 || c->nco_log2 > 24) return false;
    const uint64_t capacity = uint64_t(p->in_mask) + 1;
    if ((capacity & uint64_t(p->in_mask)) != 0 || p->count > capacity) return false;
    if ((p->flags & ~1u) != 0) return false;
    if (!std::isfinite(p->shift)) return false;
    if (!(p->clamp_radians > 0.0f) || !(double(p->clamp_radians) <= 0.5 * 3.14159265358979323846))
        return false;
    for (float k : {p->k1_acquire, p->k1_track, p->k2_acquire, p->k2_track})
        if (!(k > 0.0f) || !(k < 1.0f)) return false;
    for (float a : {p->level_alpha, p->lock_alpha})
        if (!(a > 0.0f) || !(a <= 1.0f)) return false;
    return p->hold_in != 0 && p->hold_out != 0;
}

// std::max(v, lo) then std::min(., hi)
static float clamp_to(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }

int oracle_vrx_carrier(const OCarrierConfig* c, const OCarrierParams* p, const float* nco_f,
                       uint64_t n_nco, float* ring_f, uint64_t ring_samples, uint32_t* state,
                       uint64_t state_words) {
    if (!carrier_valid(c, p)) return -1;
    if (n_nco != (uint64_t(1) << c->nco_log2)) return -1;
    if (ring_samples < uint64_t(p->in_mask) + 1) return -1;
    if (state_words != 8) return -1;
    FlushGuard flush;
    const float kTurnsPerRadian = static_cast<float>(683565275.576431632);
    const float kLevelFloor = static_cast<float>(1.0e-24);
    const float kLockRatio = static_cast<float>(2.0);
    const bool reset = (p->flags & 1u) != 0;
    const bool pll = c->power == 1;
    uint32_t phase = reset ? 0u : state[0];
    float integrator = reset ? 0.0f : bits_to_float(state[1]);
    float level = reset ? 0.0f : bits_to_float(state[2]);
    float lock_re = reset ? 0.0f : bits_to_float(state[3]);
    float lock_im = reset ? 0.0f : bits_to_float(state[4]);
    uint32_t locked = reset ? 0u : state[5];
    uint32_t held = reset ? 0u : state[6];
    uint32_t primed = reset ? 0u : state[7];
    {
        const float moved = integrator + p->shift;
        integrator = clamp_to(moved, -p->clamp_radians, p->clamp_radians);
    }
    const uint32_t down = 32u - c->nco_log2;
    for (uint32_t i = 0; i < p->count; ++i) {
        const uint64_t slot = (p->in_offset + i) & p->in_mask;
        const float zr = ring_f[2 * slot], zi = ring_f[2 * slot + 1];
        const uint64_t t = phase >> down;
        const float rr = nco_f[2 * t], ri = nco_f[2 * t + 1];
        const float yr = zr * rr - zi * ri;
        const float yi = zr * ri + zi * rr;
        ring_f[2 * slot] = yr;
        ring_f[2 * slot + 1] = yi;
        const float rsq = yr * yr, isq = yi * yi;
        const float power = rsq + isq;
        const float cross = yr * yi;
        float wr, wi, magnitude;
        if (pll) {
            wr = yr;
            wi = yi;
            magnitude = det_sqrt(power);
        } else {
            wr = rsq - isq;
            wi = cross + cross;
            magnitude = power;
        }
        if (primed == 0u) {
            level = magnitude;
            primed = 1u;
        } else {
            const float step = p->level_alpha * (magnitude - level);
            level = level + step;
        }
        const float sr = p->lock_alpha * (wr - lock_re);
        const float si = p->lock_alpha * (wi - lock_im);
        lock_re = lock_re + sr;
        lock_im = lock_im + si;
        const bool inside = lock_re > kLockRatio * std::fabs(lock_im);
        if (locked != 0u) {
            held = inside ? (held > 0u ? held - 1u : 0u) : held + 1u;
            if (held >= p->hold_out) {
                locked = 0u;
                held = 0u;
            }
        } else {
            held = inside ? held + 1u : 0u;
            if (held >= p->hold_in) {
                locked = 1u;
                held = 0u;
            }
        }
        const float inverse = det_recip(std::max(level, kLevelFloor));
        const float error = clamp_to((pll ? yi : cross) * inverse, -1.0f, 1.0f);
        const float k1 = locked ? p->k1_track : p->k1_acquire;
        const float k2 = locked ? p->k2_track : p->k2_acquire;
        const float integrated = integrator + k2 * error;
        integrator = clamp_to(integrated, -p->clamp_radians, p->clamp_radians);
        const float advance = k1 * error + integrator;
        const float turns = advance * kTurnsPerRadian;
        phase = phase + static_cast<uint32_t>(static_cast<int32_t>(turns));
    }
    state[0] = phase;
    state[1] = float_to_bits(integrator);
    state[2] = float_to_bits(level);
    state[3] = float_to_bits(lock_re);
    state[4] = float_to_bits(lock_im);
    state[5] = locked;
    state[6] = held;
    state[7] = primed;
    return 0;
}

}  // extern "C"
