// Receiver planning on the host: where a receiver sits on the grid and what
// its fine stage and demodulator are. FP64 design rounded once to float;
// every rate and phase is an exact integer or rational.
//
//   rvx_place           core/engine/vrx_place.cpp:216-340 (nearest channel by
//                       round(centre*M/rate), exact rational residual)
//   rvx_plan_vrx        core/dsp/vrx_reference.cpp:2243-2662 (design_vrx +
//                       plan_vrx) for the receivers the chain runs: AM, NFM,
//                       WFM mono, USB, LSB, DSB, CW, SAM and raw, SAM and DSB
//                       with their carrier loop. Not planned here:
//                       de-emphasis, FM stereo and the digital-voice taps.
//   rvx_fine_block      vrx_reference.cpp:2786-2871 (resample_block)
//   rvx_demod_block     vrx_reference.cpp:2873-2909
//   rvx_carrier_*       vrx_reference.cpp:918-1200 and :2911-2933 (the loop's
//                       design: gains, one-poles, holds, reach, retune, block)
// Tap tables: design_fine_taps (:1808), design_audio_taps (:1921),
// design_dc_weights (:1983).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "common.h"
#include "plan.h"

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kTwoPi = 2.0 * kPi;
constexpr int64_t kDefaultAudioRate = 48000;
constexpr double kFineAttenuationDb = 80.0;
constexpr double kAudioAttenuationDb = 80.0;
constexpr int64_t kAmDcCornerHz = 200;
constexpr double kFmAudioCeilingHz = 15000.0;
constexpr double kFmPilotHz = 19000.0;
constexpr int64_t kCompositeAudioRateHz = 114000;
// The carrier loop's design figures (core/dsp/vrx_reference.h:787-820): noise
// bandwidths acquiring and tracking, damping 1/sqrt(2), the level and lock
// one-poles' time constants, and the lock test's holds.
constexpr double kCarrierAcquireHz = 400.0;
constexpr double kCarrierTrackHz = 30.0;
constexpr double kCarrierDamping = 0.70710678118654752440;
constexpr double kCarrierLevelSeconds = 0.005;
constexpr double kCarrierLockSeconds = 0.020;
constexpr double kCarrierHoldInSeconds = 0.010;
constexpr double kCarrierHoldOutSeconds = 0.200;
constexpr double kCarrierClampCeiling = 0.5 * kPi;

int64_t iabs(int64_t v) { return v < 0 ? -v : v; }

void reduce(int64_t& num, int64_t& den) {
    if (den < 0) {
        num = -num;
        den = -den;
    }
    const int64_t g = std::gcd(num, den);
    if (g > 1) {
        num /= g;
        den /= g;
    }
}

int64_t divide_nearest(int64_t num, int64_t den) {
    const int64_t q = num / den, r = num % den;
    if (2 * iabs(r) < iabs(den)) return q;
    return r < 0 ? q - 1 : q + 1;
}

double i0(double x) {
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 256; ++k) {
        term *= q / (double(k) * double(k));
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}

double kaiser_shape(double a) {
    if (a > 50.0) return 0.1102 * (a - 8.7);
    if (a >= 21.0) return 0.5842 * std::pow(a - 21.0, 0.4) + 0.07886 * (a - 21.0);
    return 0.0;
}

// Kaiser's order estimate for a transition given as a fraction of the rate.
uint32_t kaiser_taps(double attenuation_db, double transition_fraction) {
    if (!(transition_fraction > 0.0)) return 0;
    const double radians = kTwoPi * transition_fraction;
    const double order = (attenuation_db - 8.0) / (2.285 * radians);
    if (!(order > 0.0)) return 2;
    return uint32_t(std::ceil(order)) + 1u;
}

double sinc_pi(double z) {
    if (z == 0.0) return 1.0;
    const double a = kPi * z;
    return std::sin(a) / a;
}

double kaiser_window(size_t i, size_t len, double beta, double norm) {
    if (len < 2) return 1.0;
    const double x = 2.0 * double(i) / double(len - 1) - 1.0;
    return i0(beta * std::sqrt(std::max(0.0, 1.0 - x * x))) / norm;
}

struct Band {
    int64_t low = 0, high = 0;
    int64_t width() const { return high - low; }
    int64_t reach() const { return std::max(-low, high); }
};

Band default_band(uint32_t mode) {
    switch (mode) {
        case RVX_DEMOD_RAW: return {-6000, 6000};
        case RVX_DEMOD_AM: return {-5000, 5000};
        case RVX_DEMOD_NFM: return {-8000, 8000};
        case RVX_DEMOD_WFM: return {-100000, 100000};
        case RVX_DEMOD_USB: return {300, 2700};
        case RVX_DEMOD_LSB: return {-2700, -300};
        case RVX_DEMOD_DSB: return {-3000, 3000};
        case RVX_DEMOD_CW: return {-600, 600};
        case RVX_DEMOD_SAM: return {-5000, 5000};
        default: return {};
    }
}

int32_t resolve_band(const rvx_vrx_request& r, Band& out, const char* who) {
    if (r.passband_low != 0 || r.passband_high != 0) {
        if (r.passband_low >= r.passband_high)
            return rvx::fail(RVX_EINVAL, "%s: passband %lld..%lld Hz is empty or inverted", who,
                             (long long)r.passband_low, (long long)r.passband_high);
        out = {r.passband_low, r.passband_high};
        return RVX_OK;
    }
    if (r.bandwidth_hz < 0) return rvx::fail(RVX_EINVAL, "%s: negative bandwidth", who);
    if (r.bandwidth_hz == 0) {
        out = default_band(r.mode);
        if (out.width() <= 0) return rvx::fail(RVX_EINVAL, "%s: mode %u has no default", who, r.mode);
        return RVX_OK;
    }
    const int64_t half = r.bandwidth_hz / 2;
    if (r.mode == RVX_DEMOD_USB)
        out = {0, r.bandwidth_hz};
    else if (r.mode == RVX_DEMOD_LSB)
        out = {-r.bandwidth_hz, 0};
    else
        out = {-half, half};
    return RVX_OK;
}

int64_t widest(const rvx_vrx_placement& p) {
    if (p.channel_rate <= 0 || p.residual_den <= 0) return 0;
    const int64_t scaled = p.channel_rate * p.residual_den - 2 * iabs(p.residual_num);
    return scaled <= 0 ? 0 : scaled / p.residual_den;
}

Band clamp_band(const rvx_vrx_placement& p, Band b) {
    const int64_t w = widest(p);
    if (w <= 0) return {};
    const int64_t lim = w / 2;
    return {std::clamp(b.low, -lim, lim), std::clamp(b.high, -lim, lim)};
}

// Refuses a request one channel cannot deliver (vrx_reference.cpp:1440-1465):
// the channel carries nothing at all (the receiver sits half a channel or
// more from the channel's centre), or nothing of this band (it lies wholly
// past the limit). The message gives the request's edges against the limit.
int32_t channel_carries(const rvx_vrx_placement& p, Band req, const char* who) {
    const int64_t w = widest(p);
    if (w <= 0)
        return rvx::fail(RVX_EINVAL,
                         "%s: a %lld S/s channel with residual %lld/%lld Hz can carry nothing "
                         "for this receiver (it is half a channel or more off the channel "
                         "centre)",
                         who, (long long)p.channel_rate, (long long)p.residual_num,
                         (long long)p.residual_den);
    if (clamp_band(p, req).width() > 0) return RVX_OK;
    const int64_t lim = w / 2;
    return rvx::fail(RVX_EINVAL,
                     "%s: passband %lld..%lld Hz has nothing inside the %lld..%lld Hz one "
                     "%lld S/s channel carries here (residual %lld/%lld Hz)",
                     who, (long long)req.low, (long long)req.high, (long long)-lim,
                     (long long)lim, (long long)p.channel_rate, (long long)p.residual_num,
                     (long long)p.residual_den);
}

Band mix_frame(Band b, uint32_t mode, int64_t pitch) {
    if (mode != RVX_DEMOD_CW) return b;
    return {b.low + pitch, b.high + pitch};
}

int64_t minimum_rate(uint32_t mode, Band b) {
    const int64_t width = b.width();
    if (width <= 0) return 0;
    const int64_t floor_rate = std::max((3 * width + 1) / 2, 2 * b.reach());
    if (mode == RVX_DEMOD_AM) return std::max(floor_rate, 2 * width);
    return floor_rate;
}

// The demodulation rate plan_vrx builds (vrx_reference.cpp:1670
// demod_rate_for): the fitted band, moved by the CW pitch (a negative pitch
// read as zero), needs minimum_rate; the rate is the next whole multiple of
// the audio rate. Refuses what plan_vrx refuses for the band.
int32_t demod_rate_of(const rvx_vrx_request& r, const rvx_vrx_placement& pl, int64_t audio_rate,
                      Band& band, Band& mixed, int64_t& rate, const char* who) {
    if (audio_rate <= 0)
        return rvx::fail(RVX_EINVAL, "%s: audio rate %lld", who, (long long)audio_rate);
    Band req;
    if (int32_t rc = resolve_band(r, req, who)) return rc;
    if (int32_t rc = channel_carries(pl, req, who)) return rc;
    band = clamp_band(pl, req);
    mixed = mix_frame(band, r.mode, std::max<int64_t>(0, r.cw_pitch));
    const int64_t required = minimum_rate(r.mode, mixed);
    rate = std::max<int64_t>(1, (required + audio_rate - 1) / audio_rate) * audio_rate;
    return RVX_OK;
}

bool complex_output(uint32_t mode) { return mode == RVX_DEMOD_RAW; }

// core/engine/vrx.h:302 resolve_deemphasis: an explicit curve is kept except
// on the complex tap; the default is 75 us for WFM programme audio (below
// the composite rate) and none for every other mode.
uint32_t resolve_deemphasis(uint32_t mode, uint32_t requested, int64_t audio_rate) {
    if (requested != RVX_DEEMPH_DEFAULT)
        return complex_output(mode) ? uint32_t(RVX_DEEMPH_NONE) : requested;
    if (mode == RVX_DEMOD_WFM && audio_rate < kCompositeAudioRateHz) return RVX_DEEMPH_US75;
    return RVX_DEEMPH_NONE;
}

double deemphasis_seconds(uint32_t curve) {
    return curve == RVX_DEEMPH_US75 ? 75e-6 : curve == RVX_DEEMPH_EU50 ? 50e-6 : 0.0;
}

constexpr uint32_t kMaxDeemphasisTaps = 1024;  // vrx_reference.h:471
// FM stereo's pilot bandpass (vrx_reference.h:476-502).
constexpr uint32_t kMaxPilotTaps = 2048;
constexpr int64_t kStereoPilotHz = 19000;
constexpr double kStereoPilotHalfWidthHz = 1000.0;
constexpr double kStereoPilotTransitionHz = 3000.0;
constexpr double kStereoPilotAttenuationDb = 70.0;

// core/engine/vrx.h:381 resolve_stereo: a request, kept only by a WFM
// receiver delivering programme audio (below the composite rate).
bool resolve_stereo(uint32_t mode, bool requested, int64_t audio_rate) {
    return requested && mode == RVX_DEMOD_WFM && audio_rate < kCompositeAudioRateHz;
}

// vrx_reference.cpp:2129 stereo_pilot_taps: a 70 dB Kaiser bandpass with a
// 3 kHz transition at the demodulation rate, odd, in [16, 2048].
uint32_t stereo_pilot_taps(int64_t demod_rate) {
    if (demod_rate <= 0) return 0;
    const double transition = kStereoPilotTransitionHz / double(demod_rate);
    uint32_t taps = std::clamp(kaiser_taps(kStereoPilotAttenuationDb, transition), 16u,
                               kMaxPilotTaps);
    if ((taps % 2u) == 0u) ++taps;
    return std::min(taps, kMaxPilotTaps);
}

// vrx_reference.cpp:2147 design_stereo_tables. The pilot filter: a 1 kHz
// half-width Kaiser lowpass of unit DC gain modulated by exp(+j 2 pi 19k k /
// rate) on the convolution's own k, so a 19 kHz tone comes out with the
// prototype's gain and no delay (pilot_taps complex pairs). The rotation:
// twice the cosine and sine of the 38 kHz subcarrier's advance from the
// audio filter's centre tap (audio_taps pairs), which carries the pilot
// phase measured there to every tap.
void design_stereo_tables(int64_t demod_rate, uint32_t audio_taps, uint32_t pilot_taps,
                          float* pilot, float* rotation) {
    const double rate = double(demod_rate);
    const double cutoff = kStereoPilotHalfWidthHz / rate;
    const double centre = (double(pilot_taps) - 1.0) / 2.0;
    const double beta = kaiser_shape(kStereoPilotAttenuationDb), norm = i0(beta);
    std::vector<double> proto(pilot_taps);
    double sum = 0.0;
    for (uint32_t k = 0; k < pilot_taps; ++k) {
        proto[k] = 2.0 * cutoff * sinc_pi(2.0 * cutoff * (double(k) - centre)) *
                   kaiser_window(k, pilot_taps, beta, norm);
        sum += proto[k];
    }
    const double scale = std::abs(sum) < 1e-12 ? 1.0 : 1.0 / sum;
    const double pilot_turns = double(kStereoPilotHz) / rate;
    for (uint32_t k = 0; k < pilot_taps; ++k) {
        const double angle = kTwoPi * std::fmod(pilot_turns * double(k), 1.0);
        const double w = proto[k] * scale;
        pilot[2 * k] = float(w * std::cos(angle));
        pilot[2 * k + 1] = float(w * std::sin(angle));
    }
    const double sub_turns = 2.0 * pilot_turns;
    const int64_t pivot = int64_t(audio_taps / 2u);
    for (uint32_t t = 0; t < audio_taps; ++t) {
        const double offset = double(int64_t(t) - pivot);
        const double angle = kTwoPi * std::fmod(sub_turns * offset, 1.0);
        rotation[2 * t] = float(2.0 * std::cos(angle));
        rotation[2 * t + 1] = float(2.0 * std::sin(angle));
    }
}
constexpr uint32_t kMaxAudioTaps = 4096;       // vrx_reference.h:462

// vrx_reference.cpp:2011 design_deemphasis_taps: the impulse-invariant
// one-pole a = exp(-1/(tau*rate)) as its geometric series a^k, cut where a
// longer series cannot move a float (a^N <= 2^-24 (1 - a)), at most 1024
// terms, renormalised to unit sum in double and rounded once. No curve is
// one unit tap.
std::vector<float> deemphasis_taps(uint32_t curve, int64_t audio_rate) {
    const double tau = deemphasis_seconds(curve);
    if (!(tau > 0.0) || audio_rate <= 0) return {1.0f};
    const double pole = std::exp(-(1.0 / double(audio_rate)) / tau);
    const double exact = std::log((1.0 / 16777216.0) * (1.0 - pole)) / std::log(pole);
    const uint32_t n =
        std::min(uint32_t(std::ceil(std::max(1.0, exact))) + 1u, kMaxDeemphasisTaps);
    std::vector<double> series(n);
    double sum = 0.0, term = 1.0;
    for (uint32_t k = 0; k < n; ++k) {
        series[k] = term;
        sum += term;
        term *= pole;
    }
    std::vector<float> taps(n);
    const double scale = 1.0 / sum;
    for (uint32_t k = 0; k < n; ++k) taps[k] = float(series[k] * scale);
    return taps;
}

// Taps of the decimation filter with the curve folded in at the audio stride
// (vrx_reference.cpp:2079 fold_deemphasis).
uint64_t folded_taps(uint32_t decimation_taps, uint32_t curve_taps, uint32_t decimation) {
    return uint64_t(decimation_taps) + uint64_t(curve_taps - 1) * decimation;
}

}  // namespace

// Kaiser-windowed lowpass of cutoff half_width over phases * taps points,
// each phase normalised to unit DC gain, then modulated to centre_num /
// centre_den Hz about the filter's own centre with the angle reduced in
// integers (vrx_reference.cpp:1808 design_fine_taps). out: 2 * (length + 1)
// floats, the last pair zero.
static int32_t design_fine_table(const rvx_vrx_fine_config& fine, int64_t channel_rate,
                          int64_t half_width, int64_t centre_num, int64_t centre_den,
                          double attenuation_db, float* out, const char* who) {
    const size_t phases = fine.phases, taps = fine.taps, length = phases * taps;
    if (half_width <= 0 || 2 * half_width >= channel_rate || centre_den == 0)
        return rvx::fail(RVX_EINVAL, "%s: half width %lld at %lld S/s", who,
                         (long long)half_width, (long long)channel_rate);
    const double centre = (double(length) - 1.0) / 2.0;
    const double cutoff = double(half_width) / double(channel_rate);
    const double beta = kaiser_shape(attenuation_db), norm = i0(beta);
    std::vector<double> proto(length);
    for (size_t i = 0; i < length; ++i) {
        const double pos = (double(i) - centre) / double(phases);
        proto[i] = 2.0 * cutoff * sinc_pi(2.0 * cutoff * pos) *
                   kaiser_window(i, length, beta, norm);
    }
    for (size_t ph = 0; ph < phases; ++ph) {
        double sum = 0.0;
        for (size_t k = 0; k < taps; ++k) sum += proto[ph + k * phases];
        if (std::abs(sum) < 1e-12) continue;
        for (size_t k = 0; k < taps; ++k) proto[ph + k * phases] /= sum;
    }
    const int64_t period = 2 * iabs(centre_den) * int64_t(phases) * channel_rate;
    const int64_t num = centre_den < 0 ? -centre_num : centre_num;
    for (size_t i = 0; i < length; ++i) {
        const int64_t twice = 2 * int64_t(i) - (int64_t(length) - 1);
        int64_t turns = (num * twice) % period;
        if (turns < 0) turns += period;
        const double ang = kTwoPi * double(turns) / double(period);
        out[2 * i] = float(proto[i] * std::cos(ang));
        out[2 * i + 1] = float(proto[i] * std::sin(ang));
    }
    out[2 * length] = 0.0f;
    out[2 * length + 1] = 0.0f;
    return RVX_OK;
}

// The exact resampler recurrence for outputs [first_output, first_output +
// count) from Fc to Fd: n0 = floor(j0 Fc / Fd) and its remainder without
// forming j0 Fc, and the NCO phase j0 * delta mod 2^64.
static int32_t resample_block(const rvx_vrx_fine_config& fine, int64_t channel_rate, int64_t out_rate,
                       uint64_t nco_delta, uint32_t channel_base, uint32_t channel_mask,
                       uint32_t out_mask, uint64_t first_output, uint32_t count,
                       rvx_vrx_fine_params* out, uint64_t* first_input, const char* who) {
    if (channel_rate <= 0 || out_rate <= 0)
        return rvx::fail(RVX_EINVAL, "%s: the plan carries no rates", who);
    if (count == 0) return rvx::fail(RVX_EINVAL, "%s: zero outputs", who);
    const uint64_t fc = uint64_t(channel_rate), fd = uint64_t(out_rate);
    const uint64_t sw = fc / fd, sr = fc % fd;
    const uint64_t whole = first_output / fd, part = first_output % fd;
    const uint64_t cp = part * sr;
    const uint64_t n0 = first_output * sw + whole * sr + cp / fd;
    rvx_vrx_fine_params p{};
    p.chan_base = channel_base;
    p.chan_mask = channel_mask;
    p.in_offset = uint32_t(n0 & channel_mask);
    p.out_offset = uint32_t(first_output & out_mask);
    p.out_mask = out_mask;
    p.count = count;
    p.step_whole = uint32_t(sw);
    p.step_rem = uint32_t(sr);
    p.out_rate = uint32_t(fd);
    p.frac0 = uint32_t(cp % fd);
    p.inv_out_rate = 1.0f / float(fd);
    const uint64_t phase = nco_delta * first_output;  // wraps mod 2^64
    p.nco_phase_high = uint32_t(phase >> 32);
    p.nco_phase_low = uint32_t(phase);
    p.nco_delta_high = uint32_t(nco_delta >> 32);
    p.nco_delta_low = uint32_t(nco_delta);
    if (int32_t rc = rvx::vrx_fine_valid(fine, p, who)) return rc;
    *out = p;
    if (first_input) *first_input = n0;
    return RVX_OK;
}

extern "C" {

uint32_t rvx_channel_count_for(int64_t rate, int64_t widest_receiver_hz) {
    // The largest power-of-two channel count whose channels are each at least
    // as wide as the widest receiver: rate / M >= widest. Two is the floor (a
    // single channel channelizes nothing), also for a zero or negative width.
    if (rate <= 0 || widest_receiver_hz <= 0) return 2;
    const uint64_t fits = uint64_t(rate) / uint64_t(widest_receiver_hz);
    if (fits < 2) return 2;
    uint64_t m = 1;
    while (m <= fits / 2 && m < (uint64_t(1) << 31)) m <<= 1;
    return uint32_t(m);
}

int32_t rvx_channel_centre(const rvx_grid* grid, int64_t rate, uint32_t channel,
                           int64_t* numerator, int64_t* denominator) {
    static const char* who = "rvx_channel_centre";
    if (!grid || !numerator || !denominator) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    int64_t num = 0, den = 1;
    const int64_t M = grid->channels;
    if (M != 0) {
        // k*rate/M folded so that channels from M/2 up are negative frequencies,
        // reduced so that equal frequencies compare equal (DC is 0/1).
        const int64_t k = int64_t(channel % uint32_t(M));
        num = (k < M / 2 ? k : k - M) * rate;
        den = M;
        reduce(num, den);
    }
    *numerator = num;
    *denominator = den;
    return RVX_OK;
}

uint64_t rvx_channel_sample_to_input(const rvx_grid* grid, uint64_t group_delay_samples,
                                     uint64_t base, uint64_t m) {
    // Channel sample m is taken at input base + m*D and the linear-phase
    // prototype delays by its group delay, so it is centred on
    // base + m*D - delay. Modulo 2^64: a sample from before the filter filled
    // maps near 2^64, visibly impossible, rather than clamping to zero.
    const uint64_t d = grid ? grid->decimation : 0;
    return base + m * d - group_delay_samples;
}

int32_t rvx_place(const rvx_grid* grid, int64_t rate, const rvx_vrx_request* request,
                  rvx_vrx_placement* out) {
    static const char* who = "rvx_place";
    if (!grid || !request || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    const rvx_grid g = *grid;
    if (g.channels < 2 || !rvx::pow2(g.channels) || g.decimation == 0 ||
        g.channels % g.decimation != 0)
        return rvx::fail(RVX_EINVAL, "%s: invalid grid", who);
    if (rate <= 0) return rvx::fail(RVX_EINVAL, "%s: rate %lld", who, (long long)rate);
    if (rate % int64_t(g.decimation) != 0)
        return rvx::fail(RVX_EINVAL, "%s: rate %lld not divisible by decimation %u", who,
                         (long long)rate, g.decimation);
    const int64_t nyq = rate / 2;
    if (request->center_hz > nyq || request->center_hz < -nyq)
        return rvx::fail(RVX_EINVAL, "%s: centre %lld Hz outside +-%lld", who,
                         (long long)request->center_hz, (long long)nyq);
    Band req;
    if (int32_t rc = resolve_band(*request, req, who)) return rc;
    const int64_t M = g.channels;
    int64_t idx = divide_nearest(request->center_hz * M, rate) % M;
    if (idx < 0) idx += M;
    rvx_vrx_placement p{};
    p.channel = uint32_t(idx);
    // Channel k's centre k*rate/M folded into [-rate/2, rate/2), reduced.
    rvx_channel_centre(&g, rate, p.channel, &p.centre_num, &p.centre_den);
    p.channel_rate = rate / int64_t(g.decimation);
    p.residual_num = request->center_hz * p.centre_den - p.centre_num;
    p.residual_den = p.centre_den;
    const int64_t span = rate * p.residual_den;
    while (2 * p.residual_num > span) p.residual_num -= span;
    while (2 * p.residual_num <= -span) p.residual_num += span;
    reduce(p.residual_num, p.residual_den);
    const Band granted = clamp_band(p, req);
    p.granted_low = granted.low;
    p.granted_high = granted.high;
    p.bandwidth_clamped = (granted.low != req.low || granted.high != req.high) ? 1u : 0u;
    // An empty grant is reported, not refused: plan_vrx refuses it. But a
    // discriminator fed part of an FM channel recovers the wrong audio, not
    // less of it, so an FM receiver granted less than its mode's own channel
    // plan is refused here, naming the channel count that would carry it
    // (vrx_place.cpp:398-440).
    const Band plan_band = default_band(request->mode);
    if (p.bandwidth_clamped && (request->mode == RVX_DEMOD_NFM || request->mode == RVX_DEMOD_WFM) &&
        granted.width() < plan_band.width()) {
        const uint32_t carry = rvx_channel_count_for(rate, plan_band.width());
        const int64_t guaranteed = rate / int64_t(carry);
        return rvx::fail(RVX_EINVAL,
                         "%s: an FM receiver needs %lld Hz and one channel of this %u-channel "
                         "grid carries %lld Hz of it (%lld..%lld); a cut FM channel demodulates "
                         "to the wrong audio. %u channels guarantee %lld Hz%s",
                         who, (long long)plan_band.width(), g.channels,
                         (long long)granted.width(), (long long)granted.low,
                         (long long)granted.high, carry, (long long)guaranteed,
                         guaranteed >= plan_band.width() ? "" : " (the source rate is the limit)");
    }
    *out = p;
    return RVX_OK;
}

int32_t rvx_resolve_passband(const rvx_vrx_request* request, int64_t* low, int64_t* high) {
    static const char* who = "rvx_resolve_passband";
    if (!request || !low || !high) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    Band b;
    if (int32_t rc = resolve_band(*request, b, who)) return rc;
    *low = b.low;
    *high = b.high;
    return RVX_OK;
}

int32_t rvx_default_passband(uint32_t mode, int64_t* low, int64_t* high) {
    if (!low || !high) return rvx::fail(RVX_EINVAL, "rvx_default_passband: null argument");
    const Band b = default_band(mode);
    *low = b.low;
    *high = b.high;
    return b.width() > 0 ? RVX_OK
                         : rvx::fail(RVX_EINVAL, "rvx_default_passband: mode %u has none", mode);
}

int64_t rvx_max_channel_bandwidth(const rvx_vrx_placement* placement) {
    return placement ? widest(*placement) : 0;
}

int32_t rvx_clamp_to_channel(const rvx_vrx_placement* placement, int64_t* low, int64_t* high) {
    if (!placement || !low || !high)
        return rvx::fail(RVX_EINVAL, "rvx_clamp_to_channel: null argument");
    const Band b = clamp_band(*placement, Band{*low, *high});
    *low = b.low;
    *high = b.high;
    return RVX_OK;
}

int64_t rvx_minimum_demod_rate(uint32_t mode, int64_t bandwidth, int64_t cw_pitch) {
    if (bandwidth <= 0) return 0;
    const int64_t half = bandwidth / 2;
    Band b = mode == RVX_DEMOD_USB ? Band{0, bandwidth}
             : mode == RVX_DEMOD_LSB ? Band{-bandwidth, 0}
                                     : Band{-half, half};
    return minimum_rate(mode, mix_frame(b, mode, std::max<int64_t>(0, cw_pitch)));
}

int32_t rvx_demod_rate_for(const rvx_vrx_request* request, const rvx_vrx_placement* placement,
                           int64_t audio_rate, int64_t* out) {
    static const char* who = "rvx_demod_rate_for";
    if (!request || !placement || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    Band band, mixed;
    int64_t rate = 0;
    if (int32_t rc = demod_rate_of(*request, *placement, audio_rate, band, mixed, rate, who))
        return rc;
    *out = rate;
    return RVX_OK;
}

int32_t rvx_plan_vrx(const rvx_grid* grid, int64_t rate, const rvx_vrx_request* request,
                     const rvx_vrx_placement* placement, rvx_vrx_plan* out) {
    static const char* who = "rvx_plan_vrx";
    if (!grid || !request || !placement || !out)
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    const rvx_vrx_request r = *request;
    const rvx_vrx_placement pl = *placement;
    if (rate <= 0 || pl.channel_rate <= 0 || pl.residual_den <= 0)
        return rvx::fail(RVX_EINVAL, "%s: placement carries no rates", who);
    if (r.mode > RVX_DEMOD_CW && r.mode != RVX_DEMOD_SAM)
        return rvx::fail(RVX_EINVAL, "%s: mode %u is not planned by this build", who, r.mode);
    if (r.deemphasis > RVX_DEEMPH_EU50)
        return rvx::fail(RVX_EINVAL, "%s: de-emphasis %u is not a curve", who, r.deemphasis);
    rvx_vrx_plan plan{};
    plan.placement = pl;
    plan.mode = r.mode;
    plan.channel_rate = pl.channel_rate;
    plan.audio_rate = r.audio_rate > 0 ? r.audio_rate : kDefaultAudioRate;
    Band band, mixed;
    if (int32_t rc = demod_rate_of(r, pl, plan.audio_rate, band, mixed, plan.demod_rate, who))
        return rc;
    plan.passband_low = band.low;
    plan.passband_high = band.high;
    plan.bandwidth = band.width();
    const int64_t pitch = std::max<int64_t>(0, r.cw_pitch);
    const uint32_t decimation = uint32_t(plan.demod_rate / plan.audio_rate);

    const double stop_edge = 0.5 * double(std::min(plan.demod_rate, plan.channel_rate));
    const double limit =
        2.0 * std::min(stop_edge - double(mixed.high), stop_edge + double(mixed.low));
    const double transition = std::min(limit, 0.5 * double(plan.bandwidth));
    if (!(transition > 0.0))
        return rvx::fail(RVX_EINVAL, "%s: no transition band left at %lld S/s", who,
                         (long long)plan.demod_rate);
    plan.fine.phases = 256;
    plan.fine.nco_log2 = 16;
    plan.fine.taps = std::clamp(
        kaiser_taps(kFineAttenuationDb, transition / double(plan.channel_rate)), 8u, 256u);

    plan.filter_num = 2 * pl.residual_num + pl.residual_den * (band.low + band.high);
    plan.filter_den = 2 * pl.residual_den;
    plan.mix_num = pl.residual_num;
    plan.mix_den = pl.residual_den;
    if (r.mode == RVX_DEMOD_CW) plan.mix_num = pl.residual_num - pl.residual_den * pitch;
    reduce(plan.filter_num, plan.filter_den);
    reduce(plan.mix_num, plan.mix_den);
    if (int32_t rc = rvx_nco_delta(plan.mix_num, plan.mix_den, plan.demod_rate,
                                   &plan.fine_nco_delta))
        return rc;

    const bool cx = complex_output(r.mode);
    plan.demod.mode = r.mode;
    plan.demod.decimation = cx ? 1u : decimation;
    plan.output_rate = cx ? plan.demod_rate : plan.audio_rate;
    const bool programme = r.mode == RVX_DEMOD_WFM && plan.audio_rate < kCompositeAudioRateHz;
    const double nyq = 0.5 * double(plan.audio_rate);
    plan.audio_pass_hz = programme ? std::min(kFmAudioCeilingHz, 0.8 * nyq) : 0.8 * nyq;
    plan.audio_stop_hz = programme ? std::min(kFmPilotHz, nyq) : nyq;
    if (plan.demod.decimation == 1u) {
        plan.audio_decimation_taps = 1;
        plan.audio_pass_hz = plan.audio_stop_hz = 0.0;
    } else {
        const double frac = (plan.audio_stop_hz - plan.audio_pass_hz) / double(plan.demod_rate);
        uint32_t taps = std::clamp(kaiser_taps(kAudioAttenuationDb, frac), 3u, 1024u);
        if ((taps % 2u) == 0u) ++taps;
        plan.audio_decimation_taps = std::min(taps, 1024u);
    }
    // The de-emphasis curve is folded into the audio filter at the audio
    // stride, so the kernel still runs one FIR (vrx_reference.cpp:2462-2509).
    plan.deemphasis = resolve_deemphasis(r.mode, r.deemphasis, plan.audio_rate);
    const uint64_t audio_taps =
        folded_taps(plan.audio_decimation_taps,
                    uint32_t(deemphasis_taps(plan.deemphasis, plan.audio_rate).size()),
                    plan.demod.decimation);
    if (audio_taps > kMaxAudioTaps)
        return rvx::fail(RVX_EINVAL,
                         "%s: a %u-tap audio filter with its de-emphasis folded in at a "
                         "decimation of %u is %llu taps; the demodulator holds %u",
                         who, plan.audio_decimation_taps, plan.demod.decimation,
                         (unsigned long long)audio_taps, kMaxAudioTaps);
    plan.demod.audio_taps = uint32_t(audio_taps);
    // Stereo (vrx_reference.cpp:2515-2535): two channels and a pilot filter.
    plan.stereo = resolve_stereo(r.mode, r.mono == 0, plan.audio_rate) ? 1u : 0u;
    plan.demod.channels = (cx || plan.stereo) ? 2u : 1u;
    plan.demod.pilot_taps = plan.stereo ? stereo_pilot_taps(plan.demod_rate) : 0u;
    if (plan.stereo && (plan.demod.pilot_taps == 0 || 2 * kStereoPilotHz >= plan.demod_rate))
        return rvx::fail(RVX_EINVAL,
                         "%s: a stereo receiver needs a 19 kHz pilot filter and a 38 kHz "
                         "subcarrier, and %lld S/s carries neither",
                         who, (long long)plan.demod_rate);
    plan.demod.dc_taps =
        (r.mode == RVX_DEMOD_AM || r.mode == RVX_DEMOD_SAM)
            ? uint32_t(std::clamp<int64_t>(plan.demod_rate / kAmDcCornerHz, 16, 4096))
            : 1u;
    if (r.mode == RVX_DEMOD_NFM)
        plan.deviation = std::max<int64_t>(1, plan.bandwidth / 5);
    else if (r.mode == RVX_DEMOD_WFM)
        plan.deviation = std::max<int64_t>(1, (3 * plan.bandwidth) / 8);
    plan.demod_gain = 1.0f;
    if ((r.mode == RVX_DEMOD_NFM || r.mode == RVX_DEMOD_WFM) && plan.deviation > 0)
        plan.demod_gain = float(double(plan.demod_rate) / (kTwoPi * double(plan.deviation)));
    // The carrier loop (vrx_reference.cpp:2550-2561): its NCO is the fine
    // stage's table, and the integrator reaches as far as the passband does.
    plan.carrier.power = rvx_carrier_power(r.mode);
    plan.carrier.nco_log2 = plan.fine.nco_log2;
    if (plan.carrier.power != RVX_CARRIER_NONE) {
        if (int32_t rc = rvx_design_carrier_loop(plan.demod_rate, band.reach(), &plan.carrier_loop))
            return rc;
    }
    *out = plan;
    return RVX_OK;
}

int32_t rvx_design_deemphasis(uint32_t curve, int64_t audio_rate, float* out, uint32_t max_taps,
                              uint32_t* n) {
    static const char* who = "rvx_design_deemphasis";
    if (!n) return rvx::fail(RVX_EINVAL, "%s: null length", who);
    if (curve > RVX_DEEMPH_EU50 || curve == RVX_DEEMPH_DEFAULT)
        return rvx::fail(RVX_EINVAL, "%s: %u is not a resolved curve", who, curve);
    if (deemphasis_seconds(curve) > 0.0 && audio_rate <= 0)
        return rvx::fail(RVX_EINVAL, "%s: audio rate must be positive, got %lld", who,
                         (long long)audio_rate);
    const std::vector<float> taps = deemphasis_taps(curve, audio_rate);
    *n = uint32_t(taps.size());
    if (out != nullptr) {
        if (max_taps < taps.size())
            return rvx::fail(RVX_EINVAL, "%s: %zu taps do not fit %u", who, taps.size(), max_taps);
        std::copy(taps.begin(), taps.end(), out);
    }
    return RVX_OK;
}

int32_t rvx_plan_vrx_tables(const rvx_vrx_plan* plan, float* fine_taps, uint64_t n_fine_floats,
                            float* weights, uint64_t n_weights) {
    static const char* who = "rvx_plan_vrx_tables";
    if (!plan) return rvx::fail(RVX_EINVAL, "%s: null plan", who);
    const rvx_vrx_plan& p = *plan;
    const size_t phases = p.fine.phases, taps = p.fine.taps, length = phases * taps;
    if (fine_taps != nullptr) {
        if (n_fine_floats != 2 * (length + 1))
            return rvx::fail(RVX_EINVAL, "%s: fine table wants %zu floats", who, 2 * (length + 1));
        const int64_t half_width = p.bandwidth / 2;
        if (int32_t rc = design_fine_table(p.fine, p.channel_rate, half_width, p.filter_num,
                                           p.filter_den, kFineAttenuationDb, fine_taps, who))
            return rc;
    }
    if (weights != nullptr) {
        const uint32_t at = p.demod.audio_taps, dc = p.demod.dc_taps, pt = p.demod.pilot_taps;
        // audio taps, DC window, then (stereo) pilot pairs and rotation pairs
        const uint64_t want = uint64_t(at) + dc + (pt ? 2ull * pt + 2ull * at : 0ull);
        if (n_weights != want)
            return rvx::fail(RVX_EINVAL, "%s: weights want %llu floats", who,
                             (unsigned long long)want);
        if (pt != 0 && (pt > kMaxPilotTaps || 2 * kStereoPilotHz >= p.demod_rate))
            return rvx::fail(RVX_EINVAL, "%s: %u pilot taps at %lld S/s", who, pt,
                             (long long)p.demod_rate);
        if (p.deemphasis > RVX_DEEMPH_EU50)
            return rvx::fail(RVX_EINVAL, "%s: de-emphasis %u is not a curve", who, p.deemphasis);
        const std::vector<float> curve = deemphasis_taps(p.deemphasis, p.audio_rate);
        const uint32_t dt = p.audio_decimation_taps;
        if (dt == 0 || folded_taps(dt, uint32_t(curve.size()), p.demod.decimation) != at)
            return rvx::fail(RVX_EINVAL,
                             "%s: %u audio taps are not %u decimation taps with a %zu-tap "
                             "curve folded in at %u",
                             who, at, dt, curve.size(), p.demod.decimation);
        // Audio decimation filter: Kaiser lowpass, unit DC gain, rounded to
        // float (design_audio_taps :1921).
        std::vector<float> decim(dt);
        if (dt == 1) {
            decim[0] = 1.0f;
        } else {
            const double cut = (p.audio_pass_hz + p.audio_stop_hz) * 0.5 / double(p.demod_rate);
            const double centre = (double(dt) - 1.0) / 2.0;
            const double beta = kaiser_shape(kAudioAttenuationDb), norm = i0(beta);
            std::vector<double> d(dt);
            double sum = 0.0;
            for (uint32_t i = 0; i < dt; ++i) {
                d[i] = 2.0 * cut * sinc_pi(2.0 * cut * (double(i) - centre)) *
                       kaiser_window(i, dt, beta, norm);
                sum += d[i];
            }
            const double scale = std::abs(sum) < 1e-12 ? 1.0 : 1.0 / sum;
            for (uint32_t i = 0; i < dt; ++i) decim[i] = float(d[i] * scale);
        }
        // The curve folded in (fold_deemphasis :2079): the unit tap leaves the
        // floats as they are; otherwise the products of the two float tables
        // are summed in double at the audio stride and rounded once.
        if (curve.size() == 1 && curve[0] == 1.0f) {
            std::copy(decim.begin(), decim.end(), weights);
        } else {
            std::vector<double> folded(at, 0.0);
            for (size_t k = 0; k < curve.size(); ++k) {
                const double w = double(curve[k]);
                const size_t offset = k * p.demod.decimation;
                for (uint32_t j = 0; j < dt; ++j) folded[offset + j] += w * double(decim[j]);
            }
            for (uint32_t i = 0; i < at; ++i) weights[i] = float(folded[i]);
        }
        // DC window: raised cosine over dc taps, unit sum.
        if (dc <= 1) {
            weights[at] = 1.0f;
        } else {
            std::vector<double> wdc(dc);
            double sum = 0.0;
            for (uint32_t i = 0; i < dc; ++i) {
                wdc[i] = 0.5 - 0.5 * std::cos(kTwoPi * double(i + 1) / double(dc + 1));
                sum += wdc[i];
            }
            const double scale = sum < 1e-12 ? 1.0 : 1.0 / sum;
            for (uint32_t i = 0; i < dc; ++i) weights[at + i] = float(wdc[i] * scale);
        }
        if (pt != 0)
            design_stereo_tables(p.demod_rate, at, pt, weights + at + dc,
                                 weights + at + dc + 2ull * pt);
    }
    return RVX_OK;
}

int32_t rvx_fine_block(const rvx_vrx_plan* plan, uint32_t channel_base, uint32_t channel_mask,
                       uint32_t fine_mask, uint64_t first_output, uint32_t count,
                       rvx_vrx_fine_params* out, uint64_t* first_input) {
    static const char* who = "rvx_fine_block";
    if (!plan || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    return resample_block(plan->fine, plan->channel_rate, plan->demod_rate, plan->fine_nco_delta,
                          channel_base, channel_mask, fine_mask, first_output, count, out,
                          first_input, who);
}

int32_t rvx_demod_block(const rvx_vrx_plan* plan, uint32_t fine_mask, uint64_t first_audio,
                        uint32_t count, rvx_vrx_demod_params* out) {
    static const char* who = "rvx_demod_block";
    if (!plan || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (count == 0) return rvx::fail(RVX_EINVAL, "%s: zero outputs", who);
    const uint64_t first_fine = first_audio * plan->demod.decimation;
    rvx_vrx_demod_params p{};
    p.in_mask = fine_mask;
    p.in_offset = uint32_t(first_fine & fine_mask);
    p.count = count;
    p.gain = plan->demod_gain;
    if (int32_t rc = rvx::vrx_demod_valid(plan->demod, p, who)) return rc;
    *out = p;
    return RVX_OK;
}

uint32_t rvx_demod_history(const rvx_vrx_demod_config* config) {
    return config ? rvx::demod_fine_history(*config) : 0u;
}

uint32_t rvx_carrier_power(uint32_t mode) {
    if (mode == RVX_DEMOD_SAM) return RVX_CARRIER_PLL;
    if (mode == RVX_DEMOD_DSB) return RVX_CARRIER_COSTAS;
    return RVX_CARRIER_NONE;
}

int32_t rvx_carrier_loop_gains(double noise_bandwidth_hz, double damping, int64_t rate,
                               double* k1, double* k2) {
    if (!k1 || !k2) return rvx::fail(RVX_EINVAL, "rvx_carrier_loop_gains: null output");
    *k1 = *k2 = 0.0;
    if (!(noise_bandwidth_hz > 0.0) || !(damping > 0.0) || rate <= 0) return RVX_OK;
    // Gardner's second-order loop at sample period T: theta = Bn*T/(zeta +
    // 1/(4 zeta)), K1 = 4 zeta theta / D, K2 = 4 theta^2 / D with D = 1 + 2
    // zeta theta + theta^2.
    const double period = 1.0 / double(rate);
    const double theta = noise_bandwidth_hz * period / (damping + 1.0 / (4.0 * damping));
    const double denominator = 1.0 + 2.0 * damping * theta + theta * theta;
    *k1 = 4.0 * damping * theta / denominator;
    *k2 = 4.0 * theta * theta / denominator;
    return RVX_OK;
}

int32_t rvx_design_carrier_loop(int64_t demod_rate, int64_t reach_hz,
                                rvx_vrx_carrier_params* out) {
    static const char* who = "rvx_design_carrier_loop";
    if (!out) return rvx::fail(RVX_EINVAL, "%s: null output", who);
    if (demod_rate <= 0)
        return rvx::fail(RVX_EINVAL, "%s: demodulation rate must be positive, got %lld", who,
                         (long long)demod_rate);
    if (reach_hz <= 0)
        return rvx::fail(RVX_EINVAL, "%s: a passband reaching %lld Hz from its centre holds no "
                         "carrier", who, (long long)reach_hz);
    const double fs = double(demod_rate);
    double a1, a2, t1, t2;
    rvx_carrier_loop_gains(kCarrierAcquireHz, kCarrierDamping, demod_rate, &a1, &a2);
    rvx_carrier_loop_gains(kCarrierTrackHz, kCarrierDamping, demod_rate, &t1, &t2);
    const auto one_pole = [fs](double seconds) {
        return float(1.0 - std::exp(-1.0 / (seconds * fs)));
    };
    const auto samples = [fs](double seconds) {
        return uint32_t(std::max(1.0, std::round(seconds * fs)));
    };
    rvx_vrx_carrier_params p{};
    p.k1_acquire = float(a1);
    p.k2_acquire = float(a2);
    p.k1_track = float(t1);
    p.k2_track = float(t2);
    p.level_alpha = one_pole(kCarrierLevelSeconds);
    p.lock_alpha = one_pole(kCarrierLockSeconds);
    p.clamp_radians = float(std::min(kTwoPi * double(reach_hz) / fs, kCarrierClampCeiling));
    p.hold_in = samples(kCarrierHoldInSeconds);
    p.hold_out = samples(kCarrierHoldOutSeconds);
    *out = p;
    return RVX_OK;
}

int32_t rvx_carrier_retune(int64_t moved_hz, int64_t granted_width, int64_t demod_rate,
                           int32_t* reset, double* shift) {
    if (!reset || !shift) return rvx::fail(RVX_EINVAL, "rvx_carrier_retune: null output");
    *reset = 0;
    *shift = 0.0;
    // Strictly further than the granted width either way is another signal.
    if (moved_hz > granted_width || -moved_hz > granted_width) {
        *reset = 1;
        return RVX_OK;
    }
    // Inside it the carrier moves the other way in the fine stream by exactly
    // the move, in radians per sample at the demodulation rate.
    if (demod_rate > 0) *shift = -kTwoPi * double(moved_hz) / double(demod_rate);
    return RVX_OK;
}

int32_t rvx_carrier_block(const rvx_vrx_plan* plan, uint32_t fine_mask, uint64_t first_fine,
                          uint32_t count, int32_t reset, float shift,
                          rvx_vrx_carrier_params* out) {
    static const char* who = "rvx_carrier_block";
    if (!plan || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (plan->carrier.power == RVX_CARRIER_NONE)
        return rvx::fail(RVX_EINVAL, "%s: a mode %u receiver runs no carrier loop", who,
                         plan->mode);
    rvx_vrx_carrier_params p = plan->carrier_loop;
    p.in_mask = fine_mask;
    p.in_offset = uint32_t(first_fine & uint64_t(fine_mask));
    p.count = count;
    p.flags = reset ? RVX_CARRIER_FLAG_RESET : 0u;
    p.shift = reset ? 0.0f : shift;  // a restarted loop has no frequency to carry
    if (int32_t rc = rvx::vrx_carrier_valid(plan->carrier, p, who)) return rc;
    *out = p;
    return RVX_OK;
}

}  // extern "C"

// ---- The display tap (vrx_reference.cpp:2938-3060) ---------------------------

namespace {
constexpr uint32_t kDisplayTaps = 256;
constexpr uint32_t kMaxFinePhasesDisplay = 256;
constexpr double kDisplayAttenuationDb = 80.0;
constexpr double kDisplayAttenuationCeilingDb = 100.0;

// Kaiser's order estimate read the other way: the depth `taps` reach across
// a transition given as a fraction of the rate.
double attenuation_reachable(uint32_t taps, double transition_fraction) {
    if (taps < 2 || !(transition_fraction > 0.0)) return 0.0;
    return 2.285 * (kTwoPi * transition_fraction) * double(taps - 1u) + 8.0;
}
}  // namespace

extern "C" {

uint32_t rvx_display_max_decimation(void) {
    // the transition is Fdisp / 2 = Fc / (2R), a falling fraction of the rate
    uint32_t best = 1;
    for (uint32_t r = 1; r <= kMaxFinePhasesDisplay; ++r) {
        if (attenuation_reachable(kDisplayTaps, 1.0 / (2.0 * double(r))) < kDisplayAttenuationDb)
            break;
        best = r;
    }
    return best;
}

uint32_t rvx_display_ladder(int64_t channel_rate, uint32_t* out, uint32_t max_rungs) {
    uint32_t rungs[64];
    uint32_t n = 0;
    rungs[n++] = 1;
    if (channel_rate > 0) {
        const uint32_t cap = rvx_display_max_decimation();
        for (;;) {
            uint32_t next = 0;
            for (uint32_t d = 2 * rungs[n - 1]; d <= cap; ++d)
                if (channel_rate % int64_t(d) == 0) {
                    next = d;
                    break;
                }
            if (next == 0 || n == 64) break;
            rungs[n++] = next;
        }
    }
    if (out)
        for (uint32_t i = 0; i < std::min(n, max_rungs); ++i) out[i] = rungs[i];
    return n;
}

uint32_t rvx_display_decimation(int64_t channel_rate, int64_t reach_hz) {
    if (channel_rate <= 0 || reach_hz <= 0) return 1;
    // Fdisp / 4 >= 2 reach: the pane is at least four reaches across
    const int64_t wanted = channel_rate / (8 * reach_hz);
    uint32_t rungs[64];
    const uint32_t n = rvx_display_ladder(channel_rate, rungs, 64);
    uint32_t chosen = 1;
    for (uint32_t i = 0; i < n; ++i)
        if (int64_t(rungs[i]) <= wanted) chosen = rungs[i];
    return chosen;
}

int32_t rvx_plan_vrx_display(const rvx_vrx_plan* plan, rvx_vrx_display_plan* out, float* taps,
                             uint64_t n_floats) {
    static const char* who = "rvx_plan_vrx_display";
    if (!plan || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    const rvx_vrx_plan& p = *plan;
    if (p.channel_rate <= 0) return rvx::fail(RVX_EINVAL, "%s: the plan carries no channel rate",
                                              who);
    if (p.mix_den == 0 || p.placement.residual_den == 0)
        return rvx::fail(RVX_EINVAL, "%s: the plan carries no mix frequency", who);
    // the passband in the frame the fine stage mixes to DC: its edges are about
    // the tuned frequency, (residual - mix) above the mix (the pitch for CW)
    const double shift = double(p.placement.residual_num) / double(p.placement.residual_den) -
                         double(p.mix_num) / double(p.mix_den);
    const double reach = std::max(std::abs(double(p.passband_low) + shift),
                                  std::abs(double(p.passband_high) + shift));
    rvx_vrx_display_plan d{};
    d.channel_rate = p.channel_rate;
    d.reach_hz = std::max<int64_t>(1, int64_t(std::ceil(reach)));
    d.decimation = rvx_display_decimation(p.channel_rate, d.reach_hz);
    d.display_rate = p.channel_rate / int64_t(d.decimation);
    d.fine.taps = kDisplayTaps;
    d.fine.phases = 1;
    d.fine.nco_log2 = p.fine.nco_log2;
    d.group_delay_channel_samples = (double(kDisplayTaps) - 1.0) / 2.0;
    // flat to the pane's edge Fdisp / 4, stopband by 3 Fdisp / 4 (the nearest
    // fold into the pane), or Fc / 2 at R = 1
    const double fc = double(p.channel_rate), fdisp = double(d.display_rate);
    d.pass_hz = 0.25 * fdisp;
    d.stop_hz = std::min(0.75 * fdisp, 0.5 * fc);
    const double reachable = attenuation_reachable(kDisplayTaps, (d.stop_hz - d.pass_hz) / fc);
    if (reachable < kDisplayAttenuationDb)
        return rvx::fail(RVX_EINVAL, "%s: %u taps reach %.1f dB across %.0f to %.0f Hz at %lld "
                         "S/s, short of %.0f dB", who, kDisplayTaps, reachable, d.pass_hz,
                         d.stop_hz, (long long)p.channel_rate, kDisplayAttenuationDb);
    d.attenuation_db = std::min(kDisplayAttenuationCeilingDb, reachable);
    if (int32_t rc = rvx_nco_delta(p.mix_num, p.mix_den, d.display_rate, &d.nco_delta)) return rc;
    if (taps) {
        if (n_floats != 2ull * (kDisplayTaps + 1))
            return rvx::fail(RVX_EINVAL, "%s: the display table wants %u floats", who,
                             2 * (kDisplayTaps + 1));
        // the cutoff midway between the edges
        const int64_t half_width = int64_t(0.5 * (d.pass_hz + d.stop_hz));
        if (int32_t rc = design_fine_table(d.fine, p.channel_rate, half_width, p.mix_num,
                                           p.mix_den, d.attenuation_db, taps, who))
            return rc;
    }
    *out = d;
    return RVX_OK;
}

int32_t rvx_display_block(const rvx_vrx_display_plan* plan, uint32_t channel_base,
                          uint32_t channel_mask, uint32_t display_mask, uint64_t first_output,
                          uint32_t count, rvx_vrx_fine_params* out, uint64_t* first_input) {
    static const char* who = "rvx_display_block";
    if (!plan || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    return resample_block(plan->fine, plan->channel_rate, plan->display_rate, plan->nco_delta,
                          channel_base, channel_mask, display_mask, first_output, count, out,
                          first_input, who);
}

}  // extern "C"
