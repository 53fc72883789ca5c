// The adaptive noise profile (core/dsp/noise_profile.h, noise_profile.cpp):
// one receiver's steady noise (hum, whine, birdies) learned and taken out of
// its mono audio on the host, after the audio has been read back.
//
// Audio flows through a short-time transform: frames of `size` samples every
// `hop` = size / 2, a periodic sqrt-Hann window on the way in and again on the
// way out, so the two windows sum to one across the overlap and unit gain
// gives back the input `size` samples late. Per frame and bin:
//   power     |X|^2, smoothed frame to frame (keep 0.7);
//   estimate  learned (the mean power over one second after a learn request,
//             kept until a reset), else the quiet average (a one-pole average
//             over frames taken with the squelch shut, time constant 0.5 s,
//             kept through open passages), else minimum statistics (the
//             smallest smoothed power over two seconds in eight parts, times
//             1.5 for the minimum's bias); the last two only with auto-learn;
//   gain      Wiener, snr / (1 + snr) with snr = power / noise - 1 (not below
//             zero), not below the floor 10^(-strength / 20); smoothed frame to
//             frame (keep 0.5); then [1 2 1] / 4 across bins, except that a bin
//             at the floor keeps its own gain, so smoothing never lets a noise
//             bin rise above the floor; clamped to [floor, 1].
// The constants and rules are the reference's definition of the stage's
// output (noise_profile.cpp:14-38 and its run_frame); the transform, buffers
// and code here are this build's.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <new>
#include <vector>

#include "common.h"

namespace {

using cd = std::complex<double>;

constexpr double kPowerKeep = 0.7;
constexpr double kQuietSeconds = 0.5;
constexpr double kMinSeconds = 2.0;
constexpr uint32_t kMinParts = 8;
constexpr double kMinBias = 1.5;
constexpr double kGainKeep = 0.5;
constexpr double kLearnSeconds = 1.0;
constexpr double kPi = 3.14159265358979323846;

// Radix-2 complex transform of one fixed size: the input permuted by a
// bit-reversal table, then log2(n) stages of butterflies against exp(-2 pi i k / n).
class Transform {
public:
    void resize(uint32_t n) {
        n_ = n;
        uint32_t bits = 0;
        while ((1u << bits) < n) ++bits;
        rev_.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t r = 0;
            for (uint32_t b = 0; b < bits; ++b) r |= ((i >> b) & 1u) << (bits - 1 - b);
            rev_[i] = r;
        }
        w_.resize(n / 2);
        for (uint32_t k = 0; k < n / 2; ++k) w_[k] = std::polar(1.0, -2.0 * kPi * k / n);
    }
    void forward(std::vector<cd>& x) const {
        for (uint32_t i = 0; i < n_; ++i)
            if (i < rev_[i]) std::swap(x[i], x[rev_[i]]);
        for (uint32_t half = 1; half < n_; half *= 2) {
            const uint32_t stride = n_ / (2 * half);
            for (uint32_t base = 0; base < n_; base += 2 * half)
                for (uint32_t j = 0; j < half; ++j) {
                    const cd t = w_[j * stride] * x[base + j + half];
                    x[base + j + half] = x[base + j] - t;
                    x[base + j] += t;
                }
        }
    }

private:
    uint32_t n_ = 0;
    std::vector<uint32_t> rev_;
    std::vector<cd> w_;
};

}  // namespace

struct rvx_noise_profile {
    // framing
    double rate = 0.0;
    uint32_t size = 0, hop = 0, bins = 0;
    Transform transform;
    std::vector<double> window;
    std::vector<cd> spec;
    std::vector<double> frame_in;  // the last `size` input samples
    std::vector<double> overlap;   // overlap-add of the synthesised frames
    std::vector<double> ready;     // `hop` output samples being played out
    uint32_t fill = 0, played = 0;

    // estimates
    std::vector<double> power, quiet_avg, part_min, part_mins, gain, noise;
    bool have_quiet = false, have_noise = false;
    uint32_t part_frames = 0, part_slot = 0, parts_done = 0;
    std::vector<double> learned, learn_sum;
    uint32_t learn_left = 0, learn_frames = 0;
    bool have_learned = false;

    // settings as last seen
    bool enabled_before = false, baseline = false;
    uint32_t learn_count = 0, reset_count = 0;

    void configure(double r) {
        rate = r;
        size = r >= 32000.0 ? 1024u : (r >= 16000.0 ? 512u : 256u);
        hop = size / 2;
        bins = size / 2 + 1;
        transform.resize(size);
        window.resize(size);
        for (uint32_t i = 0; i < size; ++i)
            window[i] = std::sqrt(0.5 - 0.5 * std::cos(2.0 * kPi * double(i) / double(size)));
        spec.assign(size, cd());
        learned.assign(bins, 0.0);
        learn_sum.assign(bins, 0.0);
        have_learned = false;
        learn_left = 0;
        restart();
    }

    // Everything but the learned profile: the stream and the automatic estimates.
    void restart() {
        frame_in.assign(size, 0.0);
        overlap.assign(size, 0.0);
        ready.assign(hop, 0.0);
        fill = size - hop;
        played = 0;
        power.assign(bins, 0.0);
        quiet_avg.assign(bins, 0.0);
        have_quiet = false;
        part_min.assign(bins, HUGE_VAL);
        part_mins.assign(size_t(bins) * kMinParts, HUGE_VAL);
        part_frames = part_slot = parts_done = 0;
        gain.assign(bins, 1.0);
        noise.assign(bins, 0.0);
        have_noise = false;
    }

    void frame(bool quiet, bool auto_learn, double floor) {
        for (uint32_t i = 0; i < size; ++i) spec[i] = cd(frame_in[i] * window[i], 0.0);
        transform.forward(spec);

        const bool first = parts_done == 0 && part_frames == 0;
        const double quiet_keep = std::exp(-double(hop) / (rate * kQuietSeconds));
        const bool take_quiet = auto_learn && quiet;
        for (uint32_t k = 0; k < bins; ++k) {
            const double p = std::norm(spec[k]);
            power[k] = first ? p : kPowerKeep * power[k] + (1.0 - kPowerKeep) * p;
            if (learn_left) learn_sum[k] += p;
            if (take_quiet)
                quiet_avg[k] = have_quiet ? quiet_keep * quiet_avg[k] + (1.0 - quiet_keep) * p
                                          : power[k];
            part_min[k] = std::min(part_min[k], power[k]);
        }
        if (take_quiet) have_quiet = true;
        if (learn_left && --learn_left == 0) {
            for (uint32_t k = 0; k < bins; ++k) learned[k] = learn_sum[k] / learn_frames;
            have_learned = true;
        }
        const long frames_per_part = std::lround(kMinSeconds * rate / hop / double(kMinParts));
        if (++part_frames >= uint32_t(std::max(1L, frames_per_part))) {
            std::copy(part_min.begin(), part_min.end(), part_mins.begin() + size_t(part_slot) * bins);
            std::fill(part_min.begin(), part_min.end(), HUGE_VAL);
            part_slot = (part_slot + 1) % kMinParts;
            part_frames = 0;
            ++parts_done;
        }

        have_noise = true;
        if (have_learned) {
            noise = learned;
        } else if (auto_learn && have_quiet) {
            noise = quiet_avg;
        } else if (auto_learn && parts_done > 0) {
            for (uint32_t k = 0; k < bins; ++k) {
                double m = part_min[k];
                for (uint32_t u = 0; u < kMinParts; ++u) m = std::min(m, part_mins[size_t(u) * bins + k]);
                noise[k] = kMinBias * m;
            }
        } else {
            have_noise = false;
        }

        for (uint32_t k = 0; k < bins; ++k) {
            double want = 1.0;
            if (have_noise && noise[k] > 0.0) {
                const double snr = std::max(power[k] / noise[k] - 1.0, 0.0);
                want = std::max(snr / (1.0 + snr), floor);
            }
            gain[k] = kGainKeep * gain[k] + (1.0 - kGainKeep) * want;
        }
        for (uint32_t k = 0; k < bins; ++k) {
            const double left = gain[k == 0 ? 0 : k - 1];
            const double right = gain[k + 1 < bins ? k + 1 : k];
            const double own = gain[k];
            double g = own <= floor * 1.0001 ? own : 0.25 * left + 0.5 * own + 0.25 * right;
            g = std::min(std::max(g, floor), 1.0);
            spec[k] *= g;
            if (k != 0 && k != size / 2) spec[size - k] *= g;
        }

        // inverse: conjugate, forward, conjugate (the real part is all that is kept)
        for (cd& b : spec) b = std::conj(b);
        transform.forward(spec);
        const double scale = 1.0 / double(size);
        for (uint32_t i = 0; i < size; ++i) overlap[i] += spec[i].real() * scale * window[i];
    }
};

extern "C" {

int32_t rvx_noise_profile_create(rvx_noise_profile** out) {
    if (!out) return rvx::fail(RVX_EINVAL, "rvx_noise_profile_create: null output");
    *out = new (std::nothrow) rvx_noise_profile();
    if (!*out) return rvx::fail(RVX_ENOMEM, "rvx_noise_profile_create: out of memory");
    return RVX_OK;
}

void rvx_noise_profile_destroy(rvx_noise_profile* profile) { delete profile; }

int32_t rvx_noise_profile_process(rvx_noise_profile* np, float* audio, uint64_t n, double rate,
                                  int32_t quiet, const rvx_noise_profile_settings* settings) {
    static const char* who = "rvx_noise_profile_process";
    if (!np || !settings || (n && !audio)) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (!std::isfinite(settings->strength_db))
        return rvx::fail(RVX_EINVAL, "%s: strength %g dB", who, settings->strength_db);
    if (!(rate > 0.0) || !std::isfinite(rate)) return RVX_OK;  // no rate: nothing to do
    rvx_noise_profile& p = *np;
    if (rate != p.rate) p.configure(rate);

    if (!p.baseline) {
        p.baseline = true;
        p.learn_count = settings->learn_count;
        p.reset_count = settings->reset_count;
    }
    if (settings->reset_count != p.reset_count) {
        p.reset_count = settings->reset_count;
        p.have_learned = false;
        p.learn_left = 0;
        p.restart();
    }
    if (settings->learn_count != p.learn_count) {
        p.learn_count = settings->learn_count;
        std::fill(p.learn_sum.begin(), p.learn_sum.end(), 0.0);
        p.learn_frames = std::max<uint32_t>(1, uint32_t(std::ceil(kLearnSeconds * rate / p.hop)));
        p.learn_left = p.learn_frames;
    }
    if (!settings->enabled) {
        if (p.enabled_before) {  // back on later, the stream restarts clean
            p.enabled_before = false;
            p.restart();
        }
        return RVX_OK;
    }
    p.enabled_before = true;

    const double strength =
        std::min(std::max(settings->strength_db, 0.0), RVX_NOISE_PROFILE_MAX_STRENGTH_DB);
    const double floor = std::pow(10.0, -strength / 20.0);
    for (uint64_t i = 0; i < n; ++i) {
        p.frame_in[p.fill++] = double(audio[i]);
        audio[i] = float(p.ready[p.played++]);
        if (p.fill == p.size) {
            p.frame(quiet != 0, settings->auto_learn != 0, floor);
            std::copy(p.overlap.begin(), p.overlap.begin() + p.hop, p.ready.begin());
            std::copy(p.overlap.begin() + p.hop, p.overlap.end(), p.overlap.begin());
            std::fill(p.overlap.begin() + (p.size - p.hop), p.overlap.end(), 0.0);
            std::copy(p.frame_in.begin() + p.hop, p.frame_in.end(), p.frame_in.begin());
            p.fill = p.size - p.hop;
            p.played = 0;
        }
    }
    return RVX_OK;
}

int32_t rvx_noise_profile_retune(rvx_noise_profile* np) {
    if (!np) return rvx::fail(RVX_EINVAL, "rvx_noise_profile_retune: null profile");
    if (np->size) np->restart();
    return RVX_OK;
}

int32_t rvx_noise_profile_state(const rvx_noise_profile* np, int32_t* learning, int32_t* learned,
                                uint32_t* latency, uint32_t* bins) {
    if (!np) return rvx::fail(RVX_EINVAL, "rvx_noise_profile_state: null profile");
    if (learning) *learning = np->learn_left > 0;
    if (learned) *learned = np->have_learned;
    if (latency) *latency = np->size;
    if (bins) *bins = np->have_noise ? np->bins : 0u;
    return RVX_OK;
}

int32_t rvx_noise_profile_estimate(const rvx_noise_profile* np, double* out, uint32_t max_out,
                                   uint32_t* n_out) {
    if (!np || !n_out || (max_out && !out))
        return rvx::fail(RVX_EINVAL, "rvx_noise_profile_estimate: null argument");
    const uint32_t n = np->have_noise ? std::min(np->bins, max_out) : 0u;
    std::copy(np->noise.begin(), np->noise.begin() + n, out);
    *n_out = n;
    return RVX_OK;
}

}  // extern "C"
