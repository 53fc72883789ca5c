// Detection events: from a plain decision's accepted bands to candidates, the
// residual of a stopped signal withheld, and candidates followed from decision
// to decision as tracks whose first and last sample indices are the event's
// boundaries.
//
//   rvx_candidates        core/detect/detector.cpp:1168-1334 (find_candidates' tail,
//                         emit_candidate): the occupied band, the interior-gap split,
//                         the SNR test in the 2500 Hz reference bandwidth, centre and
//                         width in hertz from the bin edges
//   rvx_tracker_update    :1336-1583 reject_residual, then :1585-1900 update_tracks
//
// The reference runs these on the host after each decision, over tens of
// candidates; so does this build. They work in double with the reference's
// operation order (sequential prefix sums, the same comparisons), so the bins,
// hertz and sample indices they produce are the reference's integers.
// Not built: the band shape measurement, margin confidence, emitter memory,
// spur cancelling, channel association and per-track level statistics.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <vector>

#include "common.h"

namespace {

constexpr double kReferenceBandwidthHz = 2500.0;       // detector.h:128
constexpr float kDbToLog2 = 0.332192809488736234787F;   // detector.cpp:26
constexpr double kPowerFloor = double(1.0e-20F);        // spectrum_reference.h:104
constexpr double kDbPerNeper = 4.342944819032518;       // 10 / ln 10
constexpr uint32_t kParts = 8;                          // detector.h:151
constexpr double kPartFloor = 0.5;                      // detector.h:177
constexpr uint32_t kUnassigned = 0xFFFFFFFFu;
// The spur rule's numbers (detector.h:1302-1314, refine.h:179): a line is at
// most 500 Hz or 3 bins wide, its frame-to-frame level sigma under 1.5 dB over
// 50 frames, 500 Hz clear of DC, continuous for 5 s; sightings match within
// 30 Hz of offset, tunings differ by 2 kHz, and both are kept 300 s.
constexpr double kSpurMaxBandwidthHz = 500.0;
constexpr double kSpurMaxBins = 3.0;
constexpr double kSpurLevelSigmaDb = 1.5;
constexpr uint32_t kSpurMinFrames = 50;
constexpr double kSpurDcGuardHz = 500.0;
constexpr double kContinuousSeconds = 5.0;
constexpr double kLoSpurToleranceHz = 30.0;
constexpr double kLoSpurMinRetuneHz = 2000.0;
constexpr double kSpurMemorySeconds = 300.0;
// The emitter memory's (refine.h:184-189, 277): coming live after a second of
// an emitter's silence is a key-up; an emitter is kept 1800 s, at most 512.
constexpr double kKeyupGapSeconds = 1.0;
constexpr double kEmitterMemorySeconds = 1800.0;
constexpr size_t kEmitterCapacity = 512;
constexpr int64_t kEmitterMatchMinHz = 1000;
// Hybrid digital sidebands beside wideband FM (hd_hybrid.h): the analogue
// signal within 100 kHz of its centroid, the digital carriers 129-198 kHz out
// on each side, the guard gap 105-120 kHz out; each side 3 dB over the floor,
// 1.5 dB over the gap, tilted by at most 3 dB, within 6 dB of the other;
// tracks of at least 50 kHz, confirmed after 2 decisions, released after 3.
constexpr double kHdInnerHz = 129000.0, kHdOuterHz = 198000.0, kHdAnalogueHalfHz = 100000.0;
constexpr double kHdGapInnerHz = 105000.0, kHdGapOuterHz = 120000.0;
constexpr double kHdOverFloorDb = 3.0, kHdOverGapDb = 1.5, kHdMaxTiltDb = 3.0,
                 kHdMaxAsymmetryDb = 6.0;
constexpr int64_t kHdMinTrackHz = 50000;
constexpr uint32_t kHdConfirm = 2, kHdRelease = 3;
constexpr double kHdLevelSmoothing = 0.3;  // detector.cpp:1947

// Power and floor summed over the bins whose centres lie in [low, high].
struct HdSums {
    double power = 0.0, floor = 0.0;
    size_t bins = 0;
    double excess() const { return std::max(0.0, power - floor); }
    double density() const { return bins ? power / double(bins) : 0.0; }
};

struct HdSpectrum {
    const double *power, *floor;
    size_t n;
    double zero, width;

    HdSums sum(double low, double high) const {
        HdSums out;
        const double first = std::ceil((low - zero) / width), last = std::floor((high - zero) / width);
        if (!(first >= 0.0) || !(last < double(n)) || last < first) return out;
        for (size_t i = size_t(first); i <= size_t(last); ++i) {
            out.power += power[i];
            out.floor += floor[i];
            ++out.bins;
        }
        return out;
    }
    // Excess-weighted mean frequency within half_hz of centre.
    double centroid(double centre, double half_hz) const {
        const double first = std::max(0.0, std::ceil((centre - half_hz - zero) / width));
        const double last = std::min(double(n) - 1.0, std::floor((centre + half_hz - zero) / width));
        double weight = 0.0, moment = 0.0;
        for (double b = first; b <= last; b += 1.0) {
            const size_t i = size_t(b);
            const double e = std::max(0.0, power[i] - floor[i]);
            weight += e;
            moment += e * (zero + b * width);
        }
        return weight > 0.0 ? moment / weight : centre;
    }
};

double hd_ratio_db(double a, double b) {
    if (!(a > 0.0) || !(b > 0.0)) return a > 0.0 ? 99.0 : -99.0;
    return 10.0 * std::log10(a / b);
}

// measure_hd_sidebands (hd_hybrid.cpp): the analogue centre by two centroid
// passes (clamped within 100 kHz), then the sidebands' level against the
// analogue excess, the floor and the gap, and each side's inner-over-outer tilt.
rvx_hd_reading measure_hd(const double* power, const double* floor, size_t n, double zero,
                          double width, double centre_hz) {
    rvx_hd_reading out{};
    if (n == 0 || !(width > 0.0) || !std::isfinite(centre_hz) || !std::isfinite(zero)) return out;
    const HdSpectrum sp{power, floor, n, zero, width};
    double centre = centre_hz;
    for (int pass = 0; pass < 2; ++pass) centre = sp.centroid(centre, kHdAnalogueHalfHz);
    centre = std::clamp(centre, centre_hz - kHdAnalogueHalfHz, centre_hz + kHdAnalogueHalfHz);
    out.centre_hz = centre;
    const HdSums analogue = sp.sum(centre - kHdAnalogueHalfHz, centre + kHdAnalogueHalfHz);
    const HdSums lower = sp.sum(centre - kHdOuterHz, centre - kHdInnerHz);
    const HdSums upper = sp.sum(centre + kHdInnerHz, centre + kHdOuterHz);
    const HdSums lower_gap = sp.sum(centre - kHdGapOuterHz, centre - kHdGapInnerHz);
    const HdSums upper_gap = sp.sum(centre + kHdGapInnerHz, centre + kHdGapOuterHz);
    const double middle = 0.5 * (kHdInnerHz + kHdOuterHz);
    const HdSums lower_inner = sp.sum(centre - middle, centre - kHdInnerHz);
    const HdSums lower_outer = sp.sum(centre - kHdOuterHz, centre - middle);
    const HdSums upper_inner = sp.sum(centre + kHdInnerHz, centre + middle);
    const HdSums upper_outer = sp.sum(centre + middle, centre + kHdOuterHz);
    if (!analogue.bins || !lower.bins || !upper.bins || !lower_gap.bins || !upper_gap.bins ||
        !(analogue.excess() > 0.0))
        return out;
    out.measured = 1;
    out.lower_dbc = hd_ratio_db(lower.excess(), analogue.excess());
    out.upper_dbc = hd_ratio_db(upper.excess(), analogue.excess());
    out.lower_over_floor_db = hd_ratio_db(lower.power, lower.floor);
    out.upper_over_floor_db = hd_ratio_db(upper.power, upper.floor);
    out.lower_over_gap_db = hd_ratio_db(lower.density(), lower_gap.density());
    out.upper_over_gap_db = hd_ratio_db(upper.density(), upper_gap.density());
    out.lower_tilt_db = hd_ratio_db(lower_inner.density(), lower_outer.density());
    out.upper_tilt_db = hd_ratio_db(upper_inner.density(), upper_outer.density());
    out.hybrid = out.lower_over_floor_db >= kHdOverFloorDb && out.upper_over_floor_db >= kHdOverFloorDb &&
                 out.lower_over_gap_db >= kHdOverGapDb && out.upper_over_gap_db >= kHdOverGapDb &&
                 std::abs(out.lower_tilt_db) <= kHdMaxTiltDb &&
                 std::abs(out.upper_tilt_db) <= kHdMaxTiltDb &&
                 std::abs(out.lower_dbc - out.upper_dbc) <= kHdMaxAsymmetryDb;
    return out;
}

double from_db(double db) { return std::exp2(db * double(kDbToLog2)); }
double to_db(double linear) { return 10.0 * std::log10(linear > kPowerFloor ? linear : kPowerFloor); }

// Running sums over a decision's spectrum, in bin order (the reference's).
struct Sums {
    std::vector<double> excess, floor;  // bins + 1 each; excess is signed
    // Per bin, the excess a band's shape is read from: a bin under its floor
    // is noise there, not a dip in the signal, so it counts as 0 (the sums
    // keep it signed: they answer how much power a band holds).
    std::vector<double> clamped;
    void build(const double* avg, const double* flo, uint32_t bins) {
        excess.assign(size_t(bins) + 1, 0.0);
        floor.assign(size_t(bins) + 1, 0.0);
        clamped.assign(bins, 0.0);
        for (uint32_t i = 0; i < bins; ++i) {
            excess[i + 1] = excess[i] + (avg[i] - flo[i]);
            floor[i + 1] = floor[i] + flo[i];
            clamped[i] = std::max(0.0, avg[i] - flo[i]);
        }
    }
    double excess_over(size_t first, size_t last) const { return excess[last + 1] - excess[first]; }
    double floor_over(size_t first, size_t last) const { return floor[last + 1] - floor[first]; }
};

// characterise::margin_confidence (transform.cpp:272-279): 0 under the
// threshold, then rising from one half toward one with a 6 dB scale.
double margin_confidence(double margin_db, double threshold_db) {
    if (margin_db < threshold_db) return 0.0;
    return 1.0 - 0.5 * std::exp(-(margin_db - threshold_db) / 6.0);
}

// round(n / d) with halves away from zero, in integers (detector.cpp:82-88).
int64_t divide_nearest(int64_t n, int64_t d) {
    if (d == 0) return 0;
    const int64_t half = d / 2;
    return n >= 0 ? (n + half) / d : (n - half) / d;
}

int32_t check_config(const rvx_event_config& c, const char* who) {
    if (!(c.bin_width_hz > 0.0) || !std::isfinite(c.bin_zero_hz))
        return rvx::fail(RVX_EINVAL, "%s: bin width %g Hz", who, c.bin_width_hz);
    if (!(c.occupied_power_fraction > 0.0 && c.occupied_power_fraction <= 1.0))
        return rvx::fail(RVX_EINVAL, "%s: occupied power fraction %g", who,
                         c.occupied_power_fraction);
    if (c.birth_hits == 0 || c.max_tracks == 0)
        return rvx::fail(RVX_EINVAL, "%s: birth_hits and max_tracks must be at least 1", who);
    if (!(c.average_seconds > 0.0) || !(c.confidence_half_life_seconds > 0.0) ||
        c.source_rate <= 0)
        return rvx::fail(RVX_EINVAL, "%s: time constants and the source rate must be positive",
                         who);
    if (!(c.channel_hysteresis >= 0.0) || !std::isfinite(c.channel_hysteresis))
        return rvx::fail(RVX_EINVAL, "%s: channel hysteresis %g", who, c.channel_hysteresis);
    return RVX_OK;
}

// measure_band (shape.h:275-372) over the clamped excess of [first, last],
// looking for skirts no further than [low, high) and twice the band's width.
void measure_shape(const std::vector<double>& x, size_t first, size_t last, size_t low,
                   size_t high, rvx_candidate& k) {
    if (x.empty() || last < first || last >= x.size() || low > first || high < last + 1 ||
        high > x.size())
        return;
    const size_t width = last - first + 1;
    double total = 0.0, peak = 0.0;
    for (size_t i = first; i <= last; ++i) {
        total += x[i];
        peak = std::max(peak, x[i]);
    }
    if (!(total > 0.0)) return;
    const double mean = total / double(width);
    k.shape_peak_to_mean = peak / mean;
    if (width <= 3) {
        k.shape_concentration = 1.0;
    } else {
        double run = x[first] + x[first + 1] + x[first + 2], best = run;
        for (size_t i = first + 3; i <= last; ++i) {
            run += x[i] - x[i - 3];
            best = std::max(best, run);
        }
        k.shape_concentration = std::min(1.0, best / total);
    }
    const size_t half = width / 2;
    double lower = 0.0, upper = 0.0;
    for (size_t i = 0; i < half; ++i) {
        lower += x[first + i];
        upper += x[last - i];
    }
    k.shape_lower_fraction = lower + upper > 0.0 ? lower / (lower + upper) : 0.5;
    const double end_level = mean * 0.1;               // kSkirtEndFraction
    const size_t reach = size_t(double(width) * 2.0);  // kSkirtReachWidths
    const size_t low_stop = std::min(first - low, reach);
    const size_t high_stop = std::min(high - (last + 1), reach);
    size_t below = 0, above = 0;
    while (below < low_stop && x[first - below - 1] > end_level) ++below;
    while (above < high_stop && x[last + above + 1] > end_level) ++above;
    k.shape_skirt_bins_available = uint32_t(low_stop + high_stop);
    k.shape_skirt_fraction = double(below + above) / double(width);
    k.shape_measured = 1;
}

// emit_candidate (:1277-1334) for bins [first, last]: false when the band
// carries no excess or does not clear the threshold. [own_low, own_high) is
// the region between the neighbouring decision bands, where its skirts may be.
bool emit(const rvx_event_config& c, const Sums& s, const double* avg, uint32_t bins,
          size_t first, size_t last, size_t own_low, size_t own_high,
          std::vector<rvx_candidate>& out) {
    if (last < first || last >= bins) return false;
    const double count = double(last - first + 1);
    const double excess = s.excess_over(first, last);
    const double noise = s.floor_over(first, last) / count;
    if (excess <= 0.0 || noise <= 0.0) return false;
    const double snr = excess * c.bin_width_hz / (noise * kReferenceBandwidthHz);
    if (!(snr > from_db(c.detection_threshold_db))) return false;
    double peak = 0.0;
    for (size_t i = first; i <= last; ++i) peak = std::max(peak, avg[i]);
    const double low_edge = c.bin_zero_hz + (double(first) - 0.5) * c.bin_width_hz;
    const double high_edge = c.bin_zero_hz + (double(last) + 0.5) * c.bin_width_hz;
    rvx_candidate k{};
    k.first_bin = uint32_t(first);
    k.last_bin = uint32_t(last);
    k.snr_2500_db = 10.0 * std::log10(snr);
    k.margin_confidence = margin_confidence(k.snr_2500_db, c.detection_threshold_db);
    k.noise_floor_dbfs = to_db(noise);
    k.peak_dbfs = to_db(peak);
    k.center_hz = c.source_center_hz + int64_t(std::llround(0.5 * (low_edge + high_edge)));
    k.bandwidth_hz = int64_t(std::llround(high_edge - low_edge));
    measure_shape(s.clamped, first, last, std::min(own_low, first), std::max(own_high, last + 1), k);
    out.push_back(k);
    return true;
}

// One accepted band to candidates (:1168-1275): trim to the span holding the
// occupied fraction of its excess (the rest split between the tails), then cut
// it at interior runs of at least split_gap_bins bins that sit within
// split_gap_db of the floor.
void band_candidates(const rvx_event_config& c, const Sums& s, const double* avg,
                     const double* flo, uint32_t bins, size_t begin, size_t end, size_t own_low,
                     size_t own_high, std::vector<uint8_t>& gap, std::vector<rvx_candidate>& out) {
    double total = 0.0;
    for (size_t i = begin; i < end; ++i) total += std::max(0.0, avg[i] - flo[i]);
    if (total <= 0.0) return;
    const double tail = total * (1.0 - c.occupied_power_fraction) * 0.5;
    size_t low = begin, high = end - 1;
    double acc = 0.0;
    for (size_t i = begin; i < end; ++i) {
        const double e = std::max(0.0, avg[i] - flo[i]);
        if (acc + e > tail) {
            low = i;
            break;
        }
        acc += e;
    }
    acc = 0.0;
    for (size_t i = end; i-- > begin;) {
        const double e = std::max(0.0, avg[i] - flo[i]);
        if (acc + e > tail) {
            high = i;
            break;
        }
        acc += e;
    }
    high = std::max(high, low);

    const double ratio = from_db(c.split_gap_db);
    const size_t n = high - low + 1;
    gap.assign(n, 0);
    for (size_t i = 0; i < n;) {
        if (avg[low + i] > flo[low + i] * ratio) {
            ++i;
            continue;
        }
        size_t j = i;
        while (j < n && avg[low + j] <= flo[low + j] * ratio) ++j;
        if (j - i >= c.split_gap_bins) std::fill(gap.begin() + i, gap.begin() + j, uint8_t(1));
        i = j;
    }
    bool any = false;
    for (size_t i = 0; i < n;) {
        if (gap[i]) {
            ++i;
            continue;
        }
        size_t j = i;
        while (j < n && !gap[j]) ++j;
        any = emit(c, s, avg, bins, low + i, low + j - 1, own_low, own_high, out) || any;
        i = j;
    }
    if (!any) (void)emit(c, s, avg, bins, low, high, own_low, own_high, out);
}

struct Residual {
    uint32_t first_bin = 0, last_bin = 0;      // the candidate's band this decision
    uint32_t run_first = 0, run_last = 0;      // the band the run is measured over
    double level_db = 0.0, run_level_db = 0.0, run_seconds = 0.0;
    uint32_t run = 0;
    double part_db[kParts] = {};
};

struct Pair {
    uint32_t track, candidate;
    double score;
    uint64_t first_seen, id;
};

}  // namespace

struct rvx_tracker {
    rvx_event_config cfg{};
    uint32_t bins = 0;
    uint64_t next_id = 1;
    Sums sums;
    std::vector<Residual> residuals;
    std::vector<rvx_track> tracks;
    rvx_tracker_stats stats{};
    // Narrow carriers that passed the spur rule, by baseband offset and the
    // tuning they were seen at (kept across a retune by the caller).
    std::vector<rvx_spur_sighting> book;
    // Spurs named for cancelling, by absolute centre, and when last named.
    struct Named {
        int64_t center_hz;
        uint64_t last;
    };
    std::vector<Named> cancelling;
    // Spurs the latest decision took back, absolute centres.
    std::vector<int64_t> withdrawn;
    // Emitters (refine::EmitterMemory): remembered at a track's first probe,
    // joined by tracks born near them, their key-ups counted per decision.
    struct Emitter {
        uint64_t id;
        int64_t center_hz, bandwidth_hz;
        uint64_t last_live;
        uint32_t keyups;
    };
    std::vector<Emitter> emitters;  // in the order remembered
    uint64_t next_emitter = 1;
    uint64_t last_decision = 0;

    // EmitterBook::match (refine.cpp:329-342): the nearest emitter within half
    // the narrower width, never under 1 kHz (ties: the earliest remembered).
    Emitter* match_emitter(int64_t center, int64_t bandwidth) {
        Emitter* best = nullptr;
        int64_t best_distance = INT64_MAX;
        for (Emitter& e : emitters) {
            const int64_t tolerance =
                std::max<int64_t>(kEmitterMatchMinHz, std::min(bandwidth, e.bandwidth_hz) / 2);
            const int64_t distance = std::abs(center - e.center_hz);
            if (distance <= tolerance && distance < best_distance) {
                best = &e;
                best_distance = distance;
            }
        }
        return best;
    }

    Emitter* find_emitter(uint64_t id) {
        if (id == 0) return nullptr;
        for (Emitter& e : emitters)
            if (e.id == id) return &e;
        return nullptr;
    }

    // EmitterBook::remember (:349-367): full, the longest quiet goes (ties:
    // the earliest remembered).
    Emitter& remember_emitter(int64_t center, int64_t bandwidth, uint64_t now) {
        if (emitters.size() >= kEmitterCapacity)
            emitters.erase(std::min_element(emitters.begin(), emitters.end(),
                                            [](const Emitter& a, const Emitter& b) {
                                                return a.last_live < b.last_live;
                                            }));
        emitters.push_back(Emitter{next_emitter++, center, bandwidth, now, 0});
        return emitters.back();
    }

    // detector.cpp:1910-1930: a live track keeps its emitter remembered, and
    // coming live after the emitter was quiet for over a second is a key-up
    // (unless the carrier there is being cancelled, which silences it); then
    // emitters quiet for 1800 s are forgotten.
    void count_keyups(uint64_t now) {
        const double rate = double(std::max<int64_t>(1, cfg.source_rate));
        const auto gap = uint64_t(kKeyupGapSeconds * rate);
        for (const rvx_track& tr : tracks) {
            if (tr.state != RVX_TRACK_LIVE || tr.emitter == 0) continue;
            if (Emitter* e = find_emitter(tr.emitter)) {
                if (now > e->last_live && now - e->last_live > gap && !is_cancelling(tr.center_hz))
                    ++e->keyups;
                e->last_live = now;
            }
        }
        const auto forget = uint64_t(kEmitterMemorySeconds * rate);
        emitters.erase(std::remove_if(emitters.begin(), emitters.end(),
                                      [&](const Emitter& e) {
                                          return now > e.last_live && now - e.last_live > forget;
                                      }),
                       emitters.end());
    }

    double level_db(uint32_t first, uint32_t last) const {
        const size_t lo = std::min<size_t>(first, bins - 1), hi = std::min<size_t>(last, bins - 1);
        return to_db(sums.excess_over(lo, hi) / double(hi - lo + 1));
    }
    double part_db(uint32_t first, uint32_t last, uint32_t p) const {
        const uint64_t span = uint64_t(last - first) + 1;
        const uint32_t lo = uint32_t(span * p / kParts), hi = uint32_t(span * (p + 1) / kParts);
        return level_db(first + lo, first + (hi > lo ? hi - 1 : lo));
    }

    // reject_residual (:1336-1583): a band whose mean excess has fallen, over
    // every one of its parts, at the rate an emptying average falls at
    // (4.34 dB per average time constant, within the tolerance) for
    // residual_decisions decisions in a row is the trace of a signal that has
    // stopped, and is withheld.
    void reject_residual(std::vector<rvx_candidate>& cands, double elapsed) {
        std::vector<Residual> next;
        next.reserve(cands.size());
        size_t kept = 0;
        for (const rvx_candidate& k : cands) {
            const Residual* prev = nullptr;
            uint32_t best = 0;
            for (const Residual& r : residuals) {
                if (r.last_bin < k.first_bin || r.first_bin > k.last_bin) continue;
                const uint32_t ov =
                    std::min(r.last_bin, k.last_bin) - std::max(r.first_bin, k.first_bin) + 1u;
                if (ov > best) {
                    best = ov;
                    prev = &r;
                }
            }
            Residual now;
            now.first_bin = now.run_first = k.first_bin;
            now.last_bin = now.run_last = k.last_bin;
            bool falling = false;
            if (prev) {
                const double carried = level_db(prev->run_first, prev->run_last);
                if (carried < prev->level_db) {
                    const double seconds = prev->run_seconds + elapsed;
                    const double expected = kDbPerNeper * seconds / cfg.average_seconds;
                    const double fall = prev->run_level_db - carried;
                    if (uint64_t(prev->run_last - prev->run_first) + 1 < kParts)
                        ++stats.residual_parts_degenerate;
                    bool uniform = true;
                    for (uint32_t p = 0; uniform && p < kParts; ++p)
                        uniform = prev->part_db[p] - part_db(prev->run_first, prev->run_last, p) >=
                                  kPartFloor * fall;
                    const double tol = cfg.residual_rate_tolerance;
                    if (uniform && fall >= (1.0 - tol) * expected && fall <= (1.0 + tol) * expected) {
                        falling = true;
                        now.run_first = prev->run_first;
                        now.run_last = prev->run_last;
                        now.level_db = carried;
                        now.run_level_db = prev->run_level_db;
                        now.run_seconds = seconds;
                        now.run = prev->run + 1;
                        std::copy(prev->part_db, prev->part_db + kParts, now.part_db);
                    }
                }
            }
            if (!falling) {
                now.level_db = now.run_level_db = level_db(now.run_first, now.run_last);
                for (uint32_t p = 0; p < kParts; ++p)
                    now.part_db[p] = part_db(now.run_first, now.run_last, p);
            }
            next.push_back(now);
            if (cfg.residual_decisions > 0 && now.run >= cfg.residual_decisions) {
                ++stats.candidates_residual;
                continue;
            }
            cands[kept++] = k;
        }
        cands.resize(kept);
        residuals.swap(next);
    }

    static void band_of(int64_t centre, int64_t width, double& lo, double& hi) {
        const double half = 0.5 * double(std::max<int64_t>(1, width));
        lo = double(centre) - half;
        hi = double(centre) + half;
    }

    // update_tracks (:1585-1900): gate track-candidate pairs on overlap over
    // the narrower band, rank them on overlap over the wider, assign one to one
    // best first; a candidate several born tracks were gated to goes to the
    // eldest and the others become its merged children; assigned tracks are
    // smoothed towards their candidate and their confidence rises; unassigned
    // ones are held with decaying confidence until it falls below the drop
    // level or the hold runs out; unassigned candidates start pending tracks,
    // born after birth_hits consecutive hits.
    void update(const std::vector<rvx_candidate>& cands, uint64_t now, double elapsed) {
        std::sort(tracks.begin(), tracks.end(), [](const rvx_track& a, const rvx_track& b) {
            return a.center_hz != b.center_hz ? a.center_hz < b.center_hz : a.id < b.id;
        });
        const size_t nt = tracks.size(), nc = cands.size();
        std::vector<Pair> pairs;
        for (size_t t = 0; t < nt; ++t) {
            double tlo, thi;
            band_of(tracks[t].center_hz, tracks[t].bandwidth_hz, tlo, thi);
            for (size_t c = 0; c < nc; ++c) {
                double lo, hi;
                band_of(cands[c].center_hz, cands[c].bandwidth_hz, lo, hi);
                if (hi < tlo) continue;
                if (lo > thi) break;
                const double ov = std::min(thi, hi) - std::max(tlo, lo);
                if (ov <= 0.0) continue;
                if (ov / std::min(thi - tlo, hi - lo) < cfg.association_overlap) continue;
                pairs.push_back({uint32_t(t), uint32_t(c), ov / std::max(thi - tlo, hi - lo),
                                 tracks[t].first_seen, tracks[t].id});
            }
        }
        std::sort(pairs.begin(), pairs.end(), [](const Pair& a, const Pair& b) {
            if (a.score != b.score) return a.score > b.score;
            if (a.first_seen != b.first_seen) return a.first_seen < b.first_seen;
            if (a.id != b.id) return a.id < b.id;
            return a.candidate < b.candidate;
        });
        std::vector<uint32_t> cand_of(nt, kUnassigned), track_of(nc, kUnassigned);
        for (const Pair& p : pairs) {
            if (cand_of[p.track] != kUnassigned || track_of[p.candidate] != kUnassigned) continue;
            cand_of[p.track] = p.candidate;
            track_of[p.candidate] = p.track;
        }
        merge(pairs, cand_of, track_of, now);

        const double decay = std::exp2(-elapsed / cfg.confidence_half_life_seconds);
        const double hold = cfg.bootstrap_hold_seconds * double(cfg.source_rate);
        std::vector<rvx_track> kept;
        kept.reserve(nt + nc);
        for (size_t t = 0; t < nt; ++t) {
            rvx_track tr = tracks[t];
            if (cand_of[t] != kUnassigned) {
                follow(tr, cands[cand_of[t]], now);
                kept.push_back(tr);
                continue;
            }
            if (tr.state == RVX_TRACK_PENDING) continue;  // birth needs consecutive hits
            if (tr.state != RVX_TRACK_MERGED) {
                if (tr.state == RVX_TRACK_LIVE) ++tr.held_spells;
                ++tr.misses;
                tr.state = RVX_TRACK_HELD;
                tr.confidence *= decay;
            }
            tr.last_seen = now;
            if (tr.confidence < cfg.drop_confidence || double(now - tr.last_detected) > hold) {
                ++stats.tracks_dropped;
                continue;
            }
            kept.push_back(tr);
        }
        for (size_t c = 0; c < nc; ++c) {
            if (track_of[c] != kUnassigned) continue;
            if (kept.size() >= cfg.max_tracks) {
                ++stats.births_refused;
                continue;
            }
            const rvx_candidate& k = cands[c];
            rvx_track tr{};
            tr.id = next_id++;
            tr.state = RVX_TRACK_PENDING;
            tr.center_hz = k.center_hz;
            tr.bandwidth_hz = std::max<int64_t>(1, k.bandwidth_hz);
            tr.snr_2500_db = k.snr_2500_db;
            tr.margin_confidence = k.margin_confidence;
            take_shape(tr, k);
            tr.first_bin = k.first_bin;
            tr.last_bin = k.last_bin;
            tr.level_first_bin = k.first_bin;
            tr.level_last_bin = k.last_bin;
            tr.level_bins_known = 1;
            tr.confidence = cfg.confidence_rise;
            tr.first_seen = tr.last_seen = tr.last_detected = now;
            tr.hits = 1;
            // born where an emitter was: it joins it (key-ups are counted
            // once it is live, so a pending track that dies is not one)
            if (const Emitter* e = match_emitter(tr.center_hz, tr.bandwidth_hz)) tr.emitter = e->id;
            if (tr.hits >= cfg.birth_hits) {
                tr.state = RVX_TRACK_LIVE;
                ++stats.tracks_born;
            }
            update_channel(tr);
            kept.push_back(tr);
        }
        tracks.swap(kept);
    }

    // A candidate several born tracks were gated to belongs to the eldest of
    // them (first seen, then id); the rest become its children (:1688-1756).
    void merge(const std::vector<Pair>& pairs, std::vector<uint32_t>& cand_of,
               std::vector<uint32_t>& track_of, uint64_t now) {
        std::vector<uint32_t> gated;
        for (size_t c = 0; c < track_of.size(); ++c) {
            const uint32_t owner = track_of[c];
            if (owner == kUnassigned) continue;
            gated.clear();
            if (tracks[owner].state != RVX_TRACK_PENDING) gated.push_back(owner);
            for (const Pair& p : pairs) {
                if (p.candidate != c || p.track == owner || cand_of[p.track] != kUnassigned ||
                    tracks[p.track].state == RVX_TRACK_PENDING)
                    continue;
                gated.push_back(p.track);
            }
            if (gated.empty() || (gated.size() == 1 && gated[0] == owner)) continue;
            uint32_t eldest = gated[0];
            for (uint32_t t : gated) {
                const rvx_track &a = tracks[t], &b = tracks[eldest];
                if (a.first_seen < b.first_seen || (a.first_seen == b.first_seen && a.id < b.id))
                    eldest = t;
            }
            if (eldest != owner) {
                cand_of[owner] = kUnassigned;
                cand_of[eldest] = uint32_t(c);
                track_of[c] = eldest;
            }
            const uint64_t parent = tracks[eldest].id;
            for (uint32_t t : gated) {
                if (t == eldest) continue;
                rvx_track& child = tracks[t];
                if (child.state != RVX_TRACK_MERGED || child.merged_into != parent) ++stats.merges;
                child.state = RVX_TRACK_MERGED;
                child.merged_into = parent;
                child.last_seen = now;
            }
        }
    }

    void follow(rvx_track& tr, const rvx_candidate& k, uint64_t now) {
        if (tr.state == RVX_TRACK_MERGED) ++stats.splits;
        tr.merged_into = 0;
        const double a = tr.state == RVX_TRACK_PENDING ? 1.0 : cfg.centre_smoothing;
        const double centre = double(tr.center_hz) + a * double(k.center_hz - tr.center_hz);
        const double width = double(tr.bandwidth_hz) + a * double(k.bandwidth_hz - tr.bandwidth_hz);
        tr.center_hz = int64_t(std::llround(centre));
        tr.bandwidth_hz = std::max<int64_t>(1, int64_t(std::llround(width)));
        tr.snr_2500_db += a * (k.snr_2500_db - tr.snr_2500_db);
        tr.margin_confidence = margin_confidence(tr.snr_2500_db, cfg.detection_threshold_db);
        take_shape(tr, k);  // taken, not blended
        tr.first_bin = k.first_bin;
        tr.last_bin = k.last_bin;
        // a level over other bins is another quantity: it starts again
        if (tr.level_bins_known &&
            (tr.level_first_bin != k.first_bin || tr.level_last_bin != k.last_bin)) {
            tr.level_frames = 0;
            tr.level_mean_db = tr.level_m2 = 0.0;
        }
        tr.level_first_bin = k.first_bin;
        tr.level_last_bin = k.last_bin;
        tr.level_bins_known = 1;
        ++tr.hits;
        tr.misses = 0;
        tr.last_detected = tr.last_seen = now;
        tr.confidence += (1.0 - tr.confidence) * cfg.confidence_rise;
        if (tr.state == RVX_TRACK_PENDING) {
            if (tr.hits >= cfg.birth_hits) {
                tr.state = RVX_TRACK_LIVE;
                ++stats.tracks_born;
            }
        } else {
            tr.state = RVX_TRACK_LIVE;
        }
        update_channel(tr);
    }

    static void take_shape(rvx_track& tr, const rvx_candidate& k) {
        tr.shape_peak_to_mean = k.shape_peak_to_mean;
        tr.shape_concentration = k.shape_concentration;
        tr.shape_lower_fraction = k.shape_lower_fraction;
        tr.shape_skirt_fraction = k.shape_skirt_fraction;
        tr.shape_skirt_bins_available = k.shape_skirt_bins_available;
        tr.shape_measured = k.shape_measured;
    }

    // update_channel (detector.cpp:1978-2007): the nearest grid channel, moved
    // only past 0.5 + channel_hysteresis channels from the one held (place()
    // rounds to the nearest, so a track parked on a boundary would flip with
    // every hertz its measurement moves, each flip a new channel and table).
    void update_channel(rvx_track& tr) const {
        if (cfg.grid_channels == 0) {
            tr.channel_valid = 0;
            tr.channel = 0;
            tr.channel_index = 0;
            return;
        }
        const int64_t channels = int64_t(cfg.grid_channels);
        const int64_t baseband = tr.center_hz - cfg.source_center_hz;
        const double exact = double(baseband) * double(channels) / double(cfg.source_rate);
        const double distance = std::abs(exact - double(tr.channel_index));
        if (!tr.channel_valid || distance > 0.5 + cfg.channel_hysteresis) {
            tr.channel_index = divide_nearest(baseband * channels, cfg.source_rate);
            tr.channel_valid = 1;
        }
        int64_t wrapped = tr.channel_index % channels;
        if (wrapped < 0) wrapped += channels;
        tr.channel = uint32_t(wrapped);
    }

    // SpurBook::note (detector.cpp:648-663): one sighting per tuning and
    // offset (refreshed in place); with the book full, the longest unseen goes.
    void note_sighting(int64_t baseband, int64_t centre, uint64_t now) {
        for (rvx_spur_sighting& e : book) {
            if (e.center_hz == centre && std::abs(double(e.baseband_hz - baseband)) <= kLoSpurToleranceHz) {
                e.baseband_hz = baseband;
                e.last_seen = now;
                return;
            }
        }
        if (book.size() >= RVX_SPUR_BOOK_CAPACITY)
            book.erase(std::min_element(book.begin(), book.end(),
                                        [](const rvx_spur_sighting& a, const rvx_spur_sighting& b) {
                                            return a.last_seen < b.last_seen;
                                        }));
        book.push_back(rvx_spur_sighting{baseband, centre, now});
    }

    // Whether a live track stands at the absolute frequency `at`.
    bool live_at(int64_t at) const {
        for (const rvx_track& o : tracks)
            if (o.state == RVX_TRACK_LIVE &&
                std::abs(double(o.center_hz - at)) <= kLoSpurToleranceHz + 0.5 * double(o.bandwidth_hz))
                return true;
        return false;
    }

    // SpurBook::stayed_at (:665-672): a sighting at this offset under a tuning
    // at least 2 kHz away, whose old absolute frequency nothing occupies now.
    bool stayed_at(int64_t baseband, int64_t centre) const {
        for (const rvx_spur_sighting& e : book)
            if (std::abs(double(e.center_hz - centre)) >= kLoSpurMinRetuneHz &&
                std::abs(double(e.baseband_hz - baseband)) <= kLoSpurToleranceHz &&
                !live_at(e.baseband_hz + e.center_hz))
                return true;
        return false;
    }

    double near_hz() const { return std::max(kLoSpurToleranceHz, cfg.bin_width_hz); }

    bool is_cancelling(int64_t center) const {
        for (const Named& n : cancelling)
            if (std::abs(double(n.center_hz - center)) <= near_hz()) return true;
        return false;
    }

    void withdraw(const rvx_track& tr) {
        withdrawn.push_back(tr.center_hz);
        const double near = near_hz();
        cancelling.erase(std::remove_if(cancelling.begin(), cancelling.end(),
                                        [&](const Named& n) {
                                            return std::abs(double(n.center_hz - tr.center_hz)) <= near;
                                        }),
                         cancelling.end());
    }

    // mark_spurs (detector.cpp:687-780), at every decision. A spur is a live,
    // narrow track off DC that a probe called a bare carrier (not double
    // sideband; no answer yet leaves it open), continuous for 5 s with no held
    // spell and at most one key-up on its emitter, its level steady. Or, once
    // steady, a lo_spur: it sits at the baseband offset of a carrier that
    // passed the rule under another tuning, whose old frequency stands empty.
    // A named spur stays named while it is being cancelled (what is left of it
    // proves nothing about keying or level); a probe naming another family,
    // or the track widening, withdraws it. Only by-rule carriers are noted.
    void mark_spurs(uint64_t now) {
        const double rate = double(std::max<int64_t>(1, cfg.source_rate));
        const double widest = std::max(kSpurMaxBandwidthHz, kSpurMaxBins * cfg.bin_width_hz);
        withdrawn.clear();
        std::vector<size_t> order(tracks.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            const rvx_track &x = tracks[a], &y = tracks[b];
            return x.center_hz != y.center_hz ? x.center_hz < y.center_hz : x.id < y.id;
        });
        for (const size_t i : order) {
            rvx_track& tr = tracks[i];
            const Emitter* emitter = find_emitter(tr.emitter);
            tr.keyups = emitter ? emitter->keyups : 0;
            const bool was = tr.spur != 0;
            const bool carrier = tr.classification == RVX_FAMILY_UNMODULATED && !tr.double_sideband;
            const bool unnamed = tr.classification == RVX_FAMILY_UNKNOWN;
            const int64_t baseband = tr.center_hz - cfg.source_center_hz;
            const bool narrow = double(tr.bandwidth_hz) <= widest &&
                                std::abs(double(baseband)) > kSpurDcGuardHz;
            if (!(carrier || unnamed) || !narrow) {
                tr.spur = tr.lo_spur = 0;
                if (was && tr.state != RVX_TRACK_PENDING) withdraw(tr);
                continue;
            }
            if (tr.state != RVX_TRACK_LIVE) {
                tr.spur = 0;  // held or merged: quiet because cancelled, not withdrawn
                continue;
            }
            const double age = double(now - tr.first_seen) / rate;
            const bool unkeyed = tr.held_spells == 0 && tr.keyups <= 1 && age >= kContinuousSeconds;
            const double sigma =
                tr.level_frames < 2 ? 0.0 : std::sqrt(tr.level_m2 / double(tr.level_frames - 1));
            const bool steady = tr.level_frames >= kSpurMinFrames && sigma <= kSpurLevelSigmaDb;
            const bool by_rule = carrier && unkeyed && steady;
            if (!tr.lo_spur && steady && stayed_at(baseband, cfg.source_center_hz)) tr.lo_spur = 1;
            if (by_rule) note_sighting(baseband, cfg.source_center_hz, now);
            const bool kept = was && is_cancelling(tr.center_hz);
            tr.spur = (kept || by_rule || (tr.lo_spur && steady)) ? 1u : 0u;
            if (was && !tr.spur) withdraw(tr);
            if (tr.spur) {
                auto named = std::find_if(cancelling.begin(), cancelling.end(), [&](const Named& n) {
                    return std::abs(double(n.center_hz - tr.center_hz)) <= near_hz();
                });
                if (named == cancelling.end())
                    cancelling.push_back(Named{tr.center_hz, now});
                else
                    named->last = now;
            }
        }
        const auto forget = uint64_t(kSpurMemorySeconds * rate);
        book.erase(std::remove_if(book.begin(), book.end(),
                                  [&](const rvx_spur_sighting& e) {
                                      return now > e.last_seen && now - e.last_seen > forget;
                                  }),
                   book.end());
        cancelling.erase(std::remove_if(cancelling.begin(), cancelling.end(),
                                        [&](const Named& n) { return now > n.last && now - n.last > forget; }),
                         cancelling.end());
    }

    // mark_hd_hybrids (detector.cpp:1943-1976): every live track at least
    // 50 kHz wide detected at this decision is measured; passing readings
    // smooth its two levels (the first taken whole) and confirm it after two
    // in a row, failing ones release it after three.
    void mark_hd_hybrids(const double* average, const double* floor, uint64_t now) {
        for (rvx_track& tr : tracks) {
            if (tr.state != RVX_TRACK_LIVE || tr.last_detected != now || tr.bandwidth_hz < kHdMinTrackHz)
                continue;
            const rvx_hd_reading r = measure_hd(average, floor, bins, cfg.bin_zero_hz, cfg.bin_width_hz,
                                                double(tr.center_hz - cfg.source_center_hz));
            if (!r.measured) continue;
            if (r.hybrid) {
                tr.hd_fails = 0;
                const double w = tr.hd_passes == 0 && !tr.hd_hybrid ? 1.0 : kHdLevelSmoothing;
                tr.hd_lower_dbc += w * (r.lower_dbc - tr.hd_lower_dbc);
                tr.hd_upper_dbc += w * (r.upper_dbc - tr.hd_upper_dbc);
                if (++tr.hd_passes >= kHdConfirm) tr.hd_hybrid = 1;
            } else {
                tr.hd_passes = 0;
                if (++tr.hd_fails >= kHdRelease) tr.hd_hybrid = 0;
            }
        }
    }

    rvx_track* find(uint64_t id) {
        for (rvx_track& tr : tracks)
            if (tr.id == id) return &tr;
        return nullptr;
    }

    // measure_levels (detector.cpp:629-646): every live track's band level in
    // this frame, summed in linear power (each bin's dB times log2(10) / 10,
    // in float, through exp2), into its running mean and squared deviations.
    void measure_levels(const float* power, uint32_t n) {
        constexpr float kDbToLog2 = 0.332192809488736234787f;
        for (rvx_track& tr : tracks) {
            if (tr.state != RVX_TRACK_LIVE || !tr.level_bins_known || tr.level_last_bin >= n)
                continue;
            double sum = 0.0;
            for (uint32_t i = tr.level_first_bin; i <= tr.level_last_bin; ++i)
                sum += double(std::exp2(power[i] * kDbToLog2));
            const double level = 10.0 * std::log10(std::max(sum, 1e-30));
            ++tr.level_frames;
            const double delta = level - tr.level_mean_db;
            tr.level_mean_db += delta / double(tr.level_frames);
            tr.level_m2 += delta * (level - tr.level_mean_db);
        }
    }
};

extern "C" {

void rvx_event_config_defaults(rvx_event_config* c) {
    if (!c) return;
    *c = rvx_event_config{};
    c->detection_threshold_db = 6.0;       // detector.h:600
    c->occupied_power_fraction = 0.99;     // :857
    c->split_gap_bins = 8;                 // :909
    c->split_gap_db = 1.0;                 // :910
    c->birth_hits = 3;                     // :918
    c->max_tracks = 512;                   // :938
    c->bootstrap_hold_seconds = 3.0;       // :955
    c->average_seconds = 1.0;              // :643
    c->residual_rate_tolerance = 0.25;     // :1183
    c->residual_decisions = 4;             // :1184
    c->association_overlap = 0.3;          // :1188
    c->confidence_rise = 0.35;             // :1194
    c->confidence_half_life_seconds = 1.0; // :1195
    c->drop_confidence = 0.05;             // :1196
    c->centre_smoothing = 0.4;             // :1200
    c->grid_channels = 0;                  // :591
    c->channel_hysteresis = 0.15;          // :1205
}

int32_t rvx_candidates(const rvx_event_config* config, const double* average, const double* floor,
                       uint32_t bins, const rvx_detect_band* bands, uint32_t n_bands,
                       rvx_candidate* out, uint32_t max_out, uint32_t* n_out) {
    static const char* who = "rvx_candidates";
    if (!config || !average || !floor || !n_out || (n_bands && !bands) || (max_out && !out))
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (int32_t rc = check_config(*config, who)) return rc;
    for (uint32_t i = 0; i < n_bands; ++i) {
        if (bands[i].width == 0 || uint64_t(bands[i].start) + bands[i].width > bins)
            return rvx::fail(RVX_EINVAL, "%s: band %u [%u, +%u) outside %u bins", who, i,
                             bands[i].start, bands[i].width, bins);
        if (i && bands[i].start < bands[i - 1].start + bands[i - 1].width)
            return rvx::fail(RVX_EINVAL, "%s: bands must be disjoint and ascending", who);
    }
    try {
        Sums s;
        s.build(average, floor, bins);
        std::vector<rvx_candidate> got;
        std::vector<uint8_t> gap;
        for (uint32_t i = 0; i < n_bands; ++i) {
            // the run between the neighbouring bands is this one's to look in
            const size_t own_low = i == 0 ? 0 : size_t(bands[i - 1].start) + bands[i - 1].width;
            const size_t own_high = i + 1 == n_bands ? size_t(bins) : size_t(bands[i + 1].start);
            band_candidates(*config, s, average, floor, bins, bands[i].start,
                            size_t(bands[i].start) + bands[i].width, own_low, own_high, gap, got);
        }
        *n_out = uint32_t(got.size());
        if (got.size() > max_out)
            return rvx::fail(RVX_EINVAL, "%s: %zu candidates, room for %u", who, got.size(),
                             max_out);
        std::copy(got.begin(), got.end(), out);
    } catch (const std::bad_alloc&) {
        return rvx::fail(RVX_ENOMEM, "%s: out of host memory", who);
    }
    return RVX_OK;
}

int32_t rvx_tracker_create(const rvx_event_config* config, uint32_t bins, rvx_tracker** out) {
    static const char* who = "rvx_tracker_create";
    if (!config || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    *out = nullptr;
    if (bins == 0) return rvx::fail(RVX_EINVAL, "%s: zero bins", who);
    if (int32_t rc = check_config(*config, who)) return rc;
    auto* t = new (std::nothrow) rvx_tracker;
    if (!t) return rvx::fail(RVX_ENOMEM, "%s: out of host memory", who);
    t->cfg = *config;
    t->bins = bins;
    *out = t;
    return RVX_OK;
}

void rvx_tracker_destroy(rvx_tracker* tracker) { delete tracker; }

int32_t rvx_tracker_measure_levels(rvx_tracker* t, const float* power_db, uint32_t bins) {
    if (!t || (bins && !power_db))
        return rvx::fail(RVX_EINVAL, "rvx_tracker_measure_levels: null argument");
    t->measure_levels(power_db, bins);
    return RVX_OK;
}

int32_t rvx_tracker_update(rvx_tracker* t, const double* average, const double* floor,
                           const rvx_candidate* candidates, uint32_t n, uint64_t now,
                           double elapsed_seconds, uint32_t* n_kept) {
    static const char* who = "rvx_tracker_update";
    if (!t || !average || !floor || (n && !candidates))
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (!(elapsed_seconds >= 0.0))
        return rvx::fail(RVX_EINVAL, "%s: elapsed %g s", who, elapsed_seconds);
    for (uint32_t i = 0; i < n; ++i)
        if (candidates[i].last_bin >= t->bins || candidates[i].first_bin > candidates[i].last_bin)
            return rvx::fail(RVX_EINVAL, "%s: candidate %u's bins are outside the spectrum", who,
                             i);
    try {
        t->sums.build(average, floor, t->bins);
        std::vector<rvx_candidate> cands(candidates, candidates + n);
        t->reject_residual(cands, elapsed_seconds);
        t->update(cands, now, elapsed_seconds);
        t->count_keyups(now);
        t->mark_spurs(now);
        t->mark_hd_hybrids(average, floor, now);
        t->last_decision = now;
        if (n_kept) *n_kept = uint32_t(cands.size());
        ++t->stats.decisions;
        t->stats.candidates += cands.size();
    } catch (const std::bad_alloc&) {
        return rvx::fail(RVX_ENOMEM, "%s: out of host memory", who);
    }
    return RVX_OK;
}

int32_t rvx_tracker_tracks(const rvx_tracker* t, rvx_track* out, uint32_t max_out, uint32_t* n) {
    if (!t || !n || (max_out && !out)) return rvx::fail(RVX_EINVAL, "rvx_tracker_tracks: null argument");
    *n = uint32_t(t->tracks.size());
    if (t->tracks.size() > max_out)
        return rvx::fail(RVX_EINVAL, "rvx_tracker_tracks: %u tracks, room for %u", *n, max_out);
    std::copy(t->tracks.begin(), t->tracks.end(), out);
    return RVX_OK;
}

int32_t rvx_tracker_record_probe(rvx_tracker* t, uint64_t id, uint32_t family,
                                 uint32_t double_sideband) {
    static const char* who = "rvx_tracker_record_probe";
    if (!t) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (family > RVX_FAMILY_OTHER) return rvx::fail(RVX_EINVAL, "%s: family %u", who, family);
    rvx_track* tr = t->find(id);
    if (!tr) return rvx::fail(RVX_EINVAL, "%s: no track %llu", who, (unsigned long long)id);
    try {
        // the track's first probe remembers its emitter (refine_track, :405-411)
        if (!t->find_emitter(tr->emitter))
            tr->emitter = t->remember_emitter(tr->center_hz, tr->bandwidth_hz, t->last_decision).id;
    } catch (const std::bad_alloc&) {
        return rvx::fail(RVX_ENOMEM, "%s: out of host memory", who);
    }
    tr->classification = family;
    tr->double_sideband = double_sideband ? 1u : 0u;
    return RVX_OK;
}

int32_t rvx_tracker_withdrawn_spurs(const rvx_tracker* t, int64_t* out, uint32_t max_out,
                                    uint32_t* n) {
    if (!t || !n || (max_out && !out))
        return rvx::fail(RVX_EINVAL, "rvx_tracker_withdrawn_spurs: null argument");
    *n = uint32_t(t->withdrawn.size());
    if (t->withdrawn.size() > max_out)
        return rvx::fail(RVX_EINVAL, "rvx_tracker_withdrawn_spurs: %u withdrawn, room for %u", *n,
                         max_out);
    std::copy(t->withdrawn.begin(), t->withdrawn.end(), out);
    return RVX_OK;
}

int32_t rvx_tracker_spur_book(const rvx_tracker* t, rvx_spur_sighting* out, uint32_t max_out,
                              uint32_t* n) {
    if (!t || !n || (max_out && !out))
        return rvx::fail(RVX_EINVAL, "rvx_tracker_spur_book: null argument");
    *n = uint32_t(t->book.size());
    if (t->book.size() > max_out)
        return rvx::fail(RVX_EINVAL, "rvx_tracker_spur_book: %u sightings, room for %u", *n, max_out);
    std::copy(t->book.begin(), t->book.end(), out);
    return RVX_OK;
}

int32_t rvx_tracker_adopt_spur_book(rvx_tracker* t, const rvx_spur_sighting* book, uint32_t n) {
    static const char* who = "rvx_tracker_adopt_spur_book";
    if (!t || (n && !book)) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (n > RVX_SPUR_BOOK_CAPACITY)
        return rvx::fail(RVX_EINVAL, "%s: %u sightings, the book holds %u", who, n,
                         RVX_SPUR_BOOK_CAPACITY);
    try {
        t->book.assign(book, book + n);
    } catch (const std::bad_alloc&) {
        return rvx::fail(RVX_ENOMEM, "%s: out of host memory", who);
    }
    return RVX_OK;
}

int32_t rvx_measure_hd_sidebands(const double* power, const double* floor, uint32_t bins,
                                 double bin_zero_hz, double bin_width_hz, double centre_hz,
                                 rvx_hd_reading* out) {
    if (!out || (bins && (!power || !floor)))
        return rvx::fail(RVX_EINVAL, "rvx_measure_hd_sidebands: null argument");
    *out = measure_hd(power, floor, bins, bin_zero_hz, bin_width_hz, centre_hz);
    return RVX_OK;
}

int32_t rvx_tracker_stats_get(const rvx_tracker* t, rvx_tracker_stats* out) {
    if (!t || !out) return rvx::fail(RVX_EINVAL, "rvx_tracker_stats_get: null argument");
    *out = t->stats;
    return RVX_OK;
}

}  // extern "C"
