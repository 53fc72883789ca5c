// CPU oracle for the plain detector's decision: a restatement of
// core/detect/detector.cpp's estimate_noise_floor (:794-876) and
// find_candidates (:879-1120, through the accepted, grown bands), with the
// derived constants of Detector::create (:300-385) and normal_quantile
// (:46-70).
//
// TEST INFRASTRUCTURE. Only the test suite and the benchmark's CPU-baseline
// leg load this library, and only to check or time the GPU path.
//
// Input is the averaged linear power per bin. The reference averages dB
// frames back to linear power on the host (accumulate, :621-627); this build's
// averaged spectrum is already linear (the chain's |X|^2 average), so the
// decision starts from it. Everything else keeps the reference's order:
// sequential prefix sums, ascending excision sums, the ladder narrowest rung
// first, the best-N peak set, the strongest-first greedy growth.
//
// Pinning: tests/test_oracle_detect.py checks the knot floors against a
// brute-force sort and mean, the deflection of a single box against its
// closed form, and that boxes of known width and level come back as bands
// covering them on flat noise.

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstdint>
#include <tuple>
#include <utility>
#include <vector>

namespace {

constexpr double kReferenceBandwidthHz = 2500.0;      // detector.h:128
constexpr float kDbToLog2 = 0.332192809488736234787F;  // detector.cpp:26
constexpr float kSpectrumPowerFloor = 1.0e-20F;        // spectrum_reference.h:104

double normal_quantile(double p) {
    constexpr double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                             -2.759285104469687e+02, 1.383577518672690e+02,
                             -3.066479806614716e+01, 2.506628277459239e+00};
    constexpr double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                             -1.556989798598866e+02, 6.680131188771972e+01,
                             -1.328068155288572e+01};
    constexpr double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                             -2.400758277161838e+00, -2.549732539343734e+00,
                             4.374664141464968e+00,  2.938163982698783e+00};
    constexpr double d[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                             2.445134137142996e+00, 3.754408661907416e+00};
    constexpr double lo = 0.02425;
    if (p <= 0.0 || p >= 1.0) return 0.0;
    if (p < lo) {
        const double q = std::sqrt(-2.0 * std::log(p));
        return (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
               ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    }
    if (p > 1.0 - lo) {
        const double q = std::sqrt(-2.0 * std::log(1.0 - p));
        return -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
               ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    }
    const double q = p - 0.5, r = q * q;
    return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
           (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
}

}  // namespace

extern "C" {

struct ODetectConfig {
    uint32_t bins;
    uint32_t noise_knots;
    uint32_t noise_window_knots;
    uint32_t max_sum_bins;
    uint32_t max_peaks;
    uint32_t split_gap_bins;
    double bin_width_hz;
    double detection_threshold_db;
    double noise_percentile;
    double noise_percentile_high;
    double noise_excision_sigma;
    double edge_floor_fraction;
    double edge_floor_sigma;
};

struct OBand {
    uint32_t start;
    uint32_t width;
    double deflection;
};

// Returns 0, or -1 for a refused config / too small an output.
int oracle_detect(const ODetectConfig* cfg, const double* average, double* floor_out,
                  double* sigma_out, OBand* bands, uint32_t max_bands, uint32_t* n_bands,
                  uint32_t* n_peaks) {
    const ODetectConfig& c = *cfg;
    const size_t bins = c.bins;
    if (bins == 0 || c.noise_knots == 0 || c.noise_window_knots == 0) return -1;
    // Detector::create's derived constants
    const double threshold_linear = std::exp2(c.detection_threshold_db * double(kDbToLog2));
    const double low_q = normal_quantile(c.noise_percentile);
    const double high_q = normal_quantile(c.noise_percentile_high);
    const double quantile_span = high_q - low_q;
    const size_t knots = std::min<size_t>(c.noise_knots, bins);
    const size_t spacing = std::max<size_t>(1, bins / knots);
    const size_t window_bins = std::min<size_t>(bins, spacing * c.noise_window_knots);
    std::vector<double> knot_position(knots), knot_floor(knots), knot_sigma(knots);
    for (size_t k = 0; k < knots; ++k)
        knot_position[k] = (double(2 * k + 1) * double(bins)) / double(2 * knots);
    uint32_t widest = c.max_sum_bins;
    if (widest == 0) widest = uint32_t(std::bit_floor(std::max<size_t>(1, bins / 8)));
    widest = std::min<uint32_t>(widest, uint32_t(bins));
    std::vector<uint32_t> widths;
    for (uint32_t w = 1; w <= widest; w *= 2) widths.push_back(w);
    const size_t budget = c.max_peaks ? c.max_peaks : bins;

    // estimate_noise_floor
    const size_t half = window_bins / 2;
    std::vector<double> window;
    for (size_t k = 0; k < knots; ++k) {
        const size_t centre = size_t(knot_position[k]);
        const size_t low = centre > half ? centre - half : 0;
        const size_t high = std::min(bins, centre + half + 1);
        window.assign(average + low, average + high);
        const size_t count = window.size();
        size_t low_rank = size_t(c.noise_percentile * double(count));
        size_t high_rank = size_t(c.noise_percentile_high * double(count));
        low_rank = std::min(low_rank, count - 1);
        high_rank = std::min(std::max(high_rank, low_rank), count - 1);
        std::nth_element(window.begin(), window.begin() + ptrdiff_t(low_rank), window.end());
        const double lp = window[low_rank];
        double hp = lp;
        if (high_rank > low_rank) {
            std::nth_element(window.begin() + ptrdiff_t(low_rank) + 1,
                             window.begin() + ptrdiff_t(high_rank), window.end());
            hp = window[high_rank];
        }
        const double deviation = std::max(0.0, (hp - lp) / quantile_span);
        const double level = hp - high_q * deviation;
        const double ceiling = level + c.noise_excision_sigma * deviation;
        double sum = 0.0;
        size_t kept = 0;
        for (size_t i = low; i < high; ++i) {
            if (average[i] <= ceiling) {
                sum += average[i];
                ++kept;
            }
        }
        const double estimate = kept > 0 ? sum / double(kept) : level;
        knot_floor[k] = std::max(estimate, double(kSpectrumPowerFloor));
        knot_sigma[k] = deviation;
    }
    std::vector<double> floor_(bins), sigma_(bins);
    if (knots == 1) {
        std::fill(floor_.begin(), floor_.end(), knot_floor[0]);
        std::fill(sigma_.begin(), sigma_.end(), knot_sigma[0]);
    } else {
        size_t k = 0;
        for (size_t i = 0; i < bins; ++i) {
            const double position = double(i);
            while (k + 2 < knots && knot_position[k + 1] < position) ++k;
            const double span = knot_position[k + 1] - knot_position[k];
            const double t =
                span > 0.0 ? std::clamp((position - knot_position[k]) / span, 0.0, 1.0) : 0.0;
            floor_[i] = knot_floor[k] + t * (knot_floor[k + 1] - knot_floor[k]);
            sigma_[i] = knot_sigma[k] + t * (knot_sigma[k + 1] - knot_sigma[k]);
        }
    }
    if (floor_out) std::copy(floor_.begin(), floor_.end(), floor_out);
    if (sigma_out) std::copy(sigma_.begin(), sigma_.end(), sigma_out);

    // find_candidates: prefix sums, the ladder, the best-N peaks
    std::vector<double> ec(bins + 1), fc(bins + 1), sc(bins + 1);
    for (size_t i = 0; i < bins; ++i) {
        const double excess = average[i] - floor_[i];
        ec[i + 1] = ec[i] + excess;
        fc[i + 1] = fc[i] + floor_[i];
        sc[i + 1] = sc[i] + sigma_[i];
    }
    struct Peak {
        uint32_t start, width;
        double deflection;
    };
    const auto stronger = [](const Peak& a, const Peak& b) {
        if (a.deflection != b.deflection) return a.deflection > b.deflection;
        if (a.start != b.start) return a.start < b.start;
        return a.width < b.width;
    };
    std::vector<Peak> peaks;
    std::vector<float> defl(bins);
    uint32_t found_total = 0;
    for (const uint32_t width : widths) {
        if (size_t(width) > bins) break;
        const size_t last = bins - width;
        const double root = std::sqrt(double(width));
        const double threshold = threshold_linear * kReferenceBandwidthHz / (root * c.bin_width_hz);
        for (size_t s = 0; s <= last; ++s) {
            const double excess = ec[s + width] - ec[s];
            const double summed_floor = fc[s + width] - fc[s];
            defl[s] = summed_floor > 0.0 ? float(excess * root / summed_floor) : 0.0F;
        }
        for (size_t s = 0; s <= last; ++s) {
            if (!(double(defl[s]) > threshold)) continue;
            const float here = defl[s];
            if (s > 0 && defl[s - 1] > here) continue;
            if (s < last && defl[s + 1] > here) continue;
            if (s > 0 && defl[s - 1] == here) continue;
            size_t end = s;
            while (end < last && defl[end + 1] == here) ++end;
            const Peak found{uint32_t(s + (end - s) / 2), width, double(here)};
            ++found_total;
            if (peaks.size() < budget) {
                peaks.push_back(found);
                if (peaks.size() == budget) std::make_heap(peaks.begin(), peaks.end(), stronger);
                continue;
            }
            if (stronger(found, peaks.front())) {
                std::pop_heap(peaks.begin(), peaks.end(), stronger);
                peaks.back() = found;
                std::push_heap(peaks.begin(), peaks.end(), stronger);
            }
        }
    }
    std::sort(peaks.begin(), peaks.end(), stronger);

    // strongest first, each winner grown to its shoulders before the next
    std::vector<Peak> accepted;
    for (const Peak& peak : peaks) {
        size_t begin = peak.start;
        size_t end = begin + peak.width;
        auto at = std::lower_bound(accepted.begin(), accepted.end(), begin,
                                   [](const Peak& held, size_t value) { return held.start < value; });
        bool clashes = at != accepted.end() && size_t(at->start) < end;
        if (!clashes && at != accepted.begin()) {
            const Peak& before = *(at - 1);
            clashes = size_t(before.start) + before.width > begin;
        }
        if (clashes) continue;
        const double seed_floor = fc[end] - fc[begin];
        const double seed_sigma = sc[end] - sc[begin];
        const double level =
            std::max(c.edge_floor_fraction * seed_floor, c.edge_floor_sigma * seed_sigma) /
            double(peak.width);
        const size_t reach = peak.width;
        const size_t low_limit = at == accepted.begin() ? 0 : size_t((at - 1)->start) + (at - 1)->width;
        const size_t high_limit = at == accepted.end() ? bins : at->start;
        const size_t low_stop = std::max(low_limit, begin > reach ? begin - reach : 0);
        while (begin > low_stop) {
            if (average[begin - 1] - floor_[begin - 1] > level) {
                --begin;
                continue;
            }
            const size_t limit = std::min(size_t(c.split_gap_bins), begin - low_stop);
            bool resumed = false;
            for (size_t back = 2; back <= limit; ++back) {
                if (average[begin - back] - floor_[begin - back] > level) {
                    begin -= back;
                    resumed = true;
                    break;
                }
            }
            if (!resumed) break;
        }
        const size_t high_stop = std::min(high_limit, std::min(bins, end + reach));
        while (end < high_stop) {
            if (average[end] - floor_[end] > level) {
                ++end;
                continue;
            }
            const size_t limit = std::min(size_t(c.split_gap_bins), high_stop - end);
            bool resumed = false;
            for (size_t ahead = 1; ahead < limit; ++ahead) {
                if (average[end + ahead] - floor_[end + ahead] > level) {
                    end += ahead + 1;
                    resumed = true;
                    break;
                }
            }
            if (!resumed) break;
        }
        accepted.insert(at, Peak{uint32_t(begin), uint32_t(end - begin), peak.deflection});
    }
    if (n_peaks) *n_peaks = found_total;
    *n_bands = uint32_t(accepted.size());
    if (accepted.size() > max_bands) return -1;
    for (size_t i = 0; i < accepted.size(); ++i)
        bands[i] = OBand{accepted[i].start, accepted[i].width, accepted[i].deflection};
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Detection events: core/detect/detector.cpp's find_candidates tail
// (:1168-1275) and emit_candidate (:1277-1334), reject_residual (:1336-1583)
// and update_tracks (:1585-1900), restated in the reference's statement order
// without the band shape, margin confidence, emitter memory, spur cancelling,
// channel association and level statistics.
//
// Pinning (tests/test_oracle_events.py): a box's candidate against its
// occupied-band and edge closed forms; a band cut at a floor-level gap;
// tracks of bursts switched on and off at known decisions against the birth,
// hold and drop rules worked by hand (first_seen, last_detected, ids); two
// tracks merging into one candidate and splitting again; an exponentially
// emptying average withheld after residual_decisions decisions while a steady
// band is not.
// ---------------------------------------------------------------------------

namespace {

constexpr double kDbPerNeperO = 4.342944819032518;
constexpr std::size_t kResidualPartsO = 8;
constexpr double kResidualPartFloorO = 0.5;

double db_to_linear_o(double db) { return std::exp2(db * static_cast<double>(kDbToLog2)); }
double linear_to_db_o(double linear) {
    const double floored = linear > static_cast<double>(kSpectrumPowerFloor)
                               ? linear
                               : static_cast<double>(kSpectrumPowerFloor);
    return 10.0 * std::log10(floored);
}

}  // namespace

extern "C" {

struct OEventConfig {
    double bin_width_hz, bin_zero_hz;
    int64_t source_center_hz, source_rate;
    double detection_threshold_db, occupied_power_fraction, split_gap_db;
    uint32_t split_gap_bins, residual_decisions;
    double average_seconds, residual_rate_tolerance, association_overlap, confidence_rise,
        confidence_half_life_seconds, drop_confidence, centre_smoothing, bootstrap_hold_seconds;
    uint32_t birth_hits, max_tracks;
};

struct OCandidate {
    uint32_t first_bin, last_bin;
    int64_t center, bandwidth;
    double snr_2500_db, noise_floor_dbfs, peak_dbfs;
};

struct OTrack {
    uint64_t id;
    uint32_t state, hits, misses, held_spells;
    int64_t center, bandwidth;
    double snr_2500_db, confidence;
    uint64_t first_seen, last_seen, last_detected, merged_into;
    uint32_t first_bin, last_bin;
};

}  // extern "C"

namespace {

enum : uint32_t { kPending = 0, kLive = 1, kHeld = 2, kMerged = 3 };

struct EventDecision {
    std::vector<double> excess_cumulative, floor_cumulative;
    void build(const double* average, const double* floor, std::size_t bins) {
        excess_cumulative.assign(bins + 1, 0.0);
        floor_cumulative.assign(bins + 1, 0.0);
        for (std::size_t i = 0; i < bins; ++i) {
            const double excess = average[i] - floor[i];
            excess_cumulative[i + 1] = excess_cumulative[i] + excess;
            floor_cumulative[i + 1] = floor_cumulative[i] + floor[i];
        }
    }
};

bool emit_candidate_o(const OEventConfig& c, const EventDecision& d, const double* average,
                      std::size_t bins, std::size_t first, std::size_t last,
                      std::vector<OCandidate>& out) {
    if (last < first || last >= bins) return false;
    const auto count = static_cast<double>(last - first + 1);
    const double excess = d.excess_cumulative[last + 1] - d.excess_cumulative[first];
    const double noise = (d.floor_cumulative[last + 1] - d.floor_cumulative[first]) / count;
    if (excess <= 0.0 || noise <= 0.0) return false;
    const double snr = excess * c.bin_width_hz / (noise * kReferenceBandwidthHz);
    if (!(snr > db_to_linear_o(c.detection_threshold_db))) return false;
    double peak = 0.0;
    for (std::size_t i = first; i <= last; ++i) peak = std::max(peak, average[i]);
    const double low_edge = c.bin_zero_hz + (static_cast<double>(first) - 0.5) * c.bin_width_hz;
    const double high_edge = c.bin_zero_hz + (static_cast<double>(last) + 0.5) * c.bin_width_hz;
    OCandidate candidate{};
    candidate.first_bin = static_cast<uint32_t>(first);
    candidate.last_bin = static_cast<uint32_t>(last);
    candidate.snr_2500_db = 10.0 * std::log10(snr);
    candidate.noise_floor_dbfs = linear_to_db_o(noise);
    candidate.peak_dbfs = linear_to_db_o(peak);
    candidate.center = c.source_center_hz +
                       static_cast<int64_t>(std::llround(0.5 * (low_edge + high_edge)));
    candidate.bandwidth = static_cast<int64_t>(std::llround(high_edge - low_edge));
    out.push_back(candidate);
    return true;
}

struct OracleTracker {
    OEventConfig config;
    std::size_t bins = 0;
    uint64_t next_id = 1;
    EventDecision decision;
    struct Decaying {
        uint32_t first_bin = 0, last_bin = 0, run_first_bin = 0, run_last_bin = 0;
        double level_db = 0.0, run_level_db = 0.0, run_seconds = 0.0;
        uint32_t run = 0;
        double run_part_db[kResidualPartsO] = {};
    };
    std::vector<Decaying> decaying;
    std::vector<OTrack> all;

    void reject_residual(std::vector<OCandidate>& candidates, double elapsed_seconds) {
        const auto level_of = [this](uint32_t first, uint32_t last) {
            const std::size_t low = std::min<std::size_t>(first, bins - 1);
            const std::size_t high = std::min<std::size_t>(last, bins - 1);
            const auto count = static_cast<double>(high - low + 1);
            const double excess =
                decision.excess_cumulative[high + 1] - decision.excess_cumulative[low];
            return linear_to_db_o(excess / count);
        };
        const auto part_of = [&level_of](uint32_t first, uint32_t last, std::size_t p) {
            const auto span = static_cast<uint64_t>(last - first + 1U);
            const auto low_offset = static_cast<uint32_t>(span * p / kResidualPartsO);
            const auto high_offset = static_cast<uint32_t>(span * (p + 1) / kResidualPartsO);
            const uint32_t last_offset = high_offset > low_offset ? high_offset - 1U : low_offset;
            return level_of(first + low_offset, first + last_offset);
        };
        std::vector<Decaying> next;
        std::size_t kept = 0;
        for (std::size_t c = 0; c < candidates.size(); ++c) {
            const OCandidate candidate = candidates[c];
            const Decaying* matched = nullptr;
            uint32_t best_overlap = 0;
            for (const Decaying& held : decaying) {
                if (held.last_bin < candidate.first_bin || held.first_bin > candidate.last_bin)
                    continue;
                const uint32_t overlap = std::min(held.last_bin, candidate.last_bin) -
                                         std::max(held.first_bin, candidate.first_bin) + 1U;
                if (overlap > best_overlap) {
                    best_overlap = overlap;
                    matched = &held;
                }
            }
            Decaying now;
            now.first_bin = candidate.first_bin;
            now.last_bin = candidate.last_bin;
            now.run_first_bin = candidate.first_bin;
            now.run_last_bin = candidate.last_bin;
            bool falling = false;
            if (matched != nullptr) {
                const double carried = level_of(matched->run_first_bin, matched->run_last_bin);
                if (carried < matched->level_db) {
                    const double seconds = matched->run_seconds + elapsed_seconds;
                    const double expected = kDbPerNeperO * seconds / config.average_seconds;
                    const double fall = matched->run_level_db - carried;
                    bool uniform = true;
                    for (std::size_t p = 0; uniform && p < kResidualPartsO; ++p) {
                        const double part_fall =
                            matched->run_part_db[p] -
                            part_of(matched->run_first_bin, matched->run_last_bin, p);
                        uniform = part_fall >= kResidualPartFloorO * fall;
                    }
                    if (uniform && fall >= (1.0 - config.residual_rate_tolerance) * expected &&
                        fall <= (1.0 + config.residual_rate_tolerance) * expected) {
                        falling = true;
                        now.run_first_bin = matched->run_first_bin;
                        now.run_last_bin = matched->run_last_bin;
                        now.level_db = carried;
                        now.run_level_db = matched->run_level_db;
                        now.run_seconds = seconds;
                        now.run = matched->run + 1;
                        for (std::size_t p = 0; p < kResidualPartsO; ++p)
                            now.run_part_db[p] = matched->run_part_db[p];
                    }
                }
            }
            if (!falling) {
                now.level_db = level_of(now.run_first_bin, now.run_last_bin);
                now.run_level_db = now.level_db;
                for (std::size_t p = 0; p < kResidualPartsO; ++p)
                    now.run_part_db[p] = part_of(now.run_first_bin, now.run_last_bin, p);
            }
            next.push_back(now);
            if (config.residual_decisions > 0 && now.run >= config.residual_decisions) continue;
            candidates[kept] = candidate;
            ++kept;
        }
        candidates.resize(kept);
        decaying.swap(next);
    }

    void update_tracks(const std::vector<OCandidate>& candidates, uint64_t now,
                       double elapsed_seconds) {
        std::sort(all.begin(), all.end(), [](const OTrack& a, const OTrack& b) {
            if (a.center != b.center) return a.center < b.center;
            return a.id < b.id;
        });
        const std::size_t track_count = all.size(), candidate_count = candidates.size();
        const auto band_of = [](int64_t center, int64_t bandwidth) {
            const double half = 0.5 * static_cast<double>(std::max<int64_t>(1, bandwidth));
            return std::pair<double, double>{static_cast<double>(center) - half,
                                             static_cast<double>(center) + half};
        };
        struct P {
            uint32_t track, candidate;
            double score;
            uint64_t first_seen, id;
        };
        std::vector<P> pairs;
        for (std::size_t t = 0; t < track_count; ++t) {
            const auto [track_low, track_high] = band_of(all[t].center, all[t].bandwidth);
            const double track_width = track_high - track_low;
            for (std::size_t c = 0; c < candidate_count; ++c) {
                const auto [low, high] = band_of(candidates[c].center, candidates[c].bandwidth);
                if (high < track_low) continue;
                if (low > track_high) break;
                const double width = high - low;
                const double overlap = std::min(track_high, high) - std::max(track_low, low);
                if (overlap <= 0.0) continue;
                if (overlap / std::min(track_width, width) < config.association_overlap) continue;
                pairs.push_back(P{static_cast<uint32_t>(t), static_cast<uint32_t>(c),
                                  overlap / std::max(track_width, width), all[t].first_seen,
                                  all[t].id});
            }
        }
        std::sort(pairs.begin(), pairs.end(), [](const P& a, const P& b) {
            if (a.score != b.score) return a.score > b.score;
            if (a.first_seen != b.first_seen) return a.first_seen < b.first_seen;
            if (a.id != b.id) return a.id < b.id;
            return a.candidate < b.candidate;
        });
        constexpr uint32_t kNone = 0xFFFFFFFFu;
        std::vector<uint32_t> candidate_of_track(track_count, kNone);
        std::vector<uint32_t> track_of_candidate(candidate_count, kNone);
        for (const P& pair : pairs) {
            if (candidate_of_track[pair.track] != kNone) continue;
            if (track_of_candidate[pair.candidate] != kNone) continue;
            candidate_of_track[pair.track] = pair.candidate;
            track_of_candidate[pair.candidate] = pair.track;
        }
        std::vector<uint32_t> gated;
        for (std::size_t c = 0; c < candidate_count; ++c) {
            const uint32_t owner = track_of_candidate[c];
            if (owner == kNone) continue;
            gated.clear();
            if (all[owner].state != kPending) gated.push_back(owner);
            for (const P& pair : pairs) {
                if (pair.candidate != c || pair.track == owner) continue;
                if (candidate_of_track[pair.track] != kNone) continue;
                if (all[pair.track].state == kPending) continue;
                gated.push_back(pair.track);
            }
            if (gated.empty() || (gated.size() == 1 && gated.front() == owner)) continue;
            uint32_t eldest = gated.front();
            for (const uint32_t t : gated) {
                const OTrack& a = all[t];
                const OTrack& b = all[eldest];
                if (std::tie(a.first_seen, a.id) < std::tie(b.first_seen, b.id)) eldest = t;
            }
            if (eldest != owner) {
                candidate_of_track[owner] = kNone;
                candidate_of_track[eldest] = static_cast<uint32_t>(c);
                track_of_candidate[c] = eldest;
            }
            const uint64_t parent = all[eldest].id;
            for (const uint32_t t : gated) {
                if (t == eldest) continue;
                OTrack& child = all[t];
                child.state = kMerged;
                child.merged_into = parent;
                child.last_seen = now;
            }
        }
        const double decay = std::exp2(-elapsed_seconds / config.confidence_half_life_seconds);
        const double hold_samples =
            config.bootstrap_hold_seconds * static_cast<double>(config.source_rate);
        std::vector<OTrack> survivors;
        for (std::size_t t = 0; t < track_count; ++t) {
            OTrack track = all[t];
            const uint32_t assigned = candidate_of_track[t];
            if (assigned != kNone) {
                const OCandidate& candidate = candidates[assigned];
                track.merged_into = 0;
                const double smoothing = track.state == kPending ? 1.0 : config.centre_smoothing;
                const double centre =
                    static_cast<double>(track.center) +
                    smoothing * static_cast<double>(candidate.center - track.center);
                const double width =
                    static_cast<double>(track.bandwidth) +
                    smoothing * static_cast<double>(candidate.bandwidth - track.bandwidth);
                track.center = static_cast<int64_t>(std::llround(centre));
                track.bandwidth = std::max<int64_t>(1, static_cast<int64_t>(std::llround(width)));
                track.snr_2500_db += smoothing * (candidate.snr_2500_db - track.snr_2500_db);
                track.first_bin = candidate.first_bin;
                track.last_bin = candidate.last_bin;
                ++track.hits;
                track.misses = 0;
                track.last_detected = now;
                track.last_seen = now;
                track.confidence += (1.0 - track.confidence) * config.confidence_rise;
                if (track.state == kPending) {
                    if (track.hits >= config.birth_hits) track.state = kLive;
                } else {
                    track.state = kLive;
                }
                survivors.push_back(track);
                continue;
            }
            if (track.state == kPending) continue;
            const bool merged = track.state == kMerged;
            if (!merged) {
                if (track.state == kLive) ++track.held_spells;
                ++track.misses;
                track.state = kHeld;
                track.confidence *= decay;
            }
            track.last_seen = now;
            const auto silent = static_cast<double>(now - track.last_detected);
            if (track.confidence < config.drop_confidence || silent > hold_samples) continue;
            survivors.push_back(track);
        }
        for (std::size_t c = 0; c < candidate_count; ++c) {
            if (track_of_candidate[c] != kNone) continue;
            if (survivors.size() >= config.max_tracks) continue;
            const OCandidate& candidate = candidates[c];
            OTrack track{};
            track.id = next_id++;
            track.state = kPending;
            track.center = candidate.center;
            track.bandwidth = std::max<int64_t>(1, candidate.bandwidth);
            track.snr_2500_db = candidate.snr_2500_db;
            track.first_bin = candidate.first_bin;
            track.last_bin = candidate.last_bin;
            track.confidence = config.confidence_rise;
            track.first_seen = now;
            track.last_seen = now;
            track.last_detected = now;
            track.hits = 1;
            if (track.hits >= config.birth_hits) track.state = kLive;
            survivors.push_back(track);
        }
        all.swap(survivors);
    }
};

}  // namespace

extern "C" {

int oracle_candidates(const OEventConfig* cfg, const double* average, const double* floor,
                      uint32_t bins, const OBand* bands, uint32_t n_bands, OCandidate* out,
                      uint32_t max_out, uint32_t* n_out) {
    if (!cfg || !average || !floor || !n_out) return -1;
    const OEventConfig& c = *cfg;
    EventDecision d;
    d.build(average, floor, bins);
    const double gap_ratio = db_to_linear_o(c.split_gap_db);
    std::vector<OCandidate> candidates;
    std::vector<uint8_t> separator;
    for (uint32_t index = 0; index < n_bands; ++index) {
        const std::size_t begin = bands[index].start;
        const std::size_t end = begin + bands[index].width;
        double total = 0.0;
        for (std::size_t i = begin; i < end; ++i) total += std::max(0.0, average[i] - floor[i]);
        if (total <= 0.0) continue;
        const double tail = total * (1.0 - c.occupied_power_fraction) * 0.5;
        std::size_t low = begin;
        double running = 0.0;
        for (std::size_t i = begin; i < end; ++i) {
            const double excess = std::max(0.0, average[i] - floor[i]);
            if (running + excess > tail) {
                low = i;
                break;
            }
            running += excess;
        }
        std::size_t high = end - 1;
        running = 0.0;
        for (std::size_t i = end; i-- > begin;) {
            const double excess = std::max(0.0, average[i] - floor[i]);
            if (running + excess > tail) {
                high = i;
                break;
            }
            running += excess;
        }
        if (high < low) high = low;
        const std::size_t count = high - low + 1;
        separator.assign(count, 0);
        std::size_t i = 0;
        while (i < count) {
            if (average[low + i] > floor[low + i] * gap_ratio) {
                ++i;
                continue;
            }
            std::size_t j = i;
            while (j < count && average[low + j] <= floor[low + j] * gap_ratio) ++j;
            if (j - i >= c.split_gap_bins)
                std::fill(separator.begin() + static_cast<std::ptrdiff_t>(i),
                          separator.begin() + static_cast<std::ptrdiff_t>(j), uint8_t{1});
            i = j;
        }
        bool emitted = false;
        i = 0;
        while (i < count) {
            if (separator[i] != 0) {
                ++i;
                continue;
            }
            std::size_t j = i;
            while (j < count && separator[j] == 0) ++j;
            emitted = emit_candidate_o(c, d, average, bins, low + i, low + j - 1, candidates) ||
                      emitted;
            i = j;
        }
        if (!emitted) static_cast<void>(emit_candidate_o(c, d, average, bins, low, high, candidates));
    }
    *n_out = static_cast<uint32_t>(candidates.size());
    if (candidates.size() > max_out) return -1;
    std::copy(candidates.begin(), candidates.end(), out);
    return 0;
}

void* oracle_tracker_create(const OEventConfig* cfg, uint32_t bins) {
    auto* t = new OracleTracker;
    t->config = *cfg;
    t->bins = bins;
    return t;
}

void oracle_tracker_destroy(void* t) { delete static_cast<OracleTracker*>(t); }

int oracle_tracker_update(void* handle, const double* average, const double* floor,
                          const OCandidate* candidates, uint32_t n, uint64_t now,
                          double elapsed_seconds, uint32_t* n_kept) {
    auto* t = static_cast<OracleTracker*>(handle);
    t->decision.build(average, floor, t->bins);
    std::vector<OCandidate> c(candidates, candidates + n);
    t->reject_residual(c, elapsed_seconds);
    t->update_tracks(c, now, elapsed_seconds);
    if (n_kept) *n_kept = static_cast<uint32_t>(c.size());
    return 0;
}

int oracle_tracker_tracks(void* handle, OTrack* out, uint32_t max_out, uint32_t* n) {
    auto* t = static_cast<OracleTracker*>(handle);
    *n = static_cast<uint32_t>(t->all.size());
    if (t->all.size() > max_out) return -1;
    std::copy(t->all.begin(), t->all.end(), out);
    return 0;
}

}  // extern "C"
