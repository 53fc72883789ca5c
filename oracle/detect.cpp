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
