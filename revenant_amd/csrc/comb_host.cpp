// The comb canceller's host half (core/dsp/comb_cancel.h): finding combs of
// evenly spaced interference lines in a time-averaged spectrum, tracking each
// comb's offset c and spacing d from the project pass's per-line projections,
// and planning the lines (phase, step, amplitude) the project and subtract
// kernels run with. Line k of a comb sits at c + k d.
//
// Detection (comb_detect.cpp). The spectrum in dB; a local floor at knots a
// quarter window apart, each the window's 20th percentile lifted by 0.84
// sigma (sigma from the 10th and 30th percentiles), so a height is measured
// from where the noise's median would be even where most bins are line; peaks
// standing max(4 dB, 4.5 sigma) over the floor and 3 dB over the lowest bin 2
// to 6 bins either side, centred by a parabola through three dB values; the
// spacing scored by the phase coherence of the peak centres,
// S(f) = |sum_j exp(2 pi i f x_j)|, for f from 7 / extent to 1 / 4 cycles per
// bin; each strong maximum of S fitted as a grid (nearest k, tolerance
// tightening, least squares through (k, x)); the largest d whose fit keeps 90%
// of the best fit's lines, refused under 40% occupancy or 8 lines. Several
// combs: a found comb is held as its even and odd halves at twice the
// spacing, each halved again while its own halves sit more than a quarter bin
// (or 25 Hz) from it; its peaks and sidebands leave the search and the rest
// is searched again. A comb already held (`active`) has its peaks removed
// first, and a grid still mostly on it is passed over.
//
// Tracking (comb_tracker.cpp). c and d are exact fixed-point polynomials in
// the sample index (phase, frequency and rate of change, cycles x 2^64,
// wrapping), so advancing them is integer arithmetic. A plan anchors each
// line at the block's middle sample. An observation is read against the
// record of its plan: per line, D = sum p_{s+1} conj(p_s) gives the frequency
// against the planned tone, P = sum p_s the phase at the middle, and the
// current model retrodicted to that middle gives the innovation. Acquisition
// runs a frequency-locked stage (freq += 0.5 e, rate += e / 16 per block) on
// the fitted frequency errors until they stay under 0.1 cycles per block for
// three blocks, then a third-order phase-locked stage (all poles at
// exp(-T / 0.05 s)) on the fitted phase errors, after a block that seeds each
// line's complex amplitude. Lines' amplitudes average the derotated block
// projection with weight 1/16. Lock: eight phase-stage blocks with the
// weighted RMS phase residual under 0.25 rad, the fitted innovation under
// 0.03 cycles, at least 8 lines standing out (6 dB over the noise in a
// block) and at most half benched (a carrier out of the fit by 0.8 rad for 4
// blocks leaves it until 8 blocks in step). Six bad blocks while locked:
// recovering (no subtraction, model kept); not locked within 8 s: lost;
// acquisition not locked within 200 blocks: lost.
//
// The bank (comb_bank.cpp) runs up to 8 trackers behind one plan and splits
// each observation back by the line counts it handed out.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <numeric>
#include <vector>

#include "common.h"

namespace {

using Cd = std::complex<double>;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kTwo32 = 4294967296.0;
constexpr double kTwo64 = 18446744073709551616.0;

// ---------------------------------------------------------------------------
// Detection
// ---------------------------------------------------------------------------

constexpr uint64_t kMinBins = 64;
constexpr double kMinSpacingBins = 4.0;
constexpr double kMinHeightDb = 4.0;
constexpr double kHeightSigmas = 4.5;
constexpr double kProminenceDb = 3.0;
constexpr int kValleyNear = 2, kValleyFar = 6;
constexpr size_t kMaxPeaks = 4 * RVX_COMB_MAX_LINES;
constexpr size_t kMaxCandidates = 64;
constexpr double kCandidateFraction = 0.5;
constexpr double kKeepFraction = 0.9;
constexpr double kMinOccupancy = 0.4;
constexpr double kSplitBins = 0.25, kSplitHz = 25.0;
constexpr double kTakenBins = 0.6;
constexpr double kEmptyParity = 0.15;

struct Peak {
    double x;       // centre, bins
    double height;  // dB over the floor
    double power;   // linear, at the peak bin
};

struct Grid {
    double c = 0.0, d = 0.0;  // bins
    int64_t k_lo = 0, k_hi = -1;
    std::vector<int64_t> ks;
    std::vector<size_t> which;  // peak of each found line
    uint32_t found() const { return uint32_t(ks.size()); }
    double occupancy() const {
        return k_hi < k_lo ? 0.0 : double(ks.size()) / double(k_hi - k_lo + 1);
    }
};

double quantile(std::vector<double>& v, double q) {
    const size_t i = size_t(q * double(v.size() - 1));
    std::nth_element(v.begin(), v.begin() + ptrdiff_t(i), v.end());
    return v[i];
}

// Floor and sigma per bin, interpolated between knots.
void local_floor(const std::vector<double>& db, std::vector<double>& floor_db,
                 std::vector<double>& sigma) {
    const size_t n = db.size();
    size_t w = std::clamp<size_t>(n / 32, 65, 513) | 1;
    w = std::min(w, n | 1);
    const size_t step = std::max<size_t>(1, w / 4);
    std::vector<size_t> knots;
    for (size_t i = 0; i < n; i += step) knots.push_back(i);
    if (knots.back() != n - 1) knots.push_back(n - 1);
    std::vector<double> kf(knots.size()), ks(knots.size()), win;
    for (size_t j = 0; j < knots.size(); ++j) {
        const size_t lo = knots[j] >= w / 2 ? knots[j] - w / 2 : 0;
        const size_t hi = std::min(n, lo + w);
        win.assign(db.begin() + ptrdiff_t(hi > w ? hi - w : 0), db.begin() + ptrdiff_t(hi));
        const double p10 = quantile(win, 0.10), p20 = quantile(win, 0.20), p30 = quantile(win, 0.30);
        const double s = std::max((p30 - p10) / 0.7572, 1e-3);  // normal: z(0.3) - z(0.1)
        ks[j] = s;
        kf[j] = p20 + 0.8416 * s;
    }
    floor_db.resize(n);
    sigma.resize(n);
    size_t j = 0;
    for (size_t i = 0; i < n; ++i) {
        while (j + 1 < knots.size() && knots[j + 1] < i) ++j;
        const size_t a = knots[j], b = knots[std::min(j + 1, knots.size() - 1)];
        const double t = b > a ? double(i - a) / double(b - a) : 0.0;
        floor_db[i] = kf[j] + t * (kf[std::min(j + 1, knots.size() - 1)] - kf[j]);
        sigma[i] = ks[j] + t * (ks[std::min(j + 1, knots.size() - 1)] - ks[j]);
    }
}

std::vector<Peak> find_peaks(const std::vector<double>& db, const std::vector<float>& power) {
    std::vector<double> fl, sg;
    local_floor(db, fl, sg);
    const int64_t n = int64_t(db.size());
    std::vector<Peak> out;
    for (int64_t i = 1; i + 1 < n; ++i) {
        if (!(db[size_t(i)] > db[size_t(i - 1)] && db[size_t(i)] >= db[size_t(i + 1)])) continue;
        const double h = db[size_t(i)] - fl[size_t(i)];
        if (h < std::max(kMinHeightDb, kHeightSigmas * sg[size_t(i)])) continue;
        double lo_l = std::numeric_limits<double>::infinity(), lo_r = lo_l;
        for (int64_t j = i - kValleyFar; j <= i - kValleyNear; ++j)
            if (j >= 0) lo_l = std::min(lo_l, db[size_t(j)]);
        for (int64_t j = i + kValleyNear; j <= i + kValleyFar; ++j)
            if (j < n) lo_r = std::min(lo_r, db[size_t(j)]);
        if (db[size_t(i)] - lo_l < kProminenceDb || db[size_t(i)] - lo_r < kProminenceDb) continue;
        const double a = db[size_t(i - 1)], b = db[size_t(i)], c = db[size_t(i + 1)];
        const double den = a - 2.0 * b + c;
        const double delta = den < 0.0 ? std::clamp(0.5 * (a - c) / den, -0.5, 0.5) : 0.0;
        out.push_back({double(i) + delta, h, double(power[size_t(i)])});
    }
    if (out.size() > kMaxPeaks) {
        std::nth_element(out.begin(), out.begin() + ptrdiff_t(kMaxPeaks), out.end(),
                         [](const Peak& p, const Peak& q) { return p.height > q.height; });
        out.resize(kMaxPeaks);
        std::sort(out.begin(), out.end(), [](const Peak& p, const Peak& q) { return p.x < q.x; });
    }
    return out;
}

// Least-squares grid through the peaks nearest c + k d, tolerance tightening
// (the given schedule, else 0.3 d, 0.2 d, then the taken tolerance).
Grid fit_grid(const std::vector<Peak>& peaks, const std::vector<char>& live, double c, double d,
              std::vector<double> tols = {}) {
    Grid g;
    g.c = c;
    g.d = d;
    if (tols.empty()) tols = {0.3 * d, 0.2 * d, std::min(0.2 * d, kTakenBins)};
    for (double tol : tols) {
        // nearest peak to each k, the tallest when several are in tolerance
        std::vector<std::pair<int64_t, size_t>> best;
        for (size_t j = 0; j < peaks.size(); ++j) {
            if (!live[j]) continue;
            const double kk = std::round((peaks[j].x - g.c) / g.d);
            if (std::abs(peaks[j].x - (g.c + kk * g.d)) > tol) continue;
            best.push_back({int64_t(kk), j});
        }
        std::sort(best.begin(), best.end(), [&](const auto& p, const auto& q) {
            return p.first != q.first ? p.first < q.first
                                      : peaks[p.second].height > peaks[q.second].height;
        });
        g.ks.clear();
        g.which.clear();
        for (size_t i = 0; i < best.size(); ++i)
            if (i == 0 || best[i].first != best[i - 1].first) {
                g.ks.push_back(best[i].first);
                g.which.push_back(best[i].second);
            }
        if (g.ks.size() < 2) break;
        double sk = 0, sx = 0, skk = 0, skx = 0;
        const double m = double(g.ks.size());
        for (size_t i = 0; i < g.ks.size(); ++i) {
            const double k = double(g.ks[i]), x = peaks[g.which[i]].x;
            sk += k;
            sx += x;
            skk += k * k;
            skx += k * x;
        }
        const double den = m * skk - sk * sk;
        if (den <= 0) break;
        g.d = (m * skx - sk * sx) / den;
        g.c = (sx - g.d * sk) / m;
        if (g.d < kMinSpacingBins) break;
    }
    if (!g.ks.empty()) {
        g.k_lo = g.ks.front();
        g.k_hi = g.ks.back();
    }
    return g;
}

// The comb's grid in the live peaks, or a grid with no lines.
Grid search_grid(const std::vector<Peak>& peaks, const std::vector<char>& live, size_t bins) {
    std::vector<double> xs;
    for (size_t j = 0; j < peaks.size(); ++j)
        if (live[j]) xs.push_back(peaks[j].x);
    Grid none;
    if (xs.size() < RVX_COMB_MIN_LINES) return none;
    const double extent = xs.back() - xs.front();
    if (extent <= 7.0 * kMinSpacingBins) return none;
    const double f_lo = 7.0 / extent, f_hi = 1.0 / kMinSpacingBins, df = 1.0 / (2.0 * extent);
    const size_t nf = size_t((f_hi - f_lo) / df) + 1;
    // S(f) by rotating each peak's phasor one step of df at a time,
    // re-anchored exactly every 256 steps.
    std::vector<double> s(nf, 0.0);
    std::vector<Cd> sum(nf);
    for (double x : xs) {
        const Cd stepr = std::polar(1.0, kTwoPi * df * x);
        for (size_t i0 = 0; i0 < nf; i0 += 256) {
            Cd z = std::polar(1.0, kTwoPi * (f_lo + double(i0) * df) * x);
            const size_t i1 = std::min(nf, i0 + 256);
            for (size_t i = i0; i < i1; ++i) {
                sum[i] += z;
                z *= stepr;
            }
        }
    }
    double smax = 0;
    for (size_t i = 0; i < nf; ++i) {
        s[i] = std::abs(sum[i]);
        smax = std::max(smax, s[i]);
    }
    std::vector<size_t> cands;
    for (size_t i = 1; i + 1 < nf; ++i)
        if (s[i] >= s[i - 1] && s[i] > s[i + 1] && s[i] >= kCandidateFraction * smax) cands.push_back(i);
    std::sort(cands.begin(), cands.end(), [&](size_t a, size_t b) { return s[a] > s[b]; });
    if (cands.size() > kMaxCandidates) cands.resize(kMaxCandidates);
    std::vector<Grid> fits;
    for (size_t i : cands) {
        // refine f by a parabola through S, then c from the phase of the sum
        double f = f_lo + double(i) * df;
        const double den = s[i - 1] - 2 * s[i] + s[i + 1];
        if (den < 0) f += df * std::clamp(0.5 * (s[i - 1] - s[i + 1]) / den, -0.5, 0.5);
        Cd z(0, 0);
        for (double x : xs) z += std::polar(1.0, kTwoPi * f * x);
        const double d = 1.0 / f;
        const double c = std::arg(z) / (kTwoPi * f);
        Grid g = fit_grid(peaks, live, c, d);
        if (g.d >= kMinSpacingBins && g.found() >= RVX_COMB_MIN_LINES &&
            g.occupancy() >= kMinOccupancy)
            fits.push_back(std::move(g));
    }
    if (fits.empty()) return none;
    uint32_t best = 0;
    for (const Grid& g : fits) best = std::max(best, g.found());
    const Grid* pick = nullptr;
    for (const Grid& g : fits)
        if (double(g.found()) >= kKeepFraction * double(best) && (!pick || g.d > pick->d)) pick = &g;
    Grid out = *pick;
    (void)bins;
    return out;
}

// k renumbered so that line 0 is the one nearest DC (bin bins / 2).
void centre_k(Grid& g, size_t bins) {
    const double dc = double(bins / 2);
    const int64_t shift = int64_t(std::llround((dc - g.c) / g.d));
    g.c += double(shift) * g.d;
    g.k_lo -= shift;
    g.k_hi -= shift;
    for (auto& k : g.ks) k -= shift;
}

// One parity (0 even, 1 odd) of a grid refitted at twice its spacing: first
// through the joint grid's lines of that parity, then through every live
// peak near the half's own lines (a half whose spacing differs from the
// joint grid's strays from it toward the ends, where the joint fit lost it).
Grid half_of(const std::vector<Peak>& peaks, const std::vector<char>& live, const Grid& g,
             int parity) {
    std::vector<char> mine(peaks.size(), 0);
    for (size_t i = 0; i < g.ks.size(); ++i)
        if (((g.ks[i] % 2) + 2) % 2 == parity) mine[g.which[i]] = 1;
    const double c = g.c + double(parity) * g.d;
    Grid h = fit_grid(peaks, mine, c, 2.0 * g.d);
    if (h.found() < 2) return h;
    for (int pass = 0; pass < 3; ++pass) {
        const uint32_t before = h.found();
        h = fit_grid(peaks, live, h.c, h.d, {std::min(0.25 * h.d, 1.5), kTakenBins});
        if (h.found() <= before) break;
    }
    return h;
}

// Whether a half's end lines sit off the joint grid by more than the split
// tolerance.
bool half_disagrees(const Grid& whole, const Grid& half, int parity, double bins_per_hz) {
    const double tol = std::max(kSplitBins, kSplitHz * bins_per_hz);
    for (int64_t hk : {half.k_lo, half.k_hi}) {
        const double x_half = half.c + double(hk) * half.d;
        const double kw = double(2 * hk + parity);
        const double x_whole = whole.c + kw * whole.d;
        if (std::abs(x_half - x_whole) > tol) return true;
    }
    return false;
}

rvx_comb_estimate to_estimate(const Grid& g, const std::vector<Peak>& peaks, size_t bins,
                              double hz_per_bin) {
    rvx_comb_estimate e{};
    const double dc = double(bins / 2);
    e.offset_hz = (g.c - dc) * hz_per_bin;
    e.spacing_hz = g.d * hz_per_bin;
    e.k_lo = int32_t(g.k_lo);
    e.k_hi = int32_t(g.k_hi);
    e.lines_found = g.found();
    std::vector<double> h;
    for (size_t j : g.which) h.push_back(peaks[j].height);
    e.median_height_db = h.empty() ? 0.0f : float(quantile(h, 0.5));
    return e;
}

// Ripple sidebands: an offset r (bins) at which peaks stand beside at least
// half the found lines on both sides, and how many multiples of it do.
void find_sidebands(const std::vector<Peak>& peaks, const Grid& g, double& ripple_bins,
                    int32_t& orders) {
    ripple_bins = 0.0;
    orders = 0;
    std::vector<char> on_line(peaks.size(), 0);
    for (size_t j : g.which) on_line[j] = 1;
    std::vector<double> offs;
    for (size_t j = 0; j < peaks.size(); ++j) {
        if (on_line[j]) continue;
        const double kk = std::round((peaks[j].x - g.c) / g.d);
        if (kk < double(g.k_lo) || kk > double(g.k_hi)) continue;
        const double r = peaks[j].x - (g.c + kk * g.d);
        if (std::abs(r) > kTakenBins && std::abs(r) < 0.45 * g.d) offs.push_back(std::abs(r));
    }
    if (offs.size() < RVX_COMB_MIN_LINES) return;
    std::sort(offs.begin(), offs.end());
    // the densest run of offsets within 2 x the taken tolerance
    size_t best_n = 0, best_a = 0;
    for (size_t a = 0, b = 0; a < offs.size(); ++a) {
        while (b < offs.size() && offs[b] - offs[a] <= 2 * kTakenBins) ++b;
        if (b - a > best_n) {
            best_n = b - a;
            best_a = a;
        }
    }
    if (double(best_n) < 0.5 * 2.0 * double(g.found())) return;  // both sides of half the lines
    double r = 0.0;
    for (size_t i = best_a; i < best_a + best_n; ++i) r += offs[i];
    r /= double(best_n);
    ripple_bins = r;
    orders = 1;
    for (int m = 2; m <= 8; ++m) {
        size_t n = 0;
        for (double o : offs)
            if (std::abs(o - m * r) <= kTakenBins) ++n;
        if (double(n) < 0.3 * 2.0 * double(g.found())) break;
        orders = m;
    }
}

// Peaks on the grid's lines or their sidebands leave the search.
void take_peaks(const std::vector<Peak>& peaks, std::vector<char>& live, double c, double d,
                double k_lo, double k_hi, double ripple, int32_t orders) {
    for (size_t j = 0; j < peaks.size(); ++j) {
        if (!live[j]) continue;
        const double kk = std::round((peaks[j].x - c) / d);
        if (kk < k_lo - 1 || kk > k_hi + 1) continue;
        const double r = peaks[j].x - (c + kk * d);
        bool taken = std::abs(r) <= kTakenBins;
        for (int m = 1; m <= orders && !taken; ++m)
            taken = std::abs(std::abs(r) - m * ripple) <= kTakenBins;
        if (taken) live[j] = 0;
    }
}

double overlap_with(const rvx_comb_estimate& e, const rvx_comb_estimate& held, double tol_hz) {
    if (e.k_hi < e.k_lo) return 0.0;
    uint32_t on = 0;
    for (int32_t k = e.k_lo; k <= e.k_hi; ++k) {
        const double f = e.offset_hz + double(k) * e.spacing_hz + double(k) * k * e.curvature_hz;
        if (held.spacing_hz <= 0) {
            if (std::abs(f - held.offset_hz) <= tol_hz) ++on;
            continue;
        }
        // the held grid's line nearest f: round against the straight grid, then
        // correct once for the curvature at that k
        const double k0 = std::round((f - held.offset_hz) / held.spacing_hz);
        const double kk =
            std::round((f - held.offset_hz - held.curvature_hz * k0 * k0) / held.spacing_hz);
        if (kk < held.k_lo || kk > held.k_hi) continue;
        const double g = held.offset_hz + kk * held.spacing_hz + kk * kk * held.curvature_hz;
        if (std::abs(f - g) <= tol_hz) ++on;
    }
    return double(on) / double(e.k_hi - e.k_lo + 1);
}

int32_t check_spectrum(const float* power, uint64_t bins, int64_t rate, const char* who) {
    if (!power || bins < kMinBins)
        return rvx::fail(RVX_EINVAL, "%s: a spectrum of %llu bins (at least %llu)", who,
                         (unsigned long long)bins, (unsigned long long)kMinBins);
    if (rate <= 0) return rvx::fail(RVX_EINVAL, "%s: sample rate %lld", who, (long long)rate);
    for (uint64_t i = 0; i < bins; ++i)
        if (!std::isfinite(power[i]) || power[i] < 0.0f)
            return rvx::fail(RVX_EINVAL, "%s: bin %llu holds %g, not a power", who,
                             (unsigned long long)i, double(power[i]));
    return RVX_OK;
}

}  // namespace

extern "C" int32_t rvx_comb_detect(const float* power, uint64_t bins, int64_t rate,
                                   const rvx_comb_estimate* active, uint32_t n_active,
                                   uint32_t halves, rvx_comb_estimate* out, uint32_t max_out,
                                   uint32_t* n_out) {
    static const char* who = "rvx_comb_detect";
    if (!n_out || (max_out && !out) || (n_active && !active))
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    *n_out = 0;
    if (int32_t rc = check_spectrum(power, bins, rate, who)) return rc;
    const double hz_per_bin = double(rate) / double(bins);
    std::vector<float> pw(power, power + bins);
    std::vector<double> db(bins);
    for (uint64_t i = 0; i < bins; ++i) db[i] = 10.0 * std::log10(std::max(double(pw[i]), 1e-30));
    const std::vector<Peak> peaks = find_peaks(db, pw);
    std::vector<char> live(peaks.size(), 1);
    const double dc = double(bins / 2);
    // peaks on a held comb's lines and sidebands leave first
    for (uint32_t a = 0; a < n_active; ++a) {
        const rvx_comb_estimate& h = active[a];
        const double c = dc + h.offset_hz / hz_per_bin;
        const double d = h.spacing_hz > 0 ? h.spacing_hz / hz_per_bin : double(bins);
        take_peaks(peaks, live, c, d, h.k_lo, h.k_hi, h.ripple_hz / hz_per_bin, h.sideband_orders);
    }
    std::vector<rvx_comb_estimate> found;
    std::vector<double> strength;
    const uint32_t cap = std::min<uint32_t>(max_out, RVX_COMB_MAX_COMBS);
    for (int pass = 0; pass < 4 * RVX_COMB_MAX_COMBS && found.size() < cap; ++pass) {
        Grid g = search_grid(peaks, live, bins);
        if (g.found() < RVX_COMB_MIN_LINES) break;
        double ripple = 0.0;
        int32_t orders = 0;
        find_sidebands(peaks, g, ripple, orders);
        // the grid split into halves: once at the top when asked, then while
        // a half's own halves disagree with it
        std::vector<Grid> parts;
        std::vector<Grid> todo{g};
        bool top = halves != 0;
        while (!todo.empty()) {
            Grid w = todo.back();
            todo.pop_back();
            Grid h0 = half_of(peaks, live, w, 0), h1 = half_of(peaks, live, w, 1);
            const bool both = h0.found() >= RVX_COMB_MIN_LINES && h1.found() >= RVX_COMB_MIN_LINES;
            const double slots = double(w.k_hi - w.k_lo + 1) / 2.0;
            const bool empty0 = double(h0.found()) < kEmptyParity * slots;
            const bool empty1 = double(h1.found()) < kEmptyParity * slots;
            if (both && (top || half_disagrees(w, h0, 0, 1.0 / hz_per_bin) ||
                         half_disagrees(w, h1, 1, 1.0 / hz_per_bin))) {
                todo.push_back(h0);
                todo.push_back(h1);
            } else if (empty0 != empty1 && (empty0 ? h1 : h0).found() >= RVX_COMB_MIN_LINES) {
                parts.push_back(empty0 ? h1 : h0);  // one parity only: the comb is at 2 d
            } else {
                parts.push_back(w);
            }
            top = false;
        }
        // the grid's peaks and sidebands leave the search, and each part's
        take_peaks(peaks, live, g.c, g.d, double(g.k_lo), double(g.k_hi), ripple, orders);
        for (size_t j : g.which) live[j] = 0;
        for (const Grid& p : parts) {
            take_peaks(peaks, live, p.c, p.d, double(p.k_lo), double(p.k_hi), ripple, orders);
            for (size_t j : p.which) live[j] = 0;
        }
        bool counted = false;
        for (Grid& p : parts) {
            centre_k(p, bins);
            rvx_comb_estimate e = to_estimate(p, peaks, bins, hz_per_bin);
            e.ripple_hz = ripple * hz_per_bin;
            e.sideband_orders = orders;
            bool held = false;
            for (uint32_t a = 0; a < n_active && !held; ++a)
                held = overlap_with(e, active[a], RVX_COMB_SAME_LINE_HZ) >= RVX_COMB_OVERLAP_FRACTION;
            if (held || found.size() >= cap) continue;  // a held comb's stray pass does not count
            double sum = 0.0;
            for (size_t j : p.which) sum += peaks[j].power;
            found.push_back(e);
            strength.push_back(sum);
            counted = true;
        }
        (void)counted;
    }
    std::vector<size_t> order(found.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return strength[a] > strength[b]; });
    for (size_t i = 0; i < order.size(); ++i) out[i] = found[order[i]];
    *n_out = uint32_t(order.size());
    return RVX_OK;
}

// ---------------------------------------------------------------------------
// Tracking
// ---------------------------------------------------------------------------

namespace {

constexpr double kStandOut = 4.0;          // 6 dB: a line's tone over its noise in a block
constexpr double kLoopSeconds = 0.05;      // phase loop time constant
constexpr double kLockResidual = 0.25;     // rad
constexpr double kLockInnovation = 0.03;   // cycles
constexpr double kBadResidual = 0.6;       // rad
constexpr double kBadFrequency = 0.25;     // cycles per block
constexpr double kStageExit = 0.1;         // cycles per block
constexpr double kBenchRad = 0.8;
constexpr uint32_t kLockBlocks = 8, kBadBlocks = 6, kStageBlocks = 3;
constexpr uint32_t kBenchBlocks = 4, kUnbenchBlocks = 8, kAcquireBlocks = 200;
constexpr double kRecoverSeconds = 8.0;
constexpr double kRecentreSeconds = 1.0;  // a recovery left to its loops this long before a detection recentres it
constexpr double kAmpWeight = 1.0 / 16.0;
constexpr uint32_t kPlanSubBlocks = 16;     // the most sub-blocks a plan asks for
constexpr double kMinSubBlockSeconds = 0.005;  // a narrowed pass's shortest sub-block
constexpr double kNarrowCycles = 1.0;          // frequency error a block under which a pass narrows

// One parameter (c or d) as an exact polynomial in the sample index: phase
// (cycles), frequency (cycles per sample) and its rate of change, x 2^64.
struct Axis {
    uint64_t phase = 0, freq = 0, rate = 0;
};

Axis advanced(const Axis& a, int64_t n) {
    const uint64_t un = uint64_t(n);
    const uint64_t tri = uint64_t((n * (n - 1)) / 2);  // |n| < 2^31 here
    return {a.phase + un * a.freq + tri * a.rate, a.freq + un * a.rate, a.rate};
}

uint64_t fixed(double cycles) { return uint64_t(int64_t(std::llround(cycles * kTwo64))); }
double signed_cycles(uint64_t v) { return double(int64_t(v)) / kTwo64; }
double wrap_half(double x) { return x - std::round(x); }

// Adds (dp, df, dr) applied at index `at` to an axis anchored at `anchor`.
void correct(Axis& a, uint64_t anchor, uint64_t at, double dp, double df, double dr) {
    const double n = double(int64_t(anchor - at));
    a.phase += fixed(wrap_half(dp + df * n + 0.5 * dr * n * n));
    a.freq += fixed(df + dr * n);
    a.rate += fixed(dr);
}

struct Line {
    int32_t k = 0;
    int32_t m = 0;  // ripple sideband order (0: the carrier)
    Cd amp{};
    bool seeded = false;
    bool strong = false;
    bool benched = false;
    uint32_t outlier = 0, inlier = 0;
    double before = 0.0, after = 0.0, noise = 0.0;  // per-sample powers, averaged
};

struct Planned {
    int32_t k, m;
    uint32_t phase0, step;
    Cd amp;
};

struct Record {
    uint64_t start = 0;
    uint32_t count = 0;
    uint32_t subs = 0;  // the sub-blocks this tracker asked for
    uint64_t gen = 0;
    bool subtract = false;
    std::vector<Planned> lines;
};

// Weighted least squares y = a + b k; b forced to 0 when `one` (a spur or a
// single distinct k). Two passes of Huber reweighting past 2.5 x the median
// absolute residual.
void robust_fit(const std::vector<double>& k, const std::vector<double>& y,
                std::vector<double> w, bool one, double& a, double& b) {
    a = b = 0.0;
    for (int pass = 0; pass < 3; ++pass) {
        double sw = 0, sk = 0, sy = 0, skk = 0, sky = 0;
        for (size_t i = 0; i < k.size(); ++i) {
            sw += w[i];
            sk += w[i] * k[i];
            sy += w[i] * y[i];
            skk += w[i] * k[i] * k[i];
            sky += w[i] * k[i] * y[i];
        }
        if (sw <= 0) return;
        const double den = sw * skk - sk * sk;
        if (one || den <= 1e-12 * sw * sw) {
            b = 0.0;
            a = sy / sw;
        } else {
            b = (sw * sky - sk * sy) / den;
            a = (sy - b * sk) / sw;
        }
        if (pass == 2) break;
        std::vector<double> r;
        for (size_t i = 0; i < k.size(); ++i)
            if (w[i] > 0) r.push_back(std::abs(y[i] - a - b * k[i]));
        if (r.size() < 3) break;
        const double med = quantile(r, 0.5);
        if (med <= 0) break;
        for (size_t i = 0; i < k.size(); ++i) {
            const double ri = std::abs(y[i] - a - b * k[i]);
            if (ri > 2.5 * med) w[i] *= 2.5 * med / ri;
        }
    }
}

// Weighted least squares y = a + b k + e k^2 (e = 0 unless `curved`), with the
// same Huber reweighting as robust_fit; rms is the weighted RMS residual of
// the last pass.
void robust_fit3(const std::vector<double>& k, const std::vector<double>& y,
                 std::vector<double> w, bool curved, double& a, double& b, double& e,
                 double& rms) {
    a = b = e = rms = 0.0;
    if (!curved) {
        robust_fit(k, y, w, false, a, b);
    } else {
        for (int pass = 0; pass < 3; ++pass) {
            double S[5] = {0, 0, 0, 0, 0}, T[3] = {0, 0, 0};
            for (size_t i = 0; i < k.size(); ++i) {
                double kp = 1.0;
                for (int j = 0; j < 5; ++j, kp *= k[i]) {
                    S[j] += w[i] * kp;
                    if (j < 3) T[j] += w[i] * kp * y[i];
                }
            }
            if (S[0] <= 0) return;
            double A[3][4] = {{S[0], S[1], S[2], T[0]}, {S[1], S[2], S[3], T[1]},
                              {S[2], S[3], S[4], T[2]}};
            bool ok = true;
            for (int c = 0; c < 3 && ok; ++c) {  // partial pivoting
                int piv = c;
                for (int r = c + 1; r < 3; ++r)
                    if (std::abs(A[r][c]) > std::abs(A[piv][c])) piv = r;
                if (std::abs(A[piv][c]) <= 1e-12 * S[0]) ok = false;
                if (!ok) break;
                for (int j = 0; j < 4; ++j) std::swap(A[c][j], A[piv][j]);
                for (int r = 0; r < 3; ++r)
                    if (r != c) {
                        const double f = A[r][c] / A[c][c];
                        for (int j = 0; j < 4; ++j) A[r][j] -= f * A[c][j];
                    }
            }
            if (!ok) {  // too few distinct k for a curve
                robust_fit(k, y, w, false, a, b);
                e = 0.0;
                break;
            }
            a = A[0][3] / A[0][0];
            b = A[1][3] / A[1][1];
            e = A[2][3] / A[2][2];
            if (pass == 2) break;
            std::vector<double> r;
            for (size_t i = 0; i < k.size(); ++i)
                if (w[i] > 0) r.push_back(std::abs(y[i] - a - b * k[i] - e * k[i] * k[i]));
            if (r.size() < 3) break;
            const double med = quantile(r, 0.5);
            if (med <= 0) break;
            for (size_t i = 0; i < k.size(); ++i) {
                const double ri = std::abs(y[i] - a - b * k[i] - e * k[i] * k[i]);
                if (ri > 2.5 * med) w[i] *= 2.5 * med / ri;
            }
        }
    }
    double sw = 0.0, s2 = 0.0;
    for (size_t i = 0; i < k.size(); ++i) {
        const double ri = y[i] - a - b * k[i] - e * k[i] * k[i];
        sw += w[i];
        s2 += w[i] * ri * ri;
    }
    rms = sw > 0 ? std::sqrt(s2 / sw) : 0.0;
}

constexpr uint32_t kGrowCarriers = 16;  // lines a locked pass may grow by in one block
// The curved model c + k d + k^2 e (comb_tracker.cpp choose_model): taken after
// 4 blocks in a row in which the quadratic fit of the carriers' frequency
// errors has under half the straight fit's mean square and bends the outermost
// line by over 0.05 cycles a block (with at least 24 lines over 16 k); let go
// after 64 blocks in which the curvature held bends it by under 0.02.
constexpr uint32_t kCurveMinLines = 24;
constexpr double kCurveMinSpan = 16.0;
constexpr double kCurveGain = 0.5;
constexpr double kCurveOnCycles = 0.05;
constexpr double kCurveOffCycles = 0.02;
constexpr uint32_t kCurveVotes = 4;
constexpr uint32_t kStraightVotes = 64;

class Tracker {
public:
    Tracker(int64_t rate, uint64_t id) : rate_(rate), id_(id) {}

    uint64_t id() const { return id_; }
    uint32_t lock() const { return lock_; }
    bool spur() const { return spur_; }
    // Where the detector put it (a spur's hold-off and withdrawal go by this).
    double detected_offset_hz() const { return est_.offset_hz; }
    // A spur tracker that has ended since its bank last looked.
    bool ended_spur() const { return spur_ && started_ && !active() && !end_noted_; }
    void note_end() { end_noted_ = true; }
    uint64_t order = 0;  // when its bank started it (newest spurs give way first)
    bool active() const {
        return lock_ == RVX_COMB_ACQUIRING || lock_ == RVX_COMB_LOCKED || lock_ == RVX_COMB_RECOVERING;
    }
    uint32_t lines() const { return uint32_t(lines_.size()); }

    void start(const rvx_comb_estimate& e, uint32_t max_lines) {
        est_ = e;
        spur_ = e.spur != 0;
        started_ = true;
        end_noted_ = false;
        min_lines_ = spur_ ? 1u : uint32_t(RVX_COMB_MIN_LINES);
        lock_ = RVX_COMB_ACQUIRING;
        stage_ = kFrequency;
        ++gen_;
        anchored_ = false;
        c_ = {0, fixed(e.offset_hz / double(rate_)), 0};
        d_ = {0, spur_ ? 0 : fixed(e.spacing_hz / double(rate_)), 0};
        lines_.clear();
        // Ripple sidebands only with a ripple to put them at, and only as many
        // orders as leave room for RVX_COMB_MIN_LINES carriers.
        orders_ = 0;
        r_ = {};
        e_ = {};
        curved_ = false;
        curve_votes_ = straight_votes_ = 0;
        if (!spur_ && e.ripple_hz > 0.0 && e.sideband_orders > 0) {
            orders_ = std::min<int32_t>(e.sideband_orders,
                                        int32_t((max_lines / RVX_COMB_MIN_LINES - 1) / 2));
            if (orders_ > 0) r_ = {0, fixed(e.ripple_hz / double(rate_)), 0};
        }
        if (spur_) {
            lines_.push_back(Line{});
        } else {
            // the central carriers of the detected range, each followed by its
            // sidebands (m = -orders .. orders), up to max_lines; the rest of
            // the range is grown into once the pass is locked
            const int64_t n = int64_t(e.k_hi) - e.k_lo + 1;
            const int64_t take = std::min<int64_t>(n, max_lines / per_line());
            const int64_t lo = e.k_lo + (n - take) / 2;
            for (int64_t k = lo; k < lo + take; ++k) add_group(int32_t(k));
            range_lo_ = e.k_lo;
            range_hi_ = e.k_hi;
            next_lo_ = int32_t(lo) - 1;
            next_hi_ = int32_t(lo + take);
            take_hi_ = true;
        }
        good_ = false;
        fitted_ = benched_ = 0;
        wide_ = true;
        records_.clear();
        observed_ = settle_ = bad_ = stage_settle_ = 0;
        seed_pending_ = false;
        suppression_ = 0.0f;
        cancelled_ = 0;
    }

    // Sub-blocks of a block's projections: the most (16, at least 16 samples
    // each) while the pass pulls in wide, a frequency step of up to half a
    // cycle per sub-block; once its error is under a cycle a block, the most
    // that keep each at least 5 ms long, for less noise per sub-block.
    uint32_t pick_sub_blocks(uint32_t count) const {
        const uint32_t shortest =
            wide_ ? 16u : std::max<uint32_t>(16u, uint32_t(kMinSubBlockSeconds * double(rate_)));
        for (uint32_t s = kPlanSubBlocks; s > 1; s /= 2)
            if (count % s == 0 && count / s >= shortest) return s;
        return 1;
    }

    void stop(uint32_t lock) {
        lock_ = lock;
        lines_.clear();
        records_.clear();
        cancelled_ = 0;
    }

    // Growth (comb_tracker.cpp wants_lines / grow / shrink_to / extend_range):
    // a healthy locked pass with more of its comb's range to reach asks for up
    // to kGrowCarriers lines a block, taken alternately above and below; a
    // pass past a fair share gives the excess back, keeping its central lines.
    uint32_t per_line() const { return uint32_t(2 * orders_ + 1); }
    // A carrier and its sidebands, the carrier in the middle (slot of order m
    // minus m is its carrier's).
    void add_group(int32_t k) {
        for (int32_t m = -orders_; m <= orders_; ++m) {
            Line l;  // unseeded: out of the phase fit until its first locked block
            l.k = k;
            l.m = m;
            lines_.push_back(l);
        }
    }
    uint32_t wants_lines() const {
        if (spur_ || lock_ != RVX_COMB_LOCKED || !good_ || 4 * benched_ > fitted_) return 0;
        const int64_t left = std::max<int64_t>(0, int64_t(range_hi_) - next_hi_ + 1) +
                             std::max<int64_t>(0, int64_t(next_lo_) - range_lo_ + 1);
        return uint32_t(std::min<int64_t>(left, kGrowCarriers)) * per_line();
    }
    uint32_t grow(uint32_t slots) {
        uint32_t taken = 0;
        while (taken + per_line() <= slots) {
            const bool hi_left = next_hi_ <= range_hi_, lo_left = next_lo_ >= range_lo_;
            if (!hi_left && !lo_left) break;
            const bool hi = hi_left && (take_hi_ || !lo_left);
            add_group(hi ? next_hi_++ : next_lo_--);  // appended: plans already out still line up
            take_hi_ = !hi;
            taken += per_line();
        }
        return taken;
    }
    uint32_t shrink_to(uint32_t slots) {
        const size_t floor = std::max<size_t>(slots, size_t(RVX_COMB_MIN_LINES) * per_line());
        if (spur_ || lines_.size() <= floor) return 0;
        const size_t before = lines_.size();
        std::vector<int32_t> ks;
        for (const Line& l : lines_) ks.push_back(l.k);
        std::sort(ks.begin(), ks.end());
        const size_t first = (ks.size() - floor) / 2;
        const int32_t lo = ks[first], hi = ks[first + floor - 1];
        std::erase_if(lines_, [&](const Line& l) { return l.k < lo || l.k > hi; });
        next_lo_ = lo - 1;
        next_hi_ = hi + 1;
        ++gen_;  // plans already out no longer name these lines
        records_.clear();
        return uint32_t(before - lines_.size());
    }
    // A detection of the comb this pass holds widens the range it grows into:
    // the detection's end lines mapped onto this pass's own k.
    void extend_range(const rvx_comb_estimate& e) {
        if (spur_ || e.spur || !active() || e.k_hi < e.k_lo) return;
        const double rate = double(rate_);
        const double c = anchored_ ? signed_cycles(c_.freq) * rate : est_.offset_hz;
        const double d = anchored_ ? signed_cycles(d_.freq) * rate : est_.spacing_hz;
        if (!(d > 0.0)) return;
        const double cv = curved_ ? signed_cycles(e_.freq) * rate : 0.0;
        auto to_k = [&](int32_t k) {
            const double f = e.offset_hz + double(k) * e.spacing_hz + double(k) * k * e.curvature_hz;
            const double k0 = std::round((f - c) / d);
            return int32_t(std::lround((f - c - cv * k0 * k0) / d));
        };
        const int32_t a = to_k(e.k_lo), b = to_k(e.k_hi);
        range_lo_ = std::min({range_lo_, a, b});
        range_hi_ = std::max({range_hi_, a, b});
    }

    // Lines of the block [start, start + count), phases at its first sample.
    void plan(uint64_t start, uint32_t count, std::vector<rvx_comb_line>& out, bool& subtract,
              uint32_t& subs) {
        if (!active() || lines_.empty()) return;
        const uint64_t m = start + count / 2;
        if (!anchored_) {
            anchor_ = m;
            anchored_ = true;
        }
        // re-anchor at the newest middle, so advances stay short
        c_ = advanced(c_, int64_t(m - anchor_));
        d_ = advanced(d_, int64_t(m - anchor_));
        r_ = advanced(r_, int64_t(m - anchor_));
        e_ = advanced(e_, int64_t(m - anchor_));
        anchor_ = m;
        Record r;
        r.start = start;
        r.count = count;
        r.subs = pick_sub_blocks(count);
        subs = std::max(subs, r.subs);
        r.gen = gen_;
        r.subtract = lock_ == RVX_COMB_LOCKED;
        cancelled_ = 0;
        for (size_t slot = 0; slot < lines_.size(); ++slot) {
            const Line& l = lines_[slot];
            const uint64_t uk = uint64_t(int64_t(l.k)), um = uint64_t(int64_t(l.m));
            const uint64_t uk2 = uint64_t(int64_t(l.k) * int64_t(l.k));
            const uint64_t ph = c_.phase + uk * d_.phase + uk2 * e_.phase + um * r_.phase;
            const uint64_t fr = c_.freq + uk * d_.freq + uk2 * e_.freq + um * r_.freq;
            const uint32_t step = uint32_t((fr + (uint64_t(1) << 31)) >> 32);
            const uint32_t mid = uint32_t((ph + (uint64_t(1) << 31)) >> 32);
            const uint32_t phase0 = mid - (count / 2) * step;
            // a sideband follows its carrier onto the bench, and is cut only
            // while its carrier stands out
            const Line& carrier = lines_[slot - size_t(ptrdiff_t(l.m))];
            const bool cut = r.subtract && l.seeded && carrier.strong && !carrier.benched;
            const Cd amp = cut ? l.amp : Cd(0.0, 0.0);
            cancelled_ += cut ? 1u : 0u;
            r.lines.push_back({l.k, l.m, phase0, step, amp});
            rvx_comb_line cl{};
            cl.phase0 = phase0;
            cl.step = step;
            cl.amplitude[0] = float(amp.real());
            cl.amplitude[1] = float(amp.imag());
            out.push_back(cl);
        }
        subtract = subtract || r.subtract;
        records_.push_back(std::move(r));
        while (records_.size() > 64) records_.erase(records_.begin());
    }

    // The projections (line-major, lines x subs complex) of a planned block.
    void observe(uint64_t start, uint32_t count, uint32_t subs, const std::complex<float>* proj,
                 uint32_t n_lines) {
        size_t at = records_.size();
        for (size_t i = 0; i < records_.size(); ++i)
            if (records_[i].start == start && records_[i].count == count) at = i;
        if (at == records_.size()) return;
        Record r = std::move(records_[at]);
        records_.erase(records_.begin(), records_.begin() + ptrdiff_t(at) + 1);
        // A record planned before the pass grew names its first lines only.
        if (r.gen != gen_ || !active() || n_lines != r.lines.size() || subs < 2 ||
            r.lines.size() > lines_.size())
            return;
        // Projections taken over more sub-blocks than this pass asked for (the
        // bank plans with the most any pass wants): adjacent ones summed back.
        std::vector<std::complex<float>> merged;
        if (r.subs >= 2 && subs > r.subs && subs % r.subs == 0) {
            const uint32_t f = subs / r.subs;
            merged.assign(size_t(n_lines) * r.subs, std::complex<float>(0, 0));
            for (uint32_t i = 0; i < n_lines; ++i)
                for (uint32_t s = 0; s < subs; ++s)
                    merged[size_t(i) * r.subs + s / f] += proj[size_t(i) * subs + s];
            proj = merged.data();
            subs = r.subs;
        }
        ++observed_;
        const uint64_t m = start + count / 2;
        const Axis mc = advanced(c_, int64_t(m - anchor_)), md = advanced(d_, int64_t(m - anchor_));
        const Axis mr = advanced(r_, int64_t(m - anchor_));
        const Axis me = advanced(e_, int64_t(m - anchor_));
        const double L = double(count) / double(subs);
        const size_t n = r.lines.size();
        std::vector<double> ks(n), ef(n), ep(n), w(n), snr_of(n);
        std::vector<Cd> vs_model(n);
        uint32_t strong = 0, benched = 0;
        // Block sums; with sidebands, each line's sum loses what the other
        // lines of its carrier put into it: line j's amplitude (its sum over
        // count) times the closed-form sum of tone_j conj(tone_i) over the block.
        std::vector<Cd> block(n, Cd(0, 0));
        for (size_t i = 0; i < n; ++i)
            for (uint32_t s = 0; s < subs; ++s) block[i] += Cd(proj[i * subs + s]);
        std::vector<Cd> clean(block);
        if (orders_ > 0) {
            const double N = double(count);
            for (size_t i = 0; i < n; ++i)
                for (size_t j = 0; j < n; ++j) {
                    if (j == i || r.lines[j].k != r.lines[i].k) continue;
                    const double dphi = double(int32_t(r.lines[j].phase0 - r.lines[i].phase0)) / kTwo32;
                    const double dw = double(int32_t(r.lines[j].step - r.lines[i].step)) / kTwo32;
                    Cd gain = std::abs(dw) < 1e-12
                                  ? Cd(N, 0.0)
                                  : (1.0 - std::polar(1.0, kTwoPi * dw * N)) /
                                        (1.0 - std::polar(1.0, kTwoPi * dw));
                    gain *= std::polar(1.0, kTwoPi * dphi);
                    clean[i] -= block[j] / N * gain;
                }
        }
        for (size_t i = 0; i < n; ++i) {
            const Planned& pl = r.lines[i];
            Line& l = lines_[i];
            Cd P = clean[i], D(0, 0);
            for (uint32_t s = 0; s + 1 < subs; ++s)
                D += Cd(proj[i * subs + s + 1]) * std::conj(Cd(proj[i * subs + s]));
            const Cd z = std::abs(D) > 0 ? D / std::abs(D) : Cd(1, 0);
            Cd cc(0, 0), zs(1, 0);
            for (uint32_t s = 0; s < subs; ++s, zs *= std::conj(z)) cc += Cd(proj[i * subs + s]) * zs;
            cc /= double(subs);
            double noise = 0.0;
            zs = Cd(1, 0);
            for (uint32_t s = 0; s < subs; ++s, zs *= z) noise += std::norm(Cd(proj[i * subs + s]) - cc * zs);
            const double np = noise / double(subs - 1) / L;  // per sample
            const double tp = std::norm(cc) / (L * L);
            const double snr = np > 0 ? tp * double(count) / np : 1e12;
            snr_of[i] = snr;
            l.strong = pl.m == 0 && snr > kStandOut;  // a sideband rides on its carrier's
            l.noise += kAmpWeight * (np - l.noise);
            const uint64_t uk = uint64_t(int64_t(pl.k));
            const double f_plan = double(int32_t(pl.step)) / kTwo32;
            const double f_meas = f_plan + std::arg(D) / (kTwoPi * L);
            const uint64_t uk2 = uint64_t(int64_t(pl.k) * int64_t(pl.k));
            const double f_model = signed_cycles(mc.freq + uk * md.freq + uk2 * me.freq);
            ef[i] = wrap_half(f_meas - f_model);
            const double ph_plan = double(uint32_t(pl.phase0 + (count / 2) * pl.step)) / kTwo32;
            const uint64_t um = uint64_t(int64_t(pl.m));
            const double ph_model =
                double(mc.phase + uk * md.phase + uk2 * me.phase + um * mr.phase) / kTwo64;
            const Cd a = P / double(count);
            vs_model[i] = a * std::polar(1.0, kTwoPi * (ph_plan - ph_model));
            ep[i] = l.seeded ? std::arg(vs_model[i] / l.amp) / kTwoPi : 0.0;
            if (r.subtract && std::abs(pl.amp) > 0) {
                l.before += kAmpWeight * (std::norm(a) - l.before);
                l.after += kAmpWeight * (std::norm(a - pl.amp) - l.after);
            }
            ks[i] = double(pl.k);
            // the carriers alone steer c and d and decide the lock
            w[i] = pl.m == 0 && l.strong && !l.benched ? snr : 0.0;
            if (pl.m != 0) ef[i] = 0.0;
            strong += l.strong ? 1u : 0u;
            benched += pl.m == 0 && l.benched ? 1u : 0u;
        }
        const double T = double(count);
        const bool one = spur_;
        double kbar = 0, sw = 0, kmin = 1e300, kmax = -1e300;
        for (size_t i = 0; i < n; ++i)
            if (w[i] > 0) {
                kbar += w[i] * ks[i];
                sw += w[i];
                kmin = std::min(kmin, ks[i]);
                kmax = std::max(kmax, ks[i]);
            }
        kbar = sw > 0 ? kbar / sw : 0.0;
        const double span = sw > 0 ? kmax - kmin : 0.0;
        if (!one) choose_model(ks, ef, w, kmin, kmax, T);
        double fa, fb, fe = 0.0, frms = 0.0;
        if (one)
            robust_fit(ks, ef, w, true, fa, fb);
        else
            robust_fit3(ks, ef, w, curved_, fa, fb, fe, frms);
        const double f_centre = std::abs(fa + fb * kbar + fe * kbar * kbar) * T;
        const double f_span = (std::abs(fb) * span + std::abs(fe) * 0.25 * span * span) * T;
        if (stage_ == kFrequency) {
            // frequency-locked: freq += 0.5 e, rate += e / 16 per block
            correct(c_, anchor_, m, 0.0, 0.5 * fa, fa / (16.0 * T));
            if (!one) correct(d_, anchor_, m, 0.0, 0.5 * fb, fb / (16.0 * T));
            if (curved_) correct(e_, anchor_, m, 0.0, 0.5 * fe, fe / (16.0 * T));
            if (wide_ && sw > 0 && f_centre + f_span < kNarrowCycles) wide_ = false;
            stage_settle_ = (sw > 0 && f_centre < kStageExit && f_span < kStageExit &&
                             strong >= min_lines_)
                                ? stage_settle_ + 1
                                : 0;
            if (stage_settle_ >= kStageBlocks) {
                stage_ = kPhase;
                seed_pending_ = true;
            }
        } else if (seed_pending_) {
            for (size_t i = 0; i < n; ++i) {
                lines_[i].amp = vs_model[i];
                lines_[i].seeded = true;
            }
            seed_pending_ = false;
        } else {
            // lines grown since the seeding have no amplitude yet: out of the fit
            std::vector<double> wp(w);
            uint32_t fitted = 0;
            for (size_t i = 0; i < n; ++i) {
                if (!lines_[i].seeded) wp[i] = 0.0;
                fitted += wp[i] > 0 ? 1u : 0u;
            }
            double pa, pb, pe = 0.0, prms = 0.0;
            if (one)
                robust_fit(ks, ep, wp, true, pa, pb);
            else
                robust_fit3(ks, ep, wp, curved_, pa, pb, pe, prms);
            double res = 0.0, inno = 0.0;
            for (size_t i = 0; i < n; ++i) {
                if (!lines_[i].strong || !lines_[i].seeded) continue;
                const double fit = pa + pb * ks[i] + pe * ks[i] * ks[i];
                const double off = std::abs(wrap_half(ep[i] - fit)) * kTwoPi;
                Line& l = lines_[i];
                if (off > kBenchRad) {
                    l.inlier = 0;
                    if (++l.outlier >= kBenchBlocks) l.benched = true;
                } else {
                    l.outlier = 0;
                    if (++l.inlier >= kUnbenchBlocks) l.benched = false;
                }
                if (wp[i] > 0) {
                    res += wp[i] * off * off;
                    inno = std::max(inno, std::abs(fit));
                }
            }
            double swp = 0.0;
            for (size_t i = 0; i < n; ++i) swp += wp[i];
            res = swp > 0 ? std::sqrt(res / swp) : 1e9;
            // third-order loop, all poles at exp(-T / 0.05 s)
            const double rho = std::exp(-T / (double(rate_) * kLoopSeconds));
            const double alpha = 1 - rho * rho * rho;
            const double beta = 1.5 * (1 - rho) * (1 - rho) * (1 + rho);
            const double gamma = (1 - rho) * (1 - rho) * (1 - rho);
            correct(c_, anchor_, m, alpha * pa, beta * pa / T, gamma * pa / (T * T));
            if (!one) correct(d_, anchor_, m, alpha * pb, beta * pb / T, gamma * pb / (T * T));
            if (curved_) correct(e_, anchor_, m, alpha * pe, beta * pe / T, gamma * pe / (T * T));
            // The ripple: a sideband's phase error past what the carriers' fit
            // gives its k is m times the ripple's; the weighted least-squares
            // slope through the origin in m, weighted by the carrier's SNR.
            double ripple = 0.0;
            if (orders_ > 0 && fitted >= min_lines_) {
                double swy = 0.0, sww = 0.0;
                uint32_t used = 0;
                for (size_t i = 0; i < n; ++i) {
                    const Line& l = lines_[i];
                    if (r.lines[i].m == 0 || !l.seeded || std::abs(l.amp) <= 0.0) continue;
                    const size_t ci = size_t(ptrdiff_t(i) - r.lines[i].m);
                    if (ci >= n || snr_of[ci] <= kStandOut) continue;
                    const double mm = double(r.lines[i].m);
                    const double y = wrap_half(ep[i] - (pa + pb * ks[i] + pe * ks[i] * ks[i]));
                    swy += snr_of[ci] * y * mm;
                    sww += snr_of[ci] * mm * mm;
                    ++used;
                }
                if (used >= min_lines_ && sww > 0.0) {
                    ripple = swy / sww;
                    correct(r_, anchor_, m, alpha * ripple, beta * ripple / T,
                            gamma * ripple / (T * T));
                }
            }
            for (size_t i = 0; i < n; ++i) {
                Line& l = lines_[i];
                const Cd derot = vs_model[i] * std::polar(
                    1.0, -kTwoPi * (pa + pb * ks[i] + pe * ks[i] * ks[i] +
                                    double(r.lines[i].m) * ripple));
                if (l.seeded) {
                    l.amp += kAmpWeight * (derot - l.amp);
                } else {  // a grown line: seeded from its first locked block
                    l.amp = derot;
                    l.seeded = true;
                }
            }
            const bool many = strong >= min_lines_ && 2 * benched <= n;
            const bool good = many && res < kLockResidual && inno < kLockInnovation;
            good_ = good;
            fitted_ = fitted;
            benched_ = benched;
            const bool bad = !many || res > kBadResidual || f_centre > kBadFrequency ||
                             f_span > kBadFrequency;
            if (lock_ == RVX_COMB_LOCKED) {
                bad_ = bad ? bad_ + 1 : 0;
                if (bad_ >= kBadBlocks) begin_recovery(m);
            } else {
                settle_ = good ? settle_ + 1 : 0;
                if (settle_ >= kLockBlocks) {
                    lock_ = RVX_COMB_LOCKED;
                    bad_ = 0;
                }
            }
            residual_ = res;
        }
        if (lock_ == RVX_COMB_RECOVERING &&
            double(m - recovering_since_) > kRecoverSeconds * double(rate_))
            stop(RVX_COMB_LOST);
        else if (lock_ == RVX_COMB_ACQUIRING && observed_ >= kAcquireBlocks)
            stop(RVX_COMB_LOST);
        // suppression over the lines being cancelled
        double sum = 0.0;
        uint32_t nsum = 0;
        for (const Line& l : lines_)
            if (l.before > 0 && l.after > 0 && l.seeded && l.strong && !l.benched) {
                const double nz = l.noise / double(count);
                const double b = std::max(l.before - nz, 1e-30), a = std::max(l.after - nz, 1e-30);
                sum += std::clamp(10.0 * std::log10(b / a), 0.0, 60.0);
                ++nsum;
            }
        if (nsum) suppression_ = float(sum / nsum);
    }

    // A recovering pass pulls in from the model it held, which is right after
    // a burst but too narrow once the comb has moved while the pass could not
    // see it. A detection names where the comb is now: mapped onto this pass's
    // own k it gives c and d, and the frequency stage pulls in from there.
    // Once per recovery after the first kRecentreSeconds, unless the comb is
    // known to have moved (`moved`: a detection on this spacing that overlaps
    // none of the pass's lines).
    bool same_spacing(const rvx_comb_estimate& e) const {
        if (spur_ || e.spur || !anchored_ || e.k_hi < e.k_lo) return false;
        const double d = signed_cycles(d_.freq) * double(rate_);
        const double half = 0.5 * double(int64_t(e.k_hi) - e.k_lo + 1);
        return d > 0.0 && std::abs(e.spacing_hz - d) * half < 0.25 * d;
    }
    bool recentre(const rvx_comb_estimate& e, bool moved) {
        if (spur_ || e.spur || lock_ != RVX_COMB_RECOVERING || !anchored_ || e.k_hi < e.k_lo ||
            !(e.spacing_hz > 0.0))
            return false;
        const double rate = double(rate_);
        if (!moved && (recentred_ || double(int64_t(anchor_ - recovering_since_)) <
                                         kRecentreSeconds * rate))
            return false;
        const double c = signed_cycles(c_.freq) * rate, d = signed_cycles(d_.freq) * rate;
        if (!(d > 0.0)) return false;
        // this pass's k at the estimate's 0: rounded against the straight grid,
        // then corrected once for the curvature the pass holds at that k
        const double cv = curved_ ? signed_cycles(e_.freq) * rate : 0.0;
        const double k0 = std::round((e.offset_hz - c) / d);
        const double j0 = std::round((e.offset_hz - c - cv * k0 * k0) / d);
        c_ = {c_.phase, fixed((e.offset_hz - j0 * e.spacing_hz) / rate), 0};
        d_ = {d_.phase, fixed(e.spacing_hz / rate), 0};
        recentred_ = true;
        stage_ = kFrequency;
        wide_ = true;  // the detection is coarser than the model was
        stage_settle_ = settle_ = bad_ = 0;
        ++gen_;  // plans already out were made against the old model
        records_.clear();
        for (Line& l : lines_) l.seeded = false;
        return true;
    }

    // Straight or curved (see kCurveVotes), judged on the carriers' frequency
    // errors each block: the curve is taken when its fit wins decisively and
    // bends the outer lines, and dropped, set to exactly zero, once what it
    // holds bends them negligibly (judged on the model, not the innovations,
    // which go quiet once the curvature is tracked).
    void choose_model(const std::vector<double>& ks, const std::vector<double>& ys,
                      const std::vector<double>& w, double kmin, double kmax, double T) {
        uint32_t lines = 0;
        for (double x : w) lines += x > 0 ? 1u : 0u;
        if (spur_ || lines < kCurveMinLines || !(kmax - kmin >= kCurveMinSpan)) {
            curve_votes_ = 0;
            return;
        }
        const double half = 0.5 * (kmax - kmin);
        if (!curved_) {
            double a, b, e, rms_s, rms_c;
            robust_fit3(ks, ys, w, false, a, b, e, rms_s);
            robust_fit3(ks, ys, w, true, a, b, e, rms_c);
            const double bend = std::abs(e) * half * half * T;
            const bool better = rms_c * rms_c < kCurveGain * rms_s * rms_s && bend > kCurveOnCycles;
            curve_votes_ = better ? curve_votes_ + 1 : 0;
            if (curve_votes_ >= kCurveVotes) {
                curved_ = true;
                curve_votes_ = straight_votes_ = 0;
            }
            return;
        }
        const double bend = std::abs(signed_cycles(e_.freq)) * half * half * T;
        straight_votes_ = bend < kCurveOffCycles ? straight_votes_ + 1 : 0;
        if (straight_votes_ >= kStraightVotes) {
            e_ = {};
            curved_ = false;
            straight_votes_ = 0;
        }
    }

    void begin_recovery(uint64_t m) {
        lock_ = RVX_COMB_RECOVERING;
        stage_ = kFrequency;
        stage_settle_ = settle_ = bad_ = 0;
        c_.rate = d_.rate = r_.rate = e_.rate = 0;  // frequencies held, rates dropped
        for (Line& l : lines_) l.seeded = false;
        recovering_since_ = m;
        recentred_ = false;
        wide_ = false;  // the model it held is close: long sub-blocks
    }

    rvx_comb_estimate estimate() const {
        rvx_comb_estimate e = est_;
        e.offset_hz = signed_cycles(c_.freq) * double(rate_);
        e.spacing_hz = spur_ ? 0.0 : signed_cycles(d_.freq) * double(rate_);
        if (!spur_ && !lines_.empty()) {  // the lines held now, grown ones included
            e.k_lo = e.k_hi = lines_.front().k;
            for (const Line& l : lines_) {
                e.k_lo = std::min(e.k_lo, l.k);
                e.k_hi = std::max(e.k_hi, l.k);
            }
            e.lines_found = uint32_t(lines_.size()) / per_line();
            e.ripple_hz = orders_ > 0 ? signed_cycles(r_.freq) * double(rate_) : 0.0;
            e.curvature_hz = curved_ ? signed_cycles(e_.freq) * double(rate_) : 0.0;
            e.sideband_orders = orders_;
        }
        return e;
    }

    // Fraction of the estimate's lines within tol_hz of one of this tracker's.
    double overlap(const rvx_comb_estimate& e, double tol_hz) const {
        const rvx_comb_estimate mine = estimate();
        if (e.spur) {
            return holds(e.offset_hz, tol_hz) ? 1.0 : 0.0;
        }
        return overlap_with(e, mine, tol_hz);  // over the k range held now
    }

    bool holds(double f_hz, double tol_hz) const {
        const rvx_comb_estimate mine = estimate();
        for (const Line& l : lines_)
            if (std::abs(mine.offset_hz + double(l.k) * mine.spacing_hz +
                         double(l.k) * l.k * mine.curvature_hz + double(l.m) * mine.ripple_hz -
                         f_hz) <= tol_hz)
                return true;
        return false;
    }

    bool agrees(const rvx_comb_estimate& e) const {
        if (!active()) return false;
        const rvx_comb_estimate mine = estimate();
        if (spur_ || e.spur) return spur_ == (e.spur != 0) && std::abs(e.offset_hz - mine.offset_hz) <= RVX_COMB_SAME_LINE_HZ;
        const double span = std::max(1.0, double(std::max(std::abs(e.k_lo), std::abs(e.k_hi))));
        return std::abs(e.spacing_hz - mine.spacing_hz) * span <= RVX_COMB_SAME_LINE_HZ &&
               overlap(e, RVX_COMB_SAME_LINE_HZ) >= 0.5;
    }

    void status(rvx_comb_status& s) const {
        s.lock = lock_;
        s.comb = estimate();
        s.lines_cancelled = lock_ == RVX_COMB_LOCKED ? cancelled_ : 0;
        s.suppression_db = lock_ == RVX_COMB_LOCKED ? suppression_ : 0.0f;
        s.spacing_drift_hz_per_s =
            float(signed_cycles(spur_ ? c_.rate : d_.rate) * double(rate_) * double(rate_));
        s.residual_rad = float(residual_);
    }

private:
    bool started_ = false, end_noted_ = false;
    int32_t range_lo_ = 0, range_hi_ = 0, next_lo_ = 0, next_hi_ = 0;
    bool take_hi_ = true, good_ = false;
    uint32_t fitted_ = 0, benched_ = 0;
    enum Stage { kFrequency, kPhase };
    int64_t rate_;
    uint64_t id_;
    rvx_comb_estimate est_{};
    bool spur_ = false;
    uint32_t min_lines_ = RVX_COMB_MIN_LINES;
    uint32_t lock_ = RVX_COMB_IDLE;
    Stage stage_ = kFrequency;
    uint64_t gen_ = 0;
    bool anchored_ = false;
    uint64_t anchor_ = 0;
    Axis c_{}, d_{}, r_{};  // offset, spacing, ripple (r_ zero without sidebands)
    Axis e_{};              // curvature: exactly zero while the straight model is in use
    bool curved_ = false;
    uint32_t curve_votes_ = 0, straight_votes_ = 0;
    int32_t orders_ = 0;
    std::vector<Line> lines_;
    std::vector<Record> records_;
    uint32_t observed_ = 0, settle_ = 0, bad_ = 0, stage_settle_ = 0;
    bool seed_pending_ = false;
    uint64_t recovering_since_ = 0;
    bool recentred_ = false;
    bool wide_ = true;
    float suppression_ = 0.0f;
    double residual_ = 0.0;
    uint32_t cancelled_ = 0;
};

struct Split {
    uint64_t start = 0;
    uint32_t count = 0;
    std::vector<std::pair<uint64_t, uint32_t>> rows;  // (tracker id, lines)
};

}  // namespace

struct rvx_comb_bank {
    int64_t rate = 0;
    std::vector<Tracker> trackers;
    std::vector<Split> splits;
    uint64_t next_id = 1;
    // A spur that went (lost its line, withdrawn or dropped) is not started
    // again near the same offset until `until`: a spur that will not hold
    // would otherwise restart on every detection.
    struct HoldOff {
        double offset_hz;
        uint64_t until;
    };
    std::vector<HoldOff> hold_offs;
    uint64_t now = 0;     // the latest block planned
    uint64_t starts = 0;  // trackers started, for Tracker::order

    void note_ended_spurs() {
        const uint64_t span = uint64_t(RVX_COMB_SPUR_HOLD_OFF_S * double(rate));
        for (Tracker& t : trackers)
            if (t.ended_spur()) {
                hold_offs.push_back({t.detected_offset_hz(), now + span});
                t.note_end();
            }
    }
    bool held_off(double f_hz) const {
        for (const HoldOff& h : hold_offs)
            if (std::abs(h.offset_hz - f_hz) < RVX_COMB_SAME_LINE_HZ) return true;
        return false;
    }
    template <typename Pred>
    void stop_spurs_if(Pred drop) {
        for (Tracker& t : trackers)
            if (t.spur() && t.active() && drop(t.detected_offset_hz())) t.stop(RVX_COMB_IDLE);
        note_ended_spurs();
    }

    uint32_t lines_held() const {
        uint32_t n = 0;
        for (const Tracker& t : trackers)
            if (t.active()) n += t.lines();
        return n;
    }
    Tracker* free_slot(bool spur) {
        for (Tracker& t : trackers)
            if (!t.active() && t.spur() == spur) return &t;
        uint32_t count = 0;
        for (const Tracker& t : trackers) count += t.spur() == spur ? 1u : 0u;
        if (count >= (spur ? RVX_COMB_MAX_SPURS : RVX_COMB_MAX_COMBS)) return nullptr;
        trackers.emplace_back(rate, next_id++);
        return &trackers.back();
    }
};

extern "C" {

int32_t rvx_comb_bank_create(int64_t rate, rvx_comb_bank** out) {
    static const char* who = "rvx_comb_bank_create";
    if (!out) return rvx::fail(RVX_EINVAL, "%s: null output", who);
    *out = nullptr;
    if (rate <= 0) return rvx::fail(RVX_EINVAL, "%s: sample rate %lld", who, (long long)rate);
    auto* b = new (std::nothrow) rvx_comb_bank();
    if (!b) return rvx::fail(RVX_ENOMEM, "%s: out of memory", who);
    b->rate = rate;
    *out = b;
    return RVX_OK;
}

void rvx_comb_bank_destroy(rvx_comb_bank* bank) { delete bank; }

int32_t rvx_comb_bank_reset(rvx_comb_bank* bank) {
    if (!bank) return rvx::fail(RVX_EINVAL, "rvx_comb_bank_reset: null bank");
    bank->trackers.clear();
    bank->splits.clear();
    bank->hold_offs.clear();
    return RVX_OK;
}

int32_t rvx_comb_bank_detected(rvx_comb_bank* bank, const rvx_comb_estimate* estimates,
                               uint32_t n) {
    static const char* who = "rvx_comb_bank_detected";
    if (!bank || (n && !estimates)) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    // Held by a running comb tracker (spurs are not counted: one line of a
    // comb's is not the comb); asked again before each start, so a comb this
    // call started counts too.
    auto held = [&](const rvx_comb_estimate& e) {
        for (const Tracker& t : bank->trackers)
            if (!t.spur() && t.active() &&
                (t.agrees(e) || t.overlap(e, RVX_COMB_SAME_LINE_HZ) >= RVX_COMB_OVERLAP_FRACTION))
                return true;
        return false;
    };
    auto lines_of = [](const rvx_comb_estimate& e) { return uint32_t(e.k_hi - e.k_lo + 1); };
    // What a comb asks for: its lines, each with its sidebands.
    auto need = [&](const rvx_comb_estimate& e) {
        const uint32_t orders = e.ripple_hz > 0.0 ? uint32_t(std::max(e.sideband_orders, 0)) : 0u;
        return lines_of(e) * (2 * orders + 1);
    };
    // A detection of a held comb widens the range that comb's pass grows into,
    // so a pass started on part of a comb reaches the rest of it.
    for (uint32_t i = 0; i < n; ++i) {
        const rvx_comb_estimate& e = estimates[i];
        if (e.spur || e.k_hi < e.k_lo || !(e.spacing_hz > 0) || !std::isfinite(e.offset_hz))
            continue;
        for (Tracker& t : bank->trackers)
            if (!t.spur() && t.active() &&
                t.overlap(e, RVX_COMB_SAME_LINE_HZ) >= RVX_COMB_OVERLAP_FRACTION) {
                t.extend_range(e);
                t.recentre(e, false);
                break;
            }
    }
    // A comb on a recovering pass's spacing that overlaps none of its lines is
    // that comb, moved while the pass could not see it: the pass takes it
    // rather than a second pass starting beside a dying one.
    auto moved = [&](const rvx_comb_estimate& e) {
        for (Tracker& t : bank->trackers)
            if (!t.spur() && t.lock() == RVX_COMB_RECOVERING && t.same_spacing(e) &&
                t.recentre(e, true)) {
                t.extend_range(e);
                return true;
            }
        return false;
    };
    std::vector<const rvx_comb_estimate*> fresh;
    for (uint32_t i = 0; i < n; ++i) {
        const rvx_comb_estimate& e = estimates[i];
        if (e.spur || e.k_hi < e.k_lo || lines_of(e) < RVX_COMB_MIN_LINES ||
            !(e.spacing_hz > 0) || !std::isfinite(e.offset_hz))
            continue;
        if (!held(e) && !moved(e)) fresh.push_back(&e);
    }
    if (fresh.empty()) return RVX_OK;
    {
        // Passes that grew past a fair share give the excess back when new
        // combs want slots that are not free: growth only borrows.
        uint32_t held_slots = 0, running = 0;
        for (const Tracker& t : bank->trackers)
            if (!t.spur() && t.active()) {
                held_slots += t.lines();
                ++running;
            }
        const uint32_t incoming = std::min<uint32_t>(
            uint32_t(fresh.size()), RVX_COMB_MAX_COMBS > running ? RVX_COMB_MAX_COMBS - running : 0u);
        if (incoming > 0) {
            const uint32_t fair = RVX_COMB_MAX_LINES / (running + incoming);
            if (RVX_COMB_MAX_LINES - std::min(held_slots, uint32_t(RVX_COMB_MAX_LINES)) < fair * incoming)
                for (Tracker& t : bank->trackers)
                    if (!t.spur() && t.active() && t.lines() > fair) t.shrink_to(fair);
        }
    }
    for (size_t i = 0; i < fresh.size(); ++i) {
        // Combs have first claim: lines spurs hold count as free here, and the
        // newest spurs give back what a comb that starts takes.
        uint32_t used = 0, combs = 0;
        for (const Tracker& t : bank->trackers)
            if (!t.spur() && t.active()) {
                used += t.lines();
                ++combs;
            }
        if (combs >= RVX_COMB_MAX_COMBS || used >= RVX_COMB_MAX_LINES) break;
        // Strongest first takes what it needs if the rest still fit, else an
        // even share of what is free.
        const uint32_t free = RVX_COMB_MAX_LINES - used;
        uint32_t rest = 0;
        for (size_t j = i + 1; j < fresh.size(); ++j) rest += need(*fresh[j]);
        const uint32_t even = free / uint32_t(fresh.size() - i);
        const uint32_t room = rest < free ? free - rest : 0u;
        const uint32_t share = std::min(need(*fresh[i]), std::max(even, room));
        if (share < RVX_COMB_MIN_LINES) break;
        if (held(*fresh[i])) continue;
        Tracker* t = bank->free_slot(false);
        if (!t) break;
        t->start(*fresh[i], share);
        t->order = ++bank->starts;
        uint32_t total = bank->lines_held();
        while (total > RVX_COMB_MAX_LINES) {
            Tracker* newest = nullptr;
            for (Tracker& s : bank->trackers)
                if (s.spur() && s.active() && (!newest || s.order > newest->order)) newest = &s;
            if (!newest) break;
            total -= newest->lines();
            newest->stop(RVX_COMB_IDLE);
        }
        bank->note_ended_spurs();
    }
    return RVX_OK;
}

int32_t rvx_comb_bank_spurs(rvx_comb_bank* bank, const rvx_comb_estimate* estimates, uint32_t n) {
    static const char* who = "rvx_comb_bank_spurs";
    if (!bank || (n && !estimates)) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    bank->note_ended_spurs();
    for (uint32_t i = 0; i < n; ++i) {
        rvx_comb_estimate e = estimates[i];
        e.spur = 1;
        e.k_lo = e.k_hi = 0;
        e.spacing_hz = 0.0;
        if (bank->held_off(e.offset_hz)) continue;  // went a moment ago
        bool held = false;
        for (const Tracker& t : bank->trackers)
            held = held || (t.active() && t.holds(e.offset_hz, RVX_COMB_SAME_LINE_HZ));
        if (held || bank->lines_held() >= RVX_COMB_MAX_LINES) continue;
        Tracker* t = bank->free_slot(true);
        if (!t) break;
        t->start(e, 1);
        t->order = ++bank->starts;
    }
    return RVX_OK;
}

int32_t rvx_comb_bank_drop_spurs(rvx_comb_bank* bank) {
    if (!bank) return rvx::fail(RVX_EINVAL, "rvx_comb_bank_drop_spurs: null bank");
    bank->stop_spurs_if([](double) { return true; });
    return RVX_OK;
}

int32_t rvx_comb_bank_withdraw_spurs(rvx_comb_bank* bank, const double* frequencies_hz,
                                     uint32_t n) {
    if (!bank || (n && !frequencies_hz))
        return rvx::fail(RVX_EINVAL, "rvx_comb_bank_withdraw_spurs: null argument");
    bank->stop_spurs_if([&](double at) {
        for (uint32_t i = 0; i < n; ++i)
            if (std::abs(frequencies_hz[i] - at) < RVX_COMB_SAME_LINE_HZ) return true;
        return false;
    });
    return RVX_OK;
}

int32_t rvx_comb_bank_drop_spurs_in(rvx_comb_bank* bank, const rvx_spur_band* bands, uint32_t n) {
    if (!bank || (n && !bands))
        return rvx::fail(RVX_EINVAL, "rvx_comb_bank_drop_spurs_in: null argument");
    bank->stop_spurs_if([&](double at) {
        for (uint32_t i = 0; i < n; ++i)
            if (at >= bands[i].low_hz && at <= bands[i].high_hz) return true;
        return false;
    });
    return RVX_OK;
}

int32_t rvx_comb_bank_plan(rvx_comb_bank* bank, uint64_t start, uint32_t count,
                           rvx_comb_line* lines, uint32_t max_lines, uint32_t* n_lines,
                           uint32_t* sub_blocks, int32_t* subtract) {
    static const char* who = "rvx_comb_bank_plan";
    if (!bank || !n_lines || !subtract || (max_lines && !lines))
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (count < 2) return rvx::fail(RVX_EINVAL, "%s: a block of %u samples", who, count);
    bank->now = start;
    std::erase_if(bank->hold_offs, [&](const rvx_comb_bank::HoldOff& h) { return h.until <= start; });
    bank->note_ended_spurs();  // a spur that lost its line since the last block
    // Slots nobody holds go to healthy locked passes with more of their comb
    // to reach, a few lines each per block; spurs keep theirs.
    {
        uint32_t held_slots = bank->lines_held();
        for (Tracker& t : bank->trackers) {
            if (held_slots >= RVX_COMB_MAX_LINES) break;
            const uint32_t want = std::min(t.wants_lines(), RVX_COMB_MAX_LINES - held_slots);
            if (want) held_slots += t.grow(want);
        }
    }
    std::vector<rvx_comb_line> all;
    bool sub = false;
    Split split;
    split.start = start;
    split.count = count;
    uint32_t subs = 1;
    for (Tracker& t : bank->trackers) {
        const size_t before = all.size();
        t.plan(start, count, all, sub, subs);
        if (all.size() > before) split.rows.push_back({t.id(), uint32_t(all.size() - before)});
    }
    if (all.size() > max_lines)
        return rvx::fail(RVX_EINVAL, "%s: the plan holds %zu lines, the buffer %u", who,
                         all.size(), max_lines);
    std::copy(all.begin(), all.end(), lines);
    *n_lines = uint32_t(all.size());
    *subtract = sub ? 1 : 0;
    if (sub_blocks) *sub_blocks = all.empty() ? kPlanSubBlocks : subs;
    bank->splits.push_back(std::move(split));
    while (bank->splits.size() > 64) bank->splits.erase(bank->splits.begin());
    return RVX_OK;
}

int32_t rvx_comb_bank_observe(rvx_comb_bank* bank, uint64_t start, uint32_t count,
                              uint32_t sub_blocks, const float* projections, uint64_t n_floats) {
    static const char* who = "rvx_comb_bank_observe";
    if (!bank || (n_floats && !projections)) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    size_t at = bank->splits.size();
    for (size_t i = 0; i < bank->splits.size(); ++i)
        if (bank->splits[i].start == start && bank->splits[i].count == count) at = i;
    if (at == bank->splits.size()) return RVX_OK;  // a block nobody planned lines for
    Split sp = std::move(bank->splits[at]);
    bank->splits.erase(bank->splits.begin(), bank->splits.begin() + ptrdiff_t(at) + 1);
    uint64_t lines = 0;
    for (const auto& r : sp.rows) lines += r.second;
    if (sub_blocks == 0 || n_floats < 2 * lines * sub_blocks)
        return rvx::fail(RVX_EINVAL, "%s: %llu floats for %llu lines x %u sub-blocks", who,
                         (unsigned long long)n_floats, (unsigned long long)lines, sub_blocks);
    const auto* p = reinterpret_cast<const std::complex<float>*>(projections);
    uint64_t row = 0;
    for (const auto& r : sp.rows) {
        for (Tracker& t : bank->trackers)
            if (t.id() == r.first) t.observe(start, count, sub_blocks, p + row * sub_blocks, r.second);
        row += r.second;
    }
    return RVX_OK;
}

int32_t rvx_comb_bank_status(const rvx_comb_bank* bank, rvx_comb_status* out) {
    if (!bank || !out) return rvx::fail(RVX_EINVAL, "rvx_comb_bank_status: null argument");
    *out = rvx_comb_status{};
    double supp = 0.0;
    uint32_t best_lines = 0;
    bool any_acquiring = false, any_recovering = false, any_lost = false;
    for (const Tracker& t : bank->trackers) {
        rvx_comb_status s{};
        t.status(s);
        if (t.spur()) {
            out->spurs_cancelled += s.lock == RVX_COMB_LOCKED ? 1u : 0u;
            continue;
        }
        out->combs_active += t.active() ? 1u : 0u;
        any_acquiring |= s.lock == RVX_COMB_ACQUIRING;
        any_recovering |= s.lock == RVX_COMB_RECOVERING;
        any_lost |= s.lock == RVX_COMB_LOST;
        if (s.lock == RVX_COMB_LOCKED) {
            out->lock = RVX_COMB_LOCKED;
            out->lines_cancelled += s.lines_cancelled;
            supp += double(s.suppression_db) * s.lines_cancelled;
            if (s.lines_cancelled >= best_lines) {
                best_lines = s.lines_cancelled;
                out->comb = s.comb;
                out->spacing_drift_hz_per_s = s.spacing_drift_hz_per_s;
                out->residual_rad = s.residual_rad;
            }
        } else if (out->lock != RVX_COMB_LOCKED && best_lines == 0) {
            out->comb = s.comb;
            out->residual_rad = s.residual_rad;
        }
    }
    if (out->lock != RVX_COMB_LOCKED)
        out->lock = any_recovering ? RVX_COMB_RECOVERING
                    : any_acquiring ? RVX_COMB_ACQUIRING
                    : any_lost      ? RVX_COMB_LOST
                                    : RVX_COMB_IDLE;
    if (out->lines_cancelled) out->suppression_db = float(supp / out->lines_cancelled);
    return RVX_OK;
}

int32_t rvx_comb_bank_active(const rvx_comb_bank* bank, rvx_comb_estimate* out, uint32_t max_out,
                             uint32_t* n_out) {
    if (!bank || !n_out || (max_out && !out))
        return rvx::fail(RVX_EINVAL, "rvx_comb_bank_active: null argument");
    uint32_t n = 0;
    for (const Tracker& t : bank->trackers)
        if (t.active() && !t.spur()) {
            if (n < max_out) out[n] = t.estimate();
            ++n;
        }
    *n_out = std::min(n, max_out);
    return RVX_OK;
}

}  // extern "C"
