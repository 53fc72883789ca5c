// The plain detector's decision on the GPU: core/detect/detector.cpp
// estimate_noise_floor (:794-876) and find_candidates (:879-1120).
//
//   knot_floor_kernel  one workgroup per knot: the two percentiles by radix
//                      select on the doubles' bit patterns (non-negative
//                      doubles order as unsigned integers, so rank r of the
//                      window is exact, as nth_element's), then the ascending
//                      excision sum by one wave, each value added in index
//                      order as the reference's loop adds it;
//   floor_interp_kernel per bin, the knots' linear interpolation;
//   prefix_kernel      the three running sums (excess, floor, spread) in the
//                      reference's sequential order, one wave per sum;
//   ladder_kernel      every (rung, position): the deflection, the threshold,
//                      the local-maximum and plateau-middle rules; peaks are
//                      appended to a list whose order does not matter, since
//                      the host ranks them by the reference's total order.
// The host keeps the strongest max_peaks (the best-N set the reference's heap
// keeps) and grows them strongest first, exactly as find_candidates does.
//
// Doubles use only +, -, *, / and comparisons on the device (IEEE, correctly
// rounded); square roots, exp2 and the normal quantile are computed on the
// host, where the reference computes them.

#pragma clang fp contract(off)

#include <algorithm>
#include <bit>
#include <cmath>
#include <vector>

#include "common.h"

namespace {

constexpr double kReferenceBandwidthHz = 2500.0;      // detector.h:128
constexpr float kDbToLog2 = 0.332192809488736234787F;  // detector.cpp:26
constexpr float kSpectrumPowerFloor = 1.0e-20F;        // spectrum_reference.h:104

struct KnotWindow {
    uint32_t low, high, low_rank, high_rank;
};

struct Peak {
    uint32_t start, width;
    double deflection;
};

double normal_quantile(double p) {  // Acklam, detector.cpp:46-70
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                                -2.759285104469687e+02, 1.383577518672690e+02,
                                -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                                -1.556989798598866e+02, 6.680131188771972e+01,
                                -1.328068155288572e+01};
    static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                                -2.400758277161838e+00, -2.549732539343734e+00,
                                4.374664141464968e+00,  2.938163982698783e+00};
    static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                                2.445134137142996e+00, 3.754408661907416e+00};
    const double split = 0.02425;
    if (!(p > 0.0 && p < 1.0)) return 0.0;
    if (p < split || p > 1.0 - split) {
        const double q = std::sqrt(-2.0 * std::log(p < split ? p : 1.0 - p));
        const double v = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
                         ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
        return p < split ? v : -v;
    }
    const double q = p - 0.5, r = q * q;
    return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
           (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
}

__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }

__global__ void __launch_bounds__(256)
knot_floor_kernel(const double* __restrict__ avg, const KnotWindow* __restrict__ knots,
                  double quantile_span, double high_q, double excision_sigma,
                  double* __restrict__ knot_floor, double* __restrict__ knot_sigma) {
    __shared__ uint32_t hist[2][256];
    __shared__ unsigned long long prefix[2];
    __shared__ uint32_t rank[2];
    const uint32_t tid = threadIdx.x;
    const KnotWindow kw = knots[blockIdx.x];
    const uint32_t count = kw.high - kw.low;
    const double* w = avg + kw.low;
    if (tid < 2) {
        prefix[tid] = 0;
        rank[tid] = tid == 0 ? kw.low_rank : kw.high_rank;
    }
    unsigned long long mask = 0;
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        for (uint32_t i = tid; i < 512; i += 256) hist[i >> 8][i & 255] = 0;
        __syncthreads();
        const unsigned long long p0 = prefix[0], p1 = prefix[1];
        for (uint32_t i = tid; i < count; i += 256) {
            const unsigned long long v = (unsigned long long)__double_as_longlong(w[i]);
            const uint32_t bucket = uint32_t(v >> shift) & 255u;
            if ((v & mask) == p0) atomicAdd(&hist[0][bucket], 1u);
            if ((v & mask) == p1) atomicAdd(&hist[1][bucket], 1u);
        }
        __syncthreads();
        if (tid < 2) {
            uint32_t below = 0, sel = 255;
            const uint32_t r = rank[tid];
            for (uint32_t bucket = 0; bucket < 256; ++bucket) {
                const uint32_t h = hist[tid][bucket];
                if (below + h > r) {
                    sel = bucket;
                    break;
                }
                below += h;
            }
            rank[tid] = r - below;
            prefix[tid] |= (unsigned long long)sel << shift;
        }
        mask |= 255ull << shift;
        __syncthreads();
    }
    if (tid >= 64) return;  // one wave sums
    const double lp = __longlong_as_double((long long)prefix[0]);
    const double hp = kw.high_rank > kw.low_rank ? __longlong_as_double((long long)prefix[1]) : lp;
    const double deviation = dmax(0.0, (hp - lp) / quantile_span);
    const double spread = high_q * deviation;
    const double level = hp - spread;
    const double reach = excision_sigma * deviation;
    const double ceiling = level + reach;
    double sum = 0.0;
    uint32_t kept = 0;
    for (uint32_t base = 0; base < count; base += 64) {
        const uint32_t i = base + tid;
        const double v = i < count ? w[i] : 0.0;
        unsigned long long in = __ballot(i < count && v <= ceiling);
        kept += __popcll(in);
        while (in) {  // ascending index order, as the reference's loop
            const int j = __ffsll((long long)in) - 1;
            in &= in - 1;
            sum = sum + __shfl(v, j);
        }
    }
    if (tid == 0) {
        const double estimate = kept > 0 ? sum / double(kept) : level;
        knot_floor[blockIdx.x] = dmax(estimate, double(kSpectrumPowerFloor));
        knot_sigma[blockIdx.x] = deviation;
    }
}

__global__ void floor_interp_kernel(const double* __restrict__ pos, const double* __restrict__ kf,
                                    const double* __restrict__ ks, uint32_t knots, uint32_t bins,
                                    double* __restrict__ floor_, double* __restrict__ sigma_) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bins) return;
    if (knots == 1) {
        floor_[i] = kf[0];
        sigma_[i] = ks[0];
        return;
    }
    const double position = double(i);
    uint32_t k = 0;
    while (k + 2 < knots && pos[k + 1] < position) ++k;
    const double span = pos[k + 1] - pos[k];
    double t = 0.0;
    if (span > 0.0) {
        t = (position - pos[k]) / span;
        t = t < 0.0 ? 0.0 : (1.0 < t ? 1.0 : t);
    }
    const double df = kf[k + 1] - kf[k], ds = ks[k + 1] - ks[k];
    const double tf = t * df, ts = t * ds;
    floor_[i] = kf[k] + tf;
    sigma_[i] = ks[k] + ts;
}

// Wave a of three: running sum of excess (a = 0), floor (1) or spread (2),
// out[0] = 0 and out[i + 1] = out[i] + x[i], one addition at a time.
__global__ void __launch_bounds__(192)
prefix_kernel(const double* __restrict__ avg, const double* __restrict__ floor_,
              const double* __restrict__ sigma_, uint32_t bins, double* __restrict__ ec,
              double* __restrict__ fc, double* __restrict__ sc) {
    const uint32_t a = threadIdx.x / 64, lane = threadIdx.x % 64;
    double* out = a == 0 ? ec : (a == 1 ? fc : sc);
    if (lane == 0) out[0] = 0.0;
    double acc = 0.0;
    for (uint32_t base = 0; base < bins; base += 64) {
        const uint32_t i = base + lane;
        double x = 0.0;
        if (i < bins) x = a == 0 ? avg[i] - floor_[i] : (a == 1 ? floor_[i] : sigma_[i]);
        double mine = 0.0;
        const uint32_t n = min(64u, bins - base);
        for (uint32_t j = 0; j < n; ++j) {
            acc = acc + __shfl(x, int(j));
            if (lane == j) mine = acc;
        }
        if (i < bins) out[i + 1] = mine;
    }
}

__device__ __forceinline__ float deflection(const double* ec, const double* fc, uint32_t s,
                                            uint32_t w, double root) {
    const double excess = ec[s + w] - ec[s];
    const double summed_floor = fc[s + w] - fc[s];
    if (!(summed_floor > 0.0)) return 0.0f;
    const double scaled = excess * root;
    return float(scaled / summed_floor);
}

struct Rung {
    uint32_t width;
    uint32_t reserved;
    double root;
    double threshold;
};

__global__ void __launch_bounds__(256)
ladder_kernel(const double* __restrict__ ec, const double* __restrict__ fc,
              const Rung* __restrict__ rungs, uint32_t bins, const int32_t* __restrict__ allow,
              Peak* __restrict__ peaks, uint32_t capacity, uint32_t* __restrict__ count) {
    const Rung r = rungs[blockIdx.y];
    const uint32_t w = r.width;
    const uint32_t last = bins - w;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > last) return;
    const float here = deflection(ec, fc, s, w, r.root);
    if (!(double(here) > r.threshold)) return;
    if (allow && allow[s + w] - allow[s] == 0) return;
    if (s > 0) {
        const float before = deflection(ec, fc, s - 1, w, r.root);
        if (before > here || before == here) return;  // not a maximum, or not the plateau's start
    }
    if (s < last && deflection(ec, fc, s + 1, w, r.root) > here) return;
    uint32_t end = s;
    while (end < last && deflection(ec, fc, end + 1, w, r.root) == here) ++end;
    const uint32_t idx = atomicAdd(count, 1u);
    if (idx < capacity) peaks[idx] = Peak{s + (end - s) / 2, w, double(here)};
}

__global__ void zero_counter(uint32_t* c) { *c = 0u; }

bool stronger(const Peak& a, const Peak& b) {
    if (a.deflection != b.deflection) return a.deflection > b.deflection;
    if (a.start != b.start) return a.start < b.start;
    return a.width < b.width;
}

}  // namespace

struct rvx_detector {
    rvx_ctx* ctx = nullptr;
    rvx_detect_config cfg{};
    uint32_t knots = 0;
    double quantile_span = 0, high_q = 0;
    std::vector<double> pos_h;
    std::vector<Rung> rungs_h;
    uint32_t budget = 0, capacity = 0;
    KnotWindow* windows = nullptr;
    double* pos = nullptr;
    double *kf = nullptr, *ks = nullptr, *floor_ = nullptr, *sigma_ = nullptr;
    double *ec = nullptr, *fc = nullptr, *sc = nullptr;
    Rung* rungs = nullptr;
    Peak* peaks = nullptr;
    uint32_t* counter = nullptr;
    std::vector<double> avg_h, floor_h, fc_h, sc_h;
    std::vector<Peak> peaks_h;

    ~rvx_detector() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (void* p : {(void*)windows, (void*)pos, (void*)kf, (void*)ks, (void*)floor_,
                        (void*)sigma_, (void*)ec, (void*)fc, (void*)sc, (void*)rungs,
                        (void*)peaks, (void*)counter})
            if (p) (void)hipFree(p);
    }
};

extern "C" {

int32_t rvx_detector_create(rvx_ctx* ctx, const rvx_detect_config* config, rvx_detector** out) {
    static const char* who = "rvx_detector_create";
    RVX_ENTER(ctx, who);
    if (!config || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    *out = nullptr;
    const rvx_detect_config c = *config;
    if (c.bins < 2 || c.bins > (1u << 24) || c.noise_knots == 0 || c.noise_window_knots == 0 ||
        !(c.bin_width_hz > 0.0))
        return rvx::fail(RVX_EINVAL, "%s: invalid geometry (bins %u, knots %u)", who, c.bins,
                         c.noise_knots);
    if (!(c.noise_percentile >= 0.001 && c.noise_percentile <= 0.998 &&
          c.noise_percentile_high > c.noise_percentile && c.noise_percentile_high <= 0.999))
        return rvx::fail(RVX_EINVAL, "%s: percentiles %g, %g", who, c.noise_percentile,
                         c.noise_percentile_high);
    auto d = new rvx_detector();
    d->ctx = ctx;
    d->cfg = c;
    const size_t bins = c.bins;
    // Detector::create's derived constants (detector.cpp:300-385)
    const double threshold_linear = std::exp2(c.detection_threshold_db * double(kDbToLog2));
    const double low_q = normal_quantile(c.noise_percentile);
    d->high_q = normal_quantile(c.noise_percentile_high);
    d->quantile_span = d->high_q - low_q;
    const size_t knots = std::min<size_t>(c.noise_knots, bins);
    d->knots = uint32_t(knots);
    const size_t spacing = std::max<size_t>(1, bins / knots);
    const size_t window_bins = std::min<size_t>(bins, spacing * c.noise_window_knots);
    const size_t half = window_bins / 2;
    d->pos_h.resize(knots);
    std::vector<KnotWindow> win(knots);
    for (size_t k = 0; k < knots; ++k) {
        d->pos_h[k] = (double(2 * k + 1) * double(bins)) / double(2 * knots);
        const size_t centre = size_t(d->pos_h[k]);
        const size_t low = centre > half ? centre - half : 0;
        const size_t high = std::min(bins, centre + half + 1);
        const size_t n = high - low;
        size_t lr = size_t(c.noise_percentile * double(n));
        size_t hr = size_t(c.noise_percentile_high * double(n));
        lr = std::min(lr, n - 1);
        hr = std::min(std::max(hr, lr), n - 1);
        win[k] = KnotWindow{uint32_t(low), uint32_t(high), uint32_t(lr), uint32_t(hr)};
    }
    uint32_t widest = c.max_sum_bins;
    if (widest == 0) widest = uint32_t(std::bit_floor(std::max<size_t>(1, bins / 8)));
    widest = std::min<uint32_t>(widest, uint32_t(bins));
    uint64_t capacity = 0;
    for (uint32_t w = 1; w <= widest; w *= 2) {
        const double root = std::sqrt(double(w));
        const double threshold = threshold_linear * kReferenceBandwidthHz / (root * c.bin_width_hz);
        d->rungs_h.push_back(Rung{w, 0, root, threshold});
        capacity += (bins - w + 2) / 2;  // strict local maxima: at most every other position
    }
    d->budget = c.max_peaks ? c.max_peaks : uint32_t(bins);
    d->capacity = uint32_t(std::min<uint64_t>(capacity, 0xFFFFFFFFull));
    hipError_t err = hipSuccess;
    auto alloc = [&](auto** p, size_t bytes) {
        if (err == hipSuccess) err = hipMalloc((void**)p, bytes);
    };
    alloc(&d->windows, knots * sizeof(KnotWindow));
    alloc(&d->pos, knots * sizeof(double));
    alloc(&d->kf, knots * sizeof(double));
    alloc(&d->ks, knots * sizeof(double));
    alloc(&d->floor_, bins * sizeof(double));
    alloc(&d->sigma_, bins * sizeof(double));
    alloc(&d->ec, (bins + 1) * sizeof(double));
    alloc(&d->fc, (bins + 1) * sizeof(double));
    alloc(&d->sc, (bins + 1) * sizeof(double));
    alloc(&d->rungs, d->rungs_h.size() * sizeof(Rung));
    alloc(&d->peaks, size_t(d->capacity) * sizeof(Peak));
    alloc(&d->counter, sizeof(uint32_t));
    if (err == hipSuccess)
        err = hipMemcpy(d->windows, win.data(), knots * sizeof(KnotWindow), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(d->pos, d->pos_h.data(), knots * sizeof(double), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(d->rungs, d->rungs_h.data(), d->rungs_h.size() * sizeof(Rung),
                        hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        delete d;
        return rvx::fail(RVX_ENOMEM, "%s: %s", who, hipGetErrorString(err));
    }
    d->avg_h.resize(bins);
    d->floor_h.resize(bins);
    d->fc_h.resize(bins + 1);
    d->sc_h.resize(bins + 1);
    *out = d;
    return RVX_OK;
}

void rvx_detector_destroy(rvx_detector* detector) { delete detector; }

const double* rvx_detector_floor(rvx_detector* d) { return d ? d->floor_ : nullptr; }
const double* rvx_detector_sigma(rvx_detector* d) { return d ? d->sigma_ : nullptr; }

int32_t rvx_detect(rvx_detector* d, const double* average, const int32_t* allow_prefix,
                   rvx_detect_band* bands, uint32_t max_bands, uint32_t* n_bands,
                   uint32_t* n_peaks) {
    static const char* who = "rvx_detect";
    if (!d) return rvx::fail(RVX_EINVAL, "%s: null detector", who);
    RVX_ENTER(d->ctx, who);
    if (!average || !n_bands || (max_bands && !bands))
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    const hipStream_t s = d->ctx->stream;
    const uint32_t bins = d->cfg.bins;
    knot_floor_kernel<<<dim3(d->knots), dim3(256), 0, s>>>(
        average, d->windows, d->quantile_span, d->high_q, d->cfg.noise_excision_sigma, d->kf,
        d->ks);
    floor_interp_kernel<<<dim3((bins + 255) / 256), dim3(256), 0, s>>>(d->pos, d->kf, d->ks,
                                                                       d->knots, bins, d->floor_,
                                                                       d->sigma_);
    prefix_kernel<<<dim3(1), dim3(192), 0, s>>>(average, d->floor_, d->sigma_, bins, d->ec, d->fc,
                                                d->sc);
    zero_counter<<<dim3(1), dim3(1), 0, s>>>(d->counter);
    ladder_kernel<<<dim3((bins + 255) / 256, uint32_t(d->rungs_h.size())), dim3(256), 0, s>>>(
        d->ec, d->fc, d->rungs, bins, allow_prefix, d->peaks, d->capacity, d->counter);
    if (int32_t rc = rvx::check_launch(who)) return rc;
    uint32_t found = 0;
    RVX_TRY_HIP(hipMemcpyAsync(&found, d->counter, sizeof(uint32_t), hipMemcpyDeviceToHost, s), who);
    RVX_TRY_HIP(hipMemcpyAsync(d->avg_h.data(), average, bins * sizeof(double),
                               hipMemcpyDeviceToHost, s),
                who);
    RVX_TRY_HIP(hipMemcpyAsync(d->floor_h.data(), d->floor_, bins * sizeof(double),
                               hipMemcpyDeviceToHost, s),
                who);
    RVX_TRY_HIP(hipMemcpyAsync(d->fc_h.data(), d->fc, (bins + 1) * sizeof(double),
                               hipMemcpyDeviceToHost, s),
                who);
    RVX_TRY_HIP(hipMemcpyAsync(d->sc_h.data(), d->sc, (bins + 1) * sizeof(double),
                               hipMemcpyDeviceToHost, s),
                who);
    RVX_TRY_HIP(hipStreamSynchronize(s), who);
    if (found > d->capacity)
        return rvx::fail(RVX_EDEVICE, "%s: %u peaks overflowed a list of %u", who, found,
                         d->capacity);
    d->peaks_h.resize(found);
    if (found)
        RVX_TRY_HIP(hipMemcpy(d->peaks_h.data(), d->peaks, found * sizeof(Peak),
                              hipMemcpyDeviceToHost),
                    who);
    if (n_peaks) *n_peaks = found;

    // The best max_peaks by the reference's order, strongest first.
    auto& peaks = d->peaks_h;
    if (peaks.size() > d->budget) {
        std::nth_element(peaks.begin(), peaks.begin() + d->budget, peaks.end(), stronger);
        peaks.resize(d->budget);
    }
    std::sort(peaks.begin(), peaks.end(), stronger);

    // Greedy growth (find_candidates :1060-1120): each winner grows onto its
    // shoulders within its seed width and the accepted neighbours, stepping
    // over gaps shorter than split_gap_bins; a peak overlapping an accepted
    // band is dropped.
    const double* avg = d->avg_h.data();
    const double* fl = d->floor_h.data();
    const double* fcum = d->fc_h.data();
    const double* scum = d->sc_h.data();
    const size_t gap = d->cfg.split_gap_bins;
    std::vector<Peak> accepted;
    auto above = [&](size_t i, double level) { return avg[i] - fl[i] > level; };
    for (const Peak& p : peaks) {
        size_t lo = p.start, hi = size_t(p.start) + p.width;
        auto at = std::lower_bound(accepted.begin(), accepted.end(), lo,
                                   [](const Peak& held, size_t v) { return held.start < v; });
        if (at != accepted.end() && size_t(at->start) < hi) continue;
        if (at != accepted.begin() && size_t((at - 1)->start) + (at - 1)->width > lo) continue;
        const double seed_floor = fcum[hi] - fcum[lo];
        const double seed_sigma = scum[hi] - scum[lo];
        const double level = std::max(d->cfg.edge_floor_fraction * seed_floor,
                                      d->cfg.edge_floor_sigma * seed_sigma) /
                             double(p.width);
        const size_t reach = p.width;
        const size_t low_limit =
            at == accepted.begin() ? 0 : size_t((at - 1)->start) + (at - 1)->width;
        const size_t high_limit = at == accepted.end() ? size_t(bins) : size_t(at->start);
        const size_t low_stop = std::max(low_limit, lo > reach ? lo - reach : 0);
        while (lo > low_stop) {
            if (above(lo - 1, level)) {
                --lo;
                continue;
            }
            const size_t limit = std::min(gap, lo - low_stop);
            size_t back = 2;
            while (back <= limit && !above(lo - back, level)) ++back;
            if (back > limit) break;
            lo -= back;
        }
        const size_t high_stop = std::min(high_limit, std::min(size_t(bins), hi + reach));
        while (hi < high_stop) {
            if (above(hi, level)) {
                ++hi;
                continue;
            }
            const size_t limit = std::min(gap, high_stop - hi);
            size_t ahead = 1;
            while (ahead < limit && !above(hi + ahead, level)) ++ahead;
            if (ahead >= limit) break;
            hi += ahead + 1;
        }
        accepted.insert(at, Peak{uint32_t(lo), uint32_t(hi - lo), p.deflection});
    }
    *n_bands = uint32_t(accepted.size());
    if (accepted.size() > max_bands)
        return rvx::fail(RVX_EINVAL, "%s: %zu bands, room for %u", who, accepted.size(), max_bands);
    for (size_t i = 0; i < accepted.size(); ++i)
        bands[i] = rvx_detect_band{accepted[i].start, accepted[i].width, accepted[i].deflection};
    return RVX_OK;
}

}  // extern "C"
