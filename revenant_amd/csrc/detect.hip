// The plain detector's decision on the GPU: core/detect/detector.cpp
// estimate_noise_floor (:794-876) and find_candidates (:879-1120).
//
//   knot_floor_kernel  one workgroup per knot: the two percentiles by radix
//                      select on the doubles' bit patterns (non-negative
//                      doubles order as unsigned integers, so rank r of the
//                      window is exact, as nth_element's), then the excision
//                      sum: the workgroup compacts the kept values into LDS
//                      in index order and one lane adds them one at a time,
//                      as the reference's loop does;
//   floor_interp_kernel per bin, the knots' linear interpolation, and the
//                      excess avg - floor;
//   (host)             the three running sums (excess, floor, spread) in the
//                      reference's sequential order: each is a chain of
//                      65,536 dependent double additions at the shipped
//                      grid, about 15 ns apiece on one GPU lane against about
//                      1 ns on a host core, so the host carries them (the
//                      three chains interleaved) and sends two back;
//   ladder_kernel      every (rung, position): the deflection, the threshold,
//                      the local-maximum and plateau-middle rules; each peak
//                      is appended as one 64-bit key whose ascending order is
//                      the reference's total order (stronger first, then the
//                      lower start, then the narrower width);
//   a radix sort of the keys (rocPRIM) puts the peaks in that order.
// The host takes the first max_peaks (the best-N set the reference's heap
// keeps) and grows them strongest first, exactly as find_candidates does,
// with a bitmap of the accepted bins for the overlap test.
//
// Doubles use only +, -, *, / and comparisons on the device (IEEE, correctly
// rounded); square roots, exp2 and the normal quantile are computed on the
// host, where the reference computes them.

#pragma clang fp contract(off)

#include <algorithm>
#include <bit>
#include <cmath>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

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

// Number of set bits of `m` below this lane.
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

constexpr uint32_t kKeepStage = 1024;  // window values per compaction stage (4 per thread)

__global__ void __launch_bounds__(256)
knot_floor_kernel(const double* __restrict__ avg, const KnotWindow* __restrict__ knots,
                  double quantile_span, double high_q, double excision_sigma,
                  double* __restrict__ knot_floor, double* __restrict__ knot_sigma) {
    __shared__ uint32_t hist[2][256];
    __shared__ unsigned long long prefix[2];
    __shared__ uint32_t rank[2];
    __shared__ uint32_t wave_kept[4][4];
    __shared__ double kept_values[kKeepStage + 16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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
        for (uint32_t i0 = tid; i0 < count; i0 += 1024) {
            unsigned long long v[4];
            for (uint32_t q = 0; q < 4; ++q)  // four loads in flight, clamped in bounds
                v[q] = (unsigned long long)__double_as_longlong(w[min(i0 + 256 * q, count - 1)]);
            for (uint32_t q = 0; q < 4; ++q) {
                if (i0 + 256 * q >= count) break;
                const uint32_t bucket = uint32_t(v[q] >> shift) & 255u;
                if ((v[q] & mask) == p0) atomicAdd(&hist[0][bucket], 1u);
                if ((v[q] & mask) == p1) atomicAdd(&hist[1][bucket], 1u);
            }
        }
        __syncthreads();
        // Waves 0 and 1 pick the bucket holding rank r of their histogram:
        // lane l owns buckets 4l..4l+3, a wave scan gives each lane the count
        // up to its last bucket, and the first lane past r resolves the rest.
        if (wave < 2) {
            const uint32_t* h = hist[wave];
            const uint32_t c[4] = {h[4 * lane], h[4 * lane + 1], h[4 * lane + 2], h[4 * lane + 3]};
            const uint32_t own = c[0] + c[1] + c[2] + c[3];
            uint32_t upto = own;
            for (uint32_t d = 1; d < 64; d <<= 1) {
                const uint32_t t = __shfl_up(upto, d);
                if (lane >= d) upto += t;
            }
            const uint32_t r = rank[wave];
            const unsigned long long past = __ballot(upto > r);
            const uint32_t pick = past ? uint32_t(__ffsll((long long)past) - 1) : 63u;
            if (lane == pick) {
                uint32_t below = upto - own, sel = 4 * lane + 3;
                for (uint32_t k = 0; k < 4; ++k) {
                    if (below + c[k] > r) {
                        sel = 4 * lane + k;
                        break;
                    }
                    below += c[k];
                }
                rank[wave] = r - below;
                prefix[wave] |= (unsigned long long)sel << shift;
            }
        }
        mask |= 255ull << shift;
        __syncthreads();
    }
    const double lp = __longlong_as_double((long long)prefix[0]);
    const double hp = kw.high_rank > kw.low_rank ? __longlong_as_double((long long)prefix[1]) : lp;
    const double deviation = dmax(0.0, (hp - lp) / quantile_span);
    const double spread = high_q * deviation;
    const double level = hp - spread;
    const double reach = excision_sigma * deviation;
    const double ceiling = level + reach;

    // The excision mean: the values at or below the ceiling, added in index
    // order. Each stage of 1024 values is compacted into LDS (element
    // base + 256 g + tid lands after every kept element of lower index), then
    // thread 0 adds them one at a time while the next stage loads.
    double sum = 0.0;
    uint32_t kept = 0;
    double v[4];
    for (uint32_t g = 0; g < 4; ++g) v[g] = w[min(g * 256 + tid, count - 1)];
    for (uint32_t base = 0; base < count; base += kKeepStage) {
        bool keep[4];
        uint32_t below[4];
        for (uint32_t g = 0; g < 4; ++g) {
            keep[g] = base + g * 256 + tid < count && v[g] <= ceiling;
            const unsigned long long m = __ballot(keep[g]);
            below[g] = lanes_below(m);
            if (lane == 0) wave_kept[g][wave] = uint32_t(__popcll(m));
        }
        __syncthreads();
        uint32_t run = 0, offset[4] = {0, 0, 0, 0};
        for (uint32_t g = 0; g < 4; ++g)
            for (uint32_t wv = 0; wv < 4; ++wv) {
                if (wv == wave) offset[g] = run;
                run += wave_kept[g][wv];
            }
        const uint32_t fill = run;
        for (uint32_t g = 0; g < 4; ++g)
            if (keep[g]) kept_values[offset[g] + below[g]] = v[g];
        __syncthreads();
        for (uint32_t g = 0; g < 4; ++g) v[g] = w[min(base + kKeepStage + g * 256 + tid, count - 1)];
        if (tid == 0) {
            // kept_values holds kKeepStage + 16 entries, so the read ahead stays in bounds
            double cur[16], next[16];
            for (uint32_t q = 0; q < 16; ++q) cur[q] = kept_values[q];
            uint32_t k = 0;
            for (; k + 16 <= fill; k += 16) {
                for (uint32_t q = 0; q < 16; ++q) next[q] = kept_values[k + 16 + q];
                for (uint32_t q = 0; q < 16; ++q) sum = sum + cur[q];
                for (uint32_t q = 0; q < 16; ++q) cur[q] = next[q];
            }
            for (; k < fill; ++k) sum = sum + kept_values[k];
        }
        kept += fill;
        __syncthreads();
    }
    if (tid == 0) {
        const double estimate = kept > 0 ? sum / double(kept) : level;
        knot_floor[blockIdx.x] = dmax(estimate, double(kSpectrumPowerFloor));
        knot_sigma[blockIdx.x] = deviation;
    }
}

__global__ void floor_interp_kernel(const double* __restrict__ pos, const double* __restrict__ kf,
                                    const double* __restrict__ ks, uint32_t knots, uint32_t bins,
                                    const double* __restrict__ avg, double* __restrict__ floor_,
                                    double* __restrict__ sigma_, double* __restrict__ excess) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bins) return;
    double f, g;
    if (knots == 1) {
        f = kf[0];
        g = ks[0];
    } else {
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
        f = kf[k] + tf;
        g = ks[k] + ts;
    }
    floor_[i] = f;
    sigma_[i] = g;
    excess[i] = avg[i] - f;
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
    uint32_t log2_width;
    double root;
    double threshold;
};

// A peak as one key: the deflection's bits inverted (a peak's deflection is
// above a positive threshold, and positive floats order as their bits), then
// the start, then log2 of the width, so ascending keys are stronger first,
// lower start first, narrower first. bins <= 2^24 leaves the start 27 bits.
__device__ __host__ inline uint64_t peak_key(float deflection, uint32_t start, uint32_t log2_width) {
    return (uint64_t(~std::bit_cast<uint32_t>(deflection)) << 32) | (uint64_t(start) << 5) |
           log2_width;
}

Peak peak_of(uint64_t key) {
    const float d = std::bit_cast<float>(~uint32_t(key >> 32));
    return Peak{uint32_t(key >> 5) & 0x7FFFFFFu, 1u << (key & 31u), double(d)};
}

__global__ void __launch_bounds__(256)
ladder_kernel(const double* __restrict__ ec, const double* __restrict__ fc,
              const Rung* __restrict__ rungs, uint32_t bins, const int32_t* __restrict__ allow,
              uint64_t* __restrict__ keys, uint32_t capacity, uint32_t* __restrict__ count) {
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
    if (idx < capacity) keys[idx] = peak_key(here, s + (end - s) / 2, r.log2_width);
}

__global__ void zero_counter(uint32_t* c) { *c = 0u; }

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
    double *ec = nullptr, *fc = nullptr, *excess = nullptr;
    Rung* rungs = nullptr;
    uint64_t *keys = nullptr, *sorted = nullptr;
    void* sort_temp = nullptr;
    size_t sort_bytes = 0;
    uint32_t* counter = nullptr;
    // pinned host mirrors
    uint32_t* found_h = nullptr;
    double *excess_h = nullptr, *floor_h = nullptr, *sigma_h = nullptr;
    double *ec_h = nullptr, *fc_h = nullptr, *sc_h = nullptr;
    uint64_t* keys_h = nullptr;
    std::vector<uint64_t> taken;        // one bit per accepted bin
    std::vector<uint64_t> taken_words;  // one bit per nonzero word of taken
    std::vector<Peak> accepted;

    ~rvx_detector() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (void* p : {(void*)windows, (void*)pos, (void*)kf, (void*)ks, (void*)floor_,
                        (void*)sigma_, (void*)ec, (void*)fc, (void*)excess,
                        (void*)rungs, (void*)keys, (void*)sorted, sort_temp, (void*)counter})
            if (p) (void)hipFree(p);
        for (void* p : {(void*)found_h, (void*)excess_h, (void*)floor_h, (void*)sigma_h,
                        (void*)ec_h, (void*)fc_h, (void*)sc_h, (void*)keys_h})
            if (p) (void)hipHostFree(p);
    }

    // Any set bit of `bits` in [lo, hi) (lo < hi). The words strictly
    // between the two end words are looked up in `summary` (a bit per word of
    // `bits`, set when that word is nonzero) when there is one, so a wide
    // window costs a few words, and the narrow case has no data-dependent
    // branch.
    static bool any_set(const uint64_t* bits, const uint64_t* summary, size_t lo, size_t hi) {
        const size_t a = lo >> 6, b = (hi - 1) >> 6;
        const uint64_t first = ~0ull << (lo & 63), last = ~0ull >> (63 - ((hi - 1) & 63));
        const uint64_t wa = bits[a] & first, wb = bits[b] & last;
        uint64_t hit = a == b ? (wa & last) : (wa | wb);
        if (b > a + 1) {
            if (summary) return hit != 0 || any_set(summary, nullptr, a + 1, b);
            for (size_t k = a + 1; k < b; ++k) hit |= bits[k];
        }
        return hit != 0;
    }

    bool overlaps_taken(size_t lo, size_t hi) const {  // any accepted bin in [lo, hi)
        return any_set(taken.data(), taken_words.data(), lo, hi);
    }

    void take(size_t lo, size_t hi) {
        for (size_t i = lo; i < hi;) {
            if ((i & 63) == 0 && i + 64 <= hi) {
                taken[i >> 6] = ~0ull;
                i += 64;
            } else {
                taken[i >> 6] |= 1ull << (i & 63);
                ++i;
            }
        }
        for (size_t w = lo >> 6; w <= (hi - 1) >> 6; ++w) taken_words[w >> 6] |= 1ull << (w & 63);
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
    for (uint32_t w = 1, lg = 0; w <= widest; w *= 2, ++lg) {
        const double root = std::sqrt(double(w));
        const double threshold = threshold_linear * kReferenceBandwidthHz / (root * c.bin_width_hz);
        d->rungs_h.push_back(Rung{w, lg, root, threshold});
        capacity += (bins - w + 2) / 2;  // strict local maxima: at most every other position
    }
    d->budget = c.max_peaks ? c.max_peaks : uint32_t(bins);
    d->capacity = uint32_t(std::min<uint64_t>(capacity, 0xFFFFFFFFull));
    const size_t kept_keys = std::min<size_t>(d->capacity, d->budget);
    hipError_t err = hipSuccess;
    auto alloc = [&](auto** p, size_t bytes) {
        if (err == hipSuccess) err = hipMalloc((void**)p, bytes);
    };
    auto pinned = [&](auto** p, size_t bytes) {
        if (err == hipSuccess) err = hipHostMalloc((void**)p, bytes, hipHostMallocDefault);
    };
    alloc(&d->windows, knots * sizeof(KnotWindow));
    alloc(&d->pos, knots * sizeof(double));
    alloc(&d->kf, knots * sizeof(double));
    alloc(&d->ks, knots * sizeof(double));
    alloc(&d->floor_, bins * sizeof(double));
    alloc(&d->sigma_, bins * sizeof(double));
    alloc(&d->ec, (bins + 1) * sizeof(double));
    alloc(&d->fc, (bins + 1) * sizeof(double));
    alloc(&d->excess, bins * sizeof(double));
    alloc(&d->rungs, d->rungs_h.size() * sizeof(Rung));
    alloc(&d->keys, size_t(d->capacity) * sizeof(uint64_t));
    alloc(&d->sorted, size_t(d->capacity) * sizeof(uint64_t));
    alloc(&d->counter, sizeof(uint32_t));
    if (err == hipSuccess)
        err = rocprim::radix_sort_keys(nullptr, d->sort_bytes, d->keys, d->sorted, d->capacity, 0,
                                       64, ctx->stream);
    alloc(&d->sort_temp, std::max<size_t>(d->sort_bytes, 16));
    pinned(&d->found_h, sizeof(uint32_t));
    pinned(&d->excess_h, bins * sizeof(double));
    pinned(&d->floor_h, bins * sizeof(double));
    pinned(&d->sigma_h, bins * sizeof(double));
    pinned(&d->ec_h, (bins + 1) * sizeof(double));
    pinned(&d->fc_h, (bins + 1) * sizeof(double));
    pinned(&d->sc_h, (bins + 1) * sizeof(double));
    pinned(&d->keys_h, std::max<size_t>(kept_keys, 1) * sizeof(uint64_t));
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
    d->taken.assign((bins + 63) / 64, 0);
    d->taken_words.assign((d->taken.size() + 63) / 64, 0);
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
    floor_interp_kernel<<<dim3((bins + 255) / 256), dim3(256), 0, s>>>(
        d->pos, d->kf, d->ks, d->knots, bins, average, d->floor_, d->sigma_, d->excess);
    if (int32_t rc = rvx::check_launch(who)) return rc;
    RVX_TRY_HIP(hipMemcpyAsync(d->excess_h, d->excess, bins * sizeof(double),
                               hipMemcpyDeviceToHost, s),
                who);
    RVX_TRY_HIP(hipMemcpyAsync(d->floor_h, d->floor_, bins * sizeof(double),
                               hipMemcpyDeviceToHost, s),
                who);
    RVX_TRY_HIP(hipMemcpyAsync(d->sigma_h, d->sigma_, bins * sizeof(double),
                               hipMemcpyDeviceToHost, s),
                who);
    RVX_TRY_HIP(hipStreamSynchronize(s), who);
    {  // the running sums, out[0] = 0 and out[i + 1] = out[i] + x[i]
        const double *ex = d->excess_h, *fl = d->floor_h, *sg = d->sigma_h;
        double *ec = d->ec_h, *fc = d->fc_h, *sc = d->sc_h;
        double e = 0.0, f = 0.0, g = 0.0;
        ec[0] = fc[0] = sc[0] = 0.0;
        for (uint32_t i = 0; i < bins; ++i) {
            e = e + ex[i];
            f = f + fl[i];
            g = g + sg[i];
            ec[i + 1] = e;
            fc[i + 1] = f;
            sc[i + 1] = g;
        }
    }
    RVX_TRY_HIP(hipMemcpyAsync(d->ec, d->ec_h, (bins + 1) * sizeof(double), hipMemcpyHostToDevice,
                               s),
                who);
    RVX_TRY_HIP(hipMemcpyAsync(d->fc, d->fc_h, (bins + 1) * sizeof(double), hipMemcpyHostToDevice,
                               s),
                who);
    zero_counter<<<dim3(1), dim3(1), 0, s>>>(d->counter);
    ladder_kernel<<<dim3((bins + 255) / 256, uint32_t(d->rungs_h.size())), dim3(256), 0, s>>>(
        d->ec, d->fc, d->rungs, bins, allow_prefix, d->keys, d->capacity, d->counter);
    if (int32_t rc = rvx::check_launch(who)) return rc;
    RVX_TRY_HIP(hipMemcpyAsync(d->found_h, d->counter, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                who);
    RVX_TRY_HIP(hipStreamSynchronize(s), who);
    const uint32_t found = *d->found_h;
    if (found > d->capacity)
        return rvx::fail(RVX_EDEVICE, "%s: %u peaks overflowed a list of %u", who, found,
                         d->capacity);
    if (n_peaks) *n_peaks = found;
    // The best max_peaks by the reference's order, strongest first: the
    // sorted keys' prefix.
    const uint32_t n = std::min(found, d->budget);
    if (n) {
        const uint64_t* ordered = d->keys;
        if (found > 1) {
            size_t bytes = d->sort_bytes;
            RVX_TRY_HIP(rocprim::radix_sort_keys(d->sort_temp, bytes, d->keys, d->sorted, found, 0,
                                                 64, s),
                        who);
            ordered = d->sorted;
        }
        RVX_TRY_HIP(hipMemcpyAsync(d->keys_h, ordered, size_t(n) * sizeof(uint64_t),
                                   hipMemcpyDeviceToHost, s),
                    who);
        RVX_TRY_HIP(hipStreamSynchronize(s), who);
    }

    // Greedy growth (find_candidates :1060-1120): each winner grows onto its
    // shoulders within its seed width and the accepted neighbours, stepping
    // over gaps shorter than split_gap_bins; a peak overlapping an accepted
    // band is dropped.
    const double* ex = d->excess_h;
    const double* fcum = d->fc_h;
    const double* scum = d->sc_h;
    const size_t gap = d->cfg.split_gap_bins;
    auto& accepted = d->accepted;
    accepted.clear();
    std::fill(d->taken.begin(), d->taken.end(), 0ull);
    std::fill(d->taken_words.begin(), d->taken_words.end(), 0ull);
    auto above = [&](size_t i, double level) { return ex[i] > level; };
    for (uint32_t k = 0; k < n; ++k) {
        const Peak p = peak_of(d->keys_h[k]);
        size_t lo = p.start, hi = size_t(p.start) + p.width;
        if (d->overlaps_taken(lo, hi)) continue;
        auto at = std::lower_bound(accepted.begin(), accepted.end(), lo,
                                   [](const Peak& held, size_t v) { return held.start < v; });
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
        d->take(lo, hi);
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
