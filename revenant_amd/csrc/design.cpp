// Host-side design of the tables the kernels read, in FP64 rounded once to
// float: the polyphase prototype, the twiddle circle, the spectrum window and
// its channel-shape correction, the NCO table, and the exact 0.64 NCO step.
//
// None of these is computed on the device: a device sin/cos would make the
// tables vendor-dependent. Kernel and oracle read the same arrays, so what
// matters is that the arrays are the reference's definitions:
//   prototype      core/dsp/pfb_design.cpp:381-510 (Kaiser-windowed sinc,
//                  length M*L+1, cutoff 1/(2M), sum normalised to 1, padded
//                  with zeros to M*(L+1))
//   twiddles       core/dsp/pfb_design.cpp:515-591 (first octant from libm,
//                  the rest by exact reflection; no -0.0)
//   window         core/dsp/spectrum_reference.cpp:195-285 (4-term
//                  Blackman-Harris, unit sum, then 1/|H|^2 of the prototype at
//                  each kept bin, capped at +12 dB)
//   NCO table      core/dsp/vrx_reference.cpp:364 (octant + reflections)
//   turn_fixed64   core/dsp/vrx_reference.cpp:308 (base-65536 long division)
// The 2048-channel ceiling of the reference (Vulkan's 16 KiB shared-memory
// floor) is lifted to 65536: gfx950 has 160 KiB of LDS per CU.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <numeric>
#include <vector>

#include "common.h"

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr uint32_t kMaxChannels = 65536;
constexpr uint64_t kMaxPrototypeTaps = uint64_t(1) << 20;

double i0(double x) {
    // Ascending series sum (x^2/4)^k / (k!)^2, stopped once a term no longer
    // moves the double sum.
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 256; ++k) {
        term *= q / (double(k) * double(k));
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}

double kaiser_shape(double attenuation_db) {
    if (attenuation_db > 50.0) return 0.1102 * (attenuation_db - 8.7);
    if (attenuation_db >= 21.0) {
        const double e = attenuation_db - 21.0;
        return 0.5842 * std::pow(e, 0.4) + 0.07886 * e;
    }
    return 0.0;
}

float unsigned_zero(double v) {
    const float f = static_cast<float>(v);
    return f == 0.0f ? 0.0f : f;
}

// Peak stopband response in dB relative to DC, from 0.75/M to Nyquist,
// sampled on a zero-padded power-of-two DFT (16x oversampled, capped).
double stopband_db(const std::vector<float>& taps, uint32_t channels) {
    size_t points = 1;
    while (points < taps.size()) points <<= 1;
    for (int d = 0; d < 4 && points < (size_t(1) << 21); ++d) points <<= 1;
    std::vector<std::complex<double>> a(points, 0.0);
    for (size_t n = 0; n < taps.size(); ++n) a[n] = double(taps[n]);
    for (size_t i = 1, j = 0; i < points; ++i) {
        size_t bit = points >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= points; len <<= 1) {
        const double ang = -2.0 * kPi / double(len);
        for (size_t k = 0; k < len / 2; ++k) {
            const std::complex<double> w(std::cos(ang * double(k)), std::sin(ang * double(k)));
            for (size_t s = 0; s < points; s += len) {
                const auto u = a[s + k];
                const auto v = a[s + k + len / 2] * w;
                a[s + k] = u + v;
                a[s + k + len / 2] = u - v;
            }
        }
    }
    const double dc = std::abs(a[0]);
    if (!(dc > 0.0)) return 0.0;
    const size_t first = size_t(std::ceil(0.75 / double(channels) * double(points)));
    double peak = 0.0;
    for (size_t i = first; i <= points / 2; ++i) peak = std::max(peak, std::abs(a[i]));
    if (!(peak > 0.0)) return -600.0;
    return std::max(-600.0, 20.0 * std::log10(peak / dc));
}

int32_t grid_valid(const rvx_grid* g, const char* who) {
    if (g == nullptr) return rvx::fail(RVX_EINVAL, "%s: null grid", who);
    if (g->channels < 2 || !rvx::pow2(g->channels) || g->channels > kMaxChannels)
        return rvx::fail(RVX_EINVAL, "%s: channel count %u is not a power of two in [2, %u]", who,
                         g->channels, kMaxChannels);
    if (g->taps_per_branch == 0)
        return rvx::fail(RVX_EINVAL, "%s: zero taps per branch", who);
    if (g->decimation == 0 || g->channels % g->decimation != 0)
        return rvx::fail(RVX_EINVAL, "%s: decimation %u does not divide %u channels", who,
                         g->decimation, g->channels);
    if (uint64_t(g->channels) * g->taps_per_branch > kMaxPrototypeTaps)
        return rvx::fail(RVX_EINVAL, "%s: prototype of %llu taps exceeds %llu", who,
                         (unsigned long long)(uint64_t(g->channels) * g->taps_per_branch),
                         (unsigned long long)kMaxPrototypeTaps);
    return RVX_OK;
}

// The prototype in double before rounding, for the window's correction too.
std::vector<float> prototype(const rvx_grid& g, double attenuation_db, double beta_override) {
    const uint32_t M = g.channels;
    const uint32_t L = g.taps_per_branch - 1;
    const uint64_t design_length = uint64_t(M) * L;
    const uint64_t used = design_length + 1;
    double beta;
    if (beta_override >= 0.0) {
        beta = beta_override;
    } else {
        // The longest the branch length can carry (Kaiser's order estimate
        // read backwards, transition pi/M): asking for more widens the
        // transition and worsens the stopband.
        const double reachable = 2.285 * kPi * double(L) + 8.0;
        beta = kaiser_shape(std::min(attenuation_db, reachable));
    }
    const double norm = i0(beta);
    const double centre = 0.5 * double(design_length);
    const double inv_m = 1.0 / double(M);
    std::vector<double> h(used, 0.0);
    double sum = 0.0;
    for (uint64_t n = 0; n < used; ++n) {
        const double off = double(n) - centre;
        const double pos = centre > 0.0 ? off / centre : 0.0;
        const double rad = std::max(0.0, 1.0 - pos * pos);
        const double win = i0(beta * std::sqrt(rad)) / norm;
        const double arg = kPi * off * inv_m;
        const double sinc = off == 0.0 ? 1.0 : std::sin(arg) / arg;
        h[n] = inv_m * sinc * win;
        sum += h[n];
    }
    std::vector<float> taps(uint64_t(M) * g.taps_per_branch, 0.0f);
    for (uint64_t n = 0; n < used; ++n) taps[n] = static_cast<float>(h[n] / sum);
    return taps;
}

// |H(f)| / H(0) of a float prototype at f = num / (den * M) cycles per sample,
// with the angle reduced exactly in integers.
double prototype_response(const std::vector<float>& taps, uint32_t M, int64_t num, int64_t den) {
    double re = 0.0, im = 0.0, dc = 0.0;
    const int64_t span = den * int64_t(M);
    for (size_t n = 0; n < taps.size(); ++n) {
        const double t = double(taps[n]);
        dc += t;
        if (t == 0.0) continue;
        int64_t r = (num * int64_t(n)) % span;
        if (r < 0) r += span;
        const double ang = -2.0 * kPi * double(r) / double(span);
        re += t * std::cos(ang);
        im += t * std::sin(ang);
    }
    const double mag = std::sqrt(re * re + im * im);
    return std::fabs(dc) > 0.0 ? mag / std::fabs(dc) : 0.0;
}

}  // namespace

extern "C" {

int32_t rvx_design_prototype(const rvx_grid* grid, double attenuation_db, double beta_override,
                             float* out, uint64_t out_len, double* stopband,
                             uint64_t* group_delay) {
    static const char* who = "rvx_design_prototype";
    if (int32_t rc = grid_valid(grid, who)) return rc;
    if (beta_override < 0.0 &&
        (!std::isfinite(attenuation_db) || attenuation_db < 0.0 || attenuation_db > 300.0))
        return rvx::fail(RVX_EINVAL, "%s: attenuation %g dB outside [0, 300]", who, attenuation_db);
    if (beta_override > 40.0 || std::isnan(beta_override))
        return rvx::fail(RVX_EINVAL, "%s: Kaiser beta %g outside [0, 40]", who, beta_override);
    const uint64_t n = uint64_t(grid->channels) * grid->taps_per_branch;
    if (out == nullptr || out_len != n)
        return rvx::fail(RVX_EINVAL, "%s: output wants %llu floats, got %llu", who,
                         (unsigned long long)n, (unsigned long long)out_len);
    const std::vector<float> taps = prototype(*grid, attenuation_db, beta_override);
    std::copy(taps.begin(), taps.end(), out);
    if (stopband) *stopband = stopband_db(taps, grid->channels);
    if (group_delay) *group_delay = uint64_t(grid->channels) * (grid->taps_per_branch - 1) / 2;
    return RVX_OK;
}

int32_t rvx_build_twiddles(uint32_t M, float* out, uint64_t out_len) {
    static const char* who = "rvx_build_twiddles";
    if (M == 0 || !rvx::pow2(M) || M > kMaxChannels)
        return rvx::fail(RVX_EINVAL, "%s: %u is not a power of two in [1, %u]", who, M, kMaxChannels);
    if (out == nullptr || out_len != 2ull * M)
        return rvx::fail(RVX_EINVAL, "%s: output wants %u floats", who, 2 * M);
    const uint32_t half = M / 2, quarter = M / 4, octant = M / 8;
    std::vector<double> oc(octant + 1), os(octant + 1), re(M, 0.0), im(M, 0.0);
    for (uint32_t i = 0; i <= octant; ++i) {
        const double a = 2.0 * kPi * double(i) / double(M);
        oc[i] = std::cos(a);
        os[i] = std::sin(a);
    }
    // First quadrant W^j = (cos, -sin); past the octant, swap through the
    // complement pi/2 - a.
    for (uint32_t j = 0; j <= quarter && j < M; ++j) {
        if (j <= octant) {
            re[j] = oc[j];
            im[j] = -os[j];
        } else {
            re[j] = os[quarter - j];
            im[j] = -oc[quarter - j];
        }
    }
    // W^(M/2 - i) = -conj(W^i)
    for (uint32_t j = quarter + 1; j <= half; ++j) {
        re[j] = -re[half - j];
        im[j] = im[half - j];
    }
    // W^(M - i) = conj(W^i)
    for (uint32_t j = half + 1; j < M; ++j) {
        re[j] = re[M - j];
        im[j] = -im[M - j];
    }
    for (uint32_t j = 0; j < M; ++j) {
        out[2 * j] = unsigned_zero(re[j]);
        out[2 * j + 1] = unsigned_zero(im[j]);
    }
    return RVX_OK;
}

int32_t rvx_build_spectrum_window(uint32_t size, uint32_t taps_per_branch, double attenuation_db,
                                  float* out, uint64_t out_len) {
    static const char* who = "rvx_build_spectrum_window";
    if (size < 4 || !rvx::pow2(size) || size > kMaxChannels)
        return rvx::fail(RVX_EINVAL, "%s: %u is not a power of two in [4, %u]", who, size,
                         kMaxChannels);
    if (out == nullptr || out_len != uint64_t(size) + size / 2)
        return rvx::fail(RVX_EINVAL, "%s: output wants %u floats", who, size + size / 2);
    std::vector<double> w(size);
    double sum = 0.0;
    for (uint32_t n = 0; n < size; ++n) {
        const double ph = 2.0 * kPi * double(n) / double(size);
        w[n] = 0.35875 - 0.48829 * std::cos(ph) + 0.14128 * std::cos(2.0 * ph) -
               0.01168 * std::cos(3.0 * ph);
        sum += w[n];
    }
    for (uint32_t n = 0; n < size; ++n) out[n] = static_cast<float>(w[n] / sum);
    // The correction inverts the shape of the channel the spectrum looks
    // through, designed once on the 64-channel, D = M/2 grid.
    const rvx_grid g{64, taps_per_branch, 32};
    if (int32_t rc = grid_valid(&g, who)) return rc;
    if (!std::isfinite(attenuation_db) || attenuation_db < 0.0 || attenuation_db > 300.0)
        return rvx::fail(RVX_EINVAL, "%s: attenuation %g dB outside [0, 300]", who, attenuation_db);
    const std::vector<float> proto = prototype(g, attenuation_db, -1.0);
    const double ceiling = std::pow(10.0, 12.0 / 10.0);
    for (uint32_t bin = 0; bin < size / 2; ++bin) {
        const int64_t num = 2 * int64_t(bin) - int64_t(size) / 2;
        const double r = prototype_response(proto, 64, num, int64_t(size));
        const double gain = 1.0 / (r * r);
        out[size + bin] = static_cast<float>((gain > ceiling || !std::isfinite(gain)) ? ceiling : gain);
    }
    return RVX_OK;
}

int32_t rvx_build_nco_table(uint32_t log2_size, float* out, uint64_t out_len) {
    static const char* who = "rvx_build_nco_table";
    if (log2_size > 24) return rvx::fail(RVX_EINVAL, "%s: log2 size %u above 24", who, log2_size);
    const uint32_t size = 1u << log2_size;
    if (out == nullptr || out_len != 2ull * size)
        return rvx::fail(RVX_EINVAL, "%s: output wants %u floats", who, 2 * size);
    if (size == 1) {
        out[0] = 1.0f;
        out[1] = 0.0f;
        return RVX_OK;
    }
    const uint32_t quarter = size / 4, octant = size / 8;
    std::vector<double> c(size, 0.0), s(size, 0.0);
    for (uint32_t i = 0; i <= octant; ++i) {
        const double a = 2.0 * kPi * double(i) / double(size);
        c[i] = std::cos(a);
        s[i] = std::sin(a);
    }
    for (uint32_t i = octant + 1; i <= quarter; ++i) {  // reflect about pi/4
        c[i] = s[quarter - i];
        s[i] = c[quarter - i];
    }
    for (uint32_t i = quarter + 1; i < size; ++i) {  // rotate by a quarter turn
        c[i] = -s[i - quarter];
        s[i] = c[i - quarter];
    }
    for (uint32_t i = 0; i < size; ++i) {
        out[2 * i] = c[i] == 0.0 ? 0.0f : static_cast<float>(c[i]);
        out[2 * i + 1] = s[i] == 0.0 ? 0.0f : static_cast<float>(-s[i]);
    }
    return RVX_OK;
}

uint64_t rvx_turn_fixed64(int64_t numerator, int64_t denominator) {
    if (denominator <= 0 || numerator < 0 || numerator >= denominator) return 0;
    // round(numerator * 2^64 / denominator) as four base-2^16 digits of long
    // division; numerator < denominator <= 2^31 keeps every shift in range.
    int64_t rem = numerator;
    uint64_t q = 0;
    for (int d = 0; d < 4; ++d) {
        rem <<= 16;
        q = (q << 16) | uint64_t(rem / denominator);
        rem %= denominator;
    }
    if (2 * rem >= denominator) ++q;
    return q;
}

int32_t rvx_nco_delta(int64_t num, int64_t den, int64_t rate, uint64_t* out) {
    static const char* who = "rvx_nco_delta";
    if (out == nullptr) return rvx::fail(RVX_EINVAL, "%s: null output", who);
    if (den == 0) return rvx::fail(RVX_EINVAL, "%s: zero denominator", who);
    if (rate <= 0) return rvx::fail(RVX_EINVAL, "%s: rate %lld not positive", who, (long long)rate);
    if (den < 0) {
        num = -num;
        den = -den;
    }
    const int64_t g = std::gcd(num, den);
    if (g > 1) {
        num /= g;
        den /= g;
    }
    const int64_t period = den * rate;
    if (period <= 0 || period > (int64_t(1) << 31))
        return rvx::fail(RVX_EINVAL, "%s: reduced period %lld exceeds 2^31", who, (long long)period);
    int64_t r = num % period;
    if (r < 0) r += period;
    *out = rvx_turn_fixed64(r, period);
    return RVX_OK;
}

}  // extern "C"

extern "C" int32_t rvx_prescale_taps(const float* taps, uint64_t n, float* out) {
    static const char* who = "rvx_prescale_taps";
    if ((n && (!taps || !out))) return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    for (uint64_t i = 0; i < n; ++i) {
        const float h = taps[i];
        if (!std::isfinite(h))
            return rvx::fail(RVX_EINVAL, "%s: tap %llu is not finite", who, (unsigned long long)i);
        if (h != 0.0f && std::fabs(h) < std::ldexp(1.0f, -111))
            return rvx::fail(RVX_EINVAL, "%s: tap %llu (%g) would scale below the normal range",
                             who, (unsigned long long)i, double(h));
        out[i] = std::ldexp(h, -15);
    }
    return RVX_OK;
}
