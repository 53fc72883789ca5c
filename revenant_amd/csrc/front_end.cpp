// The front-end estimator on the host: a block's chunk sums in, the DC and
// I/Q-imbalance estimate and the correction's four numbers out.
//
//   rvx_total_moments          core/dsp/front_end_correction.cpp:39-53
//   rvx_front_end_*            FrontEndCorrector, :55-122 (update, correction)
//   rvx_front_end_derived      FrontEndEstimate's dc_dbfs, phase_error_deg,
//                              image_rejection_db, gain_error_db (:20-37)
//
// The model (front_end_correction.h:19-45): I_r = I, Q_r = g (Q cos phi +
// I sin phi) for a proper signal whose arms have equal power and no
// correlation, so the central second moments give g = sqrt(m_qq / m_ii) and
// sin phi = m_iq / sqrt(m_ii m_qq), and the inverse is one Gram-Schmidt row:
// Q_c = cross I_r + scale Q_r with cross = -tan phi, scale = 1/(g cos phi).
// Each block's means and central moments are exponentially averaged with a
// weight 1 - exp(-n / (tau rate)) (the first block seeds them), in double.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>

#include "common.h"

struct rvx_front_end {
    rvx_front_end_config cfg{};
    rvx_front_end_estimate est{};
    double m_ii = 0.0, m_qq = 0.0, m_iq = 0.0;
};

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

double weight_of(uint64_t samples, double tau_s, int64_t rate) {
    if (tau_s <= 0.0 || rate <= 0) return 1.0;
    return 1.0 - std::exp(-double(samples) / (tau_s * double(rate)));
}

void reset_estimate(rvx_front_end& f) {
    f.est = rvx_front_end_estimate{};
    f.est.gain = 1.0;
    f.m_ii = f.m_qq = f.m_iq = 0.0;
}

}  // namespace

extern "C" {

void rvx_front_end_default_config(rvx_front_end_config* out) {
    if (!out) return;
    out->dc_time_constant_s = 0.1;
    out->iq_time_constant_s = 1.0;
    out->max_gain_error_db = 6.0;
    out->max_phase_error_deg = 30.0;
}

int32_t rvx_total_moments(const float* sums, uint64_t n_sums, uint32_t count,
                          rvx_iq_moments_total* out) {
    if (!out || (n_sums && !sums)) return rvx::fail(RVX_EINVAL, "rvx_total_moments: null argument");
    rvx_iq_moments_total t{};
    t.samples = count;
    const uint64_t chunks = std::min<uint64_t>((uint64_t(count) + RVX_IQ_MOMENTS_CHUNK - 1) /
                                                   RVX_IQ_MOMENTS_CHUNK,
                                               n_sums / RVX_IQ_MOMENTS_PER_CHUNK);
    for (uint64_t c = 0; c < chunks; ++c) {
        const float* s = sums + c * RVX_IQ_MOMENTS_PER_CHUNK;
        t.sum_i += s[0];
        t.sum_q += s[1];
        t.sum_ii += s[2];
        t.sum_qq += s[3];
        t.sum_iq += s[4];
    }
    *out = t;
    return RVX_OK;
}

int32_t rvx_front_end_create(const rvx_front_end_config* config, rvx_front_end** out) {
    if (!out) return rvx::fail(RVX_EINVAL, "rvx_front_end_create: null output");
    auto* f = new (std::nothrow) rvx_front_end();
    if (!f) return rvx::fail(RVX_ENOMEM, "rvx_front_end_create: out of memory");
    if (config)
        f->cfg = *config;
    else
        rvx_front_end_default_config(&f->cfg);
    reset_estimate(*f);
    *out = f;
    return RVX_OK;
}

void rvx_front_end_destroy(rvx_front_end* f) { delete f; }

void rvx_front_end_reset(rvx_front_end* f) {
    if (f) reset_estimate(*f);
}

int32_t rvx_front_end_update(rvx_front_end* f, const rvx_iq_moments_total* m, int64_t rate) {
    if (!f || !m) return rvx::fail(RVX_EINVAL, "rvx_front_end_update: null argument");
    if (m->samples == 0) return RVX_OK;
    const double n = double(m->samples);
    const double mean_i = m->sum_i / n;
    const double mean_q = m->sum_q / n;
    // About the block's own mean, so the imbalance estimate does not wait for
    // the DC estimate to converge.
    const double c_ii = m->sum_ii / n - mean_i * mean_i;
    const double c_qq = m->sum_qq / n - mean_q * mean_q;
    const double c_iq = m->sum_iq / n - mean_i * mean_q;
    rvx_front_end_estimate& e = f->est;
    const bool first = !e.measured;
    const double a_dc = first ? 1.0 : weight_of(m->samples, f->cfg.dc_time_constant_s, rate);
    const double a_iq = first ? 1.0 : weight_of(m->samples, f->cfg.iq_time_constant_s, rate);
    e.dc_i += a_dc * (mean_i - e.dc_i);
    e.dc_q += a_dc * (mean_q - e.dc_q);
    f->m_ii += a_iq * (c_ii - f->m_ii);
    f->m_qq += a_iq * (c_qq - f->m_qq);
    f->m_iq += a_iq * (c_iq - f->m_iq);
    e.measured = 1;
    e.samples_seen += m->samples;
    if (f->m_ii > 0.0 && f->m_qq > 0.0) {
        e.gain = std::sqrt(f->m_qq / f->m_ii);
        e.sin_phase = f->m_iq / std::sqrt(f->m_ii * f->m_qq);
        const double gain_db = std::fabs(20.0 * std::log10(e.gain));
        const double s = e.sin_phase < -1.0 ? -1.0 : e.sin_phase > 1.0 ? 1.0 : e.sin_phase;
        const double phase_deg = std::fabs(std::asin(s) * 180.0 / kPi);
        e.iq_plausible = std::isfinite(e.gain) && std::isfinite(e.sin_phase) &&
                         gain_db <= f->cfg.max_gain_error_db &&
                         phase_deg <= f->cfg.max_phase_error_deg;
    } else {
        e.gain = 1.0;
        e.sin_phase = 0.0;
        e.iq_plausible = 0;
    }
    return RVX_OK;
}

int32_t rvx_front_end_estimate_get(const rvx_front_end* f, rvx_front_end_estimate* out) {
    if (!f || !out) return rvx::fail(RVX_EINVAL, "rvx_front_end_estimate_get: null argument");
    *out = f->est;
    return RVX_OK;
}

int32_t rvx_front_end_correction(const rvx_front_end* f, int32_t dc, int32_t iq,
                                 rvx_iq_correct_params* out) {
    if (!f || !out) return rvx::fail(RVX_EINVAL, "rvx_front_end_correction: null argument");
    rvx_iq_correct_params p{};
    p.scale = 1.0f;  // the identity: dc 0, cross 0, scale 1
    const rvx_front_end_estimate& e = f->est;
    if (e.measured) {
        if (dc) {
            p.dc_i = float(e.dc_i);
            p.dc_q = float(e.dc_q);
        }
        if (iq && e.iq_plausible) {
            const double cos_phase = std::sqrt(1.0 - e.sin_phase * e.sin_phase);
            p.cross = float(-e.sin_phase / cos_phase);
            p.scale = float(1.0 / (e.gain * cos_phase));
        }
    }
    *out = p;
    return RVX_OK;
}

int32_t rvx_front_end_derived(const rvx_front_end_estimate* e, double* dc_dbfs,
                              double* gain_error_db, double* phase_error_deg,
                              double* image_rejection_db) {
    if (!e) return rvx::fail(RVX_EINVAL, "rvx_front_end_derived: null estimate");
    if (dc_dbfs) {
        const double power = e->dc_i * e->dc_i + e->dc_q * e->dc_q;
        *dc_dbfs = power > 0.0 ? 10.0 * std::log10(power) : -300.0;
    }
    if (gain_error_db) *gain_error_db = 20.0 * std::log10(e->gain);
    const double s = e->sin_phase < -1.0 ? -1.0 : e->sin_phase > 1.0 ? 1.0 : e->sin_phase;
    if (phase_error_deg) *phase_error_deg = std::asin(s) * 180.0 / kPi;
    if (image_rejection_db) {
        const double c = std::sqrt(std::fmax(0.0, 1.0 - e->sin_phase * e->sin_phase));
        const double wanted = 1.0 + 2.0 * e->gain * c + e->gain * e->gain;
        const double image = 1.0 - 2.0 * e->gain * c + e->gain * e->gain;
        *image_rejection_db = image <= 0.0 ? 300.0 : 10.0 * std::log10(wanted / image);
    }
    return RVX_OK;
}

}  // extern "C"
