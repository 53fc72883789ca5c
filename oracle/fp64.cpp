// FP64 definitions the float twins and the GPU path are measured against.
//
// TEST INFRASTRUCTURE, loaded only by the test suite and the benchmark's
// error report. These are the mathematical definitions, evaluated directly in
// double, not restatements of any kernel's operation order.

#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

extern "C" {

// The channelizer definition, core/shaders/pfb_branch.comp:9-19:
//     y_k[m] = sum_{n=0}^{N-1} h[n] x[mD - n] exp(-j 2 pi k (mD - n) / M)
// x holds complex samples from absolute index 0. Output is block-major,
// out[(m - m0) * M + k], for m in [m0, m0 + blocks).
int fp64_channelize(const double* x, uint64_t nx, const double* h, uint32_t ntaps, uint32_t M,
                    uint32_t D, uint64_t m0, uint32_t blocks, double* out) {
    const double two_pi = 6.283185307179586476925286766559;
    std::vector<std::complex<double>> w(M);
    for (uint32_t i = 0; i < M; ++i)
        w[i] = std::polar(1.0, -two_pi * double(i) / double(M));
    for (uint32_t b = 0; b < blocks; ++b) {
        const uint64_t base = (m0 + b) * D;
        if (base + 1 < ntaps || base >= nx) return -1;
        for (uint32_t k = 0; k < M; ++k) {
            std::complex<double> acc = 0.0;
            for (uint32_t n = 0; n < ntaps; ++n) {
                const uint64_t idx = base - n;
                const std::complex<double> s(x[2 * idx], x[2 * idx + 1]);
                const uint64_t ph = (uint64_t(k) * (idx % M)) % M;
                acc += h[n] * s * w[ph];
            }
            out[2 * (uint64_t(b) * M + k)] = acc.real();
            out[2 * (uint64_t(b) * M + k) + 1] = acc.imag();
        }
    }
    return 0;
}

// Brute-force DFT, X[k] = sum_n x[n] exp(-j 2 pi k n / n_points).
int fp64_dft(const double* x, uint32_t n_points, double* out) {
    const double two_pi = 6.283185307179586476925286766559;
    for (uint32_t k = 0; k < n_points; ++k) {
        std::complex<double> acc = 0.0;
        for (uint32_t n = 0; n < n_points; ++n) {
            const uint64_t ph = (uint64_t(k) * n) % n_points;
            acc += std::complex<double>(x[2 * n], x[2 * n + 1]) *
                   std::polar(1.0, -two_pi * double(ph) / double(n_points));
        }
        out[2 * k] = acc.real();
        out[2 * k + 1] = acc.imag();
    }
    return 0;
}

}  // extern "C"
