// The negative control of the conformance suite: a complex multiply with one
// sign flipped (the real part adds a.y*b.y instead of subtracting it), the
// imaginary part left correct. Every output is finite and plausible and half
// of them are wrong, which is the failure the bit diff exists to catch
// (tests/reference/shaders/cmul_wrong.comp, tests/reference/test_divergence.cpp:84-125).
// Built into its own library for the tests; the product library never links it.

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

__global__ void cmul_wrong_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                  float2* __restrict__ out, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float2 x = a[i], y = b[i];
    const float re = x.x * y.x + x.y * y.y;  // the flipped sign
    const float im = x.x * y.y + x.y * y.x;
    out[i] = make_float2(re, im);
}

}  // namespace

extern "C" int32_t cmul_wrong(const void* lhs, const void* rhs, void* out, uint64_t count,
                              void* stream) {
    if (count == 0) return 0;
    if (!lhs || !rhs || !out || count > 0xFFFFFFFFull) return -1;
    const uint32_t n = uint32_t(count);
    cmul_wrong_kernel<<<dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream)>>>(
        static_cast<const float2*>(lhs), static_cast<const float2*>(rhs),
        static_cast<float2*>(out), n);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
