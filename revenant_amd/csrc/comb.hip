// Comb cancellation on the sample ring: the device half of the comb
// canceller (a harmonic comb of spurs from a switching supply or a clock,
// tracked on the host and subtracted before the channelizer).
//
//   rvx_comb_project   comb_project.comp / comb_kernels_reference.cpp:100-141
//   rvx_comb_subtract  comb_subtract.comp / :143-180
//   rvx_comb_step, rvx_comb_table, rvx_comb_check   :44-98
//
// project: per line l and sub-block s, sum over the sub-block's samples of
// x[n] * conj(tone_l(n)), in chunks of 64 samples summed from zero in sample
// order, the chunk sums added from zero in chunk order (the kernel's two
// levels). tone_l(n) = table[(phase0_l + n * step_l) >> 20] with 32-bit
// wrapping phase. A workgroup takes one sub-block and up to 16 lines, a wave
// per line: it stages 64 chunks of samples at a time in LDS with coalesced
// loads (rows padded by one sample), each lane sums its chunk for its wave's
// line, and the wave's lane 0 folds the chunk sums in order.
//
// subtract: per sample, removed = sum over lines ascending of a_l * tone_l(n)
// (from -0, which adds to anything unchanged), cleaned = x - removed; with no
// lines the sample is copied and -0 is removed. Sixteen samples per lane over
// an LDS copy of the tone table; the line records are uniform across the wave.
// Both kernels gather the tone table from LDS.

#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

namespace {

constexpr uint32_t kTableLog2 = 12;                 // comb_cancel.h:117
constexpr uint32_t kTableSize = 1u << kTableLog2;
constexpr uint32_t kTableShift = 32u - kTableLog2;
constexpr uint32_t kMaxLines = 1024;                // :109
constexpr uint32_t kMaxSubBlocks = 16;              // :123
constexpr uint32_t kChunk = 64;                     // comb_project.comp's chunk
constexpr uint32_t kLanes = 64;
constexpr uint32_t kRow = kChunk + 1;

bool is_mask(uint32_t m) { return rvx::pow2(uint64_t(m) + 1); }
uint64_t ring_of(uint32_t m) { return uint64_t(m) + 1; }

__device__ __forceinline__ float2 tone(const float2* __restrict__ table, const rvx_comb_line& l,
                                       uint32_t n) {
    return table[(l.phase0 + n * l.step) >> kTableShift];
}

// One workgroup per (sub-block, group of `group` lines): a wave per line. The
// waves share each staged tile of 64 chunks, so the block's samples are read
// from memory once per group rather than once per line.
__global__ void __launch_bounds__(1024)
project_kernel(rvx_comb_params p, uint32_t group, const float2* __restrict__ raw,
               const rvx_comb_line* __restrict__ lines, const float2* __restrict__ table_g,
               float2* __restrict__ proj) {
    __shared__ float2 stage[kLanes * kRow];
    __shared__ float2 sums[16][kLanes];
    __shared__ float2 table[kTableSize];  // the tone table, gathered from LDS
    for (uint32_t i = threadIdx.x; i < kTableSize; i += blockDim.x) table[i] = table_g[i];
    const uint32_t groups = (p.lines + group - 1) / group;
    const uint32_t s = blockIdx.x / groups;
    const uint32_t wave = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const uint32_t l = (blockIdx.x % groups) * group + wave;
    const bool live = l < p.lines;
    const rvx_comb_line line = lines[live ? l : 0];
    const uint32_t length = p.count / p.sub_blocks;
    const uint32_t begin = s * length;
    const uint32_t chunks = (length + kChunk - 1) / kChunk;
    float total_re = 0.0f, total_im = 0.0f;  // lane 0's running fold for this wave's line
    for (uint32_t c0 = 0; c0 < chunks; c0 += kLanes) {
        const uint32_t first = c0 * kChunk;  // offset of this tile in the sub-block
        const uint32_t n_tile = min(kLanes * kChunk, length - first);
        for (uint32_t k = threadIdx.x; k < n_tile; k += blockDim.x)
            stage[(k / kChunk) * kRow + (k % kChunk)] =
                raw[(p.first + begin + first + k) & p.ring_mask];
        __syncthreads();
        const uint32_t c = c0 + lane;
        float chunk_re = 0.0f, chunk_im = 0.0f;
        if (live && c < chunks) {
            const uint32_t m0 = c * kChunk;
            const uint32_t m1 = min(m0 + kChunk, length);
            const float2* row = stage + lane * kRow;
            // eight samples' LDS reads in flight per wait; the sums still run
            // in sample order
            uint32_t m = m0;
            for (; m + 8 <= m1; m += 8) {
                float2 xs[8], ts[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    xs[u] = row[m - m0 + u];
                    ts[u] = tone(table, line, begin + m + u);
                }
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    const float re = xs[u].x * ts[u].x + xs[u].y * ts[u].y;
                    const float im = xs[u].y * ts[u].x - xs[u].x * ts[u].y;
                    chunk_re = chunk_re + re;
                    chunk_im = chunk_im + im;
                }
            }
            for (; m < m1; ++m) {
                const float2 x = row[m - m0];
                const float2 t = tone(table, line, begin + m);
                const float re = x.x * t.x + x.y * t.y;
                const float im = x.y * t.x - x.x * t.y;
                chunk_re = chunk_re + re;
                chunk_im = chunk_im + im;
            }
        }
        sums[wave][lane] = make_float2(chunk_re, chunk_im);
        __syncthreads();
        if (lane == 0 && live) {
            const uint32_t here = min(kLanes, chunks - c0);
            for (uint32_t k = 0; k < here; ++k) {
                total_re = total_re + sums[wave][k].x;
                total_im = total_im + sums[wave][k].y;
            }
        }
        __syncthreads();  // stage and sums are rewritten by the next tile
    }
    if (lane == 0 && live) proj[uint64_t(l) * p.sub_blocks + s] = make_float2(total_re, total_im);
}

// raw and cleaned may be the same ring: a lane reads and writes only its
// slots. A workgroup takes kSubPer * 256 samples (a lane's are 256 apart, so
// every pass over them is coalesced) and stages the tone table in LDS once
// for all of them; each lane runs its samples' sums side by side, lines in
// ascending order for every one.
constexpr uint32_t kSubThreads = 256, kSubPer = 16;

__global__ void __launch_bounds__(kSubThreads)
subtract_kernel(rvx_comb_params p, const float2* raw, const rvx_comb_line* __restrict__ lines,
                const float2* __restrict__ table_g, float2* cleaned,
                float2* __restrict__ removed) {
    __shared__ float2 table[kTableSize];
    for (uint32_t i = threadIdx.x; i < kTableSize; i += kSubThreads) table[i] = table_g[i];
    __syncthreads();
    const uint32_t n0 = blockIdx.x * (kSubThreads * kSubPer) + threadIdx.x;
    float sum_re[kSubPer], sum_im[kSubPer];
#pragma unroll
    for (uint32_t j = 0; j < kSubPer; ++j) sum_re[j] = sum_im[j] = -0.0f;
    for (uint32_t l = 0; l < p.lines; ++l) {
        const rvx_comb_line line = lines[l];
#pragma unroll
        for (uint32_t j = 0; j < kSubPer; ++j) {
            const uint32_t n = n0 + j * kSubThreads;
            const float2 t = table[(line.phase0 + n * line.step) >> kTableShift];
            const float re = line.amplitude[0] * t.x - line.amplitude[1] * t.y;
            const float im = line.amplitude[0] * t.y + line.amplitude[1] * t.x;
            sum_re[j] = sum_re[j] + re;
            sum_im[j] = sum_im[j] + im;
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < kSubPer; ++j) {
        const uint32_t n = n0 + j * kSubThreads;
        if (n >= p.count) break;
        const uint32_t slot = (p.first + n) & p.ring_mask;
        const float2 x = raw[slot];
        cleaned[slot] = p.lines == 0 ? x : make_float2(x.x - sum_re[j], x.y - sum_im[j]);
        removed[(p.first + n) & p.removed_mask] = make_float2(sum_re[j], sum_im[j]);
    }
}

}  // namespace

extern "C" {

uint32_t rvx_comb_step(double frequency_hz, int64_t rate) {
    if (rate <= 0 || !std::isfinite(frequency_hz)) return 0;
    const double cycle = 4294967296.0;
    double steps = std::round(frequency_hz / double(rate) * cycle);
    steps = std::fmod(steps, cycle);
    if (steps < 0.0) steps += cycle;
    return uint32_t(uint64_t(steps));
}

int32_t rvx_comb_table(float* out, uint64_t n_floats) {
    if (!out || n_floats != 2ull * kTableSize)
        return rvx::fail(RVX_EINVAL, "rvx_comb_table: wants %u floats", 2 * kTableSize);
    const double two_pi = 6.283185307179586476925286766559;
    for (uint32_t i = 0; i < kTableSize; ++i) {
        const double angle = two_pi * double(i) / double(kTableSize);
        out[2 * i] = float(std::cos(angle));
        out[2 * i + 1] = float(std::sin(angle));
    }
    return RVX_OK;
}

int32_t rvx_comb_check(const rvx_comb_params* params) {
    static const char* who = "rvx_comb_check";
    if (!params) return rvx::fail(RVX_EINVAL, "%s: null params", who);
    const rvx_comb_params& p = *params;
    if (!is_mask(p.ring_mask) || !is_mask(p.removed_mask))
        return rvx::fail(RVX_EINVAL, "%s: the ring and removed masks must each be a power of two "
                         "minus one", who);
    if (ring_of(p.removed_mask) > ring_of(p.ring_mask))
        return rvx::fail(RVX_EINVAL, "%s: the removed ring must be no larger than the sample ring",
                         who);
    if (p.lines > kMaxLines)
        return rvx::fail(RVX_EINVAL, "%s: %u lines is more than the %u the stage cancels", who,
                         p.lines, kMaxLines);
    if (p.sub_blocks < 1 || p.sub_blocks > kMaxSubBlocks)
        return rvx::fail(RVX_EINVAL, "%s: %u sub-blocks is outside [1, %u]", who, p.sub_blocks,
                         kMaxSubBlocks);
    if (p.count == 0 || p.count % p.sub_blocks != 0)
        return rvx::fail(RVX_EINVAL, "%s: %u samples do not split into %u equal non-empty "
                         "sub-blocks", who, p.count, p.sub_blocks);
    if (p.count > ring_of(p.removed_mask))
        return rvx::fail(RVX_EINVAL, "%s: %u samples do not fit the removed ring of %llu", who,
                         p.count, (unsigned long long)ring_of(p.removed_mask));
    return RVX_OK;
}

int32_t rvx_comb_project(rvx_ctx* ctx, const rvx_comb_params* params, const void* raw,
                         uint64_t raw_samples, const rvx_comb_line* lines, uint64_t n_lines,
                         const void* table, uint64_t n_table, void* projections,
                         uint64_t n_projections) {
    static const char* who = "rvx_comb_project";
    RVX_ENTER(ctx, who);
    if (int32_t rc = rvx_comb_check(params)) return rc;
    const rvx_comb_params p = *params;
    if (raw_samples < ring_of(p.ring_mask) || n_lines < p.lines || n_table != kTableSize ||
        n_projections < uint64_t(p.lines) * p.sub_blocks)
        return rvx::fail(RVX_EINVAL, "%s: a buffer is shorter than its parameters", who);
    if (p.lines == 0) return RVX_OK;
    if (!raw || !lines || !table || !projections)
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    // lines per workgroup: as many (up to 16) as keep 256+ workgroups in flight
    uint32_t group = 16;
    while (group > 1 && uint64_t((p.lines + group - 1) / group) * p.sub_blocks < 256) group /= 2;
    const uint32_t groups = (p.lines + group - 1) / group;
    project_kernel<<<dim3(groups * p.sub_blocks), dim3(group * kLanes), 0, ctx->stream>>>(
        p, group, static_cast<const float2*>(raw), lines, static_cast<const float2*>(table),
        static_cast<float2*>(projections));
    return rvx::check_launch(who);
}

int32_t rvx_comb_subtract(rvx_ctx* ctx, const rvx_comb_params* params, const void* raw,
                          uint64_t raw_samples, const rvx_comb_line* lines, uint64_t n_lines,
                          const void* table, uint64_t n_table, void* cleaned,
                          uint64_t cleaned_samples, void* removed, uint64_t removed_samples) {
    static const char* who = "rvx_comb_subtract";
    RVX_ENTER(ctx, who);
    if (int32_t rc = rvx_comb_check(params)) return rc;
    const rvx_comb_params p = *params;
    if (raw_samples < ring_of(p.ring_mask) || cleaned_samples < ring_of(p.ring_mask) ||
        n_lines < p.lines || n_table != kTableSize || removed_samples < ring_of(p.removed_mask))
        return rvx::fail(RVX_EINVAL, "%s: a buffer is shorter than its parameters", who);
    if (!raw || !table || !cleaned || !removed || (p.lines && !lines))
        return rvx::fail(RVX_EINVAL, "%s: null buffer", who);
    constexpr uint32_t per_group = kSubThreads * kSubPer;
    subtract_kernel<<<dim3((p.count + per_group - 1) / per_group), dim3(kSubThreads), 0,
                      ctx->stream>>>(
        p, static_cast<const float2*>(raw), lines, static_cast<const float2*>(table),
        static_cast<float2*>(cleaned), static_cast<float2*>(removed));
    return rvx::check_launch(who);
}

}  // extern "C"
