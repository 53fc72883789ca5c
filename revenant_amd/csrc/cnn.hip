// The learned proposal network: five 3x3 convolutions (1 -> 32 -> 32 -> 32 ->
// 32 -> 1, ReLU between) over log-power spectrogram tiles, trained and run on
// the GPU by the kernels below. No counterpart in the reference; it stands in
// front of the plain detector (core/detect), which confirms what it proposes.
//
// Every convolution is an implicit GEMM on the matrix cores:
// v_mfma_f32_16x16x16_bf16 with bf16 operands and fp32 accumulation.
//   conv_kernel<CIN, COUTP, MODE>  forward (MODE 0: bias + ReLU -> bf16;
//       MODE 1: bias -> fp32 logits) and the data gradient (MODE 2: the same
//       product with the flipped, transposed weights, times the ReLU mask of
//       the layer's input -> bf16). A workgroup owns 64 pixels of four image
//       rows; its input rows (with the 3x3 halo) sit in LDS; each wave owns
//       one row and every output channel, its weight fragments in registers.
//   wgrad_kernel<CIN, COUTP>  the weight and bias gradients: per workgroup a
//       band of image rows, dW += dY^T * im2col(X) on the matrix cores,
//       partial sums written per workgroup and added in a fixed order.
//   loss_kernel, adam_kernel, pack/flip kernels.
// Activations are NHWC bf16; a layer's K index is (ky*3 + kx)*CIN + ci.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "common.h"

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kLayers = 5;
constexpr int kWidth = 32;
constexpr int kTileW = 64;         // pixels of a row in a convolution workgroup
constexpr int kConvRows = 4;       // rows of a convolution workgroup (one per wave)
constexpr int kWgradRows = 8;      // image rows per workgroup in the weight gradient

__host__ __device__ constexpr int round16(int v) { return (v + 15) / 16 * 16; }
// K of every product is padded to the 32 of v_mfma_f32_16x16x32_bf16.
__host__ __device__ constexpr int round32(int v) { return (v + 31) / 32 * 32; }

struct LayerShape {
    int cin, cout, coutp, k, kp;
};
__host__ __device__ constexpr LayerShape layer_shape(int l) {
    const int cin = l == 0 ? 1 : kWidth;
    const int cout = l == kLayers - 1 ? 1 : kWidth;
    return LayerShape{cin, cout, round16(cout), 9 * cin, round32(9 * cin)};
}

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

__device__ __forceinline__ f32x4 mma32(uint4 a, uint4 b, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
}

// MODE 0: Y = relu(conv(X, W) + b) as bf16, COUT channels per pixel.
// MODE 1: Z = conv(X, W) + b as fp32, one channel per pixel.
// MODE 2: dX = conv(dY, W') * (mask > 0) as bf16 (W' flipped/transposed).
//
// A workgroup owns kConvRows image rows of kTileW pixels; wave w owns row w
// (four 16-pixel M tiles) and every output channel. On
// v_mfma_f32_16x16x32_bf16 a lane's eight K values are k = 32 ks + 8 g + j;
// with the K index tap * CIN + ci and CIN a multiple of 8 they are one tap's
// eight consecutive channels, a 16-byte read of the input tile in LDS (its
// pixel stride padded by 8 channels so the 16 pixels of a read hit distinct
// banks). The wave's weight fragments stay in registers for the whole tile.
// CIN = 1 (the first layer) gathers its nine taps one value at a time.
template <int CIN, int COUTP, int MODE>
__global__ void __launch_bounds__(256)
conv_kernel(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
            const float* __restrict__ bias, const uint16_t* __restrict__ mask, int cout,
            int height, int width, void* __restrict__ out) {
    constexpr int K = 9 * CIN, KP = round32(K), NT = COUTP / 16, KS = KP / 32;
    constexpr bool kVec = CIN % 8 == 0;
    constexpr int PS = kVec ? CIN + 8 : CIN;  // pixel stride in the tile, bf16
    constexpr int TW = kTileW + 2, TR = kConvRows + 2;
    __shared__ __align__(16) uint16_t tile[TR * TW * PS];
    const int n = blockIdx.z, y0 = blockIdx.y * kConvRows, x0 = blockIdx.x * kTileW;
    const int tid = threadIdx.x;
    const uint16_t* img = x + size_t(n) * height * width * CIN;
    if constexpr (kVec) {
        constexpr int V = CIN / 8;  // 16-byte pieces per pixel
        for (int e = tid; e < TR * TW * V; e += 256) {
            const int v = e % V, c = (e / V) % TW, r = e / (V * TW);
            const int yy = y0 + r - 1, xx = x0 + c - 1;
            uint4 q = make_uint4(0, 0, 0, 0);
            if (yy >= 0 && yy < height && xx >= 0 && xx < width)
                q = *reinterpret_cast<const uint4*>(img + (size_t(yy) * width + xx) * CIN + 8 * v);
            *reinterpret_cast<uint4*>(tile + (r * TW + c) * PS + 8 * v) = q;
        }
    } else {
        for (int e = tid; e < TR * TW * CIN; e += 256) {
            const int ci = e % CIN, c = (e / CIN) % TW, r = e / (CIN * TW);
            const int yy = y0 + r - 1, xx = x0 + c - 1;
            tile[(r * TW + c) * PS + ci] = (yy >= 0 && yy < height && xx >= 0 && xx < width)
                                               ? img[(size_t(yy) * width + xx) * CIN + ci]
                                               : uint16_t(0);
        }
    }
    const int wave = tid / 64, lane = tid % 64;
    const int row = lane % 16, grp = lane / 16;
    // this wave's weight fragments: column co = 16 nt + row, k = 32 ks + 8 grp + j
    uint4 wf[NT][KS];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            wf[t][ks] = *reinterpret_cast<const uint4*>(w + size_t(t * 16 + row) * KP + ks * 32 +
                                                         grp * 8);
    __syncthreads();

    const int y = y0 + wave;
    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int m = mt * 16 + row;  // pixel of this lane's A row
            uint4 a;
            if constexpr (kVec) {
                const int k0 = ks * 32 + grp * 8;
                const int tap = k0 / CIN, ci0 = k0 % CIN;
                if (k0 < K)
                    a = *reinterpret_cast<const uint4*>(
                        tile + ((wave + tap / 3) * TW + m + tap % 3) * PS + ci0);
                else
                    a = make_uint4(0, 0, 0, 0);
            } else {
                uint16_t v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = ks * 32 + grp * 8 + j;
                    v[j] = 0;
                    if (k < K) {
                        const int ci = k % CIN, tap = k / CIN;
                        v[j] = tile[((wave + tap / 3) * TW + m + tap % 3) * PS + ci];
                    }
                }
                a = make_uint4(uint32_t(v[0]) | (uint32_t(v[1]) << 16),
                               uint32_t(v[2]) | (uint32_t(v[3]) << 16),
                               uint32_t(v[4]) | (uint32_t(v[5]) << 16),
                               uint32_t(v[6]) | (uint32_t(v[7]) << 16));
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[mt][t] = mma32(a, wf[t][ks], acc[mt][t]);
        }
    }
    if (y >= height) return;
    // D layout: lane holds rows 4*grp + j (pixels) of column `row` (channel).
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + row;
            if (co >= cout) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int xx = x0 + mt * 16 + grp * 4 + j;
                if (xx >= width) continue;
                const size_t p = (size_t(n) * height + y) * width + xx;
                float v = acc[mt][t][j];
                if constexpr (MODE == 0) {
                    v = v + bias[co];
                    static_cast<uint16_t*>(out)[p * cout + co] = f2bf(v > 0.f ? v : 0.f);
                } else if constexpr (MODE == 1) {
                    static_cast<float*>(out)[p] = v + bias[co];
                } else {
                    const bool on = bf2f(mask[p * cout + co]) > 0.f;
                    static_cast<uint16_t*>(out)[p * cout + co] = f2bf(on ? v : 0.f);
                }
            }
        }
}

// Partial weight/bias gradients of one band of kWgradRows image rows:
// part[COUTP][KP] = sum over the band's pixels of dY[p][co] * im2col(X)[p][k],
// bias_part[COUTP] = sum of dY[p][co]. dY is bf16 with COUTP channels.
template <int CIN, int COUTP>
__global__ void __launch_bounds__(256)
wgrad_kernel(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy, int height,
             int width, float* __restrict__ part, float* __restrict__ bias_part) {
    constexpr int K = 9 * CIN, KP = round32(K), MT = COUTP / 16, NTILES = KP / 16;
    constexpr int TILES = MT * NTILES;
    constexpr int PER_WAVE = (TILES + 3) / 4;
    __shared__ uint16_t xin[3][kTileW + 2][CIN];
    __shared__ uint16_t dyl[kTileW][COUTP];
    const int n = blockIdx.y, band = blockIdx.x;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int row = lane % 16, grp = lane / 16;
    const uint16_t* img = x + size_t(n) * height * width * CIN;
    float4_t acc[PER_WAVE];
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) acc[i] = float4_t{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;  // thread tid < COUTP sums channel tid
    for (int r = 0; r < kWgradRows; ++r) {
        const int y = band * kWgradRows + r;
        if (y >= height) break;
        for (int x0 = 0; x0 < width; x0 += kTileW) {
            __syncthreads();
            for (int e = tid; e < 3 * (kTileW + 2) * CIN; e += 256) {
                const int ci = e % CIN, c = (e / CIN) % (kTileW + 2), rr = e / (CIN * (kTileW + 2));
                const int yy = y + rr - 1, xx = x0 + c - 1;
                xin[rr][c][ci] = (yy >= 0 && yy < height && xx >= 0 && xx < width)
                                     ? img[(size_t(yy) * width + xx) * CIN + ci]
                                     : uint16_t(0);
            }
            for (int e = tid; e < kTileW * COUTP; e += 256) {
                const int c = e / COUTP, co = e % COUTP;
                const int xx = x0 + c;
                dyl[c][co] = xx < width ? dy[((size_t(n) * height + y) * width + xx) * COUTP + co]
                                        : uint16_t(0);
            }
            __syncthreads();
            if (tid < COUTP)
                for (int c = 0; c < kTileW; ++c) bsum += bf2f(dyl[c][tid]);
#pragma unroll
            for (int i = 0; i < PER_WAVE; ++i) {
                const int tile_id = wave * PER_WAVE + i;
                if (tile_id >= TILES) break;
                const int mt = tile_id / NTILES, nt = tile_id % NTILES;
#pragma unroll
                for (int ks = 0; ks < kTileW / 32; ++ks) {
                    uint32_t a[4], b[4];
#pragma unroll
                    for (int j = 0; j < 8; j += 2) {
                        uint16_t av[2], bv[2];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int p = ks * 32 + grp * 8 + j + h;  // pixel (the GEMM's K)
                            av[h] = dyl[p][mt * 16 + row];
                            const int k = nt * 16 + row;
                            bv[h] = 0;
                            if (k < K) {
                                const int ci = k % CIN, tap = k / CIN;
                                bv[h] = xin[tap / 3][p + tap % 3][ci];
                            }
                        }
                        a[j / 2] = uint32_t(av[0]) | (uint32_t(av[1]) << 16);
                        b[j / 2] = uint32_t(bv[0]) | (uint32_t(bv[1]) << 16);
                    }
                    acc[i] = mma32(make_uint4(a[0], a[1], a[2], a[3]),
                                   make_uint4(b[0], b[1], b[2], b[3]), acc[i]);
                }
            }
        }
    }
    const size_t base = (size_t(n) * gridDim.x + band) * (COUTP * KP);
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) {
        const int tile_id = wave * PER_WAVE + i;
        if (tile_id >= TILES) break;
        const int mt = tile_id / NTILES, nt = tile_id % NTILES;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int co = mt * 16 + grp * 4 + j, k = nt * 16 + row;
            part[base + size_t(co) * KP + k] = acc[i][j];
        }
    }
    if (tid < COUTP) bias_part[(size_t(n) * gridDim.x + band) * COUTP + tid] = bsum;
}

// grad[i] = sum over parts (in order) of part[q][i]; i < count.
__global__ void reduce_parts(const float* __restrict__ part, int parts, int count,
                             float* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int q = 0; q < parts; ++q) s = s + part[size_t(q) * count + i];
    grad[i] = s;
}

// Binary cross-entropy with logits, positives weighted by pos_weight; the
// gradient (already divided by the pixel count) goes out as bf16 in channel 0
// of a 16-channel dY (the last layer's padded width).
__global__ void loss_kernel(const float* __restrict__ z, const float* __restrict__ label,
                            int count, float pos_weight, float* __restrict__ loss_part,
                            uint16_t* __restrict__ dz) {
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (i < count) {
        const float zi = z[i], yi = label[i];
        const float wgt = yi > 0.5f ? pos_weight : 1.f;
        const float az = fabsf(zi);
        l = wgt * (fmaxf(zi, 0.f) - zi * yi + log1pf(expf(-az)));
        const float sig = zi >= 0.f ? 1.f / (1.f + expf(-zi)) : expf(zi) / (1.f + expf(zi));
        dz[size_t(i) * 16] = f2bf(wgt * (sig - yi) / float(count));
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Adam on the fp32 master parameters.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int count, float lr,
                            float b1, float b2, float c1, float c2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Master weights [cout][9][cin] -> packed bf16 [coutp][kp], zero padded.
__global__ void pack_kernel(const float* __restrict__ wm, int cin, int cout, int coutp, int kp,
                            uint16_t* __restrict__ packed) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= coutp * kp) return;
    const int co = e / kp, k = e % kp;
    packed[e] = (co < cout && k < 9 * cin) ? f2bf(wm[size_t(co) * 9 * cin + k]) : uint16_t(0);
}

// Data-gradient weights: W'[ci][tap'][co] = W[co][8 - tap'][ci], packed
// [cinp][kp'] with k' = tap' * coutp + co (the gradient's channel stride is
// the padded width coutp) and kp' = round16(9 * coutp).
__global__ void flip_kernel(const float* __restrict__ wm, int cin, int cout, int coutp, int cinp,
                            int kpp, uint16_t* __restrict__ packed) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cinp * kpp) return;
    const int ci = e / kpp, k = e % kpp;
    uint16_t v = 0;
    if (ci < cin && k < 9 * coutp) {
        const int co = k % coutp, tap = k / coutp;
        if (co < cout) v = f2bf(wm[(size_t(co) * 9 + (8 - tap)) * cin + ci]);
    }
    packed[e] = v;
}

// Reduced gradients in packed layout [coutp][kp] -> master [cout][9][cin] + [cout].
__global__ void unpack_grad_kernel(const float* __restrict__ gw, const float* __restrict__ gb,
                                   int cin, int cout, int kp, float* __restrict__ w,
                                   float* __restrict__ b) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = 9 * cin;
    if (e < cout * per) w[e] = gw[size_t(e / per) * kp + e % per];
    if (e < cout) b[e] = gb[e];
}

__global__ void to_bf16_kernel(const float* __restrict__ in, int count, uint16_t* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = f2bf(in[i]);
}

template <int CIN, int COUTP, int MODE>
void conv(const uint16_t* x, const uint16_t* w, const float* b, const uint16_t* mask, int cout,
          int batch, int h, int wd, void* out, hipStream_t s) {
    conv_kernel<CIN, COUTP, MODE><<<dim3((wd + kTileW - 1) / kTileW, (h + kConvRows - 1) / kConvRows,
                                         batch),
                                    dim3(256), 0, s>>>(
        x, w, b, mask, cout, h, wd, out);
}

}  // namespace

struct rvx_net {
    rvx_ctx* ctx = nullptr;
    uint32_t h = 0, w = 0, max_batch = 0;
    uint32_t nparams = 0;
    uint32_t woff[kLayers], boff[kLayers];  // offsets of each layer's W and b in params
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* packed[kLayers] = {};
    uint16_t* flipped[kLayers] = {};
    uint16_t* act[kLayers] = {};   // act[0]: input; act[l]: output of layer l-1 (bf16)
    float* logits = nullptr;
    uint16_t* grad_act[2] = {};    // ping-pong data gradients (bf16, 32 channels)
    uint16_t* dz = nullptr;        // last layer's dY, 16 channels
    float* part = nullptr;
    float* bias_part = nullptr;
    float* loss_part = nullptr;
    uint32_t step = 0;

    ~rvx_net() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (void* p : {(void*)params, (void*)grads, (void*)m, (void*)v, (void*)logits,
                        (void*)dz, (void*)part, (void*)bias_part, (void*)loss_part,
                        (void*)grad_act[0], (void*)grad_act[1]})
            if (p) (void)hipFree(p);
        for (int l = 0; l < kLayers; ++l) {
            if (packed[l]) (void)hipFree(packed[l]);
            if (flipped[l]) (void)hipFree(flipped[l]);
            if (act[l]) (void)hipFree(act[l]);
        }
    }
};

namespace {

int32_t repack(rvx_net& n, hipStream_t s) {
    for (int l = 0; l < kLayers; ++l) {
        const LayerShape sh = layer_shape(l);
        pack_kernel<<<dim3((sh.coutp * sh.kp + 255) / 256), dim3(256), 0, s>>>(
            n.params + n.woff[l], sh.cin, sh.cout, sh.coutp, sh.kp, n.packed[l]);
        if (l > 0) {
            const int cinp = round16(sh.cin), kpp = round32(9 * sh.coutp);
            flip_kernel<<<dim3((cinp * kpp + 255) / 256), dim3(256), 0, s>>>(
                n.params + n.woff[l], sh.cin, sh.cout, sh.coutp, cinp, kpp, n.flipped[l]);
        }
    }
    return rvx::check_launch("rvx_net");
}

void forward(rvx_net& n, uint32_t batch, hipStream_t s) {
    const int h = int(n.h), w = int(n.w), b = int(batch);
    conv<1, 32, 0>(n.act[0], n.packed[0], n.params + n.boff[0], nullptr, kWidth, b, h, w, n.act[1], s);
    for (int l = 1; l < kLayers - 1; ++l)
        conv<32, 32, 0>(n.act[l], n.packed[l], n.params + n.boff[l], nullptr, kWidth, b, h, w,
                        n.act[l + 1], s);
    conv<32, 16, 1>(n.act[kLayers - 1], n.packed[kLayers - 1], n.params + n.boff[kLayers - 1],
                    nullptr, 1, b, h, w, n.logits, s);
}

}  // namespace

extern "C" {

int32_t rvx_net_create(rvx_ctx* ctx, uint32_t height, uint32_t width, uint32_t max_batch,
                       uint64_t seed, rvx_net** out) {
    static const char* who = "rvx_net_create";
    RVX_ENTER(ctx, who);
    if (!out) return rvx::fail(RVX_EINVAL, "%s: null output", who);
    *out = nullptr;
    if (height == 0 || width == 0 || max_batch == 0 || height > 4096 || width > 4096 ||
        uint64_t(height) * width * max_batch > (1ull << 26))
        return rvx::fail(RVX_EINVAL, "%s: %ux%u tiles, batch %u", who, height, width, max_batch);
    auto n = new rvx_net();
    n->ctx = ctx;
    n->h = height;
    n->w = width;
    n->max_batch = max_batch;
    uint32_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
        const LayerShape sh = layer_shape(l);
        n->woff[l] = off;
        off += uint32_t(sh.cout * 9 * sh.cin);
        n->boff[l] = off;
        off += uint32_t(sh.cout);
    }
    n->nparams = off;
    // He-normal weights, zero biases, from a seeded generator on the host.
    std::vector<float> init(off, 0.f);
    std::mt19937_64 gen(seed);
    for (int l = 0; l < kLayers; ++l) {
        const LayerShape sh = layer_shape(l);
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (9.0 * sh.cin)));
        for (int i = 0; i < sh.cout * 9 * sh.cin; ++i) init[n->woff[l] + i] = float(nd(gen));
    }
    const size_t pix = size_t(height) * width * max_batch;
    const uint32_t bands = (height + kWgradRows - 1) / kWgradRows;
    size_t part_floats = 0;  // every band's partial plus the reduced sum after them
    for (int l = 0; l < kLayers; ++l) {
        const LayerShape sh = layer_shape(l);
        part_floats = std::max(part_floats, (size_t(bands) * max_batch + 1) * sh.coutp * sh.kp);
    }
    hipError_t err = hipSuccess;
    auto alloc = [&](auto** p, size_t bytes) {
        if (err == hipSuccess) err = hipMalloc((void**)p, bytes);
    };
    alloc(&n->params, off * sizeof(float));
    alloc(&n->grads, off * sizeof(float));
    alloc(&n->m, off * sizeof(float));
    alloc(&n->v, off * sizeof(float));
    for (int l = 0; l < kLayers; ++l) {
        const LayerShape sh = layer_shape(l);
        alloc(&n->packed[l], size_t(sh.coutp) * sh.kp * 2);
        if (l > 0) alloc(&n->flipped[l], size_t(round16(sh.cin)) * round32(9 * sh.coutp) * 2);
        alloc(&n->act[l], pix * size_t(l == 0 ? 1 : kWidth) * 2);
    }
    alloc(&n->logits, pix * sizeof(float));
    alloc(&n->grad_act[0], pix * kWidth * 2);
    alloc(&n->grad_act[1], pix * kWidth * 2);
    alloc(&n->dz, pix * 16 * 2);
    alloc(&n->part, part_floats * sizeof(float));
    alloc(&n->bias_part, (size_t(bands) * max_batch + 1) * 32 * sizeof(float));
    alloc(&n->loss_part, ((pix + 255) / 256) * sizeof(float));
    if (err == hipSuccess) err = hipMemcpy(n->params, init.data(), off * sizeof(float), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(n->m, 0, off * sizeof(float));
    if (err == hipSuccess) err = hipMemset(n->v, 0, off * sizeof(float));
    if (err == hipSuccess) err = hipMemset(n->dz, 0, pix * 16 * 2);
    if (err != hipSuccess) {
        delete n;
        return rvx::fail(RVX_ENOMEM, "%s: %s", who, hipGetErrorString(err));
    }
    if (int32_t rc = repack(*n, ctx->stream)) {
        delete n;
        return rc;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        delete n;
        return rvx::fail(RVX_EDEVICE, "%s: initialisation failed", who);
    }
    *out = n;
    return RVX_OK;
}

void rvx_net_destroy(rvx_net* net) { delete net; }

uint32_t rvx_net_param_count(const rvx_net* net) { return net ? net->nparams : 0; }

int32_t rvx_net_params(rvx_net* n, float* host, uint64_t count, int32_t write) {
    static const char* who = "rvx_net_params";
    if (!n) return rvx::fail(RVX_EINVAL, "%s: null net", who);
    RVX_ENTER(n->ctx, who);
    if (!host || count != n->nparams)
        return rvx::fail(RVX_EINVAL, "%s: wants %u floats", who, n->nparams);
    RVX_TRY_HIP(hipStreamSynchronize(n->ctx->stream), who);
    if (write) {
        RVX_TRY_HIP(hipMemcpy(n->params, host, count * sizeof(float), hipMemcpyHostToDevice), who);
        if (int32_t rc = repack(*n, n->ctx->stream)) return rc;
        RVX_TRY_HIP(hipStreamSynchronize(n->ctx->stream), who);
    } else {
        RVX_TRY_HIP(hipMemcpy(host, n->params, count * sizeof(float), hipMemcpyDeviceToHost), who);
    }
    return RVX_OK;
}

int32_t rvx_net_grads(rvx_net* n, float* host, uint64_t count) {
    static const char* who = "rvx_net_grads";
    if (!n) return rvx::fail(RVX_EINVAL, "%s: null net", who);
    RVX_ENTER(n->ctx, who);
    if (!host || count != n->nparams)
        return rvx::fail(RVX_EINVAL, "%s: wants %u floats", who, n->nparams);
    RVX_TRY_HIP(hipStreamSynchronize(n->ctx->stream), who);
    RVX_TRY_HIP(hipMemcpy(host, n->grads, count * sizeof(float), hipMemcpyDeviceToHost), who);
    return RVX_OK;
}

// tiles: device bf16 [batch][h][w] (log-power, already scaled); logits:
// device fp32 [batch][h][w].
int32_t rvx_net_forward(rvx_net* n, const uint16_t* tiles, uint32_t batch, float* logits) {
    static const char* who = "rvx_net_forward";
    if (!n) return rvx::fail(RVX_EINVAL, "%s: null net", who);
    RVX_ENTER(n->ctx, who);
    if (!tiles || !logits || batch == 0 || batch > n->max_batch)
        return rvx::fail(RVX_EINVAL, "%s: batch %u (at most %u)", who, batch, n->max_batch);
    const hipStream_t s = n->ctx->stream;
    const size_t pix = size_t(n->h) * n->w * batch;
    RVX_TRY_HIP(hipMemcpyAsync(n->act[0], tiles, pix * 2, hipMemcpyDeviceToDevice, s), who);
    forward(*n, batch, s);
    RVX_TRY_HIP(hipMemcpyAsync(logits, n->logits, pix * sizeof(float), hipMemcpyDeviceToDevice, s),
                who);
    return rvx::check_launch(who);
}

// One step: forward, loss, backward, and (when lr > 0) Adam. labels: device
// fp32 0/1 [batch][h][w]. *loss: the weighted mean loss (waits for it).
int32_t rvx_net_train_step(rvx_net* n, const uint16_t* tiles, const float* labels,
                           uint32_t batch, float lr, float pos_weight, float* loss) {
    static const char* who = "rvx_net_train_step";
    if (!n) return rvx::fail(RVX_EINVAL, "%s: null net", who);
    RVX_ENTER(n->ctx, who);
    if (!tiles || !labels || batch == 0 || batch > n->max_batch)
        return rvx::fail(RVX_EINVAL, "%s: batch %u (at most %u)", who, batch, n->max_batch);
    const hipStream_t s = n->ctx->stream;
    const int h = int(n->h), w = int(n->w), b = int(batch);
    const size_t pix = size_t(h) * w * batch;
    RVX_TRY_HIP(hipMemcpyAsync(n->act[0], tiles, pix * 2, hipMemcpyDeviceToDevice, s), who);
    forward(*n, batch, s);
    const int lblocks = int((pix + 255) / 256);
    loss_kernel<<<dim3(lblocks), dim3(256), 0, s>>>(n->logits, labels, int(pix), pos_weight,
                                                    n->loss_part, n->dz);
    const int bands = (h + kWgradRows - 1) / kWgradRows;
    const int parts = bands * b;
    // backward, last layer first: dY of layer l is `dy`; its input is act[l]
    const uint16_t* dy = n->dz;
    for (int l = kLayers - 1; l >= 0; --l) {
        const LayerShape sh = layer_shape(l);
        const int wcount = sh.coutp * sh.kp;
        if (l == kLayers - 1)
            wgrad_kernel<32, 16><<<dim3(bands, b), dim3(256), 0, s>>>(n->act[l], dy, h, w, n->part,
                                                                     n->bias_part);
        else if (l > 0)
            wgrad_kernel<32, 32><<<dim3(bands, b), dim3(256), 0, s>>>(n->act[l], dy, h, w, n->part,
                                                                     n->bias_part);
        else
            wgrad_kernel<1, 32><<<dim3(bands, b), dim3(256), 0, s>>>(n->act[l], dy, h, w, n->part,
                                                                    n->bias_part);
        // fold the partials into the packed-layout scratch, then to the master
        // layout [cout][9][cin] + [cout]
        reduce_parts<<<dim3((wcount + 255) / 256), dim3(256), 0, s>>>(n->part, parts, wcount,
                                                                      n->part + size_t(parts) * wcount);
        reduce_parts<<<dim3(1), dim3(64), 0, s>>>(n->bias_part, parts, sh.coutp,
                                                  n->bias_part + size_t(parts) * sh.coutp);
        const float* gw = n->part + size_t(parts) * wcount;
        const float* gb = n->bias_part + size_t(parts) * sh.coutp;
        const int wn = sh.cout * 9 * sh.cin;
        unpack_grad_kernel<<<dim3((wn + 255) / 256), dim3(256), 0, s>>>(
            gw, gb, sh.cin, sh.cout, sh.kp, n->grads + n->woff[l], n->grads + n->boff[l]);
        if (l == 0) break;
        // data gradient into the previous layer's output, masked by its ReLU
        uint16_t* dx = n->grad_act[l % 2];
        if (l == kLayers - 1)
            conv<16, 32, 2>(dy, n->flipped[l], nullptr, n->act[l], kWidth, b, h, w, dx, s);
        else
            conv<32, 32, 2>(dy, n->flipped[l], nullptr, n->act[l], kWidth, b, h, w, dx, s);
        dy = dx;
    }
    if (lr > 0.f) {
        ++n->step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, float(n->step)), c2 = 1.f - std::pow(b2, float(n->step));
        adam_kernel<<<dim3((n->nparams + 255) / 256), dim3(256), 0, s>>>(
            n->params, n->grads, n->m, n->v, int(n->nparams), lr, b1, b2, c1, c2, 1e-8f);
        if (int32_t rc = repack(*n, s)) return rc;
    }
    if (int32_t rc = rvx::check_launch(who)) return rc;
    if (loss) {
        std::vector<float> lp(lblocks);
        RVX_TRY_HIP(hipMemcpyAsync(lp.data(), n->loss_part, lblocks * sizeof(float),
                                   hipMemcpyDeviceToHost, s),
                    who);
        RVX_TRY_HIP(hipStreamSynchronize(s), who);
        double sum = 0.0;
        for (float v : lp) sum += v;
        *loss = float(sum / double(pix));
    }
    return RVX_OK;
}

int32_t rvx_to_bf16(rvx_ctx* ctx, const float* in, uint64_t count, uint16_t* out) {
    static const char* who = "rvx_to_bf16";
    RVX_ENTER(ctx, who);
    if (!in || !out || count > 0x7FFFFFFFull) return rvx::fail(RVX_EINVAL, "%s: bad buffers", who);
    if (count == 0) return RVX_OK;
    to_bf16_kernel<<<dim3(uint32_t((count + 255) / 256)), dim3(256), 0, ctx->stream>>>(
        in, int(count), out);
    return rvx::check_launch(who);
}

}  // extern "C"
