// nrc.hip -- gfx950 kernels of the learned radiance cache (nrc.h): fused encode + MLP inference,
// fused forward + backward with gradient accumulation, Adam with a moving average of the weights.
//
// All matrix products run on the matrix cores (v_mfma_f32_16x16x16_bf16: bf16 operands, fp32
// accumulation).  For D = A (16 x K) . B (K x 16) lane l supplies A[l % 16][k0 + 4 (l / 16) .. + 3] and
// B[k0 + 4 (l / 16) .. + 3][l % 16] and receives D[4 (l / 16) + j][l % 16], j = 0..3.  Weights are
// stored [out][in], so a weight fragment is four consecutive bf16.  Inference uses the weights as A
// and the activations as B, so a layer's output is already the next layer's operand and activations
// never leave registers; training keeps a block's 64 samples' activations in LDS for the backward
// pass (144 KB).  The bf16 network (60 KB) sits in LDS for the life of a block.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrc.h"

namespace vt {
namespace {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kOffW2 = kNrcW1, kOffW3 = kOffW2 + kNrcWh, kOffW4 = kOffW3 + kNrcWh, kOffW5 = kOffW4 + kNrcWh;

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ f4v mfma(s4v a, s4v b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ s4v load4(const uint16_t *p) { return *reinterpret_cast<const s4v *>(p); }
__device__ __forceinline__ s4v gather4(const uint16_t *p, int stride) {
    s4v r;
    r.x = (short)p[0];
    r.y = (short)p[stride];
    r.z = (short)p[2 * stride];
    r.w = (short)p[3 * stride];
    return r;
}

// Copies the bf16 network into LDS (all threads of the block).
__device__ __forceinline__ void load_weights(uint16_t *w, const uint16_t *src) {
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    uint4 *d = reinterpret_cast<uint4 *>(w);
    for (int i = threadIdx.x; i < kNrcParams / 8; i += blockDim.x) d[i] = s[i];
}

// sin and cos of pi 2^b x and of pi 2^(b+1) x, packed as four bf16 (feature order sin, cos, sin, cos).
// 2^(b-1) x is exact in fp32 and so is its fractional part, so the hardware sine (v_sin_f32 /
// v_cos_f32, whose argument is in turns) works on the exactly reduced angle.
__device__ __forceinline__ s4v band_pair(float x, int b) {
    const float t1 = __builtin_amdgcn_fractf(ldexpf(x, b - 1)), t2 = __builtin_amdgcn_fractf(ldexpf(x, b));
    s4v r;
    r.x = (short)to_bf16(__builtin_amdgcn_sinf(t1));
    r.y = (short)to_bf16(__builtin_amdgcn_cosf(t1));
    r.z = (short)to_bf16(__builtin_amdgcn_sinf(t2));
    r.w = (short)to_bf16(__builtin_amdgcn_cosf(t2));
    return r;
}
// Input k of query q: inputs are [q][11] (stride 0) or [11][stride].
__device__ __forceinline__ float query_input(const float *inputs, uint32_t stride, uint32_t q, int k) {
    return stride ? inputs[(size_t)k * stride + q] : inputs[(size_t)q * kNrcInputs + k];
}

// Encodes 16 queries (rows of x, stride 272) from inputs[q0..q0 + 15] (rows >= n are zeros).
__device__ __forceinline__ void encode_tile(uint16_t *x, const float *inputs, uint32_t stride, uint32_t q0, uint32_t n,
                                            int lane) {
    for (int idx = lane; idx < 16 * kNrcInputs * (kNrcBands / 2); idx += 64) {
        const int r = idx / (kNrcInputs * kNrcBands / 2), rem = idx - r * (kNrcInputs * kNrcBands / 2);
        const int i = rem / (kNrcBands / 2), b = 2 * (rem - i * (kNrcBands / 2));
        const uint32_t q = q0 + (uint32_t)r;
        s4v v = {0, 0, 0, 0};
        if (q < n) v = band_pair(query_input(inputs, stride, q, i), b);
        *reinterpret_cast<s4v *>(x + r * kNrcFeatures + 24 * i + 2 * b) = v;
    }
    for (int idx = lane; idx < 16 * 8; idx += 64) {
        const int r = idx >> 3, j = idx & 7;
        x[r * kNrcFeatures + 264 + j] = (j == 0 && q0 + (uint32_t)r < n) ? to_bf16(1.0f) : (uint16_t)0;
    }
}

// h_out (16 x 64, bf16) = relu(x (16 x K) . W^T), W stored [64][K].
template <int K>
__device__ __forceinline__ void dense_relu(const uint16_t *x, const uint16_t *w, uint16_t *h_out, int lane) {
    f4v acc[4] = {};
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        const s4v a = load4(x + r * K + k0 + kq);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[nt] = mfma(a, load4(w + (nt * 16 + r) * K + k0 + kq), acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) h_out[(kq + j) * kNrcHidden + nt * 16 + r] = to_bf16(fmaxf(acc[nt][j], 0.0f));
}
// y (16 x 16, fp32 in the D layout) = h (16 x 64) . W5^T, W5 stored [16][64].
__device__ __forceinline__ f4v dense_out(const uint16_t *h, const uint16_t *w5, int lane) {
    f4v acc = {};
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int k0 = 0; k0 < kNrcHidden; k0 += 16) acc = mfma(load4(h + r * kNrcHidden + k0 + kq), load4(w5 + r * kNrcHidden + k0 + kq), acc);
    return acc;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- inference
// The weights are the A operand and the activations the B operand (D^T = W . X^T), so a layer's
// output tile -- lane l holds neurons 16 m + 4 (l / 16) + j of query l % 16 -- is exactly the B
// operand of k-step m of the next layer: activations stay in registers from the encoding to the
// output.  A wave handles 64 queries (4 tiles of 16) per pass, so every weight fragment read from
// LDS feeds 4 products.  LDS rows are padded (276 / 68 bf16) so the 16 rows a quarter-wave reads
// fall in distinct banks.
constexpr int kW1Pitch = 276, kWhPitch = 68;
constexpr int kLdsW1 = 0, kLdsW2 = kLdsW1 + kNrcHidden * kW1Pitch, kLdsW3 = kLdsW2 + kNrcHidden * kWhPitch,
              kLdsW4 = kLdsW3 + kNrcHidden * kWhPitch, kLdsW5 = kLdsW4 + kNrcHidden * kWhPitch;
constexpr int kInferLds = (kLdsW5 + kNrcOutPad * kWhPitch) * 2;  // 63,616 bytes
#ifndef VT_NRC_INFER_TILES
#define VT_NRC_INFER_TILES 2
#endif
#ifndef VT_NRC_RELOAD
#define VT_NRC_RELOAD 1
#endif
constexpr int kInferTiles = VT_NRC_INFER_TILES;

__device__ __forceinline__ void load_weights_padded(uint16_t *w, const uint16_t *src) {
    for (int c = threadIdx.x; c < kNrcParams / 4; c += blockDim.x) {
        const int e = 4 * c;
        int dst;
        if (e < kNrcW1) {
            const int row = e / kNrcFeatures;
            dst = kLdsW1 + row * kW1Pitch + (e - row * kNrcFeatures);
        } else {
            const int f = e - kNrcW1, row = f / kNrcHidden;  // W2..W4 then W5: rows of 64 back to back
            dst = kLdsW2 + row * kWhPitch + (f - row * kNrcHidden);
        }
        *reinterpret_cast<uint2 *>(w + dst) = *reinterpret_cast<const uint2 *>(src + e);
    }
}

// relu and bf16 of a layer's output tiles: the next layer's B operands, h[k-step][query tile].
template <int T>
__device__ __forceinline__ void relu_pack(const f4v (&acc)[4][T], s4v (&h)[4][T]) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < T; ++t) {
            s4v v;
            v.x = (short)to_bf16(fmaxf(acc[m][t][0], 0.0f));
            v.y = (short)to_bf16(fmaxf(acc[m][t][1], 0.0f));
            v.z = (short)to_bf16(fmaxf(acc[m][t][2], 0.0f));
            v.w = (short)to_bf16(fmaxf(acc[m][t][3], 0.0f));
            h[m][t] = v;
        }
}

// One hidden layer: h (64 neurons x 16 T queries, B layout) -> relu(W h) in place.
template <int T>
__device__ __forceinline__ void hidden_layer(const uint16_t *w, s4v (&h)[4][T], int r, int g) {
    f4v acc[4][T] = {};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        s4v a[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) a[m] = load4(w + (16 * m + r) * kWhPitch + 16 * ks + 4 * g);
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int m = 0; m < 4; ++m) acc[m][t] = mfma(a[m], h[ks][t], acc[m][t]);
    }
    relu_pack<T>(acc, h);
}

// A wave handles 16 T queries per pass (T tiles of 16); blocks of kInferBlock lanes share one LDS copy
// of the network.
#ifndef VT_NRC_INFER_BLOCK
#define VT_NRC_INFER_BLOCK 512
#endif
constexpr int kInferBlock = VT_NRC_INFER_BLOCK, kInferWaves = kInferBlock / 64;
template <int T>
__global__ __launch_bounds__(kInferBlock) void nrc_infer_kernel(const uint16_t *wsrc, const float *inputs, uint32_t stride,
                                                                uint32_t n, float *out, const uint32_t *count) {
    extern __shared__ __align__(16) uint16_t lds[];
    if (count) n = min(n, *count);
    if (n == 0) return;
    load_weights_padded(lds, wsrc);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const uint32_t passes = (n + 16 * T - 1) / (16 * T);
    for (uint32_t p = blockIdx.x * kInferWaves + wave; p < passes; p += gridDim.x * kInferWaves) {
        const uint32_t base = p * 16 * T;
#if VT_NRC_RELOAD
        // An offset the compiler cannot see through: the hidden layers' weight fragments are read from
        // LDS each pass instead of being hoisted into (and occupying) 104 registers per lane.
        int zero = 0;
        asm volatile("" : "+v"(zero));
        const uint16_t *wl = lds + zero;
#else
        const uint16_t *wl = lds;
#endif
        bool live[T];
#pragma unroll
        for (int t = 0; t < T; ++t) live[t] = base + 16 * t + r < n;
        // Layer 1 over the 17 k-steps of the encoding: lane l's features f = 16 ks + 4 g .. + 3 are
        // bands b, b + 1 of input i = f / 24 (f < 264), or the constant 1 and zero padding.
        f4v acc[4][T] = {};
#pragma unroll 1
        for (int ks = 0; ks < kNrcFeatures / 16; ++ks) {
            s4v a[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) a[m] = load4(lds + kLdsW1 + (16 * m + r) * kW1Pitch + 16 * ks + 4 * g);
            const int f = 16 * ks + 4 * g;
            constexpr int kLast = kNrcInputs * 2 * kNrcBands;  // 264
            const int i = f < kLast ? f / 24 : 0, b = (f - 24 * i) / 2;
#pragma unroll
            for (int t = 0; t < T; ++t) {
                s4v bv = {0, 0, 0, 0};
                if (f < kLast) {
                    const float xv = live[t] ? query_input(inputs, stride, base + 16 * t + r, i) : 0.0f;
                    bv = band_pair(xv, b);
                } else if (f == kLast) {
                    bv.x = (short)to_bf16(1.0f);
                }
#pragma unroll
                for (int m = 0; m < 4; ++m) acc[m][t] = mfma(a[m], bv, acc[m][t]);
            }
        }
        s4v h[4][T];
        relu_pack<T>(acc, h);
        hidden_layer<T>(wl + kLdsW2, h, r, g);
        hidden_layer<T>(wl + kLdsW3, h, r, g);
        hidden_layer<T>(wl + kLdsW4, h, r, g);
        f4v y[T] = {};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const s4v a = load4(wl + kLdsW5 + r * kWhPitch + 16 * ks + 4 * g);
#pragma unroll
            for (int t = 0; t < T; ++t) y[t] = mfma(a, h[ks][t], y[t]);
        }
        // Lanes 0..15 hold outputs 0..3 of query 16 t + l.
        if (g == 0) {
#pragma unroll
            for (int t = 0; t < T; ++t)
                if (live[t]) *reinterpret_cast<float4 *>(out + 4 * (size_t)(base + 16 * t + r)) = make_float4(y[t][0], y[t][1], y[t][2], y[t][3]);
        }
    }
}

// ---------------------------------------------------------------- training (forward + backward)
// LDS: weights, a0 [64][272], a1..a4 [64][64] each, da ping-pong [2][64][64], dy [64][16] (bf16).
constexpr int kTrainLds = (kNrcParams + 64 * kNrcFeatures + 4 * 64 * kNrcHidden + 2 * 64 * kNrcHidden + 64 * kNrcOutPad) * 2;

// grads[off + o * in + i] += sum over the block's 64 samples of da[q][o] * a[q][i] for one 16 x 16
// tile (o0, i0); da has `out` columns, a has `in` columns.
__device__ __forceinline__ void weight_grad_tile(const uint16_t *da, int out, const uint16_t *a, int in, int o0, int i0,
                                                 float *grads, int lane) {
    f4v acc = {};
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int k0 = 0; k0 < 64; k0 += 16)
        acc = mfma(gather4(da + (k0 + kq) * out + o0 + r, out), gather4(a + (k0 + kq) * in + i0 + r, in), acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) atomicAdd(grads + (size_t)(o0 + kq + j) * in + i0 + r, acc[j]);
}
// da_out (the wave's 16 rows x 64) = (da_in (16 x K) . W) masked by a_prev > 0, W stored [K][64].
template <int K>
__device__ __forceinline__ void back_dense(const uint16_t *da_in, const uint16_t *w, const uint16_t *a_prev, uint16_t *da_out,
                                           int lane) {
    f4v acc[4] = {};
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        const s4v a = load4(da_in + r * K + k0 + kq);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[nt] = mfma(a, gather4(w + (k0 + kq) * kNrcHidden + nt * 16 + r, kNrcHidden), acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = kq + j, col = nt * 16 + r;
            const float g = from_bf16(a_prev[row * kNrcHidden + col]) > 0.0f ? acc[nt][j] : 0.0f;
            da_out[row * kNrcHidden + col] = to_bf16(g);
        }
}

// The samples [offset, offset + n) of a slice (NrcSlice), n capped by the launch's n.
__device__ __forceinline__ uint32_t slice_bounds(const NrcSlice &sl, uint32_t n, uint32_t &offset) {
    offset = 0;
    if (!sl.count) return n;
    const uint32_t c = min(*sl.count, sl.limit);
    const uint32_t per = min((c + sl.parts - 1) / sl.parts, n);
    offset = sl.part * per;
    return offset < c ? min(per, c - offset) : 0u;
}

__global__ __launch_bounds__(256) void nrc_backward_kernel(const uint16_t *wsrc, const float *inputs, uint32_t stride,
                                                           const float *targets, uint32_t n, float norm, float *grads,
                                                           float *loss_sum, NrcSlice slice) {
    extern __shared__ __align__(16) uint16_t lds[];
    uint32_t offset;
    n = slice_bounds(slice, n, offset);
    if (n == 0) return;
    inputs += stride ? (size_t)offset : (size_t)offset * kNrcInputs;
    targets += (size_t)offset * kNrcOut;
    if (!(norm > 0.0f)) norm = (float)n * kNrcOut;
    uint16_t *w = lds;
    uint16_t *a0 = w + kNrcParams;
    uint16_t *a1 = a0 + 64 * kNrcFeatures, *a2 = a1 + 64 * kNrcHidden, *a3 = a2 + 64 * kNrcHidden, *a4 = a3 + 64 * kNrcHidden;
    uint16_t *da0 = a4 + 64 * kNrcHidden, *da1 = da0 + 64 * kNrcHidden;
    uint16_t *dy = da1 + 64 * kNrcHidden;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    load_weights(w, wsrc);
    __syncthreads();
    const uint32_t groups = (n + 63) / 64;
    for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
        const uint32_t q0 = g * 64 + wave * 16;
        const int rows = wave * 16;
        // Forward for this wave's 16 samples, keeping every layer's input.
        encode_tile(a0 + rows * kNrcFeatures, inputs, stride, q0, n, lane);
        wave_sync();
        dense_relu<kNrcFeatures>(a0 + rows * kNrcFeatures, w, a1 + rows * kNrcHidden, lane);
        wave_sync();
        dense_relu<kNrcHidden>(a1 + rows * kNrcHidden, w + kOffW2, a2 + rows * kNrcHidden, lane);
        wave_sync();
        dense_relu<kNrcHidden>(a2 + rows * kNrcHidden, w + kOffW3, a3 + rows * kNrcHidden, lane);
        wave_sync();
        dense_relu<kNrcHidden>(a3 + rows * kNrcHidden, w + kOffW4, a4 + rows * kNrcHidden, lane);
        wave_sync();
        const f4v y = dense_out(a4 + rows * kNrcHidden, w + kOffW5, lane);
        // d loss / d y (relative L2, y constant in the denominator), zero for padded outputs and rows.
        const int c = lane & 15;
        float lsum = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = (lane >> 4) * 4 + j;
            const uint32_t q = q0 + (uint32_t)row;
            float d = 0.0f;
            if (c < kNrcOut && q < n) {
                const float t = targets[(size_t)q * kNrcOut + c], den = y[j] * y[j] + 0.01f;
                d = 2.0f * (y[j] - t) / den / norm;
                lsum += (y[j] - t) * (y[j] - t) / den;
            }
            dy[(rows + row) * kNrcOutPad + c] = to_bf16(d);
        }
        if (loss_sum) {
            for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
            if (lane == 0) atomicAdd(loss_sum, lsum);
        }
        __syncthreads();
        // Output layer: dW5 (16 x 64, one i-tile per wave) and da4 = dy . W5, masked.
        weight_grad_tile(dy, kNrcOutPad, a4, kNrcHidden, 0, wave * 16, grads + kOffW5, lane);
        back_dense<kNrcOutPad>(dy + rows * kNrcOutPad, w + kOffW5, a4 + rows * kNrcHidden, da0 + rows * kNrcHidden, lane);
        __syncthreads();
        // Hidden layers 4, 3, 2: dW (64 x 64: the wave's o-tile, every i-tile), then the next da.
        const uint16_t *acts[3] = {a3, a2, a1};
        const int offs[3] = {kOffW4, kOffW3, kOffW2};
        uint16_t *cur = da0, *nxt = da1;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
#pragma unroll
            for (int it = 0; it < 4; ++it) weight_grad_tile(cur, kNrcHidden, acts[l], kNrcHidden, wave * 16, it * 16, grads + offs[l], lane);
            back_dense<kNrcHidden>(cur + rows * kNrcHidden, w + offs[l], acts[l] + rows * kNrcHidden, nxt + rows * kNrcHidden, lane);
            __syncthreads();
            uint16_t *tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
        // First layer: dW1 (64 x 272): the wave's o-tile, all 17 i-tiles.
        for (int it = 0; it < kNrcFeatures / 16; ++it)
            weight_grad_tile(cur, kNrcHidden, a0, kNrcFeatures, wave * 16, it * 16, grads, lane);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void nrc_adam_kernel(float *params, float *grads, float *m, float *v, uint16_t *wbf16,
                                                       float *ema, uint16_t *wema,
                                                       float lr, float c1, float c2, uint32_t n, NrcSlice slice) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kNrcParams) return;
    uint32_t offset;
    if (slice.count && slice_bounds(slice, n, offset) == 0u) {
        grads[i] = 0.0f;
        return;
    }
    const float g = grads[i];
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.99f * v[i] + 0.01f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = params[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    params[i] = p;
    wbf16[i] = to_bf16(p);
    const float e = kNrcEmaDecay * ema[i] + (1.0f - kNrcEmaDecay) * p;
    ema[i] = e;
    wema[i] = to_bf16(e);
    grads[i] = 0.0f;
}

__global__ __launch_bounds__(256) void nrc_sync_kernel(const float *params, uint16_t *wbf16, float *ema, uint16_t *wema) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kNrcParams) return;
    const float p = params[i];
    wbf16[i] = to_bf16(p);
    ema[i] = p;
    wema[i] = to_bf16(p);
}

__global__ __launch_bounds__(256) void nrc_zero_kernel(float *p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Blocks for `work` items at `per_block` a block, at most `per_cu` blocks per compute unit.
int grid_for(uint32_t work, int per_block, int per_cu = 1) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t blocks = (work + per_block - 1) / per_block, cap = (uint32_t)(cus * per_cu);
    return (int)(blocks < cap ? (blocks > 0 ? blocks : 1) : cap);
}

}  // namespace

hipError_t launch_nrc_infer(const NrcDevice &d, const float *inputs, uint32_t n, float *out, hipStream_t s,
                            const uint32_t *count, uint32_t stride) {
    if (n == 0) return hipSuccess;
    static bool configured = false;
    if (!configured) {
        hipError_t e = hipFuncSetAttribute((const void *)nrc_infer_kernel<kInferTiles>, hipFuncAttributeMaxDynamicSharedMemorySize, kInferLds);
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL(nrc_infer_kernel<kInferTiles>, dim3(grid_for(n, 16 * kInferTiles * kInferWaves, 2)), dim3(kInferBlock),
                       kInferLds, s, d.wema, inputs, stride, n, out, count);
    return hipGetLastError();
}

hipError_t launch_nrc_backward(const NrcDevice &d, const float *inputs, const float *targets, uint32_t n, float norm,
                               float *loss_sum, hipStream_t s, const NrcSlice &slice, uint32_t stride) {
    if (n == 0) return hipSuccess;
    static bool configured = false;
    if (!configured) {
        hipError_t e = hipFuncSetAttribute((const void *)nrc_backward_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, kTrainLds);
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL(nrc_backward_kernel, dim3(grid_for(n, 64)), dim3(256), kTrainLds, s, d.wbf16, inputs, stride, targets, n, norm,
                       d.grads, loss_sum, slice);
    return hipGetLastError();
}

hipError_t launch_nrc_adam(const NrcDevice &d, uint32_t step, float lr, hipStream_t s, uint32_t n, const NrcSlice &slice) {
    const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.99f, (float)step);
    hipLaunchKernelGGL(nrc_adam_kernel, dim3((kNrcParams + 255) / 256), dim3(256), 0, s, d.params, d.grads, d.m, d.v, d.wbf16, d.ema, d.wema,
                       lr, c1, c2, n, slice);
    return hipGetLastError();
}

hipError_t launch_nrc_sync(const NrcDevice &d, hipStream_t s) {
    hipLaunchKernelGGL(nrc_sync_kernel, dim3((kNrcParams + 255) / 256), dim3(256), 0, s, d.params, d.wbf16, d.ema, d.wema);
    return hipGetLastError();
}

hipError_t launch_nrc_zero(float *p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(nrc_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace vt
