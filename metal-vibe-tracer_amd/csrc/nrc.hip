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
// v_mfma_f32_16x16x32_bf16 (gfx950: twice the K of the 16x16x16 form for the same issue cost).  The
// weights are the A operand and the activations the B operand (D^T = W . X^T).  Lane l = r + 16 g
// holds A row r / B column r and the eight k of group g; the sum over k does not depend on which
// hardware k a (group, element) pair stands for, as long as A and B agree, so each layer's inputs are
// numbered for the convenience of the data:
//   layer 1: k-step ks, group g carries feature quad Q = 9 g + ks, element e feature 8 Q + e of the
//            network's order (24 per input: sin and cos of bands 0..11), i.e. bands 4 (ks % 3) ..
//            4 (ks % 3) + 3 of input 3 g + ks / 3.  A lane thus needs three inputs of its query, and
//            the band group of every step is a compile-time constant.  Quads 33 .. 35 (lanes of
//            group 3 in steps 6 .. 8) have zero weights: the constant feature 264 is the bias,
//            which starts the accumulation instead (fp32 copy of its bf16 weights);
//   layers 2-5: k-step ks, group g, element e is input neuron 32 ks + 16 (e / 4) + 4 g + e % 4, the
//            neurons the lane already holds in its output tiles 2 ks and 2 ks + 1 (D layout: lane l
//            holds rows 4 g + j of a tile), so activations never leave registers.
// The LDS copy of each weight row is stored in that k order.  A wave handles 16 T queries per pass
// and loads the next pass's inputs while it computes the current one.
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef short s8v __attribute__((ext_vector_type(8)));

constexpr int kK1 = 288;                           // layer-1 k: 272 features padded to 9 steps of 32
constexpr int kW1Pitch = kK1 + 8, kWhPitch = 64 + 8;  // bf16; rows 592 / 144 bytes apart: the 16 rows
                                                      // a quarter-wave reads fall in distinct banks
constexpr int kLdsW1 = 0, kLdsW2 = kLdsW1 + kNrcHidden * kW1Pitch, kLdsW3 = kLdsW2 + kNrcHidden * kWhPitch,
              kLdsW4 = kLdsW3 + kNrcHidden * kWhPitch, kLdsW5 = kLdsW4 + kNrcHidden * kWhPitch;
constexpr int kLdsBias = kLdsW5 + kNrcOutPad * kWhPitch;          // 64 floats
constexpr int kInferLds = kLdsBias * 2 + kNrcHidden * 4;          // 68,096 bytes
constexpr int kInferTiles = 2;

// Input neuron of position p (0..63) of a permuted hidden-layer row.
__host__ __device__ constexpr int hidden_input(int p) {
    return 32 * (p >> 5) + 16 * ((p & 7) >> 2) + 4 * ((p >> 3) & 3) + (p & 3);
}
// The weight copy moves positions 8 k .. 8 k + 7 as two runs of four inputs.
constexpr bool hidden_runs_agree() {
    for (int p = 0; p < 64; ++p)
        if (hidden_input(p) != 32 * (p >> 5) + 4 * ((p >> 3) & 3) + 16 * ((p & 7) >> 2) + (p & 3)) return false;
    return true;
}
static_assert(hidden_runs_agree(), "hidden-layer weight permutation");

// Copies the network into LDS in the permuted orders, all loads of a thread issued before its stores.
template <int kThreads>
__device__ __forceinline__ void load_weights_permuted(uint16_t *w, const uint16_t *src) {
    // Layer 1: position 32 ks + 8 g + e of a row holds feature 8 (9 g + ks) + e (below 264), else 0;
    // the weights of feature 264 (the constant 1) are the bias, as floats.  Hidden layers: positions
    // 8 k .. 8 k + 7 of a row hold inputs b .. b + 3 and b + 16 .. b + 19, b = 32 (k / 4) + 4 (k % 4).
    constexpr int kChunks1 = kNrcHidden * (kK1 / 8), kRowsH = 3 * kNrcHidden + kNrcOutPad, kChunksH = kRowsH * 8;
    constexpr int kI1 = (kChunks1 + kThreads - 1) / kThreads, kIH = (kChunksH + kThreads - 1) / kThreads;
    const int tid = threadIdx.x;
    uint4 v1[kI1], vh[kIH];
#pragma unroll
    for (int i = 0; i < kI1; ++i) {
        const int c = tid + i * kThreads, row = c / (kK1 / 8), chunk = c - row * (kK1 / 8);
        const int q = 9 * (chunk & 3) + (chunk >> 2);
        v1[i] = make_uint4(0u, 0u, 0u, 0u);
        if (c < kChunks1 && q < 33) v1[i] = *reinterpret_cast<const uint4 *>(src + row * kNrcFeatures + 8 * q);
    }
#pragma unroll
    for (int i = 0; i < kIH; ++i) {
        const int c = tid + i * kThreads, row = c >> 3, k = c & 7;
        if (c < kChunksH) {
            const uint16_t *r = src + kNrcW1 + row * 64 + 32 * (k >> 2) + 4 * (k & 3);
            const uint2 lo = *reinterpret_cast<const uint2 *>(r), hi = *reinterpret_cast<const uint2 *>(r + 16);
            vh[i] = make_uint4(lo.x, lo.y, hi.x, hi.y);
        }
    }
    const float bias = tid < kNrcHidden ? from_bf16(src[tid * kNrcFeatures + 264]) : 0.0f;
#pragma unroll
    for (int i = 0; i < kI1; ++i) {
        const int c = tid + i * kThreads, row = c / (kK1 / 8), chunk = c - row * (kK1 / 8);
        if (c < kChunks1) *reinterpret_cast<uint4 *>(w + kLdsW1 + row * kW1Pitch + 8 * chunk) = v1[i];
    }
#pragma unroll
    for (int i = 0; i < kIH; ++i) {
        const int c = tid + i * kThreads;
        if (c < kChunksH) *reinterpret_cast<uint4 *>(w + kLdsW2 + (c >> 3) * kWhPitch + 8 * (c & 7)) = vh[i];
    }
    if (tid < kNrcHidden) reinterpret_cast<float *>(w + kLdsBias)[tid] = bias;
}

__device__ __forceinline__ f4v mfma32(bf8v a, bf8v b, f4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ bf8v lds8(const uint16_t *p) { return *reinterpret_cast<const bf8v *>(p); }

// Features of bands b0 .. b0 + 3 of input value x: sin and cos of pi 2^b x per band, as bf16.
// 2^(b-1) x is exact in fp32 and so is its fractional part, so the hardware sine and cosine
// (argument in turns) work on the exactly reduced angle.
template <int B0>
__device__ __forceinline__ bf8v band_quad(float x) {
    bf8v r;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float t = __builtin_amdgcn_fractf(x * __builtin_ldexpf(1.0f, B0 + k - 1));
        r[2 * k] = (__bf16)__builtin_amdgcn_sinf(t);
        r[2 * k + 1] = (__bf16)__builtin_amdgcn_cosf(t);
    }
    return r;
}

// relu and bf16 of output tiles 2 ks and 2 ks + 1: the B operand of the next layer's k-step ks.
// Rounded first, then the relu as a signed 16-bit max with 0 (a bf16 with its sign bit set is a
// negative short): the same bits as rounding the relu.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef short s2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_bf16x2(float a, float b) {
    const s2v v = __builtin_bit_cast(s2v, __builtin_convertvector(f2v{a, b}, bf2v));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, s2v{0, 0}));
}
__device__ __forceinline__ bf8v relu_pair(f4v lo, f4v hi) {
    const uint4 u = {relu_bf16x2(lo[0], lo[1]), relu_bf16x2(lo[2], lo[3]), relu_bf16x2(hi[0], hi[1]), relu_bf16x2(hi[2], hi[3])};
    return __builtin_bit_cast(bf8v, u);
}

// One hidden layer: 64 neurons x 16 T queries in (acc, D layout) -> W . relu(acc) in acc.
template <int T>
__device__ __forceinline__ void hidden_layer(const uint16_t *w, f4v (&acc)[4][T], int r, int g) {
    bf8v h[2][T];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int t = 0; t < T; ++t) h[ks][t] = relu_pair(acc[2 * ks][t], acc[2 * ks + 1][t]);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const bf8v a = lds8(w + (16 * m + r) * kWhPitch + 32 * ks + 8 * g);
#pragma unroll
            for (int t = 0; t < T; ++t) acc[m][t] = mfma32(a, h[ks][t], ks ? acc[m][t] : f4v{0.0f, 0.0f, 0.0f, 0.0f});
        }
}

// Inputs 3 g .. 3 g + 2 (capped at the last) of query 16 T p + 16 t + r, zeros past n.
template <int T>
__device__ __forceinline__ void load_inputs(float (&x)[T][3], const float *inputs, uint32_t stride, uint32_t n, uint32_t p,
                                            int r, int g) {
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const uint64_t q = (uint64_t)p * (16 * T) + 16 * t + r;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int i = min(3 * g + c, kNrcInputs - 1);
            const float v = query_input(inputs, stride, (uint32_t)min(q, (uint64_t)n - 1), i);
            x[t][c] = q < n ? v : 0.0f;
        }
    }
}

// Layer-1 steps 3 c .. 3 c + 2 (input slot c of each query, band groups 0, 1, 2); the first
// accumulates onto the bias.
template <int T, bool kFirst>
__device__ __forceinline__ void layer1_steps(const uint16_t *lds, f4v (&acc)[4][T], const f4v (&bias)[4],
                                             const float (&x)[T][3], int c, int r, int g) {
    float xc[T];
#pragma unroll
    for (int t = 0; t < T; ++t) xc[t] = kFirst ? x[t][0] : c == 1 ? x[t][1] : x[t][2];
#pragma unroll
    for (int bg = 0; bg < 3; ++bg) {
        const int ks = 3 * c + bg;
        bf8v a[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) a[m] = lds8(lds + kLdsW1 + (16 * m + r) * kW1Pitch + 32 * ks + 8 * g);
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const bf8v bv = bg == 0 ? band_quad<0>(xc[t]) : bg == 1 ? band_quad<4>(xc[t]) : band_quad<8>(xc[t]);
#pragma unroll
            for (int m = 0; m < 4; ++m) acc[m][t] = mfma32(a[m], bv, kFirst && bg == 0 ? bias[m] : acc[m][t]);
        }
    }
}

// Blocks of kInferBlock lanes share one LDS copy of the network.
constexpr int kInferBlock = 512, kInferWaves = kInferBlock / 64;
template <int T>
__global__ __launch_bounds__(kInferBlock) void nrc_infer_kernel(const uint16_t *wsrc, const float *inputs, uint32_t stride,
                                                                uint32_t n, float *out, const uint32_t *count) {
    extern __shared__ __align__(16) uint16_t lds[];
    if (count) n = min(n, *count);
    if (n == 0) return;
    load_weights_permuted<kInferBlock>(lds, wsrc);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const uint32_t passes = (n + 16 * T - 1) / (16 * T), step = gridDim.x * kInferWaves;
    float xn[T][3];
    uint32_t p = blockIdx.x * kInferWaves + wave;
    load_inputs<T>(xn, inputs, stride, n, p, r, g);
    for (; p < passes; p += step) {
        float x[T][3];
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int c = 0; c < 3; ++c) x[t][c] = xn[t][c];
        load_inputs<T>(xn, inputs, stride, n, p + step, r, g);
        // Layer 1: steps ks = 3 c + bg take input slot c, bands 4 bg .. 4 bg + 3, from the bias.
        f4v acc[4][T], bias[4];
        {
            // Read each pass (an opaque zero offset keeps the compiler from holding the bias in
            // registers across the loop, which it needs for the hidden layers' weights).
            int zero = 0;
            asm volatile("" : "+v"(zero));
            const float *bl = reinterpret_cast<const float *>(lds + kLdsBias) + zero;
#pragma unroll
            for (int m = 0; m < 4; ++m) bias[m] = *reinterpret_cast<const f4v *>(bl + 16 * m + 4 * g);
        }
        layer1_steps<T, true>(lds, acc, bias, x, 0, r, g);
#pragma unroll 1
        for (int c = 1; c < 3; ++c) layer1_steps<T, false>(lds, acc, bias, x, c, r, g);
        hidden_layer<T>(lds + kLdsW2, acc, r, g);
        hidden_layer<T>(lds + kLdsW3, acc, r, g);
        hidden_layer<T>(lds + kLdsW4, acc, r, g);
        f4v y[T];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf8v a = lds8(lds + kLdsW5 + r * kWhPitch + 32 * ks + 8 * g);
#pragma unroll
            for (int t = 0; t < T; ++t)
                y[t] = mfma32(a, relu_pair(acc[2 * ks][t], acc[2 * ks + 1][t]), ks ? y[t] : f4v{0.0f, 0.0f, 0.0f, 0.0f});
        }
        // Lanes 0..15 hold outputs 0..3 of query 16 t + l.
        if (g == 0) {
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const uint64_t q = (uint64_t)p * (16 * T) + 16 * t + r;
                if (q < n) *reinterpret_cast<float4 *>(out + 4 * q) = make_float4(y[t][0], y[t][1], y[t][2], y[t][3]);
            }
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
