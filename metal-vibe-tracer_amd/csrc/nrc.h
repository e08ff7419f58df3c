// nrc.h -- the learned radiance cache's network: layout, kernels' launch interface.
//
// A query is 11 numbers in [0, 1]: position in the scene box (3), direction (d + 1) / 2 (3), normal
// (n + 1) / 2 (3), roughness, wavelength number.  The encoding is 12 frequency bands per input,
// feature 24 i + 2 b + {0 sin, 1 cos} of 2^b pi x_i, then a constant 1 (feature 264, the first
// layer's bias) and zeros to 272 = 17 x 16.  The network is a 4 x 64 ReLU MLP without biases:
// y = W5 relu(W4 relu(W3 relu(W2 relu(W1 e)))), W1 64 x 272, W2..W4 64 x 64, W5 16 x 64 of which rows
// 0..3 are the four wavelength lanes' radiance.  Parameters are stored row-major in that order
// (30720 numbers): fp32 master copy, Adam moments and gradients, and a bf16 copy the training kernel
// reads.  Inference reads a bf16 copy of an exponential moving average of the parameters over the
// Adam steps (decay kNrcEmaDecay), which averages out the step-to-step noise of online training.
// Loss: mean over queries and outputs of (y - t)^2 / (y^2 + 0.01) with y constant in the denominator.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace vt {

constexpr int kNrcInputs = 11, kNrcBands = 12, kNrcFeatures = 272, kNrcHidden = 64, kNrcOutPad = 16, kNrcOut = 4;
constexpr int kNrcW1 = kNrcHidden * kNrcFeatures;           // 17408
constexpr int kNrcWh = kNrcHidden * kNrcHidden;             // 4096
constexpr int kNrcW5 = kNrcOutPad * kNrcHidden;             // 1024
constexpr int kNrcParams = kNrcW1 + 3 * kNrcWh + kNrcW5;    // 30720
constexpr float kNrcEmaDecay = 0.95f;

struct NrcDevice {
    float *params, *grads, *m, *v, *ema;  // kNrcParams each
    uint16_t *wbf16, *wema;                // kNrcParams bf16 bits: params (training), ema (inference)
};

// out[q * 4 + c] = y_c(inputs of query q) with the averaged weights, for q < n, or for q < min(*count, n)
// when count is given (a device counter, e.g. a queue's length).  Inputs are [q][11] (stride 0) or
// [11][stride] (input k of query q at k * stride + q).
hipError_t launch_nrc_infer(const NrcDevice &d, const float *inputs, uint32_t n, float *out, hipStream_t s,
                            const uint32_t *count = nullptr, uint32_t stride = 0);
// A batch chosen on the device: slice `part` of `parts` equal slices of the first min(*count, limit)
// samples, each slice at most the launch's n samples (the rest of the samples are left out).
struct NrcSlice {
    const uint32_t *count = nullptr;
    uint32_t limit = 0, parts = 1, part = 0;
};
// grads += d loss / d params over samples [0, n), or over the slice when slice.count is given, the loss
// normalised by `norm` outputs (norm <= 0: 4 x the samples present); loss_sum (if not null) += the
// per-sample losses.
hipError_t launch_nrc_backward(const NrcDevice &d, const float *inputs, const float *targets, uint32_t n, float norm,
                               float *loss_sum, hipStream_t s, const NrcSlice &slice = NrcSlice{}, uint32_t stride = 0);
// Adam step `step` (1-based) on params from grads, then ema = decay ema + (1 - decay) params; refreshes
// both bf16 copies and clears grads.  With
// slice.count given, the step is skipped (grads cleared) when the slice of at most n samples is empty.
hipError_t launch_nrc_adam(const NrcDevice &d, uint32_t step, float lr, hipStream_t s, uint32_t n = 0,
                           const NrcSlice &slice = NrcSlice{});
// ema = params and both bf16 copies (after the params are set from the host).
hipError_t launch_nrc_sync(const NrcDevice &d, hipStream_t s);
hipError_t launch_nrc_zero(float *p, uint32_t n, hipStream_t s);

}  // namespace vt

struct vt_nrc;
namespace vt {
// Device view and the next Adam step number of a cache object (nrc.cpp).
NrcDevice nrc_device(vt_nrc *n);
uint32_t nrc_next_step(vt_nrc *n);
}  // namespace vt
