// zsampler.hip -- the Z++ sampler's draws for every pixel of an image (vt_z_draws).
#include <hip/hip_runtime.h>

#include "scene.h"
#include "vt_sampler.h"
#include "zsampler.h"

namespace vt {
namespace {

__global__ __launch_bounds__(256) void z_draws_kernel(vt_z_query q, float *out) {
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= q.width * q.height) return;
    const uint32_t x = pix % q.width, y = pix / q.width;
    ZSampler s{z_pixel_key(x, y, q.frame_index, q.sample_index, q.temporal_model), 0u, 0u, 0u};
    if (q.candidate_bits) s.candidate(q.path_vertex, q.stream, q.candidate, q.candidate_bits);
    else s.event(q.path_vertex, q.stream);
    float *o = out + (size_t)pix * q.draws * q.kind;
    for (uint32_t k = 0; k < q.draws; ++k) {
        if (q.kind == 1u) {
            o[k] = s.rand1();
        } else if (q.kind == 2u) {
            s.rand2(o[2 * k], o[2 * k + 1]);
        } else {
            s.rand3(o[3 * k], o[3 * k + 1], o[3 * k + 2]);
        }
    }
}

}  // namespace
}  // namespace vt

extern "C" vt_status vt_z_draws(const vt_z_query *q, float *out, void *stream) {
    if (!q || !out) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "query and out must not be NULL");
    if (q->kind < 1u || q->kind > 3u) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "kind must be 1, 2 or 3");
    if (q->temporal_model > 3u) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "temporal_model must be 0..3");
    if (q->stream > 15u || q->draws > 32u) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "stream must be 0..15 and draws at most 32");
    if (q->candidate_bits > 18u || (q->candidate_bits && q->candidate >= (1u << q->candidate_bits)))
        return vt::fail(VT_ERROR_INVALID_ARGUMENT, "candidate must be below 2^candidate_bits, at most 2^18");
    const uint64_t n = (uint64_t)q->width * q->height;
    if (n == 0 || q->draws == 0) return VT_OK;
    if (n > (1ull << 28)) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "image too large");
    hipLaunchKernelGGL(vt::z_draws_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *q, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? VT_OK : vt::fail(VT_ERROR_DEVICE, hipGetErrorString(e));
}
