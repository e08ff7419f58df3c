// trace.hip -- gfx950 closest-hit and any-hit traversal kernels over SoA ray buffers.
//
// One lane per ray, 64-lane wavefronts, 128-lane blocks.  Ray components are read and hits
// written as separate arrays, so every wavefront access is one coalesced 256-byte line.  The
// per-ray traversal (traverse8.h) keeps its node-group stack in LDS, laid out [depth][lane] so
// every lane's push/pop touches its own bank.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trace_kernels.h"
#include "traverse8.h"

namespace vt {
namespace {

constexpr int kBlock = 128;

// A tree of depth D leaves at most D - 1 node groups pending; the launcher picks the smallest
// instantiated depth that covers the scene's tree (the builder bounds it by 24), so LDS per block,
// and with it the number of resident waves, follows the tree actually traced.
template <int kStack, int kBlk = kBlock>
struct LdsStack {
    uint32_t (*base)[kBlk];
    uint32_t (*bits)[kBlk];
    int lane;
    int sp;
    __device__ __forceinline__ void push(uint32_t b, uint32_t m) {
        base[sp][lane] = b;
        bits[sp][lane] = m;
        ++sp;
    }
    __device__ __forceinline__ bool full() const { return sp >= kStack; }
    __device__ __forceinline__ void pop(uint32_t &b, uint32_t &m) {
        --sp;
        b = base[sp][lane];
        m = bits[sp][lane];
    }
    __device__ __forceinline__ bool empty() const { return sp == 0; }
};

template <bool ANY, bool STATS, int kStack>
__global__ __launch_bounds__(kBlock) void trace_kernel(const U4 *__restrict__ nodes, const F4 *__restrict__ tris,
                                                       RayArgs rays, HitArgs hits, uint8_t *__restrict__ occluded,
                                                       uint32_t n, const uint32_t *__restrict__ count,
                                                       float root_max, unsigned long long *__restrict__ counters) {
    __shared__ uint32_t stack_base[kStack][kBlock];
    __shared__ uint32_t stack_bits[kStack][kBlock];
    const int lane = threadIdx.x;
    const uint32_t i = blockIdx.x * kBlock + lane;
    if (count) n = min(n, *count);
    if (i >= n) return;
    TraceRay r{rays.ox[i], rays.oy[i], rays.oz[i], rays.dx[i], rays.dy[i], rays.dz[i], rays.tmin[i], rays.tmax[i]};
    LdsStack<kStack> st{stack_base, stack_bits, lane, 0};
    TraceResult res = traverse<ANY, STATS>(nodes, tris, root_max, r, st);
    if (ANY) {
        occluded[i] = res.occluded ? 1 : 0;
    } else {
        hits.t[i] = res.t;
        hits.b1[i] = res.b1;
        hits.b2[i] = res.b2;
        hits.face[i] = res.face;
        if (hits.instance) hits.instance[i] = 0u;
    }
    if (STATS) {
        unsigned long long *c = counters + 2 * (blockIdx.x & 63u);
        atomicAdd(c, (unsigned long long)res.nodes);
        atomicAdd(c + 1, (unsigned long long)res.tris);
    }
}

template <bool ANY>
__global__ __launch_bounds__(256) void miss_kernel(RayArgs rays, HitArgs hits, uint8_t *occluded, uint32_t n,
                                                   const uint32_t *count) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (count) n = min(n, *count);
    if (i >= n) return;
    if (ANY) {
        occluded[i] = 0;
    } else {
        hits.t[i] = rays.tmax[i];
        hits.b1[i] = 0.0f;
        hits.b2[i] = 0.0f;
        hits.face[i] = -1;
        if (hits.instance) hits.instance[i] = 0u;
    }
}

// Batched rendering: rays of many scenes in one persistent launch.  Ray i belongs to job job[i],
// whose scene is job_trace[job[i]]; a fixed grid of kBlk-lane blocks strides over the device-side
// count.  Rays of one job are contiguous in the queues, so a wave rarely spans two scenes.
template <bool ANY, int kStack, int kBlk>
__global__ __launch_bounds__(kBlk) void batch_trace_kernel(const TraceScene *__restrict__ job_trace,
                                                           const uint32_t *__restrict__ job, RayArgs rays, HitArgs hits,
                                                           uint8_t *__restrict__ occluded, uint32_t n,
                                                           const uint32_t *__restrict__ count) {
    __shared__ uint32_t stack_base[kStack][kBlk];
    __shared__ uint32_t stack_bits[kStack][kBlk];
    const int lane = threadIdx.x;
    n = min(n, *count);
    for (uint32_t i = blockIdx.x * kBlk + lane; i < n; i += gridDim.x * kBlk) {
        const TraceScene &ts = job_trace[job[i]];
        TraceRay r{rays.ox[i], rays.oy[i], rays.oz[i], rays.dx[i], rays.dy[i], rays.dz[i], rays.tmin[i], rays.tmax[i]};
        TraceResult res;
        if (ts.node_count == 0) {
            res.t = r.tmax;
            res.b1 = res.b2 = 0.0f;
            res.face = -1;
            res.occluded = false;
        } else {
            LdsStack<kStack, kBlk> st{stack_base, stack_bits, lane, 0};
            res = traverse<ANY, false>(reinterpret_cast<const U4 *>(ts.nodes), reinterpret_cast<const F4 *>(ts.tris),
                                       ts.root_max, r, st);
        }
        if (ANY) {
            occluded[i] = res.occluded ? 1 : 0;
        } else {
            hits.t[i] = res.t;
            hits.b1[i] = res.b1;
            hits.b2[i] = res.b2;
            hits.face[i] = res.face;
        }
    }
}

}  // namespace

hipError_t launch_trace_batch(const TraceScene *job_trace, const uint32_t *job, const RayArgs &rays, const HitArgs &hits,
                              uint8_t *occluded, uint32_t n, bool any, const uint32_t *count, uint32_t max_depth,
                              uint32_t block, uint32_t grid, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const int need = max_depth > 1 ? (int)max_depth - 1 : 1;
    if (need > 24 || (block != 64 && block != 128 && block != 256) || grid == 0) return hipErrorInvalidValue;
#define VT_BATCH(S, B)                                                                                                 \
    do {                                                                                                               \
        if (any) hipLaunchKernelGGL((batch_trace_kernel<true, S, B>), dim3(grid), dim3(B), 0, stream, job_trace, job,   \
                                    rays, hits, occluded, n, count);                                                   \
        else hipLaunchKernelGGL((batch_trace_kernel<false, S, B>), dim3(grid), dim3(B), 0, stream, job_trace, job,      \
                                rays, hits, occluded, n, count);                                                       \
    } while (0)
#define VT_BATCH_STACK(S)                                                                                              \
    do {                                                                                                               \
        if (block == 64) VT_BATCH(S, 64);                                                                              \
        else if (block == 128) VT_BATCH(S, 128);                                                                       \
        else VT_BATCH(S, 256);                                                                                         \
    } while (0)
    if (need <= 8) VT_BATCH_STACK(8);
    else if (need <= 12) VT_BATCH_STACK(12);
    else if (need <= 16) VT_BATCH_STACK(16);
    else VT_BATCH_STACK(24);
#undef VT_BATCH_STACK
#undef VT_BATCH
    return hipGetLastError();
}

hipError_t launch_trace(const TraceScene &scene, const RayArgs &rays, const HitArgs &hits, uint8_t *occluded,
                        uint32_t n, bool any, unsigned long long *counters, hipStream_t stream, const uint32_t *count) {
    if (n == 0) return hipSuccess;
    if (scene.node_count == 0) {
        dim3 grid((n + 255) / 256);
        if (any) hipLaunchKernelGGL(miss_kernel<true>, grid, dim3(256), 0, stream, rays, hits, occluded, n, count);
        else hipLaunchKernelGGL(miss_kernel<false>, grid, dim3(256), 0, stream, rays, hits, occluded, n, count);
        return hipGetLastError();
    }
    dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
    const U4 *nd = reinterpret_cast<const U4 *>(scene.nodes);
    const F4 *tr = reinterpret_cast<const F4 *>(scene.tris);
    const int need = scene.depth > 1 ? (int)scene.depth - 1 : 1;
    if (need > 24) return hipErrorInvalidValue;
#define VT_LAUNCH(S)                                                                                                   \
    do {                                                                                                               \
        if (counters) {                                                                                                \
            if (any) hipLaunchKernelGGL((trace_kernel<true, true, S>), grid, block, 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, counters); \
            else hipLaunchKernelGGL((trace_kernel<false, true, S>), grid, block, 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, counters); \
        } else {                                                                                                       \
            if (any) hipLaunchKernelGGL((trace_kernel<true, false, S>), grid, block, 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, nullptr); \
            else hipLaunchKernelGGL((trace_kernel<false, false, S>), grid, block, 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, nullptr); \
        }                                                                                                              \
    } while (0)
    if (need <= 8) VT_LAUNCH(8);
    else if (need <= 12) VT_LAUNCH(12);
    else if (need <= 16) VT_LAUNCH(16);
    else VT_LAUNCH(24);
#undef VT_LAUNCH
    return hipGetLastError();
}

}  // namespace vt
