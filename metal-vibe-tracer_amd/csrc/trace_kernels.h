// trace_kernels.h -- launch interface between the host runtime and the traversal kernels.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace vt {

struct RayArgs {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
};
struct HitArgs {
    float *t, *b1, *b2;
    int32_t *face;
    uint32_t *instance;
};
struct TraceScene {
    const void *nodes;  // Node8 records (bvh8.h)
    const void *tris;   // TriRecord records
    uint32_t node_count;
    uint32_t depth;  // tree depth (root = 1); the traversal stack holds depth - 1 groups
    float root_max;
    // Device word raised (bit 0) by a ray whose node-group stack filled: the ray ends early with the
    // hit found so far.  The launcher sizes the stack to depth - 1 groups, which a traversal never
    // exceeds (one pending group per level), so only a stack capped below that (stack_cap, a test
    // hook; 0 = no cap) can raise it.
    uint32_t *fault;
    uint32_t stack_cap;
    // 0: Node8 records (the 8-wide tree); 1: MeshNode records (the reference's flat binary median
    // hierarchy, bvh8.h), walked by trace_median_kernel with a stack of kMedianStack entries.
    uint32_t kind = 0;
};
constexpr uint32_t kMedianStack = 32;

// Two-level scenes (the reference's MeshSceneLayout, Sources/Acceleration.swift:356-573): an 8-wide
// hierarchy over the instances' world bounds whose leaf slots name InstanceRecords (in leaf order),
// each tracing its asset's hierarchy with the object-space ray.
struct alignas(16) InstanceRecord {
    float local[3][4];  // rows of the world-to-object matrix: x = row.xyz . p + row.w
    const void *nodes;  // the asset's Node8 records
    const void *tris;   // the asset's TriRecord records
    float root_max;     // the asset's max |local coordinate| (its box enlargement term)
    uint32_t index;     // the instance's position in the caller's list
    uint32_t faces;     // the asset's face count (0: nothing to trace)
    uint32_t pad;
};
static_assert(sizeof(InstanceRecord) == 80, "InstanceRecord is five 16-byte words");
struct TlasScene {
    const void *nodes;                  // Node8 records over instance bounds
    const InstanceRecord *instances;    // leaf order
    uint32_t node_count, depth;         // of the instance hierarchy
    uint32_t blas_depth;                // deepest asset hierarchy
    float enlarge, reach;               // per-ray growth of the instance boxes: enlarge (|origin|max + reach)
    uint32_t *fault;                    // raised (bit 0) by a ray whose node-group stack filled
};
// Closest hit (hits.instance = the instance's index, 0xffffffff on a miss) or any hit over a
// two-level scene.  Among hits at exactly equal t the smaller instance index wins, then the smaller
// face index.
hipError_t launch_trace_tlas(const TlasScene &scene, const RayArgs &rays, const HitArgs &hits, uint8_t *occluded,
                             uint32_t n, bool any, hipStream_t stream);

// Closest hit (any = false) into `hits`, or any hit into `occluded`.  `counters` (device, 128
// u64, zeroed by the caller) selects the counting variant: 64 pairs (node fetches, triangle tests).
// `count` (device, optional) caps n at launch time: the wavefront integrator's queue length.
hipError_t launch_trace(const TraceScene &scene, const RayArgs &rays, const HitArgs &hits, uint8_t *occluded,
                        uint32_t n, bool any, unsigned long long *counters, hipStream_t stream,
                        const uint32_t *count = nullptr);

// The counting traversal of launch_trace (counters variant) with its partial counters in the
// library's static device scratch (kStatsScratchWords u64, 64 pairs) instead of an allocation:
// zeroes the scratch with a kernel, traces into the caller's hits / occluded, and sums the pairs into
// out[0] (node fetches) and out[1] (triangle tests), device memory.  Asynchronous; allocates nothing.
// scratch_init(device) must have run once with `device` current.  Calls that overlap on one device
// race on the scratch.
constexpr int kStatsScratchWords = 128;
hipError_t scratch_init(int device);
hipError_t launch_trace_stats_scratch(int device, const TraceScene &scene, const RayArgs &rays, const HitArgs &hits,
                                      uint8_t *occluded, uint32_t n, bool any, unsigned long long *out,
                                      hipStream_t stream);

// Batched rendering: ray i traces the scene job_trace[job[i]], for i < min(n, *count), on a
// persistent grid of `grid` blocks of `block` (64, 128, 256 or 512) lanes; max_depth is the deepest tree.
hipError_t launch_trace_batch(const TraceScene *job_trace, const uint32_t *job, const RayArgs &rays, const HitArgs &hits,
                              uint8_t *occluded, uint32_t n, bool any, const uint32_t *count, uint32_t max_depth,
                              uint32_t block, uint32_t grid, hipStream_t stream);

}  // namespace vt
