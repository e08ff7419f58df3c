// trace.hip -- gfx950 closest-hit and any-hit traversal kernels over SoA ray buffers.
//
// One lane per ray, 64-lane wavefronts.  Ray components are read and hits
// written as separate arrays, so every wavefront access is one coalesced 256-byte line.  The
// per-ray traversal (traverse8.h) keeps its node-group stack in LDS, laid out [depth][lane] so
// every lane's push/pop touches its own bank.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "trace_kernels.h"
#include "traverse8.h"

namespace vt {
namespace {

constexpr int kBlock = 128;
constexpr int kRefillIdle = 12;

// A tree of depth D leaves at most D - 1 node groups pending; the launcher picks the smallest
// instantiated depth that covers the scene's tree (the builder bounds it by 24), so LDS per block,
// and with it the number of resident waves, follows the tree actually traced.
template <int kStack, int kBlk = kBlock>
struct LdsStack {
    uint32_t (*base)[kBlk];
    uint32_t (*bits)[kBlk];
    int lane;
    int sp;
    int cap;  // groups the stack may hold: kStack, or less under TraceScene::stack_cap
    __device__ __forceinline__ void push(uint32_t b, uint32_t m) {
        base[sp][lane] = b;
        bits[sp][lane] = m;
        ++sp;
    }
    __device__ __forceinline__ bool full() const { return sp >= cap; }
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
                                                       float root_max, unsigned long long *__restrict__ counters,
                                                       uint32_t *__restrict__ fault, int cap) {
    __shared__ uint32_t stack_base[kStack][kBlock];
    __shared__ uint32_t stack_bits[kStack][kBlock];
    const int lane = threadIdx.x;
    const uint32_t i = blockIdx.x * kBlock + lane;
    if (count) n = min(n, *count);
    if (i >= n) return;
    TraceRay r{rays.ox[i], rays.oy[i], rays.oz[i], rays.dx[i], rays.dy[i], rays.dz[i], rays.tmin[i], rays.tmax[i]};
    LdsStack<kStack> st{stack_base, stack_bits, lane, 0, cap};
    TraceResult res = traverse<ANY, STATS>(nodes, tris, root_max, r, st);
    if (res.overflow && fault) atomicOr(fault, 1u);
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

// Persistent traversal with lane refill.  Each block owns a contiguous range of the rays (the grid is
// sized to what the chip holds at once); its waves draw batches of 64 ray indices from an LDS counter,
// and a lane whose ray finishes takes the next index at once, so a wave no longer idles until its
// slowest ray ends.  Node visits are per lane (trav_nodes); the triangles the wave's lanes reach in a
// step are tested by the whole wave together (wave_triangles).  Per ray the arithmetic, the nodes
// visited and the result are exactly the one-ray-per-lane kernel's (trav_step).
// Source: where a ray's scene comes from -- one scene for the launch (OneScene) or the ray's job
// (JobScenes, batched rendering).
struct OneScene {
    static constexpr bool kShared = true;  // every lane reads the same triangle array
    const U4 *nodes;
    const F4 *tris;
    float root_max;
    uint32_t *fault;
    int cap;
    struct Lane {};
    __device__ __forceinline__ bool begin(uint32_t, Lane &) const { return true; }
    __device__ __forceinline__ const U4 *node_ptr(const Lane &) const { return nodes; }
    __device__ __forceinline__ const F4 *tri_ptr(const Lane &) const { return tris; }
    __device__ __forceinline__ float root(const Lane &) const { return root_max; }
    __device__ __forceinline__ uint32_t *fault_word(const Lane &) const { return fault; }
    __device__ __forceinline__ int stack_cap(const Lane &, int k) const { return cap < k ? cap : k; }
};
struct JobScenes {
    static constexpr bool kShared = false;
    const TraceScene *job_trace;
    const uint32_t *job;
    struct Lane {
        const U4 *nodes;
        const F4 *tris;
        float root_max;
        uint32_t *fault;
        uint32_t cap;
    };
    // False for a scene without geometry (the ray misses).
    __device__ __forceinline__ bool begin(uint32_t ray, Lane &l) const {
        const TraceScene &ts = job_trace[job[ray]];
        l.nodes = reinterpret_cast<const U4 *>(ts.nodes);
        l.tris = reinterpret_cast<const F4 *>(ts.tris);
        l.root_max = ts.root_max;
        l.fault = ts.fault;
        l.cap = ts.stack_cap;
        return ts.node_count != 0;
    }
    __device__ __forceinline__ const U4 *node_ptr(const Lane &l) const { return l.nodes; }
    __device__ __forceinline__ const F4 *tri_ptr(const Lane &l) const { return l.tris; }
    __device__ __forceinline__ float root(const Lane &l) const { return l.root_max; }
    __device__ __forceinline__ uint32_t *fault_word(const Lane &l) const { return l.fault; }
    __device__ __forceinline__ int stack_cap(const Lane &l, int k) const {
        return l.cap && (int)l.cap < k ? (int)l.cap : k;
    }
};

// Order-preserving map of a float's bits to an unsigned integer.  -0 maps with +0, since the closest-hit
// rule compares t numerically (a hit at t = -0 ties with one at +0 and the smaller face wins).
__device__ __forceinline__ uint32_t ord_bits(float f) {
    const uint32_t u = __float_as_uint(f == 0.0f ? 0.0f : f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// Position of the k-th set bit (from 0) of m, for k < popcount(m) and m below 2^31.
__device__ __forceinline__ uint32_t nth_set_bit(uint32_t m, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t w = 16; w; w >>= 1) {
        const uint32_t low = __popc(m & ((1u << w) - 1u));
        const bool up = k >= low;
        k = up ? k - low : k;
        m = up ? m >> w : m;
        pos = up ? pos + w : pos;
    }
    return pos;
}
// Orders the wave's LDS accesses on either side (no instruction: a wave's LDS operations complete in
// issue order; this only keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Triangle half of a step for a whole wave.  After its node visit each lane holds up to 24 triangles
// (trimask over tri_base); tested one lane at a time the wave would loop as often as its fullest lane
// (5.6 times per step on incoherent bounce rays with 0.44 x 64 tests to do).  Here the wave's tests are
// numbered by a prefix sum over the lanes' counts and handed out 64 at a time: test j belongs to the
// lane whose range holds j (binary search over the prefixes) and is its k-th triangle; the worker reads
// that lane's ray through cross-lane permutes, runs the same Woop test, and folds a hit into the
// owner's LDS slot with an atomic minimum of (order-preserving t bits << 32 | face) -- the nearest t,
// equal t going to the smaller face, as the one-lane rule.  The winner of each round stores its t and
// barycentrics; the key's identity picks it since a face is tested at most once per ray.  A hit at t
// must also lie in [tmin, tmax) (the one-lane rule's first accepted hit needs t < tmax).  Any-hit
// queries only raise the owner's flag.  Must be called by all 64 lanes together.
// Exclusive prefix over the wave of per-lane counts below 32 (pre) and their sum (total): one ballot per
// bit of the count.
__device__ __forceinline__ void wave_prefix(uint32_t cnt, unsigned long long below, uint32_t &pre, uint32_t &total) {
    pre = 0;
    total = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        const unsigned long long m = __ballot((cnt >> b) & 1u);
        pre += (uint32_t)__popcll(m & below) << b;
        total += (uint32_t)__popcll(m) << b;
    }
}

template <bool ANY, int kBlk, class Source>
__device__ __forceinline__ void wave_triangles(const Source &src, const typename Source::Lane &sl, TravState &ts,
                                               uint32_t trimask, uint32_t tri_base, uint32_t lane, uint32_t pre,
                                               uint32_t total, unsigned long long *key, float (*hit_tb)[kBlk]) {
    const uint32_t cnt = __popc(trimask);
    if (total == 0u) return;
    const uint32_t wave0 = threadIdx.x & ~63u;
    unsigned long long key0 = 0ull;
    if (cnt) {
        if (!ANY) key0 = ((unsigned long long)ord_bits(ts.res.t) << 32) | (uint32_t)ts.res.face;
        key[threadIdx.x] = key0;
    }
    wave_lds_order();
    const int kpack = ts.w.kx | (ts.w.ky << 2) | (ts.w.kz << 4);
    const uint64_t tri_bits = reinterpret_cast<uint64_t>(src.tri_ptr(sl));
    for (uint32_t j0 = 0; j0 < total; j0 += 64u) {
        const uint32_t j = j0 + lane;
        uint32_t owner = 0;
#pragma unroll
        for (uint32_t step = 32; step; step >>= 1) {
            const uint32_t cand = owner + step;  // at most 63
            if ((uint32_t)__shfl((int)pre, (int)cand) <= j) owner = cand;
        }
        const int o = (int)owner;
        const uint32_t k = j - (uint32_t)__shfl((int)pre, o);
        const uint32_t omask = (uint32_t)__shfl((int)trimask, o), obase = (uint32_t)__shfl((int)tri_base, o);
        const int kp = __shfl(kpack, o);
        WoopRay w;
        w.kx = kp & 3;
        w.ky = (kp >> 2) & 3;
        w.kz = kp >> 4;
        w.sx = __shfl(ts.w.sx, o);
        w.sy = __shfl(ts.w.sy, o);
        w.sz = __shfl(ts.w.sz, o);
        const float ox = __shfl(ts.ox, o), oy = __shfl(ts.oy, o), oz = __shfl(ts.oz, o);
        const float tmin = __shfl(ts.tmin, o), tmax = __shfl(ts.tmax, o);
        const F4 *tris;
        if constexpr (Source::kShared) {
            tris = src.tri_ptr(sl);
        } else {
            const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)tri_bits, o);
            const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(tri_bits >> 32), o);
            tris = reinterpret_cast<const F4 *>(((uint64_t)hi << 32) | lo);
        }
        bool hit = false;
        unsigned long long mine = 0ull;
        float t = 0.0f, b1 = 0.0f, b2 = 0.0f;
        if (j < total) {
            const F4 *tp = tris + 3 * (size_t)(obase + nth_set_bit(omask, k));
            const F4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
            if (woop_triangle(w, ox, oy, oz, p0, p1, p2, t, b1, b2) && t >= tmin && t < tmax) {
                hit = true;
                if (ANY) {
                    key[wave0 + owner] = 1ull;
                } else {
                    mine = ((unsigned long long)ord_bits(t) << 32) | float_to_bits(p0.w);
                    atomicMin(&key[wave0 + owner], mine);
                }
            }
        }
        if (!ANY) {
            wave_lds_order();
            if (hit && key[wave0 + owner] == mine) {
                hit_tb[0][wave0 + owner] = t;
                hit_tb[1][wave0 + owner] = b1;
                hit_tb[2][wave0 + owner] = b2;
            }
        }
        wave_lds_order();
    }
    if (cnt) {
        const unsigned long long got = key[threadIdx.x];
        if (ANY) {
            if (got) ts.res.occluded = true;
        } else if (got != key0) {
            ts.res.t = hit_tb[0][threadIdx.x];
            ts.res.face = (int32_t)(uint32_t)got;
            ts.res.b1 = hit_tb[1][threadIdx.x];
            ts.res.b2 = hit_tb[2][threadIdx.x];
        }
    }
    wave_lds_order();
}

template <bool ANY, int kStack, int kBlk, class Source>
__device__ __forceinline__ void refill_trace(const Source &src, const RayArgs &rays, const HitArgs &hits,
                                             uint8_t *__restrict__ occluded, uint32_t n) {
    __shared__ uint32_t stack_base[kStack][kBlk];
    __shared__ uint32_t stack_bits[kStack][kBlk];
    __shared__ uint32_t next;
    __shared__ unsigned long long tri_key[kBlk];
    __shared__ float tri_hit[ANY ? 1 : 3][kBlk];
    constexpr uint32_t kNone = 0xffffffffu;
    const uint32_t chunk = ((n + gridDim.x - 1) / gridDim.x + 63u) & ~63u;
    const uint32_t begin = min(n, blockIdx.x * chunk), end = min(n, begin + chunk);
    if (threadIdx.x == 0) next = begin;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long below = (1ull << lane) - 1ull;
    LdsStack<kStack, kBlk> st{stack_base, stack_bits, (int)threadIdx.x, 0, kStack};
    TravState ts;
    typename Source::Lane sl{};
    uint32_t ray = kNone;
    // The lane's pending triangles (from its last node visit, not yet tested) and whether its ray has
    // no node left to visit.
    uint32_t trimask = 0, tri_base = 0;
    bool done = false;
    // Wave-uniform: the wave's current batch of 64 indices and how many of them are handed out.
    uint32_t batch = 0, used = 64;
    bool exhausted = false;
    for (;;) {
        const unsigned long long need = __ballot(ray == kNone);
        // Refilling costs the whole wave the ray set-up (loads, Woop constants, reciprocals), so idle
        // lanes wait until kRefillIdle of them can be served together (a CPU replay of the benchmark
        // rays put the best trade-off between 8 and 12; measured: bounce rays 4% faster).
        if (!exhausted && __popcll(need) >= kRefillIdle) {
            const uint32_t cnt = (uint32_t)__popcll(need), rank = (uint32_t)__popcll(need & below);
            const uint32_t avail = 64u - used;
            uint32_t idx = batch + used + rank;
            if (cnt > avail) {
                uint32_t fresh = 0;
                if (lane == 0) fresh = atomicAdd(&next, 64u);
                fresh = __shfl(fresh, 0);
                if (rank >= avail) idx = fresh + (rank - avail);
                batch = fresh;
                used = cnt - avail;
                if (fresh >= end) exhausted = true;
            } else {
                used += cnt;
            }
            if (ray == kNone && idx < end) {
                const TraceRay r{rays.ox[idx], rays.oy[idx], rays.oz[idx], rays.dx[idx], rays.dy[idx], rays.dz[idx],
                                 rays.tmin[idx], rays.tmax[idx]};
                if (src.begin(idx, sl)) {
                    ray = idx;
                    trav_init(ts, r, src.root(sl));
                    st.sp = 0;
                    st.cap = src.stack_cap(sl, kStack);
                    done = false;
                } else if (ANY) {
                    occluded[idx] = 0;
                } else {
                    hits.t[idx] = r.tmax;
                    hits.b1[idx] = 0.0f;
                    hits.b2[idx] = 0.0f;
                    hits.face[idx] = -1;
                    if (hits.instance) hits.instance[idx] = 0u;
                }
            }
        }
        const bool busy = ray != kNone;
        const unsigned long long live = __ballot(busy);
        if (!live && exhausted) break;
        // A lane holding untested triangles waits (its next node visit needs the hits they may give),
        // so triangles of several steps are tested together: the wave flushes once a quarter of its
        // live lanes wait or 64 tests are pending (a CPU replay of the benchmark rays: 0.97 -> 0.62
        // rounds per step on bounce rays, about 11% fewer issue cycles).  Nodes visited and results
        // are unchanged.
        if (busy && !done && trimask == 0u) done = trav_nodes<false>(ts, src.node_ptr(sl), st, trimask, tri_base);
        const unsigned long long waiting = __ballot(busy && (done || trimask != 0u));
        uint32_t pre, total;
        wave_prefix(__popc(trimask), below, pre, total);
        if (total >= 64u || 4u * (uint32_t)__popcll(waiting) >= (uint32_t)__popcll(live)) {
            wave_triangles<ANY, kBlk>(src, sl, ts, trimask, tri_base, lane, pre, total, tri_key, tri_hit);
            trimask = 0u;
        }
        if (busy && ((done && trimask == 0u) || (ANY && ts.res.occluded))) {
            if (ts.res.overflow) {
                uint32_t *f = src.fault_word(sl);
                if (f) atomicOr(f, 1u);
            }
            if (ANY) {
                occluded[ray] = ts.res.occluded ? 1 : 0;
            } else {
                hits.t[ray] = ts.res.t;
                hits.b1[ray] = ts.res.b1;
                hits.b2[ray] = ts.res.b2;
                hits.face[ray] = ts.res.face;
                if (hits.instance) hits.instance[ray] = 0u;
            }
            ray = kNone;
        }
    }
}

// Work-group of the persistent kernel: a block's waves share its range of rays, so larger blocks
// balance uneven rays better (512 lanes: icosphere-field camera rays 0.356 -> 0.237 ms against 128);
// deeper stacks take 256 so that LDS still holds 5 waves per SIMD.
template <int kStack>
constexpr int refill_block() {
    return kStack <= 8 ? 512 : 256;
}

template <bool ANY, int kStack>
__global__ __launch_bounds__(refill_block<kStack>()) void trace_refill_kernel(const U4 *__restrict__ nodes, const F4 *__restrict__ tris,
                                                              RayArgs rays, HitArgs hits, uint8_t *__restrict__ occluded,
                                                              uint32_t n, const uint32_t *__restrict__ count,
                                                              float root_max, uint32_t *__restrict__ fault, int cap) {
    if (count) n = min(n, *count);
    refill_trace<ANY, kStack, refill_block<kStack>()>(OneScene{nodes, tris, root_max, fault, cap}, rays, hits, occluded,
                                                      n);
}

// Two-level traversal, one ray per lane (main.swift:1843-1900 mesh_closest, 1902-1983 mesh_occluded,
// software branch).  The instance hierarchy is walked as a one-level one whose leaves name instances,
// with the boxes grown by enlarge (|origin|max + reach); a leaf's instance traces its asset with the
// object-space ray, up to the closest hit so far.  Each level keeps its node-group stack in LDS.
// Both levels run in one loop: each iteration a lane steps either the instance hierarchy or its
// current instance's asset, so lanes between instances need not wait for another lane's whole asset
// traversal, as they would with the asset walk nested inside the instance walk.
// Ties: the closest-hit rule compares (t, instance, face), so the result does not depend on the order
// instances are visited in.
template <bool ANY, int kTop, int kBot>
__global__ __launch_bounds__(kBlock) void trace_tlas_kernel(TlasScene sc, RayArgs rays, HitArgs hits,
                                                            uint8_t *__restrict__ occluded, uint32_t n) {
    __shared__ uint32_t top_base[kTop][kBlock], top_bits[kTop][kBlock];
    __shared__ uint32_t bot_base[kBot][kBlock], bot_bits[kBot][kBlock];
    const int lane = threadIdx.x;
    const uint32_t i = blockIdx.x * kBlock + lane;
    if (i >= n) return;
    const TraceRay r{rays.ox[i], rays.oy[i], rays.oz[i], rays.dx[i], rays.dy[i], rays.dz[i], rays.tmin[i], rays.tmax[i]};
    LdsStack<kTop> top{top_base, top_bits, lane, 0, kTop};
    LdsStack<kBot> bot{bot_base, bot_bits, lane, 0, kBot};
    TravState ts, bs;
    trav_init_e(ts, r, sc.enlarge * (max_abs3(r.ox, r.oy, r.oz) + sc.reach));
    float bt = r.tmax, bb1 = 0.0f, bb2 = 0.0f;
    int32_t bface = -1, sentinel = -1;
    uint32_t binst = 0xffffffffu, cur = 0u, items = 0u, base = 0u;
    bool occ = false, overflow = false, top_done = false, in_asset = false;
    const U4 *tn = reinterpret_cast<const U4 *>(sc.nodes);
    const U4 *bn = nullptr;
    const F4 *btr = nullptr;
    for (;;) {
        const bool want_top = !in_asset && items == 0u;
        if (want_top && top_done) break;
        // The instance hierarchy is stepped once at least a quarter of the wave's remaining lanes wait
        // for it; until then those lanes idle while the others walk their assets, so lanes that enter
        // their instances together stay together.  Instanced icosphere field (1024 instances, 2M
        // rays): camera 0.718 -> 0.664 ms, bounce 1.858 -> 1.229 ms against the nested walk; stepping
        // for any waiting lane gave 0.794 / 1.454 ms, waiting for half 0.674 / 1.245, three quarters
        // 0.713 / 1.432.
        const uint32_t waiting = (uint32_t)__popcll(__ballot(want_top)), live = (uint32_t)__popcll(__ballot(1));
        if (want_top) {
            if (4u * waiting >= live) {
                top_done = trav_nodes<false>(ts, tn, top, items, base);
                overflow |= ts.res.overflow;
            }
            continue;
        }
        if (!in_asset) {
            const uint32_t b = __builtin_ctz(items);
            items &= items - 1u;
            const InstanceRecord &R = sc.instances[base + b];
            if (R.faces == 0) continue;
            trav_init(bs, instance_ray(R.local, r, bt), R.root_max);
            // Equal t: a hit of this instance wins over the current one only from a smaller index.
            sentinel = (bface >= 0 && R.index < binst) ? 0x7fffffff : -1;
            bs.res.face = sentinel;
            bot.sp = 0;
            bn = reinterpret_cast<const U4 *>(R.nodes);
            btr = reinterpret_cast<const F4 *>(R.tris);
            cur = R.index;
            in_asset = true;
        }
        if (trav_step<ANY, false>(bs, bn, btr, bot)) {
            in_asset = false;
            overflow |= bs.res.overflow;
            if (ANY) {
                if (bs.res.occluded) {
                    occ = true;
                    break;
                }
            } else if (bs.res.face != sentinel) {
                bt = bs.res.t;
                bb1 = bs.res.b1;
                bb2 = bs.res.b2;
                bface = bs.res.face;
                binst = cur;
                ts.res.t = bt;
            }
        }
    }
    if (overflow && sc.fault) atomicOr(sc.fault, 1u);
    if (ANY) {
        occluded[i] = occ ? 1 : 0;
    } else {
        hits.t[i] = bt;
        hits.b1[i] = bb1;
        hits.b2[i] = bb2;
        hits.face[i] = bface;
        if (hits.instance) hits.instance[i] = binst;
    }
}

template <bool ANY>
__global__ __launch_bounds__(256) void miss_kernel(RayArgs rays, HitArgs hits, uint8_t *occluded, uint32_t n,
                                                   const uint32_t *count, uint32_t miss_instance) {
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
        if (hits.instance) hits.instance[i] = miss_instance;
    }
}

// The reference's software traversal of its flat binary hierarchy (trace_scene's mesh branch,
// main.swift:2058-2080, with mesh_box_ray / mesh_node_hit / push_mesh_children :1511-1528): one ray per
// lane, a stack of node indices in LDS, the child on the ray's side of the node's split axis first.
// Boxes are the reference's: the slabs of [lo, hi] against the origin moved by -+e, e = 2^-23 (1 + ...)
// x (|origin|max + |root|max), a direction component below 1e-8 testing the origin against the
// enlarged slab.  Triangles are the same watertight test as the 8-wide traversal, and the closest-hit
// rule the same (t, then the smaller document face index at equal t), so hits do not depend on the
// hierarchy.  A ray whose stack would overflow (never, for the depths the host admits) raises *fault.
template <bool ANY>
__global__ __launch_bounds__(256) void trace_median_kernel(const F4 *__restrict__ nodes, const F4 *__restrict__ tris,
                                                           RayArgs rays, HitArgs hits, uint8_t *__restrict__ occluded,
                                                           uint32_t n, const uint32_t *__restrict__ count, float root_max,
                                                           uint32_t *__restrict__ fault) {
    __shared__ uint32_t stack[kMedianStack][256];
    const uint32_t lane = threadIdx.x, i = blockIdx.x * 256 + lane;
    if (count) n = min(n, *count);
    if (i >= n) return;
    const float o[3] = {rays.ox[i], rays.oy[i], rays.oz[i]}, d[3] = {rays.dx[i], rays.dy[i], rays.dz[i]};
    const float tmin = rays.tmin[i], tmax = rays.tmax[i];
    const WoopRay w = woop_ray(d[0], d[1], d[2]);
    const float e = 1.1920929e-6f * (fmaxf(fabsf(o[0]), fmaxf(fabsf(o[1]), fabsf(o[2]))) + root_max);
    const float low[3] = {o[0] + e, o[1] + e, o[2] + e}, high[3] = {o[0] - e, o[1] - e, o[2] - e};
    float best = tmax, bb1 = 0.0f, bb2 = 0.0f;
    int32_t bface = -1;
    bool hit = false, overflow = false;
    uint32_t top = 0;
    stack[top++][lane] = 0u;
    while (top > 0u) {
        const uint32_t ni = stack[--top][lane];
        const F4 lo = nodes[3 * (size_t)ni], hi = nodes[3 * (size_t)ni + 1], lk = nodes[3 * (size_t)ni + 2];
        const int32_t links[4] = {(int32_t)__float_as_uint(lk.x), (int32_t)__float_as_uint(lk.y),
                                  (int32_t)__float_as_uint(lk.z), (int32_t)__float_as_uint(lk.w)};
        // mesh_node_hit against [tmin, closest hit so far].
        float near_t = tmin, far_t = ANY ? tmax : best;
        const float nlo[3] = {lo.x, lo.y, lo.z}, nhi[3] = {hi.x, hi.y, hi.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (fabsf(d[a]) < 1e-8f) {
                if (low[a] < nlo[a] || high[a] > nhi[a]) far_t = -1.0f;
            } else {
                const float ta = (nlo[a] - low[a]) / d[a], tb = (nhi[a] - high[a]) / d[a];
                near_t = fmaxf(near_t, fminf(ta, tb));
                far_t = fminf(far_t, fmaxf(ta, tb));
            }
        }
        if (!(near_t <= far_t)) continue;
        if (links[3] == 0) {
            if (top + 2u > kMedianStack) {
                overflow = true;
                continue;
            }
            const int axis = min(max(links[2], 0), 2);
            const bool left_first = d[axis] >= 0.0f;
            stack[top++][lane] = (uint32_t)(left_first ? links[1] : links[0]);
            stack[top++][lane] = (uint32_t)(left_first ? links[0] : links[1]);
            continue;
        }
        for (int32_t k = 0; k < links[3]; ++k) {
            const size_t r = 3 * (size_t)(links[2] + k);
            const F4 p0 = tris[r], p1 = tris[r + 1], p2 = tris[r + 2];
            const int32_t face = (int32_t)__float_as_uint(p0.w);
            float t, b1, b2;
            if (!woop_triangle(w, o[0], o[1], o[2], p0, p1, p2, t, b1, b2)) continue;
            if (ANY) {
                if (t >= tmin && t < tmax) {
                    hit = true;
                    break;
                }
                continue;
            }
            if (t >= tmin && (t < best || (t == best && bface >= 0 && face < bface))) {
                best = t;
                bb1 = b1;
                bb2 = b2;
                bface = face;
            }
        }
        if (ANY && hit) break;
    }
    if (overflow && fault) atomicOr(fault, 1u);
    if (ANY) {
        occluded[i] = hit ? 1 : 0;
    } else {
        hits.t[i] = best;
        hits.b1[i] = bb1;
        hits.b2[i] = bb2;
        hits.face[i] = bface;
        if (hits.instance) hits.instance[i] = 0u;
    }
}

// Batched rendering: rays of many scenes in one persistent launch with lane refill.  Ray i belongs to
// job job[i], whose scene is job_trace[job[i]]; rays of one job are contiguous in the queues, so a wave
// rarely spans two scenes.
template <bool ANY, int kStack, int kBlk>
__global__ __launch_bounds__(kBlk) void batch_trace_kernel(const TraceScene *__restrict__ job_trace,
                                                           const uint32_t *__restrict__ job, RayArgs rays, HitArgs hits,
                                                           uint8_t *__restrict__ occluded, uint32_t n,
                                                           const uint32_t *__restrict__ count) {
    n = min(n, *count);
    refill_trace<ANY, kStack, kBlk>(JobScenes{job_trace, job}, rays, hits, occluded, n);
}

}  // namespace

namespace {

// Blocks of the refill kernel that are resident at once on the current device (at least 1), capped by
// the blocks n rays need; the occupancy query is made once per kernel.
uint32_t refill_grid(const void *kernel, uint32_t n, int block) {
    struct Entry {
        const void *k;
        uint32_t blocks;
    };
    static Entry cache[8] = {};
    static std::mutex lock;
    std::lock_guard<std::mutex> guard(lock);
    uint32_t resident = 0;
    for (auto &e : cache)
        if (e.k == kernel) resident = e.blocks;
    if (!resident) {
        int dev = 0, cus = 256, per = 1;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0) != hipSuccess || per < 1) per = 1;
        resident = (uint32_t)(cus * per);
        for (auto &e : cache)
            if (!e.k) {
                e = Entry{kernel, resident};
                break;
            }
    }
    const uint32_t need = (n + (uint32_t)block - 1) / (uint32_t)block;
    return need < resident ? (need > 0 ? need : 1) : resident;
}

}  // namespace

// ---------------------------------------------------------------- allocation-free work counters
// The counting traversal's 64 partial (node fetches, triangle tests) pairs live in this file-scope
// array: 1 KiB per device, whatever the number of rays.  Host code never passes the symbol itself
// (in host code it only names a stand-in); it passes the per-device address scratch_init found.
static __device__ __align__(256) unsigned long long scratch[kStatsScratchWords];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_lock;

__global__ void zero_scratch_kernel(unsigned long long *p) {
    if (threadIdx.x < kStatsScratchWords) p[threadIdx.x] = 0ull;
}
__global__ void sum_pairs_kernel(const unsigned long long *p, unsigned long long *out) {
    if (threadIdx.x != 0) return;
    unsigned long long nodes = 0, tris = 0;
    for (int k = 0; k < kStatsScratchWords / 2; ++k) {
        nodes += p[2 * k];
        tris += p[2 * k + 1];
    }
    out[0] = nodes;
    out[1] = tris;
}
}  // namespace

hipError_t scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (g_scratch[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = p;
    return hipSuccess;
}

hipError_t launch_trace_stats_scratch(int device, const TraceScene &scene, const RayArgs &rays, const HitArgs &hits,
                                      uint8_t *occluded, uint32_t n, bool any, unsigned long long *out,
                                      hipStream_t stream) {
    if (device < 0 || device >= kMaxDevices || !g_scratch[device]) return hipErrorNotInitialized;
    unsigned long long *p = static_cast<unsigned long long *>(g_scratch[device]);
    hipLaunchKernelGGL(zero_scratch_kernel, dim3(1), dim3(kStatsScratchWords), 0, stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (n) {
        e = launch_trace(scene, rays, hits, occluded, n, any, p, stream);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(sum_pairs_kernel, dim3(1), dim3(64), 0, stream, p, out);
    return hipGetLastError();
}

hipError_t launch_trace_batch(const TraceScene *job_trace, const uint32_t *job, const RayArgs &rays, const HitArgs &hits,
                              uint8_t *occluded, uint32_t n, bool any, const uint32_t *count, uint32_t max_depth,
                              uint32_t block, uint32_t grid, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const int need = max_depth > 1 ? (int)max_depth - 1 : 1;
    if (need > 24 || (block != 64 && block != 128 && block != 256 && block != 512) || grid == 0) return hipErrorInvalidValue;
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
        else if (block == 256) VT_BATCH(S, 256);                                                                       \
        else VT_BATCH(S, 512);                                                                                         \
    } while (0)
    if (need <= 8) VT_BATCH_STACK(8);
    else if (need <= 12) VT_BATCH_STACK(12);
    else if (need <= 16) VT_BATCH_STACK(16);
    else VT_BATCH_STACK(24);
#undef VT_BATCH_STACK
#undef VT_BATCH
    return hipGetLastError();
}

hipError_t launch_trace_tlas(const TlasScene &scene, const RayArgs &rays, const HitArgs &hits, uint8_t *occluded,
                             uint32_t n, bool any, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (scene.node_count == 0) {
        dim3 grid((n + 255) / 256);
        if (any) hipLaunchKernelGGL(miss_kernel<true>, grid, dim3(256), 0, stream, rays, hits, occluded, n, nullptr, 0u);
        else hipLaunchKernelGGL(miss_kernel<false>, grid, dim3(256), 0, stream, rays, hits, occluded, n, nullptr, 0xffffffffu);
        return hipGetLastError();
    }
    const int top = scene.depth > 1 ? (int)scene.depth - 1 : 1, bot = scene.blas_depth > 1 ? (int)scene.blas_depth - 1 : 1;
    if (top > 16 || bot > 24) return hipErrorInvalidValue;
    dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
#define VT_TLAS(T, B)                                                                                                  \
    do {                                                                                                               \
        if (any) hipLaunchKernelGGL((trace_tlas_kernel<true, T, B>), grid, block, 0, stream, scene, rays, hits, occluded, n); \
        else hipLaunchKernelGGL((trace_tlas_kernel<false, T, B>), grid, block, 0, stream, scene, rays, hits, occluded, n); \
    } while (0)
    if (top <= 8) {
        if (bot <= 12) VT_TLAS(8, 12);
        else VT_TLAS(8, 24);
    } else {
        if (bot <= 12) VT_TLAS(16, 12);
        else VT_TLAS(16, 24);
    }
#undef VT_TLAS
    return hipGetLastError();
}

hipError_t launch_trace(const TraceScene &scene, const RayArgs &rays, const HitArgs &hits, uint8_t *occluded,
                        uint32_t n, bool any, unsigned long long *counters, hipStream_t stream, const uint32_t *count) {
    if (n == 0) return hipSuccess;
    if (scene.node_count == 0) {
        dim3 grid((n + 255) / 256);
        if (any) hipLaunchKernelGGL(miss_kernel<true>, grid, dim3(256), 0, stream, rays, hits, occluded, n, count, 0u);
        else hipLaunchKernelGGL(miss_kernel<false>, grid, dim3(256), 0, stream, rays, hits, occluded, n, count, 0u);
        return hipGetLastError();
    }
    if (scene.kind == 1u) {
        // The reference's flat binary hierarchy: no counting variant.
        if (counters) return hipErrorInvalidValue;
        const dim3 grid((n + 255) / 256);
        const F4 *nd = reinterpret_cast<const F4 *>(scene.nodes), *tr = reinterpret_cast<const F4 *>(scene.tris);
        if (any)
            hipLaunchKernelGGL(trace_median_kernel<true>, grid, dim3(256), 0, stream, nd, tr, rays, hits, occluded, n,
                               count, scene.root_max, scene.fault);
        else
            hipLaunchKernelGGL(trace_median_kernel<false>, grid, dim3(256), 0, stream, nd, tr, rays, hits, occluded, n,
                               count, scene.root_max, scene.fault);
        return hipGetLastError();
    }
    dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
    const U4 *nd = reinterpret_cast<const U4 *>(scene.nodes);
    const F4 *tr = reinterpret_cast<const F4 *>(scene.tris);
    const int need = scene.depth > 1 ? (int)scene.depth - 1 : 1;
    if (need > 24) return hipErrorInvalidValue;
    const int cap = scene.stack_cap ? (int)scene.stack_cap : 24;
    if (!counters) {
        // Persistent refill kernel: as many blocks as the chip holds at once (never more than the rays
        // need), each draining its own range.
#define VT_REFILL(S)                                                                                                   \
    do {                                                                                                               \
        constexpr int b = refill_block<S>();                                                                          \
        const uint32_t g = refill_grid(any ? (const void *)trace_refill_kernel<true, S> : (const void *)trace_refill_kernel<false, S>, n, b); \
        if (any) hipLaunchKernelGGL((trace_refill_kernel<true, S>), dim3(g), dim3(b), 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, scene.fault, cap < S ? cap : S); \
        else hipLaunchKernelGGL((trace_refill_kernel<false, S>), dim3(g), dim3(b), 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, scene.fault, cap < S ? cap : S); \
    } while (0)
        if (need <= 8) VT_REFILL(8);
        else if (need <= 12) VT_REFILL(12);
        else if (need <= 16) VT_REFILL(16);
        else VT_REFILL(24);
#undef VT_REFILL
        return hipGetLastError();
    }
#define VT_LAUNCH(S)                                                                                                   \
    do {                                                                                                               \
        if (counters) {                                                                                                \
            if (any) hipLaunchKernelGGL((trace_kernel<true, true, S>), grid, block, 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, counters, scene.fault, cap < S ? cap : S); \
            else hipLaunchKernelGGL((trace_kernel<false, true, S>), grid, block, 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, counters, scene.fault, cap < S ? cap : S); \
        } else {                                                                                                       \
            if (any) hipLaunchKernelGGL((trace_kernel<true, false, S>), grid, block, 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, nullptr, scene.fault, cap < S ? cap : S); \
            else hipLaunchKernelGGL((trace_kernel<false, false, S>), grid, block, 0, stream, nd, tr, rays, hits, occluded, n, count, scene.root_max, nullptr, scene.fault, cap < S ? cap : S); \
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
