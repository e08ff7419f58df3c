// bvh8.cpp -- host builder of the 8-wide BVH (layout in bvh8.h).
//
// By default a binary binned-SAH tree (Wald 2007) is collapsed into 8-wide nodes by the SAH-optimal
// dynamic programme below.  The other mode widens top-down the way the reference widens its 4-wide
// tree (Sources/Acceleration.swift:247-284): a node starts with one child holding all its items and
// repeatedly splits the child of largest surface area with a binary binned-SAH split until it
// has eight children or none is larger than a leaf.  Subtrees below a fraction of the input are
// built concurrently into private node lists and stitched in a fixed order, so the tree does
// not depend on the thread count.  A tree deeper than max_depth is rebuilt with object-median
// splits (the reference does the same, Acceleration.swift:296-299), which bounds the depth by
// log8 of the leaf count.  The wide tree is then laid out with each node's internal children
// contiguous and compressed to 80-byte nodes (bvh8.h).
#include "bvh8.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <cstdlib>
#include <exception>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <type_traits>

namespace vt {
namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float *l, const float *h) {
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], l[a]); hi[a] = std::max(hi[a], h[a]); }
    }
    void grow(const Box &b) { grow(b.lo, b.hi); }
    void grow_point(const float *p) { grow(p, p); }
    float half_area() const {
        float e[3];
        for (int a = 0; a < 3; ++a) e[a] = std::max(hi[a] - lo[a], 0.0f);
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
    }
};

struct Range {
    int32_t start = 0, end = 0;
    Box box;   // item bounds
    Box cbox;  // centroid bounds
    int32_t count() const { return end - start; }
};

struct WideNode {
    int nchild = 0;
    Box box[8];
    int32_t child[8];   // >= 0: internal (index into the owning node list); < 0: leaf -(start + 1)
    int32_t count[8];   // leaf item count
};

// An item as the builder moves it: its bounds and its face index, 32 bytes, kept in the order the
// splits put the items in (binning and partitioning then sweep memory sequentially).  The centroid is
// (lo + hi) * 0.5, recomputed where needed.
struct PrimRef {
    float lo[3];
    int32_t idx;
    float hi[3];
    float pad;
    float centroid(int a) const { return (lo[a] + hi[a]) * 0.5f; }
};
static_assert(sizeof(PrimRef) == 32, "PrimRef is 32 bytes");

int thread_count();

// Worker threads for parallel_run, started on first use and kept for the process (a build makes a few
// hundred parallel calls; starting threads for each would cost more than some of them do).
class WorkerPool {
public:
    explicit WorkerPool(int n) {
        for (int i = 1; i < n; ++i) threads_.emplace_back([this, i]() { loop(i); });
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    int size() const { return (int)threads_.size() + 1; }
    // Runs f(t) for t in [0, T), T <= size(), the calling thread taking t = 0; rethrows the first error.
    void run(int T, const std::function<void(int)> &f) {
        std::lock_guard<std::mutex> one_at_a_time(run_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            job_threads_ = T;
            pending_ = T - 1;
            errs_.assign(T, nullptr);
            ++generation_;
        }
        cv_.notify_all();
        call(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&]() { return pending_ == 0; });
        for (auto &e : errs_)
            if (e) std::rethrow_exception(e);
    }
    static thread_local bool inside;  // this thread is running a parallel_run body

private:
    void call(int t) {
        inside = true;
        try {
            (*job_)(t);
        } catch (...) {
            errs_[t] = std::current_exception();
        }
        inside = false;
    }
    void loop(int id) {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&]() { return stop_ || generation_ != seen; });
            if (stop_) return;
            seen = generation_;
            if (id >= job_threads_) continue;
            lk.unlock();
            call(id);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_, run_mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    int job_threads_ = 0, pending_ = 0;
    uint64_t generation_ = 0;
    bool stop_ = false;
    std::vector<std::exception_ptr> errs_;
};
thread_local bool WorkerPool::inside = false;

WorkerPool &worker_pool() {
    static WorkerPool pool(thread_count());
    return pool;
}

// Runs f(t) for t in [0, T) (chunk t of some work): on the worker pool, or one after another on this
// thread when T is 1 or this thread is already inside a parallel body (a nested call).  Results do
// not depend on which: callers split work by t alone.
template <class F>
void parallel_run(int T, F &&f) {
    if (T <= 1 || WorkerPool::inside) {
        for (int t = 0; t < T; ++t) f(t);
        return;
    }
    WorkerPool &pool = worker_pool();
    const int P = std::min(T, pool.size());
    const std::function<void(int)> body = [&](int id) {
        for (int t = id; t < T; t += P) f(t);
    };
    pool.run(P, body);
}

// Ranges at least this large are binned, measured and partitioned by all threads together (the top
// levels of the tree, which a single thread would otherwise sweep one after another).  The result is
// the same as the one-thread sweep's: bins merge exactly (counts add, boxes take min / max), and the
// partition is stable.
constexpr int32_t kParallelItems = 1 << 17;

// Four floats in one SSE register (GCC / Clang vector extension): a bin's box is (lo, hi) with x, y, z
// in lanes 0-2 and lane 3 unused, grown with one minimum and one maximum per item.
typedef float v4f __attribute__((vector_size(16)));
inline v4f vmin(v4f a, v4f b) { return a < b ? a : b; }
inline v4f vmax(v4f a, v4f b) { return a > b ? a : b; }
inline v4f load4(const float *p) {
    v4f v;
    std::memcpy(&v, p, 16);
    return v;
}

struct Bins {
    v4f lo[3][64], hi[3][64];
    int32_t cnt[3][64];
    void clear(int B) {
        const v4f big = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < B; ++b) { lo[a][b] = big; hi[a][b] = -big; cnt[a][b] = 0; }
    }
    void merge(const Bins &o, int B) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < B; ++b) {
                cnt[a][b] += o.cnt[a][b];
                lo[a][b] = vmin(lo[a][b], o.lo[a][b]);
                hi[a][b] = vmax(hi[a][b], o.hi[a][b]);
            }
    }
    Box box(int a, int b) const {
        Box r;
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = lo[a][b][k];
            r.hi[k] = hi[a][b][k];
        }
        return r;
    }
};

class Builder {
public:
    // tmp: scratch of at least as many items as the largest range split by all threads (kParallelItems
    // and up), or null when no range is that large.
    Builder(PrimRef *refs, PrimRef *tmp, const BuildOptions &o, bool median)
        : ref(refs), tmp(tmp), opt(o), median(median), bins(o.bins) {}

    Range measure(int32_t s, int32_t e) const {
        Range r;
        r.start = s;
        r.end = e;
        const int T = e - s >= kParallelItems ? thread_count() : 1;
        std::vector<Range> part(T);
        parallel_run(T, [&](int t) {
            const int32_t a = s + (int32_t)((int64_t)(e - s) * t / T), b = s + (int32_t)((int64_t)(e - s) * (t + 1) / T);
            for (int32_t k = a; k < b; ++k) {
                const PrimRef &q = ref[k];
                const float c[3] = {q.centroid(0), q.centroid(1), q.centroid(2)};
                part[t].box.grow(q.lo, q.hi);
                part[t].cbox.grow_point(c);
            }
        });
        for (const Range &p : part) {
            r.box.grow(p.box);
            r.cbox.grow(p.cbox);
        }
        return r;
    }

    void split(const Range &r, Range &left, Range &right) {
        float ext[3];
        for (int a = 0; a < 3; ++a) ext[a] = r.cbox.hi[a] - r.cbox.lo[a];
        float emax = std::max(ext[0], std::max(ext[1], ext[2]));
        if (!median && emax > 0) {
            if (binned_split(r, ext, left, right)) return;
        }
        int32_t mid = (r.start + r.end) / 2;
        if (emax > 0) {
            int a = ext[0] > ext[1] ? (ext[0] > ext[2] ? 0 : 2) : (ext[1] > ext[2] ? 1 : 2);
            std::nth_element(ref + r.start, ref + mid, ref + r.end, [a](const PrimRef &x, const PrimRef &y) {
                const float cx = x.centroid(a), cy = y.centroid(a);
                return cx == cy ? x.idx < y.idx : cx < cy;
            });
        }
        left = measure(r.start, mid);
        right = measure(mid, r.end);
    }

    // Builds the subtree of `root` into `nodes`; ranges smaller than `defer` go to `deferred`.
    struct Task { Range r; int32_t parent, slot, level; };
    int run(const Range &root, int level0, std::vector<WideNode> &nodes, int32_t defer, std::vector<Task> &deferred) {
        int reached = level0;
        std::vector<Task> work{{root, -1, 0, level0}};
        const int leaf = opt.leaf_size;
        while (!work.empty()) {
            Task t = work.back();
            work.pop_back();
            reached = std::max(reached, t.level);
            Range ch[8];
            int n = 1;
            ch[0] = t.r;
            while (n < 8) {
                int best = -1;
                for (int k = 0; k < n; ++k) {
                    if (ch[k].count() <= leaf) continue;
                    if (best < 0 || (median ? ch[k].count() > ch[best].count()
                                            : ch[k].box.half_area() > ch[best].box.half_area()))
                        best = k;
                }
                if (best < 0) break;
                Range a, b;
                split(ch[best], a, b);
                for (int k = n; k > best + 1; --k) ch[k] = ch[k - 1];
                ch[best] = a;
                ch[best + 1] = b;
                ++n;
            }
            WideNode node;
            node.nchild = n;
            int32_t index = (int32_t)nodes.size();
            if (t.parent >= 0) nodes[t.parent].child[t.slot] = index;
            for (int k = 0; k < n; ++k) {
                node.box[k] = ch[k].box;
                if (ch[k].count() <= leaf) {
                    node.child[k] = -(ch[k].start + 1);
                    node.count[k] = ch[k].count();
                } else {
                    node.child[k] = 0;
                    node.count[k] = 0;
                }
            }
            nodes.push_back(node);
            for (int k = n - 1; k >= 0; --k) {
                if (ch[k].count() <= leaf) continue;
                Task c{ch[k], index, k, t.level + 1};
                if (ch[k].count() < defer) deferred.push_back(c);
                else work.push_back(c);
            }
        }
        return reached;
    }

private:
    void bin_items(int32_t s, int32_t e, const Range &r, const float *scale, int B, Bins &out) const {
        out.clear(B);
        for (int32_t k = s; k < e; ++k) {
            const PrimRef &q = ref[k];
            const v4f lo = load4(q.lo), hi = load4(q.hi);  // lane 3: the index bits / padding, never read
            for (int a = 0; a < 3; ++a) {
                const int b = bin_of(q.centroid(a), r.cbox.lo[a], scale[a], B);
                out.cnt[a][b] += 1;
                out.lo[a][b] = vmin(out.lo[a][b], lo);
                out.hi[a][b] = vmax(out.hi[a][b], hi);
            }
        }
    }

    bool binned_split(const Range &r, const float *ext, Range &left, Range &right) {
        // Small ranges take fewer bins (at least 4, at most one per item): the split of a range of a
        // few items costs in proportion to its size rather than to the full bin count.
        const int B = std::min(bins, std::max(4, r.count()));
        float scale[3];
        for (int a = 0; a < 3; ++a) {
            float s = (float)B / ext[a];
            scale[a] = (ext[a] > 0 && std::isfinite(s)) ? s : 0.0f;
        }
        const bool wide = r.count() >= kParallelItems;
        const int T = wide ? thread_count() : 1;
        if (!wide) {
            bin_items(r.start, r.end, r, scale, B, local);
        } else {
            std::vector<Bins> part(T);
            parallel_run(T, [&](int t) {
                const int32_t a = r.start + (int32_t)((int64_t)r.count() * t / T);
                const int32_t b = r.start + (int32_t)((int64_t)r.count() * (t + 1) / T);
                bin_items(a, b, r, scale, B, part[t]);
            });
            for (int t = 1; t < T; ++t) part[0].merge(part[t], B);
            local = part[0];
        }
        const Bins &bn = local;
        float best_cost = FLT_MAX;
        int best_axis = -1, best_bin = 0;
        float suffix[64];
        for (int a = 0; a < 3; ++a) {
            if (!(scale[a] > 0)) continue;
            Box acc;
            int32_t c = 0;
            for (int b = B - 1; b > 0; --b) {
                if (bn.cnt[a][b] > 0) acc.grow(bn.box(a, b));
                c += bn.cnt[a][b];
                suffix[b] = c > 0 ? acc.half_area() * (float)c : 0.0f;
            }
            acc = Box();
            c = 0;
            for (int b = 0; b < B - 1; ++b) {
                if (bn.cnt[a][b] > 0) acc.grow(bn.box(a, b));
                c += bn.cnt[a][b];
                if (c == 0 || c == r.count()) continue;
                float cost = acc.half_area() * (float)c + suffix[b + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_bin = b; }
            }
        }
        if (best_axis < 0) return false;
        const int a = best_axis;
        const float s = scale[a], origin = r.cbox.lo[a];
        auto goes_left = [&](const PrimRef &q) { return bin_of(q.centroid(a), origin, s, B) <= best_bin; };
        // The children's item and centroid bounds are gathered while partitioning.
        Range sides[2];
        auto add = [](Range &d, const PrimRef &q) {
            const float c[3] = {q.centroid(0), q.centroid(1), q.centroid(2)};
            d.box.grow(q.lo, q.hi);
            d.cbox.grow_point(c);
        };
        int32_t m;
        if (!wide) {
            int32_t i = r.start, j = r.end - 1;
            for (;;) {
                while (i <= j && goes_left(ref[i])) add(sides[0], ref[i++]);
                while (i <= j && !goes_left(ref[j])) add(sides[1], ref[j--]);
                if (i >= j) break;
                std::swap(ref[i], ref[j]);
                add(sides[0], ref[i++]);
                add(sides[1], ref[j--]);
            }
            m = i;
        } else {
            // Stable partition by all threads: count each chunk's left items, then scatter both sides.
            std::vector<int32_t> lefts(T + 1, 0);
            if (!tmp) throw std::logic_error("no scratch for a parallel split");
            parallel_run(T, [&](int t) {
                const int32_t a0 = r.start + (int32_t)((int64_t)r.count() * t / T);
                const int32_t b0 = r.start + (int32_t)((int64_t)r.count() * (t + 1) / T);
                int32_t c = 0;
                for (int32_t k = a0; k < b0; ++k) c += goes_left(ref[k]) ? 1 : 0;
                lefts[t + 1] = c;
            });
            for (int t = 0; t < T; ++t) lefts[t + 1] += lefts[t];
            const int32_t total_left = lefts[T];
            std::vector<Range> got(2 * T);
            parallel_run(T, [&](int t) {
                const int32_t a0 = r.start + (int32_t)((int64_t)r.count() * t / T);
                const int32_t b0 = r.start + (int32_t)((int64_t)r.count() * (t + 1) / T);
                int32_t l = lefts[t], g = total_left + (a0 - r.start - lefts[t]);
                PrimRef *out = tmp + r.start;
                for (int32_t k = a0; k < b0; ++k) {
                    if (goes_left(ref[k])) {
                        add(got[2 * t], ref[k]);
                        out[l++] = ref[k];
                    } else {
                        add(got[2 * t + 1], ref[k]);
                        out[g++] = ref[k];
                    }
                }
            });
            for (int t = 0; t < T; ++t)
                for (int side = 0; side < 2; ++side) {
                    sides[side].box.grow(got[2 * t + side].box);
                    sides[side].cbox.grow(got[2 * t + side].cbox);
                }
            parallel_run(T, [&](int t) {
                const int32_t a0 = (int32_t)((int64_t)r.count() * t / T), b0 = (int32_t)((int64_t)r.count() * (t + 1) / T);
                std::memcpy(ref + r.start + a0, tmp + r.start + a0, sizeof(PrimRef) * (size_t)(b0 - a0));
            });
            m = r.start + total_left;
        }
        if (m <= r.start || m >= r.end) return false;
        left = sides[0];
        right = sides[1];
        left.start = r.start; left.end = m;
        right.start = m; right.end = r.end;
        return true;
    }

    static int bin_of(float c, float origin, float s, int B) {
        float f = (c - origin) * s;
        int b = (int)f;
        return std::min(B - 1, std::max(0, b));
    }

    PrimRef *ref;
    PrimRef *tmp;
    Bins local;  // the bins of the current split (reused: clearing the used bins is all a split pays)
    BuildOptions opt;
    bool median;
    int bins;
};

int thread_count() {
    unsigned h = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("VT_BUILD_THREADS")) {
        int v = std::atoi(e);
        if (v > 0) return v;
    }
    return (int)std::max(1u, std::min(h, 16u));
}

// Builds the wide tree; returns nodes (root at 0) and the depth.
int build_wide(PrimRef *refs, int32_t n, const BuildOptions &opt, bool median, std::vector<WideNode> &nodes) {
    std::unique_ptr<PrimRef[]> scratch(n >= kParallelItems ? new PrimRef[n] : nullptr);
    Builder root_builder(refs, scratch.get(), opt, median);
    std::vector<Builder::Task> deferred;
    int32_t threshold = n >= 65536 ? n / 64 : 0;
    Range root = root_builder.measure(0, n);
    int depth = root_builder.run(root, 1, nodes, threshold, deferred);
    if (deferred.empty()) return depth;
    std::vector<std::vector<WideNode>> parts(deferred.size());
    std::vector<int> depths(deferred.size(), 0);
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (;;) {
            size_t t = next.fetch_add(1);
            if (t >= deferred.size()) return;
            Builder b(refs, scratch.get(), opt, median);
            std::vector<Builder::Task> none;
            depths[t] = b.run(deferred[t].r, deferred[t].level, parts[t], 0, none);
        }
    };
    int T = std::min<int>(thread_count(), (int)deferred.size());
    std::vector<std::thread> pool;
    for (int i = 1; i < T; ++i) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
    for (size_t t = 0; t < deferred.size(); ++t) {
        int32_t offset = (int32_t)nodes.size();
        nodes[deferred[t].parent].child[deferred[t].slot] = offset;
        for (WideNode nd : parts[t]) {
            for (int k = 0; k < nd.nchild; ++k)
                if (nd.count[k] == 0) nd.child[k] += offset;  // internal link, local -> global
            nodes.push_back(nd);
        }
        depth = std::max(depth, depths[t]);
    }
    return depth;
}

// Biased exponent e of the smallest power-of-two scale with 255 * 2^(e-127) >= extent.
uint8_t grid_exponent(float extent) {
    int e = 1;
    if (extent > 0) {
        int x;
        std::frexp((double)extent / 255.0, &x);  // extent/255 = m * 2^x, m in [0.5, 1)
        e = std::max(1, x + 127);
        while (e > 1 && std::ldexp(255.0, e - 1 - 127) >= (double)extent) --e;
        while (std::ldexp(255.0, e - 127) < (double)extent) ++e;
    }
    if (e > 254) throw std::runtime_error("scene extent too large for the node grid");
    return (uint8_t)e;
}

float exp_scale(uint8_t e) {
    uint32_t bits = (uint32_t)e << 23;
    float s;
    std::memcpy(&s, &bits, 4);
    return s;
}

// SAH-optimal collapse of a binary tree into the 8-wide tree (Ylitie, Karras and Laine 2017, "Efficient
// Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", section 3.1).  The binary tree is
// built top-down with the same binned SAH split down to single triangles.  Bottom-up, C(n, k) is the
// least SAH cost of representing the subtree of n by at most k slots of a wide node: n itself (a
// leaf of at most leaf_size triangles, cost A(n) c_tri count, or a wide node, cost A(n) c_node plus
// its children's best use of 8 slots), or its two children sharing the k slots.  Greedy widening
// leaves many two-child nodes at the bottom of the tree (46% of the nodes of the 1.3M-triangle
// icosphere field); the optimal collapse fills them.
// Storage of `n` elements left uninitialised: every element is written before it is read, so the
// pages are first touched by the threads that fill them rather than by one zeroing pass.
template <class T>
std::unique_ptr<T[]> raw_array(size_t n) {
    static_assert(std::is_trivially_default_constructible<T>::value, "raw_array needs a trivial type");
    return std::unique_ptr<T[]>(new T[n]);
}

struct PlainBox {
    float lo[3], hi[3];
    float half_area() const {
        float e[3];
        for (int a = 0; a < 3; ++a) e[a] = std::max(hi[a] - lo[a], 0.0f);
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
    }
};

struct BinTree {
    std::unique_ptr<PlainBox[]> box;
    std::unique_ptr<int32_t[]> start, count, left, right;  // left < 0: a leaf (a single triangle)
    size_t size = 0;
    // A binary tree whose leaves hold single items has exactly 2n - 1 nodes; the arrays are sized
    // once and subtrees built concurrently write their own index ranges.
    void resize(size_t n) {
        box = raw_array<PlainBox>(n);
        start = raw_array<int32_t>(n);
        count = raw_array<int32_t>(n);
        left = raw_array<int32_t>(n);
        right = raw_array<int32_t>(n);
        size = n;
    }
    void set(int32_t i, const Range &r) {
        for (int a = 0; a < 3; ++a) {
            box[i].lo[a] = r.box.lo[a];
            box[i].hi[a] = r.box.hi[a];
        }
        start[i] = r.start;
        count[i] = r.count();
        left[i] = -1;
        right[i] = -1;
    }
};

// A node visit costs the traversal kernel about 300 wave instructions whatever the number of children;
// a triangle test about 3 per lane when the wave tests them together.  Replayed on the CPU, the
// icosphere field's bounce rays visit 13.04 nodes per ray with greedy widening, 12.79 with a triangle
// cost of 0.3, 12.64 with 0.1 and 13.27 with 0.6 (the 10M heightfield: 17.53 -> 16.84 at 0.1).
constexpr float kCostNode = 1.0f, kCostTri = 0.1f;

// Builds the binary tree below `root` (already stored at index root_index of t), taking new node
// indices from `next`; ranges smaller than `defer` go to `deferred` as (range, node index) and stay
// leaves to be replaced.
void build_binary(Builder &b, const Range &root, int32_t root_index, BinTree &t, int32_t &next, int32_t defer,
                  std::vector<std::pair<Range, int32_t>> *deferred) {
    std::vector<std::pair<Range, int32_t>> work{{root, root_index}};
    while (!work.empty()) {
        auto [r, i] = work.back();
        work.pop_back();
        if (r.count() <= 1) continue;
        if (deferred && r.count() < defer) {
            deferred->push_back({r, i});
            continue;
        }
        Range a, c;
        b.split(r, a, c);
        const int32_t ia = next++, ic = next++;
        t.set(ia, a);
        t.set(ic, c);
        t.left[i] = ia;
        t.right[i] = ic;
        work.push_back({c, ic});
        work.push_back({a, ia});
    }
}

struct Collapse {
    const BinTree &t;
    int leaf;
    // Written by solve for every node before any read (dint only for nodes with children).
    std::unique_ptr<float[]> cost;     // 8 per node: C(n, 1..8)
    std::unique_ptr<uint8_t[]> dec;    // 8 per node: [0] 1 = leaf; [k-1], k >= 2: 0 = as k-1 slots, else slots given to left
    std::unique_ptr<uint8_t[]> dint;   // slots given to left when n is a wide node

    Collapse(const BinTree &tree, int leaf_size)
        : t(tree), leaf(leaf_size), cost(raw_array<float>(8 * tree.size)), dec(raw_array<uint8_t>(8 * tree.size)),
          dint(raw_array<uint8_t>(tree.size)) {}

    // Children have larger indices than their parent, so a descending sweep over [lo, hi) sees them first.
    void solve(int32_t lo, int32_t hi) {
        for (int32_t n = hi - 1; n >= lo; --n) {
            const float A = t.box[n].half_area();
            float *C = &cost[8 * (size_t)n];
            uint8_t *d = &dec[8 * (size_t)n];
            const float c_leaf = t.count[n] <= leaf ? A * kCostTri * (float)t.count[n] : FLT_MAX;
            if (t.left[n] < 0) {
                for (int k = 0; k < 8; ++k) { C[k] = c_leaf; d[k] = 0; }
                d[0] = 1;
                continue;
            }
            const float *L = &cost[8 * (size_t)t.left[n]], *R = &cost[8 * (size_t)t.right[n]];
            float D[9];
            uint8_t J[9];
            for (int k = 2; k <= 8; ++k) {
                D[k] = FLT_MAX;
                J[k] = 1;
                for (int j = 1; j < k; ++j) {
                    const float v = L[j - 1] + R[k - j - 1];
                    if (v < D[k]) { D[k] = v; J[k] = (uint8_t)j; }
                }
            }
            const float c_int = A * kCostNode + D[8];
            dint[n] = J[8];
            d[0] = c_leaf <= c_int ? 1 : 0;
            C[0] = std::min(c_leaf, c_int);
            for (int k = 2; k <= 8; ++k) {
                if (D[k] < C[k - 2]) { C[k - 1] = D[k]; d[k - 1] = J[k]; }
                else { C[k - 1] = C[k - 2]; d[k - 1] = 0; }
            }
        }
    }

    // The slots (binary nodes) representing the subtree of n in at most k slots, appended to out.
    void slots(int32_t n, int k, int32_t *out, int &cnt) const {
        while (k > 1 && t.left[n] >= 0 && dec[8 * (size_t)n + k - 1] == 0) --k;
        if (k == 1 || t.left[n] < 0) {
            out[cnt++] = n;
            return;
        }
        const int j = dec[8 * (size_t)n + k - 1];
        slots(t.left[n], j, out, cnt);
        slots(t.right[n], k - j, out, cnt);
    }

    // The wide node of binary node n; its internal children (binary nodes) in `inner` as (node, slot).
    WideNode make(int32_t n, int32_t inner[8][2], int &ninner) const {
        int32_t sl[8];
        int cnt = 0;
        if (t.left[n] < 0) {
            sl[cnt++] = n;
        } else {
            slots(t.left[n], dint[n], sl, cnt);
            slots(t.right[n], 8 - dint[n], sl, cnt);
        }
        WideNode w;
        w.nchild = cnt;
        ninner = 0;
        for (int k = 0; k < cnt; ++k) {
            const int32_t m = sl[k];
            w.box[k].grow(t.box[m].lo, t.box[m].hi);
            if (dec[8 * (size_t)m] == 1) {
                w.child[k] = -(t.start[m] + 1);
                w.count[k] = t.count[m];
            } else {
                w.child[k] = 0;
                w.count[k] = 0;
                inner[ninner][0] = m;
                inner[ninner][1] = k;
                ++ninner;
            }
        }
        return w;
    }

    // Emits the wide subtree of binary node n with its root at nodes[at]; returns its depth.  Subtrees
    // of fewer than `defer` items are left to `tasks` as (binary node, parent index, slot, level).
    struct Task {
        int32_t node, parent;
        int slot, level;
    };
    int emit(int32_t n, int32_t at, std::vector<WideNode> &nodes, int level, int32_t defer, std::vector<Task> *tasks) const {
        int32_t inner[8][2];
        int ninner;
        nodes[at] = make(n, inner, ninner);
        int depth = 1;
        for (int q = 0; q < ninner; ++q) {
            const int32_t m = inner[q][0];
            if (tasks && t.count[m] < defer) {
                tasks->push_back(Task{m, at, inner[q][1], level + 1});
                continue;
            }
            const int32_t idx = (int32_t)nodes.size();
            nodes.emplace_back();
            nodes[at].child[inner[q][1]] = idx;
            depth = std::max(depth, 1 + emit(m, idx, nodes, level + 1, defer, tasks));
        }
        return depth;
    }

    // The whole wide tree into `nodes` (root at 0): the top serially, then subtrees below 1/256 of the
    // items concurrently into private lists, appended in a fixed order (the result does not depend on
    // the thread count).  Returns the depth.
    int emit_all(int32_t n, std::vector<WideNode> &nodes) const {
        nodes.clear();
        nodes.emplace_back();
        std::vector<Task> tasks;
        const int32_t defer = n >= 65536 ? n / 256 : 0;
        int depth = emit(0, 0, nodes, 1, defer, defer ? &tasks : nullptr);
        std::vector<std::vector<WideNode>> parts(tasks.size());
        std::vector<int> depths(tasks.size(), 0);
        std::atomic<size_t> cursor{0};
        parallel_run(std::min<int>(thread_count(), std::max<int>(1, (int)tasks.size())), [&](int) {
            for (;;) {
                const size_t k = cursor.fetch_add(1);
                if (k >= tasks.size()) return;
                parts[k].emplace_back();
                depths[k] = emit(tasks[k].node, 0, parts[k], tasks[k].level, 0, nullptr);
            }
        });
        for (size_t k = 0; k < tasks.size(); ++k) {
            const int32_t offset = (int32_t)nodes.size();
            nodes[tasks[k].parent].child[tasks[k].slot] = offset;
            for (WideNode w : parts[k]) {
                for (int q = 0; q < w.nchild; ++q)
                    if (w.count[q] == 0) w.child[q] += offset;  // internal link, local -> global
                nodes.push_back(w);
            }
            depth = std::max(depth, tasks[k].level - 1 + depths[k]);
        }
        return depth;
    }
};

int build_wide_optimal(PrimRef *refs, int32_t n, const BuildOptions &opt, std::vector<WideNode> &nodes) {
    std::unique_ptr<PrimRef[]> scratch(n >= kParallelItems ? new PrimRef[n] : nullptr);
    Builder top(refs, scratch.get(), opt, false);
    const Range root = top.measure(0, n);
    BinTree t;
    t.resize(2 * (size_t)n - 1);
    t.set(0, root);
    int32_t next = 1;
    std::vector<std::pair<Range, int32_t>> deferred;
    build_binary(top, root, 0, t, next, n >= 65536 ? n / 64 : 0, n >= 65536 ? &deferred : nullptr);
    const int32_t top_size = next;
    // Deferred subtrees are built concurrently, each into its own index range after the top: part k
    // of m_k items takes 2 m_k - 2 nodes (its root is the top's placeholder).
    std::vector<int32_t> offset(deferred.size() + 1, top_size);
    for (size_t k = 0; k < deferred.size(); ++k) offset[k + 1] = offset[k] + 2 * deferred[k].first.count() - 2;
    if (offset.back() != 2 * n - 1) throw std::runtime_error("binary tree size mismatch");
    // Largest parts first, so that the last ones to finish are small.
    std::vector<size_t> by_size(deferred.size());
    for (size_t k = 0; k < by_size.size(); ++k) by_size[k] = k;
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](size_t x, size_t y) { return deferred[x].first.count() > deferred[y].first.count(); });
    std::atomic<size_t> cursor{0};
    parallel_run(std::min<int>(thread_count(), std::max<int>(1, (int)deferred.size())), [&](int) {
        for (;;) {
            const size_t q = cursor.fetch_add(1);
            if (q >= deferred.size()) return;
            const size_t k = by_size[q];
            Builder b(refs, scratch.get(), opt, false);
            int32_t nx = offset[k];
            build_binary(b, deferred[k].first, deferred[k].second, t, nx, 0, nullptr);
            if (nx != offset[k + 1]) throw std::runtime_error("binary subtree size mismatch");
        }
    });
    Collapse c(t, opt.leaf_size);
    // Solve the parts concurrently, then the top (whose placeholders' children lie in the parts).
    cursor = 0;
    parallel_run(std::min<int>(thread_count(), std::max<int>(1, (int)deferred.size())), [&](int) {
        for (;;) {
            const size_t q = cursor.fetch_add(1);
            if (q >= deferred.size()) return;
            const size_t k = by_size[q];
            c.solve(offset[k], offset[k + 1]);
        }
    });
    c.solve(0, top_size);
    return c.emit_all(n, nodes);
}

// Octant slot assignment: slot s gets the child whose centre projects lowest on the sign
// vector of octant s (direction component negative where bit a of s is set).
void assign_slots(const WideNode &w, int slot_of[8]) {
    Box parent;
    for (int k = 0; k < w.nchild; ++k) parent.grow(w.box[k]);
    float pc[3];
    for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (parent.lo[a] + parent.hi[a]);
    float cost[8][8];
    for (int k = 0; k < w.nchild; ++k) {
        float c[3];
        for (int a = 0; a < 3; ++a) c[a] = 0.5f * (w.box[k].lo[a] + w.box[k].hi[a]) - pc[a];
        for (int s = 0; s < 8; ++s) {
            float v = 0;
            for (int a = 0; a < 3; ++a) v += (s >> a & 1) ? -c[a] : c[a];
            cost[k][s] = v;
        }
    }
    bool child_done[8] = {}, slot_done[8] = {};
    for (int k = 0; k < 8; ++k) slot_of[k] = -1;
    for (int round = 0; round < w.nchild; ++round) {
        int bk = -1, bs = -1;
        float bv = FLT_MAX;
        for (int k = 0; k < w.nchild; ++k) {
            if (child_done[k]) continue;
            for (int s = 0; s < 8; ++s) {
                if (slot_done[s]) continue;
                if (bk < 0 || cost[k][s] < bv) { bv = cost[k][s]; bk = k; bs = s; }
            }
        }
        child_done[bk] = true;
        slot_done[bs] = true;
        slot_of[bk] = bs;
    }
}

struct Layout {
    const std::vector<WideNode> &W;
    const int32_t *order;
    const float *vertices;
    const uint32_t *faces;
    std::vector<Node8> &out;
    std::vector<TriRecord> &tris;
    std::vector<uint32_t> *items = nullptr;  // box hierarchies: leaf slots name items, not triangles

    void tri_record(int32_t item, TriRecord &t) const {
        uint32_t f = (uint32_t)item;
        const uint32_t *fc = faces + 4 * (size_t)f;
        const float *a = vertices + 8 * (size_t)fc[0], *b = vertices + 8 * (size_t)fc[1], *c = vertices + 8 * (size_t)fc[2];
        for (int q = 0; q < 3; ++q) { t.v0[q] = a[q]; t.v1[q] = b[q]; t.v2[q] = c[q]; }
        t.face = f;
        t.pad1 = 0;
        t.pad2 = 0;
    }

    // Lays the wide tree out: every node's internal children contiguous from child_base and its
    // leaves' triangles contiguous from tri_base, in slot order, depth first from the root (node 0).
    // The placement is a cheap serial walk; slot assignment, quantisation and the triangle records
    // are computed for all nodes concurrently.
    void run(uint32_t face_count) {
        const size_t N = W.size();
        std::vector<std::array<int8_t, 8>> child_at(N);
        const int T = N >= 4096 ? thread_count() : 1;
        parallel_run(T, [&](int th) {
            for (size_t i = N * th / T; i < N * (th + 1) / T; ++i) {
                int slot_of[8];
                assign_slots(W[i], slot_of);
                for (int q = 0; q < 8; ++q) child_at[i][q] = -1;
                for (int k = 0; k < W[i].nchild; ++k) child_at[i][slot_of[k]] = (int8_t)k;
            }
        });
        std::vector<uint32_t> dst(N), cbase(N), tbase(N);
        uint32_t next_node = 1, next_tri = 0;
        std::vector<std::pair<int32_t, uint32_t>> stack{{0, 0u}};
        while (!stack.empty()) {
            const auto [src, d] = stack.back();
            stack.pop_back();
            const WideNode &w = W[src];
            dst[src] = d;
            int internal = 0;
            tbase[src] = next_tri;
            for (int q = 0; q < 8; ++q) {
                const int k = child_at[src][q];
                if (k < 0) continue;
                if (w.count[k] == 0) ++internal;
                else next_tri += (uint32_t)w.count[k];
            }
            cbase[src] = internal ? next_node : 0u;
            next_node += (uint32_t)internal;
            int rank = internal;
            for (int q = 7; q >= 0; --q) {
                const int k = child_at[src][q];
                if (k >= 0 && w.count[k] == 0) stack.push_back({w.child[k], cbase[src] + (uint32_t)--rank});
            }
        }
        if (next_node != N || next_tri != face_count) throw std::runtime_error("BVH layout size mismatch");
        out.resize(N);
        if (items) items->resize(face_count);
        else tris.resize(face_count);
        // Contiguous runs of wide nodes, whose records and triangles lie close together in the output
        // (interleaved nodes would have threads writing into the same cache lines).
        parallel_run(4 * T, [&](int th) {
            const size_t P = 4 * (size_t)T;
            for (size_t i = N * th / P; i < N * (th + 1) / P; ++i) write(i, child_at[i].data(), dst[i], cbase[i], tbase[i]);
        });
    }

    // The record of wide node `src` at out[d], its triangles from tris[tri_base].
    void write(size_t src, const int8_t *child_at, uint32_t d, uint32_t child_base, uint32_t tri_base) {
        const WideNode &w = W[src];
        Node8 node;
        std::memset(&node, 0, sizeof node);
        Box parent;
        for (int k = 0; k < w.nchild; ++k) parent.grow(w.box[k]);
        node.px = parent.lo[0]; node.py = parent.lo[1]; node.pz = parent.lo[2];
        uint8_t e[3];
        float sc[3], p[3] = {parent.lo[0], parent.lo[1], parent.lo[2]};
        for (int a = 0; a < 3; ++a) {
            e[a] = grid_exponent(parent.hi[a] - parent.lo[a]);
            sc[a] = exp_scale(e[a]);
        }
        node.ex = e[0]; node.ey = e[1]; node.ez = e[2];
        node.child_base = child_base;
        node.tri_base = tri_base;
        int tri_off = 0;
        for (int s = 0; s < 8; ++s) {
            const int k = child_at[s];
            if (k < 0) continue;
            const Box &b = w.box[k];
            for (int a = 0; a < 3; ++a) {
                node.qlo[a][s] = quant_lo(b.lo[a], p[a], sc[a]);
                node.qhi[a][s] = quant_hi(b.hi[a], p[a], sc[a]);
            }
            if (w.count[k] == 0) {
                node.imask |= (uint8_t)(1u << s);
                node.meta[s] = 0x80;
            } else {
                const int32_t first = -w.child[k] - 1;
                const int c = w.count[k];
                if (c < 1 || c > 3 || tri_off + c > 24) throw std::runtime_error("leaf does not fit the node encoding");
                node.meta[s] = (uint8_t)((tri_off << 2) | c);
                for (int q = 0; q < c; ++q) {
                    if (items) (*items)[tri_base + (uint32_t)(tri_off + q)] = (uint32_t)order[first + q];
                    else tri_record(order[first + q], tris[tri_base + (uint32_t)(tri_off + q)]);
                }
                tri_off += c;
            }
        }
        out[d] = node;
    }

    static uint8_t quant_lo(float v, float p, float s) {
        double q = std::floor(((double)v - (double)p) / (double)s);
        int qi = (int)std::max(0.0, std::min(255.0, q));
        while (qi > 0 && std::fma((float)qi, s, p) > v) --qi;
        return (uint8_t)qi;
    }
    static uint8_t quant_hi(float v, float p, float s) {
        double q = std::ceil(((double)v - (double)p) / (double)s);
        int qi = (int)std::max(0.0, std::min(255.0, q));
        while (qi < 255 && std::fma((float)qi, s, p) < v) ++qi;
        if (std::fma((float)qi, s, p) < v) throw std::runtime_error("node grid does not enclose a child");
        return (uint8_t)qi;
    }
};

}  // namespace

// The wide tree over n items whose bounds `item_bounds(refs, scene)` writes in document order, laid
// out over the triangles of `faces` (tris) or, with `items`, naming the items of each leaf slot.
template <class ItemBounds>
static Bvh8 build_over(uint32_t n, const BuildOptions &opt_in, ItemBounds &&item_bounds, const float *vertices,
                       const uint32_t *faces, std::vector<uint32_t> *items) {
    auto t0 = std::chrono::steady_clock::now();
    BuildOptions opt = opt_in;
    if (opt.leaf_size < 1 || opt.leaf_size > 3) throw std::invalid_argument("leaf_size must be 1..3");
    if (opt.bins < 4 || opt.bins > 64) throw std::invalid_argument("bins must be 4..64");
    if (opt.max_depth < 2 || opt.max_depth > 24) throw std::invalid_argument("max_depth must be 2..24");
    if (opt.collapse < 0 || opt.collapse > 1) throw std::invalid_argument("collapse must be 0 or 1");
    Bvh8 out;
    if (n == 0) return out;
    std::unique_ptr<PrimRef[]> refs = raw_array<PrimRef>(n);
    Box scene;
    item_bounds(refs.get(), scene);

    // Each attempt starts from the items in document order; the order the winning attempt leaves
    // them in is the leaf order.
    std::vector<WideNode> wide;
    int depth = opt.max_depth + 1;
    if (opt.collapse == 0) depth = build_wide_optimal(refs.get(), (int32_t)n, opt, wide);
    if (depth > opt.max_depth) {
        wide.clear();
        if (opt.collapse == 0) item_bounds(refs.get(), scene);
        depth = build_wide(refs.get(), (int32_t)n, opt, false, wide);
    }
    if (depth > opt.max_depth) {
        wide.clear();
        item_bounds(refs.get(), scene);
        depth = build_wide(refs.get(), (int32_t)n, opt, true, wide);
        if (depth > opt.max_depth) throw std::runtime_error("BVH deeper than the traversal stack");
    }
    std::vector<int32_t> order(n);
    for (uint32_t k = 0; k < n; ++k) order[k] = refs[k].idx;
    refs.reset();
    out.depth = depth;
    Layout L{wide, order.data(), vertices, faces, out.nodes, out.tris, items};
    L.run(n);
    float m = 0;
    for (int a = 0; a < 3; ++a) m = std::max(m, std::max(std::fabs(scene.lo[a]), std::fabs(scene.hi[a])));
    out.root_max = m;
    out.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

Bvh8 build_bvh8_boxes(const float *boxes, uint32_t n, const BuildOptions &opt, std::vector<uint32_t> &items) {
    auto item_bounds = [&](PrimRef *refs, Box &scene) {
        scene = Box();
        for (uint32_t i = 0; i < n; ++i) {
            PrimRef &q = refs[i];
            for (int k = 0; k < 3; ++k) {
                q.lo[k] = boxes[6 * (size_t)i + k];
                q.hi[k] = boxes[6 * (size_t)i + 3 + k];
                if (!std::isfinite(q.lo[k]) || !std::isfinite(q.hi[k]) || q.lo[k] > q.hi[k])
                    throw std::invalid_argument("invalid item box");
            }
            q.idx = (int32_t)i;
            q.pad = 0.0f;
            scene.grow(q.lo, q.hi);
        }
    };
    return build_over(n, opt, item_bounds, nullptr, nullptr, &items);
}

Bvh8 build_bvh8(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t n,
                const BuildOptions &opt_in) {
    // Item bounds (the reference's MeshSceneLayout.bounds, Sources/Acceleration.swift:445-458) in
    // document order, by all threads; the scene box.
    auto item_bounds = [&](PrimRef *refs, Box &scene) {
        const int T = n >= 65536 ? thread_count() : 1;
        std::vector<Box> part(T);
        parallel_run(T, [&](int t) {
            const size_t s0 = (size_t)n * t / T, e0 = (size_t)n * (t + 1) / T;
            for (size_t i = s0; i < e0; ++i) {
                const uint32_t *fc = faces + 4 * i;
                if (fc[0] >= vertex_count || fc[1] >= vertex_count || fc[2] >= vertex_count)
                    throw std::invalid_argument("face index out of range");
                const float *a = vertices + 8 * (size_t)fc[0], *b = vertices + 8 * (size_t)fc[1], *c = vertices + 8 * (size_t)fc[2];
                PrimRef &q = refs[i];
                for (int k = 0; k < 3; ++k) {
                    const float l = std::min(std::min(a[k], b[k]), c[k]), h = std::max(std::max(a[k], b[k]), c[k]);
                    if (!std::isfinite(l) || !std::isfinite(h)) throw std::invalid_argument("non-finite vertex");
                    q.lo[k] = l;
                    q.hi[k] = h;
                }
                q.idx = (int32_t)i;
                q.pad = 0.0f;
                part[t].grow(q.lo, q.hi);
            }
        });
        scene = Box();
        for (auto &b : part) scene.grow(b);
    };
    return build_over(n, opt_in, item_bounds, vertices, faces, nullptr);
}

}  // namespace vt
