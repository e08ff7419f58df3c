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
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>

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

struct Prims {
    std::vector<float> bounds;    // 6 per item
    std::vector<float> centroid;  // 3 per item
};

class Builder {
public:
    Builder(const Prims &p, int32_t *order, const BuildOptions &o, bool median)
        : P(p), order(order), opt(o), median(median), bins(o.bins) {}

    Range measure(int32_t s, int32_t e) const {
        Range r;
        r.start = s;
        r.end = e;
        for (int32_t k = s; k < e; ++k) {
            int32_t i = order[k];
            r.box.grow(&P.bounds[6 * (size_t)i], &P.bounds[6 * (size_t)i + 3]);
            r.cbox.grow_point(&P.centroid[3 * (size_t)i]);
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
            const float *c = P.centroid.data();
            std::nth_element(order + r.start, order + mid, order + r.end, [c, a](int32_t x, int32_t y) {
                float cx = c[3 * (size_t)x + a], cy = c[3 * (size_t)y + a];
                return cx == cy ? x < y : cx < cy;
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
    bool binned_split(const Range &r, const float *ext, Range &left, Range &right) {
        const int B = bins;
        float scale[3];
        for (int a = 0; a < 3; ++a) {
            float s = (float)B / ext[a];
            scale[a] = (ext[a] > 0 && std::isfinite(s)) ? s : 0.0f;
        }
        Box bb[3][64], cb[3][64];
        int32_t cnt[3][64];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < B; ++b) { bb[a][b] = Box(); cb[a][b] = Box(); cnt[a][b] = 0; }
        for (int32_t k = r.start; k < r.end; ++k) {
            int32_t i = order[k];
            const float *bl = &P.bounds[6 * (size_t)i];
            const float *c = &P.centroid[3 * (size_t)i];
            for (int a = 0; a < 3; ++a) {
                int b = bin_of(c[a], r.cbox.lo[a], scale[a]);
                cnt[a][b] += 1;
                bb[a][b].grow(bl, bl + 3);
                cb[a][b].grow_point(c);
            }
        }
        float best_cost = FLT_MAX;
        int best_axis = -1, best_bin = 0;
        float suffix[64];
        for (int a = 0; a < 3; ++a) {
            if (!(scale[a] > 0)) continue;
            Box acc;
            int32_t c = 0;
            for (int b = B - 1; b > 0; --b) {
                if (cnt[a][b] > 0) acc.grow(bb[a][b]);
                c += cnt[a][b];
                suffix[b] = c > 0 ? acc.half_area() * (float)c : 0.0f;
            }
            acc = Box();
            c = 0;
            for (int b = 0; b < B - 1; ++b) {
                if (cnt[a][b] > 0) acc.grow(bb[a][b]);
                c += cnt[a][b];
                if (c == 0 || c == r.count()) continue;
                float cost = acc.half_area() * (float)c + suffix[b + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_bin = b; }
            }
        }
        if (best_axis < 0) return false;
        const int a = best_axis;
        const float s = scale[a], origin = r.cbox.lo[a];
        const float *cen = P.centroid.data();
        int32_t *mid = std::partition(order + r.start, order + r.end, [&](int32_t i) {
            return bin_of(cen[3 * (size_t)i + a], origin, s) <= best_bin;
        });
        int32_t m = (int32_t)(mid - order);
        if (m <= r.start || m >= r.end) return false;
        left = Range();
        right = Range();
        left.start = r.start; left.end = m;
        right.start = m; right.end = r.end;
        for (int b = 0; b < B; ++b) {
            if (cnt[a][b] <= 0) continue;
            Range &d = b <= best_bin ? left : right;
            d.box.grow(bb[a][b]);
            d.cbox.grow(cb[a][b]);
        }
        return true;
    }

    int bin_of(float c, float origin, float s) const {
        float f = (c - origin) * s;
        int b = (int)f;
        return std::min(bins - 1, std::max(0, b));
    }

    const Prims &P;
    int32_t *order;
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
int build_wide(const Prims &P, int32_t n, int32_t *order, const BuildOptions &opt, bool median,
               std::vector<WideNode> &nodes) {
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    Builder root_builder(P, order, opt, median);
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
            Builder b(P, order, opt, median);
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
struct BinTree {
    std::vector<Box> box;
    std::vector<int32_t> start, count, left, right;  // left < 0: a leaf (a single triangle)
    int32_t add(const Range &r) {
        box.push_back(r.box);
        start.push_back(r.start);
        count.push_back(r.count());
        left.push_back(-1);
        right.push_back(-1);
        return (int32_t)box.size() - 1;
    }
};

// A node visit costs the traversal kernel about 300 wave instructions whatever the number of children;
// a triangle test about 3 per lane when the wave tests them together.  Replayed on the CPU, the
// icosphere field's bounce rays visit 13.04 nodes per ray with greedy widening, 12.79 with a triangle
// cost of 0.3, 12.64 with 0.1 and 13.27 with 0.6 (the 10M heightfield: 17.53 -> 16.84 at 0.1).
constexpr float kCostNode = 1.0f, kCostTri = 0.1f;

// Builds the binary tree of `root` into t (root at the first new index); ranges smaller than `defer`
// go to `deferred` as (range, node index) and are left as leaves to be replaced.
void build_binary(Builder &b, const Range &root, BinTree &t, int32_t defer,
                  std::vector<std::pair<Range, int32_t>> *deferred) {
    std::vector<std::pair<Range, int32_t>> work{{root, t.add(root)}};
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
        const int32_t ia = t.add(a), ic = t.add(c);
        t.left[i] = ia;
        t.right[i] = ic;
        work.push_back({c, ic});
        work.push_back({a, ia});
    }
}

struct Collapse {
    const BinTree &t;
    int leaf;
    std::vector<float> cost;     // 8 per node: C(n, 1..8)
    std::vector<uint8_t> dec;    // 8 per node: [0] 1 = leaf; [k-1], k >= 2: 0 = as k-1 slots, else slots given to left
    std::vector<uint8_t> dint;   // slots given to left when n is a wide node

    Collapse(const BinTree &tree, int leaf_size)
        : t(tree), leaf(leaf_size), cost(8 * tree.box.size()), dec(8 * tree.box.size()), dint(tree.box.size()) {}

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

    // The slots (binary nodes) representing the subtree of n in at most k slots.
    void slots(int32_t n, int k, std::vector<int32_t> &out) const {
        while (k > 1 && t.left[n] >= 0 && dec[8 * (size_t)n + k - 1] == 0) --k;
        if (k == 1 || t.left[n] < 0) {
            out.push_back(n);
            return;
        }
        const int j = dec[8 * (size_t)n + k - 1];
        slots(t.left[n], j, out);
        slots(t.right[n], k - j, out);
    }

    // Emits the wide node of binary node n (at index `at` of nodes); returns the depth below it.
    int emit(int32_t n, int32_t at, std::vector<WideNode> &nodes) const {
        std::vector<int32_t> sl;
        if (t.left[n] < 0) {
            sl.push_back(n);
        } else {
            slots(t.left[n], dint[n], sl);
            slots(t.right[n], 8 - dint[n], sl);
        }
        WideNode w;
        w.nchild = (int)sl.size();
        std::vector<std::pair<int32_t, int>> inner;
        for (int k = 0; k < w.nchild; ++k) {
            const int32_t m = sl[k];
            w.box[k] = t.box[m];
            if (dec[8 * (size_t)m] == 1) {
                w.child[k] = -(t.start[m] + 1);
                w.count[k] = t.count[m];
            } else {
                w.child[k] = 0;
                w.count[k] = 0;
                inner.push_back({m, k});
            }
        }
        nodes[at] = w;
        int depth = 1;
        for (auto &[m, k] : inner) {
            const int32_t idx = (int32_t)nodes.size();
            nodes.emplace_back();
            nodes[at].child[k] = idx;
            depth = std::max(depth, 1 + emit(m, idx, nodes));
        }
        return depth;
    }
};

int build_wide_optimal(const Prims &P, int32_t n, int32_t *order, const BuildOptions &opt, std::vector<WideNode> &nodes) {
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    Builder top(P, order, opt, false);
    const Range root = top.measure(0, n);
    BinTree t;
    t.box.reserve(2 * (size_t)n);
    std::vector<std::pair<Range, int32_t>> deferred;
    build_binary(top, root, t, n >= 65536 ? n / 64 : 0, n >= 65536 ? &deferred : nullptr);
    const int32_t top_size = (int32_t)t.box.size();
    // Deferred subtrees are built concurrently into private trees and appended in a fixed order.
    std::vector<BinTree> parts(deferred.size());
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= deferred.size()) return;
            Builder b(P, order, opt, false);
            build_binary(b, deferred[k].first, parts[k], 0, nullptr);
        }
    };
    {
        const int T = std::min<int>(thread_count(), std::max<int>(1, (int)deferred.size()));
        std::vector<std::thread> pool;
        for (int i = 1; i < T; ++i) pool.emplace_back(worker);
        worker();
        for (auto &th : pool) th.join();
    }
    std::vector<int32_t> offset(deferred.size());
    for (size_t k = 0; k < deferred.size(); ++k) {
        // The part's root replaces the placeholder; its other nodes are appended.
        const BinTree &q = parts[k];
        const int32_t base = (int32_t)t.box.size() - 1;  // part index j > 0 lands at base + j
        offset[k] = base + 1;
        const int32_t hole = deferred[k].second;
        auto map = [&](int32_t j) { return j < 0 ? j : (j == 0 ? hole : base + j); };
        t.left[hole] = map(q.left[0]);
        t.right[hole] = map(q.right[0]);
        for (size_t j = 1; j < q.box.size(); ++j) {
            t.box.push_back(q.box[j]);
            t.start.push_back(q.start[j]);
            t.count.push_back(q.count[j]);
            t.left.push_back(map(q.left[j]));
            t.right.push_back(map(q.right[j]));
        }
    }
    parts.clear();
    Collapse c(t, opt.leaf_size);
    // Parts occupy [offset[k], offset[k + 1]) after the top; solve them concurrently, then the top
    // (whose placeholders' children now lie in the parts).
    {
        std::vector<std::pair<int32_t, int32_t>> ranges;
        for (size_t k = 0; k < deferred.size(); ++k)
            ranges.push_back({offset[k], k + 1 < deferred.size() ? offset[k + 1] : (int32_t)t.box.size()});
        std::atomic<size_t> nx{0};
        auto solver = [&]() {
            for (;;) {
                const size_t k = nx.fetch_add(1);
                if (k >= ranges.size()) return;
                c.solve(ranges[k].first, ranges[k].second);
            }
        };
        const int T = std::min<int>(thread_count(), std::max<int>(1, (int)ranges.size()));
        std::vector<std::thread> pool;
        for (int i = 1; i < T; ++i) pool.emplace_back(solver);
        solver();
        for (auto &th : pool) th.join();
        c.solve(0, top_size);
    }
    nodes.clear();
    nodes.emplace_back();
    return c.emit(0, 0, nodes);
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

    void tri_record(int32_t item, TriRecord &t) const {
        uint32_t f = (uint32_t)item;
        const uint32_t *fc = faces + 4 * (size_t)f;
        const float *a = vertices + 8 * (size_t)fc[0], *b = vertices + 8 * (size_t)fc[1], *c = vertices + 8 * (size_t)fc[2];
        for (int q = 0; q < 3; ++q) { t.v0[q] = a[q]; t.v1[q] = b[q]; t.v2[q] = c[q]; }
        t.face = f;
        t.pad1 = 0;
        t.pad2 = 0;
    }

    // Fills out[dst] from wide node `src`, allocating its children and triangles.
    void emit(int32_t src, uint32_t dst) {
        const WideNode &w = W[src];
        int slot_of[8];
        assign_slots(w, slot_of);
        int child_at[8];
        for (int s = 0; s < 8; ++s) child_at[s] = -1;
        for (int k = 0; k < w.nchild; ++k) child_at[slot_of[k]] = k;

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

        // Internal children first (contiguous, slot order), then triangles (slot order).
        uint32_t child_base = (uint32_t)out.size();
        int internal = 0;
        for (int s = 0; s < 8; ++s) {
            int k = child_at[s];
            if (k >= 0 && w.count[k] == 0) ++internal;
        }
        out.resize(out.size() + internal);
        node.child_base = internal ? child_base : 0;
        node.tri_base = (uint32_t)tris.size();
        int tri_off = 0;
        std::vector<std::pair<int32_t, uint32_t>> recurse;
        int rank = 0;
        for (int s = 0; s < 8; ++s) {
            int k = child_at[s];
            if (k < 0) continue;
            const Box &b = w.box[k];
            for (int a = 0; a < 3; ++a) {
                node.qlo[a][s] = quant_lo(b.lo[a], p[a], sc[a]);
                node.qhi[a][s] = quant_hi(b.hi[a], p[a], sc[a]);
            }
            if (w.count[k] == 0) {
                node.imask |= (uint8_t)(1u << s);
                node.meta[s] = 0x80;
                recurse.push_back({w.child[k], child_base + (uint32_t)rank});
                ++rank;
            } else {
                int32_t first = -w.child[k] - 1;
                int c = w.count[k];
                if (c < 1 || c > 3 || tri_off + c > 24) throw std::runtime_error("leaf does not fit the node encoding");
                node.meta[s] = (uint8_t)((tri_off << 2) | c);
                for (int q = 0; q < c; ++q) {
                    TriRecord t;
                    tri_record(order[first + q], t);
                    tris.push_back(t);
                }
                tri_off += c;
            }
        }
        out[dst] = node;
        for (auto &r : recurse) emit(r.first, r.second);
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

Bvh8 build_bvh8(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t n,
                const BuildOptions &opt_in) {
    auto t0 = std::chrono::steady_clock::now();
    BuildOptions opt = opt_in;
    if (opt.leaf_size < 1 || opt.leaf_size > 3) throw std::invalid_argument("leaf_size must be 1..3");
    if (opt.bins < 4 || opt.bins > 64) throw std::invalid_argument("bins must be 4..64");
    if (opt.max_depth < 2 || opt.max_depth > 24) throw std::invalid_argument("max_depth must be 2..24");
    if (opt.collapse < 0 || opt.collapse > 1) throw std::invalid_argument("collapse must be 0 or 1");
    Bvh8 out;
    if (n == 0) return out;
    for (uint32_t f = 0; f < n; ++f)
        for (int q = 0; q < 3; ++q)
            if (faces[4 * (size_t)f + q] >= vertex_count) throw std::invalid_argument("face index out of range");

    Prims P;
    P.bounds.resize(6 * (size_t)n);
    P.centroid.resize(3 * (size_t)n);
    Box scene;
    {
        int T = thread_count();
        std::vector<std::thread> pool;
        std::vector<Box> part(T);
        auto work = [&](int t) {
            size_t s = (size_t)n * t / T, e = (size_t)n * (t + 1) / T;
            for (size_t i = s; i < e; ++i) {
                const uint32_t *fc = faces + 4 * i;
                const float *a = vertices + 8 * (size_t)fc[0], *b = vertices + 8 * (size_t)fc[1], *c = vertices + 8 * (size_t)fc[2];
                for (int q = 0; q < 3; ++q) {
                    float l = std::min(std::min(a[q], b[q]), c[q]), h = std::max(std::max(a[q], b[q]), c[q]);
                    if (!std::isfinite(l) || !std::isfinite(h)) throw std::invalid_argument("non-finite vertex");
                    P.bounds[6 * i + q] = l;
                    P.bounds[6 * i + 3 + q] = h;
                    P.centroid[3 * i + q] = (l + h) * 0.5f;
                }
                part[t].grow(&P.bounds[6 * i], &P.bounds[6 * i + 3]);
            }
        };
        std::vector<std::exception_ptr> errs(T);
        for (int t = 0; t < T; ++t)
            pool.emplace_back([&, t]() { try { work(t); } catch (...) { errs[t] = std::current_exception(); } });
        for (auto &th : pool) th.join();
        for (auto &e : errs) if (e) std::rethrow_exception(e);
        for (auto &b : part) scene.grow(b);
    }

    std::vector<int32_t> order(n);
    std::vector<WideNode> wide;
    int depth = opt.collapse == 0 ? build_wide_optimal(P, (int32_t)n, order.data(), opt, wide) : opt.max_depth + 1;
    if (depth > opt.max_depth) {
        wide.clear();
        depth = build_wide(P, (int32_t)n, order.data(), opt, false, wide);
    }
    if (depth > opt.max_depth) {
        wide.clear();
        depth = build_wide(P, (int32_t)n, order.data(), opt, true, wide);
        if (depth > opt.max_depth) throw std::runtime_error("BVH deeper than the traversal stack");
    }
    out.depth = depth;
    out.nodes.reserve(wide.size());
    out.tris.reserve(n);
    out.nodes.resize(1);
    Layout L{wide, order.data(), vertices, faces, out.nodes, out.tris};
    L.emit(0, 0);
    float m = 0;
    for (int a = 0; a < 3; ++a) m = std::max(m, std::max(std::fabs(scene.lo[a]), std::fabs(scene.hi[a])));
    out.root_max = m;
    out.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

}  // namespace vt
