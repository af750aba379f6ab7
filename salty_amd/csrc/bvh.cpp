// bvh.cpp -- reference-shaped SAH tree on the host and its 8-wide compressed layout.
#include "bvh.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <atomic>
#include <deque>
#include <thread>
#include <utility>

namespace s3dg {
namespace {

const float kFMax = 3.402823466e+38F;

inline Aabb empty_box() {
    Aabb b;
    for (int a = 0; a < 3; ++a) { b.lo[a] = kFMax; b.hi[a] = -kFMax; }
    return b;
}
inline void grow(Aabb &b, const Aabb &o) {
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = o.lo[a] < b.lo[a] ? o.lo[a] : b.lo[a];
        b.hi[a] = o.hi[a] > b.hi[a] ? o.hi[a] : b.hi[a];
    }
}
// SurfaceArea(const BoundingBox&) (src/include/s3d_math.h:2904-2910); an empty
// box (lo = FLT_MAX, hi = -FLT_MAX) evaluates to +inf exactly as there.
inline float surface_area(const Aabb &b) {
    float X = fabsf(b.hi[0] - b.lo[0]);
    float Y = fabsf(b.hi[1] - b.lo[1]);
    float Z = fabsf(b.hi[2] - b.lo[2]);
    return 2.0f * ((X * Y) + (X * Z) + (Y * Z));
}

class SahBuilder {
public:
    SahBuilder(const std::vector<Aabb> &boxes, RefTree &t) : boxes_(boxes), t_(t) {
        const size_t n = boxes.size();
        for (int a = 0; a < 3; ++a) ctr_[a].resize(n);
        for (size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a)   // BoundingBox ctor: (mini + maxi) * 0.5f
                ctr_[a][i] = (boxes[i].lo[a] + boxes[i].hi[a]) * 0.5f;
        t_.order.resize(n);
        for (size_t i = 0; i < n; ++i) t_.order[i] = (int32_t)i;
    }

    int32_t bvh8(int32_t begin, int32_t count) {
        if (count <= 8) return node8_fallback(begin, count);
        int32_t mid, midL, midR;
        if (!split(begin, count, 8, 16, mid)) return node8_fallback(begin, count);
        if (!split(begin, mid, 8, 16, midL)) return node8_fallback(begin, count);
        if (!split(begin + mid, count - mid, 8, 16, midR)) return node8_fallback(begin, count);
        const int32_t s[4] = {0, midL, mid, mid + midR};
        const int32_t c[4] = {midL, mid - midL, midR, count - mid - midR};
        int32_t m[4];
        for (int i = 0; i < 4; ++i)
            if (!split(begin + s[i], c[i], 8, 16, m[i])) return node8_fallback(begin, count);
        int32_t cut[9];
        for (int i = 0; i < 4; ++i) { cut[2 * i] = s[i]; cut[2 * i + 1] = s[i] + m[i]; }
        cut[8] = count;
        int32_t kids[8];
        for (int i = 0; i < 8; ++i) kids[i] = bvh8(begin + cut[i], cut[i + 1] - cut[i]);
        return inner(8, kids, 8, begin, count);
    }

    int32_t bvh4(int32_t begin, int32_t count) {
        if (count <= 4) return leaf(begin, count);
        int32_t mid, midL, midR;
        if (!split(begin, count, 4, 16, mid)) return leaf(begin, count);
        if (!split(begin, mid, 4, 16, midL)) return leaf(begin, count);
        if (!split(begin + mid, count - mid, 4, 16, midR)) return leaf(begin, count);
        const int32_t cut[5] = {0, midL, mid, mid + midR, count};
        int32_t kids[4];
        for (int i = 0; i < 4; ++i) kids[i] = bvh4(begin + cut[i], cut[i + 1] - cut[i]);
        return inner(4, kids, 4, begin, count);
    }

    int32_t bvh2(int32_t begin, int32_t count) {
        if (count <= 4) return leaf(begin, count);
        int32_t mid;
        if (!split(begin, count, 0, 12, mid)) return leaf(begin, count);
        int32_t kids[2];
        kids[0] = bvh2(begin, mid);
        kids[1] = bvh2(begin + mid, count - mid);
        return inner(2, kids, 2, begin, count);
    }

private:
    const std::vector<Aabb> &boxes_;
    RefTree &t_;
    std::vector<float> ctr_[3];

    // CreateNode of s3d_bvh8.cpp:94-110: Leaf up to 8 prims, else BVH4::Create.
    int32_t node8_fallback(int32_t begin, int32_t count) {
        return count <= 8 ? leaf(begin, count) : bvh4(begin, count);
    }

    Aabb range_box(int32_t begin, int32_t count) const {
        Aabb b = empty_box();
        for (int32_t i = 0; i < count; ++i) grow(b, boxes_[t_.order[begin + i]]);
        return b;
    }

    int32_t leaf(int32_t begin, int32_t count) {
        RefNode n;
        n.kind = 0;
        n.first_child = -1;
        n.child_count = 0;
        n.begin = begin;
        n.count = count;
        n.box = range_box(begin, count);
        t_.nodes.push_back(n);
        return (int32_t)t_.nodes.size() - 1;
    }

    int32_t inner(int kind, const int32_t *kids, int nk, int32_t begin, int32_t count) {
        RefNode n;
        n.kind = kind;
        n.first_child = (int32_t)t_.children.size();
        n.child_count = nk;
        n.begin = begin;
        n.count = count;
        n.box = empty_box();
        for (int i = 0; i < nk; ++i) {
            t_.children.push_back(kids[i]);
            grow(n.box, t_.nodes[kids[i]].box);
        }
        t_.nodes.push_back(n);
        return (int32_t)t_.nodes.size() - 1;
    }

    // Binned SAH split over order[begin, begin+count): returns false where
    // SplitSAH (s3d_bvh8.cpp:254-342) / the BVH2 body (s3d_bvh2.cpp:185-262)
    // gives up; otherwise partitions in place and sets mid (relative).
    bool split(int32_t begin, int32_t count, int32_t min_count, int nb, int32_t &mid) {
        if (count <= min_count) return false;
        int32_t *ord = t_.order.data() + begin;
        Aabb bound = empty_box();
        float cmin[3] = {kFMax, kFMax, kFMax}, cmax[3] = {-kFMax, -kFMax, -kFMax};
        for (int32_t i = 0; i < count; ++i) {
            const int32_t p = ord[i];
            grow(bound, boxes_[p]);
            for (int a = 0; a < 3; ++a) {
                const float c = ctr_[a][p];
                cmin[a] = c < cmin[a] ? c : cmin[a];
                cmax[a] = c > cmax[a] ? c : cmax[a];
            }
        }
        // GetLongestAxis (s3d_bvh8.cpp:59-69): x only if strictly longest.
        const float ex = cmax[0] - cmin[0], ey = cmax[1] - cmin[1], ez = cmax[2] - cmin[2];
        const int axis = (ex > ey && ex > ez) ? 0 : (ey > ez ? 1 : 2);
        if (cmax[axis] == cmin[axis]) return false;
        const float base = cmin[axis];
        const float extent = cmax[axis] - cmin[axis];   // CalcOffset divisor
        const float fnb = (float)nb;
        const std::vector<float> &ca = ctr_[axis];

        int32_t bcount[16];
        Aabb bbox[16];
        for (int b = 0; b < nb; ++b) { bcount[b] = 0; bbox[b] = empty_box(); }
        for (int32_t i = 0; i < count; ++i) {
            const int32_t p = ord[i];
            int idx = (int)(fnb * ((ca[p] - base) / extent));
            if (idx == nb) idx = nb - 1;
            bcount[idx]++;
            grow(bbox[idx], boxes_[p]);
        }
        // Prefix / suffix unions give the same boxes as the reference's
        // per-candidate re-merge, because min and max are exact.
        Aabb left[16], right[16];
        int32_t lc[16], rc[16];
        Aabb acc = empty_box();
        int32_t cc = 0;
        for (int b = 0; b < nb; ++b) {
            if (bcount[b] > 0) grow(acc, bbox[b]);
            cc += bcount[b];
            left[b] = acc;
            lc[b] = cc;
        }
        acc = empty_box();
        cc = 0;
        for (int b = nb - 1; b >= 0; --b) {
            if (bcount[b] > 0) grow(acc, bbox[b]);
            cc += bcount[b];
            right[b] = acc;
            rc[b] = cc;
        }
        const float sa_bound = surface_area(bound);
        float best = 0.0f;
        int best_b = 0;
        for (int b = 0; b < nb - 1; ++b) {
            const float cost = 1 + ((float)lc[b] * surface_area(left[b]) +
                                    (float)rc[b + 1] * surface_area(right[b + 1])) / sa_bound;
            if (b == 0 || cost < best) { best = cost; best_b = b; }
        }
        if (!(best < (float)count)) return false;
        // Partition predicate keeps the untruncated float (s3d_bvh8.cpp:321-334).
        const float fsplit = (float)best_b;
        auto goes_left = [&](int32_t p) {
            float idx = fnb * ((ca[p] - base) / extent);
            if (idx == fnb) idx = fnb - 1;
            return idx <= fsplit;
        };
        // Two-ended swap partition (the element order both libstdc++ and MSVC
        // std::partition produce on bidirectional iterators).
        int32_t *first = ord, *last = ord + count;
        for (;;) {
            while (first != last && goes_left(*first)) ++first;
            if (first == last) break;
            --last;
            while (first != last && !goes_left(*last)) --last;
            if (first == last) break;
            std::swap(*first, *last);
            ++first;
        }
        mid = (int32_t)(first - ord);
        return true;
    }
};

// ---------------------------------------------------------------------------
// Traversal tree: binned SAH over the prim centres on all three axes, binary,
// leaves of at most `max_leaf` prims.  Only its shape differs from the
// reference tree; which hit wins a tie is decided by the reference tree's
// depth-first rank, so the traversal tree changes the speed, never a result.
// ---------------------------------------------------------------------------
class TraversalBuilder {
public:
    static constexpr int kBins = 32;

    TraversalBuilder(const std::vector<Aabb> &boxes, int max_leaf, float node_cost, RefTree &t)
        : boxes_(boxes), max_leaf_(std::max(1, max_leaf)), node_cost_(node_cost), t_(t) {
        const size_t n = boxes.size();
        ctr_.resize(3 * n);
        for (size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) ctr_[3 * i + a] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
        t_.order.resize(n);
        for (size_t i = 0; i < n; ++i) t_.order[i] = (int32_t)i;
    }

    // The top of the tree is split on the calling thread down to nodes of at
    // most kSerialPrims prims; the subtrees below those are then
    // built on worker threads into their own node lists and appended.  Each
    // subtree partitions only its own range of `order`, so the tree is the
    // same for any thread count.
    int32_t build() {
        Part top;
        new_node(top, 0, (int32_t)boxes_.size());
        std::vector<int32_t> pending;
        expand(top, 0, boxes_.size() > (size_t)kSerialPrims ? kSerialPrims : 0, pending);
        std::vector<Part> parts(pending.size());
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t k; (k = next.fetch_add(1)) < pending.size();) {
                Part &p = parts[k];
                p.nodes.push_back(top.nodes[pending[k]]);
                std::vector<int32_t> none;
                expand(p, 0, 0, none);
            }
        };
        std::vector<std::thread> pool;
        if (pending.size() > 1)
            for (unsigned i = 1; i < hw; ++i) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
        // append every subtree; its local node 0 replaces the pending node
        t_.nodes = std::move(top.nodes);
        t_.children = std::move(top.children);
        for (size_t k = 0; k < pending.size(); ++k) {
            const Part &p = parts[k];
            const int32_t off = (int32_t)t_.nodes.size() - 1;   // local i > 0 -> off + i
            const int32_t coff = (int32_t)t_.children.size();
            for (size_t i = 1; i < p.nodes.size(); ++i) {
                RefNode n = p.nodes[i];
                if (n.child_count) n.first_child += coff;
                t_.nodes.push_back(n);
            }
            for (int32_t c : p.children) t_.children.push_back(c == 0 ? pending[k] : off + c);
            RefNode root = p.nodes[0];
            if (root.child_count) root.first_child += coff;
            t_.nodes[pending[k]] = root;
        }
        return 0;
    }

private:
    const std::vector<Aabb> &boxes_;
    const int max_leaf_;
    const float node_cost_;   // cost of a binary node relative to one prim test
    RefTree &t_;
    std::vector<float> ctr_;

    static constexpr int32_t kSerialPrims = 1 << 15;

    struct Part {
        std::vector<RefNode> nodes;
        std::vector<int32_t> children;
    };

    int32_t new_node(Part &p, int32_t begin, int32_t count) const {
        RefNode n;
        n.kind = 0;
        n.first_child = -1;
        n.child_count = 0;
        n.begin = begin;
        n.count = count;
        n.box = empty_box();
        for (int32_t i = 0; i < count; ++i) grow(n.box, boxes_[t_.order[begin + i]]);
        p.nodes.push_back(n);
        return (int32_t)p.nodes.size() - 1;
    }

    // Splits node `from` of p and its descendants depth first.  With
    // defer > 0, descendants of at most `defer` prims are not split but
    // listed in `pending`.
    void expand(Part &p, int32_t from, int32_t defer, std::vector<int32_t> &pending) {
        std::vector<int32_t> todo{from};
        while (!todo.empty()) {
            const int32_t ni = todo.back();
            todo.pop_back();
            if (defer > 0 && ni != from && p.nodes[ni].count <= defer) {
                pending.push_back(ni);
                continue;
            }
            int32_t left_count;
            if (!split(p.nodes[ni], left_count)) continue;   // stays a leaf
            const int32_t b = p.nodes[ni].begin, c = p.nodes[ni].count;
            const int32_t l = new_node(p, b, left_count), r = new_node(p, b + left_count, c - left_count);
            RefNode &n = p.nodes[ni];
            n.kind = 2;
            n.first_child = (int32_t)p.children.size();
            n.child_count = 2;
            p.children.push_back(l);
            p.children.push_back(r);
            todo.push_back(r);
            todo.push_back(l);
        }
    }

    // Chooses a split of the node's prims and partitions them (left side
    // first); false when the node stays a leaf.
    bool split(const RefNode &n, int32_t &left_count) {
        const int32_t b = n.begin, c = n.count;
        if (c <= 1) return false;
        float cmin[3] = {kFMax, kFMax, kFMax}, cmax[3] = {-kFMax, -kFMax, -kFMax};
        for (int32_t i = 0; i < c; ++i) {
            const float *p = &ctr_[3 * (size_t)t_.order[b + i]];
            for (int a = 0; a < 3; ++a) {
                cmin[a] = std::min(cmin[a], p[a]);
                cmax[a] = std::max(cmax[a], p[a]);
            }
        }
        const float node_sa = surface_area(n.box);
        float best = kFMax;
        int best_axis = -1, best_bin = 0;
        float scale[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) {
            const float ext = cmax[a] - cmin[a];
            if (!(ext > 0.0f)) continue;
            scale[a] = (float)kBins / ext * (1.0f - 1e-6f);
            Aabb bb[kBins];
            int32_t bc[kBins];
            for (int k = 0; k < kBins; ++k) { bb[k] = empty_box(); bc[k] = 0; }
            for (int32_t i = 0; i < c; ++i) {
                const int32_t pi = t_.order[b + i];
                const int k = bin_of(ctr_[3 * (size_t)pi + a], cmin[a], scale[a]);
                ++bc[k];
                grow(bb[k], boxes_[pi]);
            }
            float right_cost[kBins];
            Aabb acc = empty_box();
            int32_t cnt = 0;
            for (int k = kBins - 1; k > 0; --k) {
                grow(acc, bb[k]);
                cnt += bc[k];
                right_cost[k] = cnt ? surface_area(acc) * (float)cnt : 0.0f;
            }
            acc = empty_box();
            cnt = 0;
            for (int k = 0; k < kBins - 1; ++k) {
                grow(acc, bb[k]);
                cnt += bc[k];
                if (cnt == 0 || cnt == c) continue;
                const float cost = surface_area(acc) * (float)cnt + right_cost[k + 1];
                if (cost < best) { best = cost; best_axis = a; best_bin = k + 1; }
            }
        }
        const float split_cost = node_sa > 0.0f && best < kFMax ? node_cost_ + best / node_sa : kFMax;
        if (c <= max_leaf_ && (float)c <= split_cost) return false;
        if (best_axis < 0) {
            // every centre coincides: halve the range as it stands
            left_count = c / 2;
            return true;
        }
        const int a = best_axis;
        const float lo = cmin[a], sc = scale[a];
        int32_t *first = t_.order.data() + b;
        int32_t *mid = std::partition(first, first + c, [&](int32_t pi) {
            return bin_of(ctr_[3 * (size_t)pi + a], lo, sc) < best_bin;
        });
        left_count = (int32_t)(mid - first);
        return true;
    }

    static int bin_of(float x, float lo, float scale) {
        const int k = (int)((x - lo) * scale);
        return k < 0 ? 0 : (k >= kBins ? kBins - 1 : k);
    }
};

// ---------------------------------------------------------------------------
// Re-layout into 8-wide compressed nodes.
// ---------------------------------------------------------------------------
enum { CH_LEAF = 0, CH_TREE = 1, CH_RANGE = 2 };

struct Child {
    int kind;           // CH_LEAF: a run of prims; CH_TREE: a tree node; CH_RANGE: a run too
                        // long for one slot, expanded into its own GPU node
    int32_t ref;        // CH_TREE: tree node index
    int32_t start;      // first prim (depth-first position)
    int32_t count;
    Aabb box;
};

class Compressor {
public:
    Compressor(const RefTree &t, const std::vector<Aabb> &prim_boxes, uint32_t tri_offset, bool top,
               float margin, Compressed &out, std::vector<int32_t> &leaf_order)
        : t_(t), pbox_(prim_boxes), tri_offset_(tri_offset), top_(top), margin_(margin), out_(out),
          leaf_order_(leaf_order), run_(top ? 1 : 255) {}

    uint32_t run() {
        const uint32_t root = (uint32_t)out_.nodes.size();
        out_.nodes.emplace_back();
        struct Item { std::vector<Child> kids; uint32_t gi; int depth; };
        std::deque<Item> work;
        Child rc;
        rc.kind = CH_TREE;
        rc.ref = t_.root;
        rc.start = t_.nodes[t_.root].begin;
        rc.count = t_.nodes[t_.root].count;
        rc.box = t_.nodes[t_.root].box;
        work.push_back({children_of(rc), root, 1});
        while (!work.empty()) {
            Item item = std::move(work.front());
            work.pop_front();
            if (item.depth > out_.max_depth) out_.max_depth = item.depth;
            int slot_of[8];
            assign_slots(item.kids, slot_of);
            // Internal children are laid out contiguously in slot order.
            int by_slot[8];
            for (int s = 0; s < 8; ++s) by_slot[s] = -1;
            for (size_t k = 0; k < item.kids.size(); ++k) by_slot[slot_of[k]] = (int)k;
            const uint32_t base = (uint32_t)out_.nodes.size();
            int ni = 0;
            for (int s = 0; s < 8; ++s)
                if (by_slot[s] >= 0 && item.kids[by_slot[s]].kind != CH_LEAF) ++ni;
            out_.nodes.resize(out_.nodes.size() + ni);
            fill_node(item.gi, item.kids, by_slot, base);
            uint32_t next = base;
            for (int s = 0; s < 8; ++s) {
                if (by_slot[s] < 0 || item.kids[by_slot[s]].kind == CH_LEAF) continue;
                work.push_back({children_of(item.kids[by_slot[s]]), next++, item.depth + 1});
            }
        }
        return out_.node_offset + root;
    }

private:
    const RefTree &t_;
    const std::vector<Aabb> &pbox_;
    uint32_t tri_offset_;
    bool top_;
    float margin_;
    Compressed &out_;
    std::vector<int32_t> &leaf_order_;
    int32_t run_;   // most prims one leaf slot may hold

    Aabb prims_box(int32_t start, int32_t count) const {
        Aabb b = empty_box();
        for (int32_t i = 0; i < count; ++i) grow(b, pbox_[t_.order[start + i]]);
        return b;
    }
    Child make(int kind, int32_t ref, int32_t start, int32_t count, const Aabb &box) const {
        Child c;
        c.kind = kind;
        c.ref = ref;
        c.start = start;
        c.count = count;
        c.box = box;
        return c;
    }
    Child child_of(int32_t node) const {
        const RefNode &n = t_.nodes[node];
        if (n.kind == 0 && n.count <= run_) return make(CH_LEAF, -1, n.begin, n.count, n.box);
        return make(CH_TREE, node, n.begin, n.count, n.box);
    }
    std::vector<Child> runs_of(int32_t start, int32_t count) const {
        std::vector<Child> kids;
        const int32_t pieces = (count + run_ - 1) / run_;
        if (pieces <= 8) {
            for (int32_t s = 0; s < count; s += run_) {
                const int32_t c = std::min(run_, count - s);
                kids.push_back(make(CH_LEAF, -1, start + s, c, prims_box(start + s, c)));
            }
        } else {
            const int32_t per = (pieces + 7) / 8 * run_;
            for (int32_t s = 0; s < count; s += per) {
                const int32_t c = std::min(per, count - s);
                kids.push_back(make(CH_RANGE, -1, start + s, c, prims_box(start + s, c)));
            }
        }
        return kids;
    }
    // Children of a GPU node: the tree node's children, with grandchildren
    // pulled up while they fit in 8 slots (largest surface area first).
    std::vector<Child> children_of(const Child &c) const {
        if (c.kind == CH_RANGE) return runs_of(c.start, c.count);
        const RefNode &n = t_.nodes[c.ref];
        if (n.kind == 0) return runs_of(n.begin, n.count);
        std::vector<Child> kids;
        for (int i = 0; i < n.child_count; ++i) kids.push_back(child_of(t_.children[n.first_child + i]));
        for (;;) {
            int pick = -1;
            float best = -1.0f;
            for (size_t k = 0; k < kids.size(); ++k) {
                if (kids[k].kind != CH_TREE) continue;
                const RefNode &kn = t_.nodes[kids[k].ref];
                if (kn.kind == 0) continue;
                if ((int)kids.size() - 1 + kn.child_count > 8) continue;
                const float sa = surface_area(kn.box);
                if (sa > best) { best = sa; pick = (int)k; }
            }
            if (pick < 0) break;
            const RefNode &kn = t_.nodes[kids[pick].ref];
            std::vector<Child> repl;
            for (int i = 0; i < kn.child_count; ++i) repl.push_back(child_of(t_.children[kn.first_child + i]));
            kids.erase(kids.begin() + pick);
            kids.insert(kids.begin() + pick, repl.begin(), repl.end());
        }
        return kids;
    }

    // Greedy octant slot assignment: slot s receives the child nearest along
    // the direction whose sign bits are s, so a ray of octant o visiting slots
    // in the order k ^ o meets children roughly front to back.
    static void assign_slots(const std::vector<Child> &kids, int *slot_of) {
        const int n = (int)kids.size();
        double pc[3] = {0, 0, 0};
        for (int k = 0; k < n; ++k)
            for (int a = 0; a < 3; ++a) pc[a] += 0.5 * ((double)kids[k].box.lo[a] + kids[k].box.hi[a]);
        for (int a = 0; a < 3; ++a) pc[a] /= (double)std::max(n, 1);
        double cost[8][8];
        for (int k = 0; k < n; ++k)
            for (int s = 0; s < 8; ++s) {
                double d = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double c = 0.5 * ((double)kids[k].box.lo[a] + kids[k].box.hi[a]) - pc[a];
                    d += ((s >> a) & 1) ? -c : c;
                }
                cost[k][s] = d;
            }
        bool used_k[8] = {false, false, false, false, false, false, false, false};
        bool used_s[8] = {false, false, false, false, false, false, false, false};
        for (int it = 0; it < n; ++it) {
            int bk = -1, bs = -1;
            double best = 0.0;
            for (int k = 0; k < n; ++k) {
                if (used_k[k]) continue;
                for (int s = 0; s < 8; ++s) {
                    if (used_s[s]) continue;
                    if (bk < 0 || cost[k][s] < best) { best = cost[k][s]; bk = k; bs = s; }
                }
            }
            used_k[bk] = true;
            used_s[bs] = true;
            slot_of[bk] = bs;
        }
    }

    void fill_node(uint32_t gi, const std::vector<Child> &kids, const int *by_slot, uint32_t base) {
        GNode &g = out_.nodes[gi];
        memset(&g, 0, sizeof(GNode));
        Aabb grown[8];
        Aabb all = empty_box();
        for (size_t k = 0; k < kids.size(); ++k) {
            grown[k] = kids[k].box;
            for (int a = 0; a < 3; ++a) {
                grown[k].lo[a] -= margin_;
                grown[k].hi[a] += margin_;
            }
            grow(all, grown[k]);
        }
        double step[3];
        for (int a = 0; a < 3; ++a) {
            g.origin[a] = all.lo[a];
            const double ext = (double)all.hi[a] - (double)g.origin[a];
            int e = -126;
            if (ext > 0.0) {
                e = (int)std::ceil(std::log2(ext / 255.0));
                if (e < -126) e = -126;
                while (std::ldexp(255.0, e) < ext) ++e;
            }
            g.exp[a] = (uint8_t)(e + 127);
            step[a] = std::ldexp(1.0, e);
        }
        g.child_base = (out_.node_offset + base) | (top_ ? kTopLevelBit : 0u);
        g.leaf_base = tri_offset_ + (uint32_t)leaf_order_.size();
        for (int s = 0; s < 8; ++s) {
            for (int a = 0; a < 3; ++a) { g.qlo[a][s] = 255; g.qhi[a][s] = 0; }
            const int k = by_slot[s];
            if (k < 0) continue;
            for (int a = 0; a < 3; ++a) {
                double lo = std::floor(((double)grown[k].lo[a] - (double)g.origin[a]) / step[a]);
                double hi = std::ceil(((double)grown[k].hi[a] - (double)g.origin[a]) / step[a]);
                g.qlo[a][s] = (uint8_t)std::max(0.0, std::min(255.0, lo));
                g.qhi[a][s] = (uint8_t)std::max(0.0, std::min(255.0, hi));
            }
            if (kids[k].kind == CH_LEAF) {
                g.cnt[s] = (uint8_t)kids[k].count;
                for (int32_t i = 0; i < kids[k].count; ++i) leaf_order_.push_back(kids[k].start + i);
            } else {
                g.imask |= (uint8_t)(1u << s);
            }
        }
    }

};

}  // namespace

void build_reference_tree(const std::vector<Aabb> &boxes, int kind, RefTree &out) {
    out.nodes.clear();
    out.children.clear();
    out.order.clear();
    out.root = -1;
    if (boxes.empty()) return;
    out.nodes.reserve(boxes.size() / 2 + 16);
    SahBuilder b(boxes, out);
    const int32_t n = (int32_t)boxes.size();
    if (kind == 2) out.root = b.bvh2(0, n);
    else if (kind == 4) out.root = b.bvh4(0, n);
    else out.root = b.bvh8(0, n);
}

void build_traversal_tree(const std::vector<Aabb> &boxes, int max_leaf, RefTree &out, float node_cost) {
    out.nodes.clear();
    out.children.clear();
    out.order.clear();
    out.root = -1;
    if (boxes.empty()) return;
    out.nodes.reserve(boxes.size() / 2 + 16);
    TraversalBuilder b(boxes, max_leaf, node_cost, out);
    out.root = b.build();
}

uint32_t compress_tree(const RefTree &tree, const std::vector<Aabb> &prim_boxes, uint32_t tri_offset,
                       bool top_level, float margin, Compressed &out, std::vector<int32_t> &leaf_order) {
    leaf_order.clear();
    Compressor c(tree, prim_boxes, tri_offset, top_level, margin, out, leaf_order);
    return c.run();
}

}  // namespace s3dg
