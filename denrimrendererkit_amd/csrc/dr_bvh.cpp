// dr_bvh.cpp -- binned-SAH BVH build and conversion to the traversal layout.
//
// The reference builds a count-median BVH with 48-B nodes and a separate
// index array (Acceleration/BVHBuilder.swift:4-135, FlatBVH.swift:13-33).
// Closest-hit results with the lowest-index tie-break do not depend on the
// tree, so this backend builds its own: binned SAH (32 bins, 3 axes), leaves
// of at most `maxLeaf` triangles, collapsed into four-wide nodes (each node
// takes the cut of its binary subtree into at most four nodes of least SAH
// cost, chosen by a bottom-up dynamic program) in depth-first order, leaf
// triangles stored contiguously in leaf order.
#include "dr_bvh.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>

namespace dr {
namespace {

constexpr int kBins = 32;
constexpr float kNodeCost = 1.0f;
constexpr float kTriCost = 1.0f;

struct Box {
    float mn[3], mx[3];
    void reset() {
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::numeric_limits<float>::infinity();
            mx[a] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box &b) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], b.mn[a]);
            mx[a] = std::max(mx[a], b.mx[a]);
        }
    }
    void grow(const float *p) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], p[a]);
            mx[a] = std::max(mx[a], p[a]);
        }
    }
    float area() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        if (!(dx >= 0.f) || !(dy >= 0.f) || !(dz >= 0.f)) return 0.f;
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Prim {
    Box box;
    float c[3];
};

struct TNode {
    Box box;
    int32_t left = -1, right = -1;
    uint32_t first = 0, count = 0;
};

struct Builder {
    const std::vector<Prim> &prims;
    std::vector<uint32_t> &idx;
    std::vector<TNode> &nodes;
    std::atomic<uint32_t> next{0};
    uint32_t maxLeaf;
    std::atomic<uint32_t> maxDepth{0};

    Builder(const std::vector<Prim> &p, std::vector<uint32_t> &i, std::vector<TNode> &n, uint32_t ml)
        : prims(p), idx(i), nodes(n), maxLeaf(ml) {}

    uint32_t alloc() { return next.fetch_add(1); }

    void build(uint32_t nodeIndex, uint32_t lo, uint32_t hi, uint32_t depth) {
        uint32_t d = maxDepth.load();
        while (depth > d && !maxDepth.compare_exchange_weak(d, depth)) {
        }
        TNode &node = nodes[nodeIndex];
        Box bounds, cb;
        bounds.reset();
        cb.reset();
        for (uint32_t i = lo; i < hi; ++i) {
            const Prim &p = prims[idx[i]];
            bounds.grow(p.box);
            cb.grow(p.c);
        }
        node.box = bounds;
        const uint32_t count = hi - lo;
        auto makeLeaf = [&]() {
            node.first = lo;
            node.count = count;
            node.left = node.right = -1;
        };
        if (count <= 1 || depth + 2 >= DR_MAX_DEPTH) {
            makeLeaf();
            return;
        }
        // Binned SAH over centroid bounds.
        int bestAxis = -1;
        int bestSplit = -1;
        float bestCost = std::numeric_limits<float>::infinity();
        for (int a = 0; a < 3; ++a) {
            float ext = cb.mx[a] - cb.mn[a];
            if (!(ext > 0.f)) continue;
            float scale = float(kBins) / ext;
            Box bb[kBins];
            uint32_t bc[kBins] = {0};
            for (int b = 0; b < kBins; ++b) bb[b].reset();
            for (uint32_t i = lo; i < hi; ++i) {
                const Prim &p = prims[idx[i]];
                int b = std::min(kBins - 1, int((p.c[a] - cb.mn[a]) * scale));
                bc[b]++;
                bb[b].grow(p.box);
            }
            float rightArea[kBins];
            uint32_t rightCount[kBins];
            Box acc;
            acc.reset();
            uint32_t n = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bb[b]);
                n += bc[b];
                rightArea[b] = acc.area();
                rightCount[b] = n;
            }
            acc.reset();
            n = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                acc.grow(bb[b]);
                n += bc[b];
                if (n == 0 || rightCount[b + 1] == 0) continue;
                float cost = acc.area() * float(n) + rightArea[b + 1] * float(rightCount[b + 1]);
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = a;
                    bestSplit = b;
                }
            }
        }
        float parentArea = std::max(bounds.area(), 1e-30f);
        float leafCost = kTriCost * float(count);
        float splitCost = kNodeCost + kTriCost * bestCost / parentArea;
        if (count <= maxLeaf && (bestAxis < 0 || leafCost <= splitCost)) {
            makeLeaf();
            return;
        }
        uint32_t mid;
        if (bestAxis < 0) {
            mid = lo + count / 2;  // all centroids coincide: split by position in the list
        } else {
            float ext = cb.mx[bestAxis] - cb.mn[bestAxis];
            float scale = float(kBins) / ext;
            auto it = std::partition(idx.begin() + lo, idx.begin() + hi, [&](uint32_t k) {
                const Prim &p = prims[k];
                int b = std::min(kBins - 1, int((p.c[bestAxis] - cb.mn[bestAxis]) * scale));
                return b <= bestSplit;
            });
            mid = uint32_t(it - idx.begin());
            if (mid == lo || mid == hi) mid = lo + count / 2;
        }
        uint32_t l = alloc(), r = alloc();
        node.left = int32_t(l);
        node.right = int32_t(r);
        node.count = 0;
        if (count > 65536 && depth < 6) {
            auto f = std::async(std::launch::async, [=]() { build(l, lo, mid, depth + 1); });
            build(r, mid, hi, depth + 1);
            f.get();
        } else {
            build(l, lo, mid, depth + 1);
            build(r, mid, hi, depth + 1);
        }
    }
};

float widenDown(float v, float margin) {
    double w = double(v) - double(margin);
    float f = float(w);
    if (double(f) > w) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
    return std::nextafter(f, -std::numeric_limits<float>::infinity());
}
float widenUp(float v, float margin) {
    double w = double(v) + double(margin);
    float f = float(w);
    if (double(f) < w) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return std::nextafter(f, std::numeric_limits<float>::infinity());
}

void childBox(const Box &b, float out[6]) {
    float ext = std::max(b.mx[0] - b.mn[0], std::max(b.mx[1] - b.mn[1], b.mx[2] - b.mn[2]));
    for (int a = 0; a < 3; ++a) {
        float mag = std::max(std::fabs(b.mn[a]), std::fabs(b.mx[a]));
        float margin = 1e-6f * (mag + ext) + 1e-30f;
        out[a] = widenDown(b.mn[a], margin);
        out[3 + a] = widenUp(b.mx[a], margin);
    }
}

struct Flattener {
    const std::vector<TNode> &tn;
    const std::vector<uint32_t> &idx;
    const DrTriangle *tris;
    std::vector<DrNode4> &out;
    std::vector<float4> &pos;
    uint32_t leaves = 0;
    uint32_t depth = 0;
    double sah = 0.0;
    double rootArea = 1.0;

    // Triangles [start, start + count) of the index list, appended in leaf order.
    void emitLeafTris(const TNode &n, uint32_t start, uint32_t count, uint32_t &first) {
        first = uint32_t(pos.size() / 3);
        for (uint32_t k = 0; k < count; ++k) {
            uint32_t orig = idx[start + k];
            const DrTriangle &t = tris[orig];
            float4 v0 = t.v0, v1 = t.v1, v2 = t.v2;
            uint32_t bits = orig;
            float fb;
            std::memcpy(&fb, &bits, 4);
            pos.push_back(make_float4(v0.x, v0.y, v0.z, fb));
            pos.push_back(make_float4(v1.x, v1.y, v1.z, 0.f));
            pos.push_back(make_float4(v2.x, v2.y, v2.z, 0.f));
        }
        leaves++;
        sah += double(n.box.area()) / rootArea * kTriCost * count;
    }

    // A leaf of more than DR_LEAF_MAX triangles (a traversal stack entry
    // holds a leaf's count in 4 bits; only a large maxLeaf or the depth cap
    // makes such leaves): a node of up to four leaf chunks of <= DR_LEAF_MAX
    // triangles, nested further when needed; every chunk has the leaf's box.
    uint32_t emitBigLeaf(const TNode &leaf, uint32_t start, uint32_t count, uint32_t level) {
        depth = std::max(depth, level);
        const uint32_t me = uint32_t(out.size());
        out.push_back(DrNode4{});
        float bx[6];
        childBox(leaf.box, bx);
        uint32_t ref[4] = {0, 0, 0, 0}, cnt[4] = {0, 0, 0, 0}, valid = 0;
        const uint32_t per = count <= 4u * DR_LEAF_MAX ? DR_LEAF_MAX : (count + 3u) / 4u;
        for (int c = 0; c < 4 && count > 0; ++c) {
            const uint32_t n = std::min(per, count);
            valid |= 1u << c;
            if (n <= DR_LEAF_MAX) {
                emitLeafTris(leaf, start, n, ref[c]);
                cnt[c] = n;
            } else {
                ref[c] = emitBigLeaf(leaf, start, n, level + 1);
            }
            start += n;
            count -= n;
        }
        DrNode4 &o = out[me];
        o.lox = make_float4(bx[0], bx[0], bx[0], bx[0]);
        o.loy = make_float4(bx[1], bx[1], bx[1], bx[1]);
        o.loz = make_float4(bx[2], bx[2], bx[2], bx[2]);
        o.hix = make_float4(bx[3], bx[3], bx[3], bx[3]);
        o.hiy = make_float4(bx[4], bx[4], bx[4], bx[4]);
        o.hiz = make_float4(bx[5], bx[5], bx[5], bx[5]);
        o.child = make_uint4(ref[0], ref[1], ref[2], ref[3]);
        uint32_t leafBits = 0;
        for (int c = 0; c < 4; ++c) leafBits |= (cnt[c] ? 1u : 0u) << c;
        o.meta = make_uint4(cnt[0] | cnt[1] << 8 | cnt[2] << 16 | cnt[3] << 24, valid | leafBits << 4, 0u, 0u);
        return me;
    }

    // Cost-optimal collapse: cut[n][k - 1] is the least SAH cost of
    // covering binary subtree n with at most k children of one wide node
    // (n itself, or cuts of its two subtrees); split[n][k - 1] is the number
    // of those children given to the left subtree (0: n itself). A wide
    // node at binary node n costs A(n) kNodeCost plus its children's cut
    // for four slots; a leaf costs A(n) kTriCost per triangle.
    std::vector<std::array<float, 4>> cut;
    std::vector<std::array<uint8_t, 4>> split;

    void planCollapse(uint32_t used) {
        cut.assign(used, {});
        split.assign(used, {});
        // Children are allocated after their parent: a reverse sweep sees
        // both children of a node before the node.
        for (uint32_t i = used; i-- > 0;) {
            const TNode &n = tn[i];
            const float area = n.box.area();
            if (n.count > 0 || n.left < 0) {
                const float c = area * kTriCost * float(n.count);
                for (int k = 0; k < 4; ++k) cut[i][k] = c, split[i][k] = 0;
                continue;
            }
            const auto &L = cut[n.left], &R = cut[n.right];
            float best4 = std::numeric_limits<float>::infinity();
            for (int l = 1; l <= 3; ++l) best4 = std::min(best4, L[l - 1] + R[4 - l - 1]);
            const float self = area * kNodeCost + best4;
            cut[i][0] = self;
            split[i][0] = 0;
            for (int k = 2; k <= 4; ++k) {
                float b = self;
                uint8_t s = 0;
                for (int l = 1; l < k; ++l) {
                    const float c = L[l - 1] + R[k - l - 1];
                    if (c < b) b = c, s = uint8_t(l);
                }
                cut[i][k - 1] = b;
                split[i][k - 1] = s;
            }
        }
    }

    void expand(int32_t ni, int k, int32_t ch[4], int &n) const {
        const int l = split[ni][k - 1];
        if (l == 0) {
            ch[n++] = ni;
            return;
        }
        expand(tn[ni].left, l, ch, n);
        expand(tn[ni].right, k - l, ch, n);
    }

    // The two to four children of the wide node made at inner binary node
    // ni: the cheapest cut of its two subtrees into at most four nodes.
    int gather(int32_t ni, int32_t ch[4]) const {
        const TNode &t = tn[ni];
        const auto &L = cut[t.left], &R = cut[t.right];
        int bl = 1;
        float best = std::numeric_limits<float>::infinity();
        for (int l = 1; l <= 3; ++l)
            if (L[l - 1] + R[4 - l - 1] < best) best = L[l - 1] + R[4 - l - 1], bl = l;
        int n = 0;
        expand(t.left, bl, ch, n);
        expand(t.right, 4 - bl, ch, n);
        return n;
    }

    // Children are given as binary-tree nodes (n of them, 1...4).
    uint32_t emitNode(const int32_t *ch, int n, uint32_t level) {
        depth = std::max(depth, level);
        const uint32_t me = uint32_t(out.size());
        out.push_back(DrNode4{});
        float lo[3][4], hi[3][4];
        uint32_t ref[4] = {0, 0, 0, 0}, cnt[4] = {0, 0, 0, 0}, valid = 0;
        for (int c = 0; c < 4; ++c)
            for (int a = 0; a < 3; ++a) lo[a][c] = hi[a][c] = 0.f;
        for (int c = 0; c < n; ++c) {
            const TNode &cn = tn[ch[c]];
            float bx[6];
            childBox(cn.box, bx);
            for (int a = 0; a < 3; ++a) {
                lo[a][c] = bx[a];
                hi[a][c] = bx[3 + a];
            }
            valid |= 1u << c;
            if (cn.count > DR_LEAF_MAX) {
                ref[c] = emitBigLeaf(cn, cn.first, cn.count, level + 1);
            } else if (cn.count > 0) {
                emitLeafTris(cn, cn.first, cn.count, ref[c]);
                cnt[c] = cn.count;
            } else {
                int32_t g[4];
                const int gn = gather(ch[c], g);
                sah += double(cn.box.area()) / rootArea * kNodeCost;
                ref[c] = emitNode(g, gn, level + 1);
            }
        }
        DrNode4 &o = out[me];
        o.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
        o.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
        o.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
        o.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
        o.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
        o.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
        o.child = make_uint4(ref[0], ref[1], ref[2], ref[3]);
        uint32_t leafBits = 0;
        for (int c = 0; c < 4; ++c) leafBits |= (cnt[c] ? 1u : 0u) << c;
        o.meta = make_uint4(cnt[0] | cnt[1] << 8 | cnt[2] << 16 | cnt[3] << 24, valid | leafBits << 4, 0u, 0u);
        return me;
    }
};

}  // namespace

void buildBvh(const DrTriangle *tris, uint32_t count, uint32_t maxLeaf, std::vector<DrNode4> &nodes,
              std::vector<float4> &triPos, BvhBuildStats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    nodes.clear();
    triPos.clear();
    if (count == 0) {
        if (stats) *stats = BvhBuildStats{};
        return;
    }
    std::vector<Prim> prims(count);
    for (uint32_t i = 0; i < count; ++i) {
        const DrTriangle &t = tris[i];
        const float *v[3] = {&t.v0.x, &t.v1.x, &t.v2.x};
        Prim &p = prims[i];
        p.box.reset();
        for (int k = 0; k < 3; ++k) p.box.grow(v[k]);
        for (int a = 0; a < 3; ++a) p.c[a] = 0.5f * (p.box.mn[a] + p.box.mx[a]);
    }
    std::vector<uint32_t> idx(count);
    for (uint32_t i = 0; i < count; ++i) idx[i] = i;
    std::vector<TNode> tn(size_t(2) * count + 1);
    Builder b(prims, idx, tn, std::max(1u, maxLeaf));
    uint32_t root = b.alloc();
    b.build(root, 0, count, 0);

    triPos.reserve(size_t(count) * 3);
    nodes.reserve(count / 2 + 1);
    Flattener f{tn, idx, tris, nodes, triPos};
    f.planCollapse(b.next.load());
    f.rootArea = std::max(double(tn[root].box.area()), 1e-30);
    f.sah = kNodeCost;
    if (tn[root].count > 0) {
        // Whole scene in one leaf: a root with one leaf child.
        const int32_t only = int32_t(root);
        f.emitNode(&only, 1, 1);
    } else {
        int32_t g[4];
        const int gn = f.gather(int32_t(root), g);
        f.emitNode(g, gn, 1);
    }
    if (stats) {
        stats->nodeCount = uint32_t(nodes.size());
        stats->leafCount = f.leaves;
        stats->maxDepth = f.depth;
        stats->sahCost = f.sah;
        stats->buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

}  // namespace dr
