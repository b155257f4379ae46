// st_bvh.cpp - binned-SAH BVH topology builder (host, O(n log n)).
#include "st_bvh.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>

#include "st_common.h"

namespace st {
namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.f;
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

constexpr int kBins = 16;
constexpr float kCostTraverse = 1.0f;
constexpr float kCostIntersect = 1.2f;

struct Builder {
    const std::vector<Box>& boxes;
    std::vector<float> cen;  // 3 per element
    std::vector<int32_t> idx;
    BvhTopology& out;
    int32_t max_leaf;
    int32_t median_depth;  // from this level on split at the median (bounds the depth)

    int32_t emit_leaf(int32_t start, int32_t count) {
        for (int32_t i = 0; i < count; ++i) {
            int32_t e = idx[start + i];
            out.slot_elem[start + i] = e;
            out.elem_slot[e] = start + i;
        }
        out.max_leaf = std::max(out.max_leaf, count);
        return leaf_code(start, count);
    }

    // returns a child pointer code for the range [start, start+count)
    int32_t build(int32_t start, int32_t count, int32_t level, bool force_inner) {
        if (!force_inner && count <= max_leaf && count <= 1) return emit_leaf(start, count);
        Box bb, cb;
        for (int32_t i = start; i < start + count; ++i) {
            bb.grow(boxes[idx[i]]);
            Box c;
            for (int k = 0; k < 3; ++k) c.lo[k] = c.hi[k] = cen[3 * idx[i] + k];
            cb.grow(c);
        }
        int32_t split = -1;
        if (count > 1 && level < median_depth) {
            float best = FLT_MAX;
            int best_axis = -1, best_bin = -1;
            float leaf_cost = kCostIntersect * count;
            for (int ax = 0; ax < 3; ++ax) {
                float ext = cb.hi[ax] - cb.lo[ax];
                if (!(ext > 0)) continue;
                Box bin_box[kBins];
                int bin_cnt[kBins] = {0};
                float scale = kBins / ext;
                for (int32_t i = start; i < start + count; ++i) {
                    int b = std::min(kBins - 1, (int)((cen[3 * idx[i] + ax] - cb.lo[ax]) * scale));
                    bin_cnt[b]++;
                    bin_box[b].grow(boxes[idx[i]]);
                }
                Box left[kBins];
                int lc[kBins];
                Box acc;
                int n = 0;
                for (int b = 0; b < kBins; ++b) {
                    acc.grow(bin_box[b]);
                    n += bin_cnt[b];
                    left[b] = acc;
                    lc[b] = n;
                }
                acc = Box();
                n = 0;
                float pa = bb.area() > 0 ? bb.area() : 1.f;
                for (int b = kBins - 1; b > 0; --b) {
                    acc.grow(bin_box[b]);
                    n += bin_cnt[b];
                    if (lc[b - 1] == 0 || n == 0) continue;
                    float c = kCostTraverse + kCostIntersect * (left[b - 1].area() * lc[b - 1] + acc.area() * n) / pa;
                    if (c < best) { best = c; best_axis = ax; best_bin = b; }
                }
            }
            if (best_axis >= 0 && (force_inner || count > max_leaf || best < leaf_cost)) {
                float ext = cb.hi[best_axis] - cb.lo[best_axis];
                float scale = kBins / ext;
                auto mid = std::partition(idx.begin() + start, idx.begin() + start + count, [&](int32_t e) {
                    int b = std::min(kBins - 1, (int)((cen[3 * e + best_axis] - cb.lo[best_axis]) * scale));
                    return b < best_bin;
                });
                split = (int32_t)(mid - idx.begin()) - start;
                if (split <= 0 || split >= count) split = -1;
            } else if (best_axis < 0 || best >= leaf_cost) {
                if (!force_inner && count <= max_leaf) return emit_leaf(start, count);
            }
        }
        if (split < 0) {
            if (!force_inner && count <= max_leaf) return emit_leaf(start, count);
            if (count == 1) {
                // a lone primitive under a forced inner node: both children point at it
                int32_t me = (int32_t)out.node_level.size();
                out.node_level.push_back(level);
                out.child.push_back(0);
                out.child.push_back(0);
                int32_t leaf = emit_leaf(start, 1);
                out.child[2 * me] = leaf;
                out.child[2 * me + 1] = leaf;
                return me;
            }
            // median along the widest centroid axis
            int ax = 0;
            for (int k = 1; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > cb.hi[ax] - cb.lo[ax]) ax = k;
            split = count / 2;
            std::nth_element(idx.begin() + start, idx.begin() + start + split, idx.begin() + start + count,
                             [&](int32_t a, int32_t b) {
                                 float ca = cen[3 * a + ax], cb2 = cen[3 * b + ax];
                                 return ca < cb2 || (ca == cb2 && a < b);
                             });
        }
        int32_t me = (int32_t)out.node_level.size();
        out.node_level.push_back(level);
        out.child.push_back(0);
        out.child.push_back(0);
        int32_t l = build(start, split, level + 1, false);
        int32_t r = build(start + split, count - split, level + 1, false);
        out.child[2 * me] = l;
        out.child[2 * me + 1] = r;
        return me;
    }
};

}  // namespace

BvhTopology build_bvh(const float* lo_hi, int32_t n, int32_t max_leaf) {
    if (n <= 0) throw std::invalid_argument("empty BVH");
    if (max_leaf < 1 || max_leaf > 8) throw std::invalid_argument("max_leaf must be in [1, 8]");
    std::vector<Box> boxes((size_t)n);
    for (int32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            boxes[i].lo[k] = lo_hi[6 * i + k];
            boxes[i].hi[k] = lo_hi[6 * i + 3 + k];
        }
    BvhTopology out;
    out.slot_elem.assign((size_t)n, -1);
    out.elem_slot.assign((size_t)n, -1);
    // SAH above median_depth, medians below: depth <= median_depth + ceil(log2 n) <= 30
    int32_t lg = 0;
    while ((1ll << lg) < (long long)n) ++lg;
    int32_t median_depth = std::max(1, std::min(24, 30 - lg));
    Builder b{boxes, std::vector<float>(3 * (size_t)n), std::vector<int32_t>((size_t)n), out, max_leaf, median_depth};
    for (int32_t i = 0; i < n; ++i) {
        b.idx[i] = i;
        for (int k = 0; k < 3; ++k) b.cen[3 * i + k] = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
    }
    int32_t root = b.build(0, n, 0, true);
    if (root != 0) throw std::logic_error("BVH root must be node 0");
    int32_t deepest = 0;
    for (int32_t lv : out.node_level) deepest = std::max(deepest, lv);
    out.depth = deepest + 1;
    if (out.depth >= kStackDepth) throw std::runtime_error("BVH deeper than the traversal stack");
    return out;
}

}  // namespace st
