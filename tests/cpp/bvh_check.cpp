// Host check of the BVH topology builder (st_bvh.cpp) over random and
// degenerate box sets: every element sits in exactly one leaf slot, the two
// slot maps are inverse, leaves hold 1..max_leaf primitives and tile the
// slots, each inner node is reached once, levels count from the root and the
// depth stays within the traversal stack. Prints "ok <cases>" or the first
// violation. Built plain and under AddressSanitizer / UBSan by the tests.
#include <cstdio>
#include <random>
#include <vector>

#include "st_bvh.h"
#include "st_common.h"

using namespace st;

static int check(const std::vector<float>& b, int n, int max_leaf, const char* name) {
    const BvhTopology t = build_bvh(b.data(), n, max_leaf);
    auto bad = [&](const char* what) {
        std::printf("violation %s: %s (n %d)\n", name, what, n);
        return 1;
    };
    if ((int)t.slot_elem.size() != n || (int)t.elem_slot.size() != n) return bad("slot map sizes");
    for (int s = 0; s < n; ++s) {
        const int e = t.slot_elem[s];
        if (e < 0 || e >= n || t.elem_slot[e] != s) return bad("slot maps not inverse");
    }
    if (t.depth > kStackDepth) return bad("deeper than the traversal stack");
    std::vector<int> covered(n, 0), seen(t.n_nodes(), 0);
    if (n == 1) return 0;
    if (t.n_nodes() < 1 || (int)t.child.size() != 2 * t.n_nodes()) return bad("node arrays");
    std::vector<std::pair<int, int>> stack{{0, 0}};
    while (!stack.empty()) {
        const auto [node, level] = stack.back();
        stack.pop_back();
        if (node < 0 || node >= t.n_nodes()) return bad("child index out of range");
        if (seen[node]++) return bad("inner node reached twice");
        if (t.node_level[node] != level) return bad("node level");
        for (int k = 0; k < 2; ++k) {
            const int c = t.child[2 * node + k];
            if (c >= 0) {
                stack.push_back({c, level + 1});
            } else {
                const int s0 = leaf_start(c), cnt = leaf_count(c);
                if (cnt < 1 || cnt > max_leaf) return bad("leaf size");
                if (s0 < 0 || s0 + cnt > n) return bad("leaf range");
                for (int s = s0; s < s0 + cnt; ++s)
                    if (covered[s]++) return bad("slot in two leaves");
            }
        }
    }
    for (int s = 0; s < n; ++s)
        if (!covered[s]) return bad("slot in no leaf");
    for (int i = 0; i < t.n_nodes(); ++i)
        if (!seen[i]) return bad("unreachable inner node");
    return 0;
}

int main() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<float> U(-800.f, 800.f), S(0.01f, 12.f);
    int cases = 0, fails = 0;
    for (int n : {1, 2, 3, 4, 5, 7, 9, 33, 100, 1001, 10001}) {
        for (int mode = 0; mode < 4; ++mode) {
            std::vector<float> b(6 * (size_t)n);
            for (int i = 0; i < n; ++i) {
                float c[3] = {U(rng), U(rng), mode == 1 ? 0.f : U(rng) * 0.1f};
                if (mode == 2) c[0] = c[1] = c[2] = 5.f;        // all boxes identical
                if (mode == 3) c[0] = (float)i * 1e-3f;          // nearly coincident centroids
                for (int k = 0; k < 3; ++k) {
                    const float h = S(rng);
                    b[6 * i + k] = c[k] - h;
                    b[6 * i + 3 + k] = c[k] + h;
                }
            }
            for (int ml : {1, 2, 4}) {
                fails += check(b, n, ml, mode == 2 ? "identical" : mode == 3 ? "coincident" : "random");
                ++cases;
            }
        }
    }
    if (fails) return 1;
    std::printf("ok %d\n", cases);
    return 0;
}
