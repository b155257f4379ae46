// st_bvh.h - host builder of the binary BVH topology (binned SAH).
//
// Replaces the OptiX GAS build over custom-primitive AABBs
// (lib/geometry_manager.cpp:150-220). Only the topology is built on the host;
// node boxes are (re)fitted on the GPU from the FP64 element bounds, so
// re-aiming a field (lib/geometry_manager.cpp:223-265) never rebuilds.
#pragma once
#include <cstdint>
#include <vector>

namespace st {

struct BvhTopology {
    std::vector<int32_t> child;      // 2 per inner node (child pointer codes, st_common.h)
    std::vector<int32_t> slot_elem;  // primitive slot (leaf order) -> element index
    std::vector<int32_t> elem_slot;  // element index -> primitive slot
    std::vector<int32_t> node_level; // depth of every inner node (root = 0)
    int32_t depth = 0;               // deepest inner-node level + 1 (= stack bound)
    int32_t max_leaf = 0;
    int32_t n_nodes() const { return (int32_t)node_level.size(); }
};

// lo_hi: n boxes as (lo.x, lo.y, lo.z, hi.x, hi.y, hi.z). The tree depth is
// kept below kStackDepth (median splits take over deep down).
BvhTopology build_bvh(const float* lo_hi, int32_t n, int32_t max_leaf);

}  // namespace st
