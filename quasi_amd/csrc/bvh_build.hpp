// bvh_build.hpp -- host-side binned-SAH BVH2 build for the HIP traversal kernels.
//
// No reference counterpart: the reference intersects at most 32 quads by brute
// force (backends/metal/plugin.mm:199-212).  32 centroid bins, at most 4
// triangles per leaf, output in the 64-byte pair-node layout and the 128-byte
// 4-wide layout described in common.hpp.  Deterministic for a given input
// whatever the thread count.
#pragma once

#include <cstdint>
#include <vector>

namespace quasi_hip {

struct BvhBuildResult {
    std::vector<float> nodes;       // 16 floats (4 float4) per internal node
    std::vector<uint32_t> order;    // BVH-order position -> input triangle index
    uint32_t num_nodes = 0;         // internal (pair) nodes
    uint32_t num_leaves = 0;
    uint32_t depth = 0;             // deepest internal node, root = 1
    double sah_cost = 0;            // SAH cost of the tree (traversal cost 1, triangle cost 1)
    // 4-wide layout over the same leaves and triangle order (see common.hpp)
    std::vector<float> nodes4;      // 32 floats (8 float4) per node
    uint32_t num_nodes4 = 0;
    uint32_t depth4 = 0;
};

// p0/p1/p2: vertex positions of each triangle, 3 floats each, in input order.
BvhBuildResult build_bvh(const float* p0, const float* p1, const float* p2, uint32_t ntri,
                         unsigned threads);

}  // namespace quasi_hip
