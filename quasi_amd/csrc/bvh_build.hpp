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
// cut_depth as for convert_binary_tree.
BvhBuildResult build_bvh(const float* p0, const float* p1, const float* p2, uint32_t ntri,
                         unsigned threads, uint32_t cut_depth = 0);

// A node of a binary build tree: internal (count == 0: children left, right)
// or leaf (count 1..kLeafMax triangles prims[first .. first + count)).  Node 0
// is the root.
struct BinaryNode {
    float lo[3], hi[3];
    uint32_t left, right;
    uint32_t first, count;
};

// The pair-node and 4-wide layouts of a binary build tree (shared by the host
// build and the GPU build); prims maps the tree's triangle positions to input
// triangles.  The output is independent of how the tree's node slots were
// numbered and of `threads` (0 = the machine's, at most 16).  cut_depth: the
// depth below which subtrees are converted in parallel (0 = 12 for trees of
// 2^18 nodes or more, else one sequential walk; ~0u = always sequential); the
// layouts are the same for every cut, and the SAH cost agrees to rounding.
BvhBuildResult convert_binary_tree(const BinaryNode* nodes, uint32_t num_nodes, const uint32_t* prims,
                                   uint32_t ntri, unsigned threads = 0, uint32_t cut_depth = 0);

// True when the tree rooted at node 0 reaches every one of its num_nodes
// nodes exactly once, internal children are in range, leaves hold 1..kLeafMax
// triangles, the leaves cover tree positions 0..ntri-1 exactly once and prims
// is a permutation of 0..ntri-1.  Requires (and checks) that every child has a
// larger index than its parent, which both builders guarantee (slots are
// taken from a counter when a node is split), so the checks run in parallel.
// The GPU scene build applies the same checks on the device while it lays
// the tree out (bvh_layout.hip); this host version states them for the CPU
// tests of malformed trees.
bool validate_binary_tree(const BinaryNode* nodes, uint32_t num_nodes, const uint32_t* prims, uint32_t ntri);

}  // namespace quasi_hip
