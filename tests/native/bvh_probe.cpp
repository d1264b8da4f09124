// bvh_probe.cpp -- test-only C shim over the plugin's host BVH builder
// (quasi_amd/csrc/bvh_build.cpp), so its pair-node output can be checked on
// the CPU before any kernel walks it.  Built by tests/test_bvh_layout.py.
#include <cstring>

#include "bvh_build.hpp"

extern "C" int probe_build_cut(const float* p0, const float* p1, const float* p2, unsigned ntri, unsigned threads,
                               float* nodes_out, unsigned nodes_cap, unsigned* order_out, unsigned* info,
                               float* nodes4_out, unsigned cut_depth, double* sah_out) {
    quasi_hip::BvhBuildResult r = quasi_hip::build_bvh(p0, p1, p2, ntri, threads, cut_depth);
    if (sah_out) *sah_out = r.sah_cost;
    info[0] = r.num_nodes;
    info[1] = r.num_leaves;
    info[2] = r.depth;
    info[3] = unsigned(r.order.size());
    info[4] = r.num_nodes4;
    info[5] = r.depth4;
    if (r.num_nodes > nodes_cap) return -1;
    std::memcpy(nodes_out, r.nodes.data(), r.nodes.size() * sizeof(float));
    std::memcpy(order_out, r.order.data(), r.order.size() * sizeof(unsigned));
    if (nodes4_out) {
        if (r.num_nodes4 > nodes_cap) return -2;
        std::memcpy(nodes4_out, r.nodes4.data(), r.nodes4.size() * sizeof(float));
    }
    return 0;
}

extern "C" int probe_build(const float* p0, const float* p1, const float* p2, unsigned ntri, unsigned threads,
                           float* nodes_out, unsigned nodes_cap, unsigned* order_out, unsigned* info,
                           float* nodes4_out) {
    return probe_build_cut(p0, p1, p2, ntri, threads, nodes_out, nodes_cap, order_out, info, nodes4_out, 0u, nullptr);
}

extern "C" int probe_validate(const quasi_hip::BinaryNode* nodes, unsigned num_nodes, const unsigned* prims,
                              unsigned ntri) {
    return quasi_hip::validate_binary_tree(nodes, num_nodes, prims, ntri) ? 1 : 0;
}

