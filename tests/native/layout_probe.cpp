// layout_probe.cpp -- test-only C shim over the plugin's GPU scene build
// (quasi_amd/csrc/bvh_gpu.hip, bvh_layout.hip): lays out one binary tree both
// on the device (layout_tree_gpu) and on the host (convert_binary_tree), so
// tests can compare the two byte for byte.  Built by quasi_amd/build.py into
// tests/native/liblayoutprobe.so from the plugin's own objects.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "bvh_build.hpp"
#include "bvh_gpu.hpp"

#define PROBE_API extern "C" __attribute__((visibility("default")))

namespace {

using quasi_hip::BinaryNode;

void say(char* why, unsigned len, const char* msg) {
    if (why && len) std::snprintf(why, len, "%s", msg);
}

// info: 0 num_nodes, 1 num_leaves, 2 depth, 3 num_nodes4, 4 depth4 (GPU), then
// the same five from the host converter at 5..9; probe_layout_built adds the
// tree's node slots at 10 and its unused (reserved, unreachable) slots at 11.
int layout_both(const BinaryNode* tree, unsigned used, unsigned unused, const unsigned* prims, unsigned n,
                float* gpu_nodes,
                float* host_nodes, unsigned nodes_cap, float* gpu_nodes4, float* host_nodes4, unsigned nodes4_cap,
                unsigned* gpu_order, unsigned* host_order, unsigned* info, double* sah, char* why, unsigned why_len) {
    quasi_hip::BvhBuildResult h = quasi_hip::convert_binary_tree(tree, used, prims, n);
    info[5] = h.num_nodes;
    info[6] = h.num_leaves;
    info[7] = h.depth;
    info[8] = h.num_nodes4;
    info[9] = h.depth4;
    sah[1] = h.sah_cost;
    if (h.num_nodes > nodes_cap || h.num_nodes4 > nodes4_cap) { say(why, why_len, "capacity"); return -1; }
    std::memcpy(host_nodes, h.nodes.data(), h.nodes.size() * 4);
    std::memcpy(host_nodes4, h.nodes4.data(), h.nodes4.size() * 4);
    std::memcpy(host_order, h.order.data(), h.order.size() * 4);

    BinaryNode* dtree = nullptr;
    unsigned* dprims = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&dtree), used * sizeof(BinaryNode)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&dprims), n * sizeof(unsigned)) != hipSuccess) {
        say(why, why_len, "allocation");
        return -2;
    }
    int rc = 0;
    quasi_hip::DeviceLayout d;
    const char* reason = "";
    if (hipMemcpy(dtree, tree, used * sizeof(BinaryNode), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dprims, prims, n * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
        say(why, why_len, "upload");
        rc = -3;
    } else if (!quasi_hip::layout_tree_gpu(dtree, used, unused, dprims, n, nullptr, d, &reason)) {
        say(why, why_len, reason);
        rc = 1;   // rejected by the layout's checks
    } else {
        info[0] = d.num_nodes;
        info[1] = d.num_leaves;
        info[2] = d.depth;
        info[3] = d.num_nodes4;
        info[4] = d.depth4;
        sah[0] = d.sah_cost;
        if (d.num_nodes > nodes_cap || d.num_nodes4 > nodes4_cap) {
            say(why, why_len, "capacity");
            rc = -1;
        } else if (hipMemcpy(gpu_nodes, d.nodes, size_t(d.num_nodes) * 64, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(gpu_nodes4, d.nodes4, size_t(d.num_nodes4) * 128, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(gpu_order, d.order, size_t(n) * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            say(why, why_len, "download");
            rc = -4;
        }
    }
    d.release();
    (void)hipFree(dtree);
    (void)hipFree(dprims);
    return rc;
}

}  // namespace

// A caller-made binary tree (`used` BinaryNode records, 40 bytes each).
PROBE_API int probe_layout_tree(const void* tree, unsigned used, const unsigned* prims, unsigned n,
                                 float* gpu_nodes, float* host_nodes, unsigned nodes_cap, float* gpu_nodes4,
                                 float* host_nodes4, unsigned nodes4_cap, unsigned* gpu_order, unsigned* host_order,
                                 unsigned* info, double* sah, char* why, unsigned why_len) {
    return layout_both(static_cast<const BinaryNode*>(tree), used, 0u, prims, n, gpu_nodes, host_nodes, nodes_cap,
                       gpu_nodes4, host_nodes4, nodes4_cap, gpu_order, host_order, info, sah, why, why_len);
}

// Only the device layout of a caller-made tree: 0 when accepted, 1 when the
// layout's checks reject it (why says which), negative on a HIP failure.
PROBE_API int probe_layout_check(const void* tree, unsigned used, const unsigned* prims, unsigned n, char* why,
                                  unsigned why_len) {
    BinaryNode* dtree = nullptr;
    unsigned* dprims = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&dtree), used * sizeof(BinaryNode)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&dprims), n * sizeof(unsigned)) != hipSuccess)
        return -2;
    int rc = 0;
    quasi_hip::DeviceLayout d;
    const char* reason = "";
    if (hipMemcpy(dtree, tree, used * sizeof(BinaryNode), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dprims, prims, n * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess)
        rc = -3;
    else if (!quasi_hip::layout_tree_gpu(dtree, used, 0u, dprims, n, nullptr, d, &reason)) {
        say(why, why_len, reason);
        rc = 1;
    }
    d.release();
    (void)hipFree(dtree);
    (void)hipFree(dprims);
    return rc;
}

// The GPU builder's tree for triangles p0/p1/p2 (3 floats each), laid out both ways.
PROBE_API int probe_layout_built(const float* p0, const float* p1, const float* p2, unsigned ntri,
                                  float* gpu_nodes, float* host_nodes, unsigned nodes_cap, float* gpu_nodes4,
                                  float* host_nodes4, unsigned nodes4_cap, unsigned* gpu_order, unsigned* host_order,
                                  unsigned* info, double* sah, char* why, unsigned why_len) {
    std::vector<BinaryNode> tree;
    std::vector<unsigned> prims;
    const char* reason = "";
    unsigned unused = 0;
    if (!quasi_hip::build_tree_gpu_host(p0, p1, p2, ntri, nullptr, tree, prims, &unused, &reason)) {
        say(why, why_len, reason);
        return -5;
    }
    info[10] = unsigned(tree.size());
    info[11] = unused;
    return layout_both(tree.data(), unsigned(tree.size()), unused, prims.data(), ntri, gpu_nodes, host_nodes,
                       nodes_cap, gpu_nodes4, host_nodes4, nodes4_cap, gpu_order, host_order, info, sah, why,
                       why_len);
}
