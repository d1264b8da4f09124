// bvh_gpu.hpp -- the scene build on the GPU: binned-SAH binary tree
// (bvh_gpu.hip), its pair-node and 4-wide traversal layouts (bvh_layout.hip)
// and the triangle / material arrays in BVH order, without the scene passing
// through host memory after the mesh upload.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bvh_build.hpp"

namespace quasi_hip {

// Phases of a GPU scene build: device time of the kernels (events) and the
// host wall clock of each step.
struct BvhGpuTimes {
    double device_ms = 0;          // tree build kernels
    double layout_device_ms = 0;   // layout kernels (pair nodes, 4-wide nodes, triangles)
    double alloc_ms = 0, upload_ms = 0, split_ms = 0, build_ms = 0, layout_ms = 0;
};

// Traversal layouts of a binary tree on the device (hipMalloc'd; the caller
// owns them and frees them with hipFree).
struct DeviceLayout {
    float* nodes = nullptr;      // 16 floats per pair node
    float* nodes4 = nullptr;     // 32 floats per 4-wide node
    uint32_t* order = nullptr;   // BVH-order position -> tree triangle (prims entry)
    uint32_t num_nodes = 0, num_leaves = 0, depth = 0, num_nodes4 = 0, depth4 = 0;
    double sah_cost = 0;
    float device_ms = 0;
    void release();
};

// The layouts that convert_binary_tree writes for the same tree, byte for
// byte (the SAH cost agrees to rounding), computed on the device.  tree:
// `used` node slots, root 0 internal, every child at a larger index than its
// parent, exactly `unused` slots unreachable from the root (space a builder
// reserved and did not need); prims: n entries.  Every index is checked;
// false (with a reason) when the tree is malformed or an allocation fails.
bool layout_tree_gpu(const BinaryNode* tree, uint32_t used, uint32_t unused, const uint32_t* prims, uint32_t n,
                     hipStream_t s, DeviceLayout& out, const char** why);

// One scene built on the device: the arrays of common.hpp (triangles,
// materials, pair nodes, 4-wide nodes), hipMalloc'd and owned by the caller.
struct DeviceScene {
    void* tris = nullptr;     // 12 floats per BVH-order position (a triangle reference)
    void* mats = nullptr;     // 8 floats per BVH-order position
    float* nodes = nullptr;
    float* nodes4 = nullptr;
    uint32_t num_prims = 0;   // BVH-order positions: references of the split triangles (bvh_split.hpp)
    uint32_t num_nodes = 0, num_leaves = 0, depth = 0, num_nodes4 = 0, depth4 = 0;
    double sah_cost = 0;
    void release();
};

// Builds a mesh on the current device and stream: vertex gather, the split
// of large triangles into references (bvh_split.hpp), the binned-SAH tree over
// the references, layouts and the BVH-order triangle / material arrays (one
// copy of a triangle's record per reference).  emitters
// (n_emit input triangle ids, ascending) lists every triangle whose emission
// has a nonzero bit: only those emissions are uploaded when they are few (all
// others are zero).  query (nquery input triangle ids, each < nt) receives
// in query_pos the BVH-order position of each (of its first reference;
// emitters, the triangles queried for the light table, are never split).  False (with a reason) on any
// failure, including a vertex index out of range; the caller then builds on
// the host.
bool build_scene_gpu(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* albedo,
                     const float* emission, const uint32_t* emitters, uint32_t n_emit, const uint32_t* query,
                     uint32_t nquery, uint32_t* query_pos, hipStream_t s, DeviceScene& out, BvhGpuTimes* times,
                     const char** why);

// The binary tree that build_scene_gpu builds for triangles p0/p1/p2 (host
// arrays, 3 floats per triangle), downloaded: tree node slots (with `unused`
// unreachable ones) and prims, so that tests can lay it out both ways.  False
// on failure.
bool build_tree_gpu_host(const float* p0, const float* p1, const float* p2, uint32_t ntri, hipStream_t s,
                         std::vector<BinaryNode>& tree, std::vector<uint32_t>& prims, uint32_t* unused,
                         const char** why);

}  // namespace quasi_hip
