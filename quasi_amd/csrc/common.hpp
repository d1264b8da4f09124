// common.hpp -- POD types shared by the plugin's host code and its HIP kernels.
//
// Device data layout (all in HBM, 16-byte aligned):
//   tris   : 3 float4 per triangle in BVH leaf order
//              [0] = v0.xyz, w = input triangle index (bit pattern)
//              [1] = e1 = v1 - v0, w = 0
//              [2] = e2 = v2 - v0, w = 0
//   mats   : 2 float4 per triangle in BVH leaf order
//              [0] = albedo.rgb, w = max emission channel
//              [1] = emission.rgb, w = 0
//   nodes  : BVH2 "pair" nodes, 4 float4 (64 B, one cache line) per internal node;
//            each node stores the boxes of both children so one fetch serves
//            both slab tests:
//              [0] = (L.lo.x, L.hi.x, L.lo.y, L.hi.y)
//              [1] = (R.lo.x, R.hi.x, R.lo.y, R.hi.y)
//              [2] = (L.lo.z, L.hi.z, R.lo.z, R.hi.z)
//              [3] = (child L, child R, 0, 0) as int bits
//            child >= 0: internal node index; child < 0: leaf, ~child =
//            (first triangle << 2) | (count - 1), count 1..4.
//   lights : one LightEntry per emissive triangle
#pragma once

#include <cstdint>

namespace quasi_hip {

constexpr int kLeafMax = 4;
constexpr int kSahBins = 32;
constexpr uint32_t kMissId = 0xFFFFFFFFu;

struct LightEntry {
    uint32_t tri;    // BVH-order triangle index
    float cdf;       // inclusive cumulative selection probability (last = 1)
    float psel;      // selection probability
    float area;      // triangle area
};

struct CameraBasis {
    float origin[3];
    float llo[3];        // lower_left - origin
    float horizontal[3];
    float vertical[3];
};

struct Counters {
    unsigned long long rays_camera;
    unsigned long long rays_extension;
    unsigned long long rays_shadow;
    unsigned long long pad;
};

// Everything one sample pass needs (passed by value as a kernel argument).
struct PassParams {
    const void* tris;      // float4*
    const void* mats;      // float4*
    const void* nodes;     // float4*
    const LightEntry* lights;
    void* accum_beauty;    // float4 per pixel (sums)
    void* accum_albedo;
    void* accum_normal;
    void* accum_depth;
    uint32_t* ids;         // primary ids (may be null)
    Counters* counters;
    CameraBasis cam;
    uint32_t width, height;
    uint32_t sample_begin, sample_count;
    uint32_t max_depth, rr_start;
    uint32_t seed, rng_mode, integrator;
    uint32_t num_lights;
    uint32_t num_nodes;
    uint32_t record_ids, id_sample;
};

}  // namespace quasi_hip
