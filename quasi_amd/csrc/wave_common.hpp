// wave_common.hpp -- device helpers shared by the wavefront kernels
// (kernels_wave.hip) and the single-launch path kernel (kernels_paths.hip).
#pragma once

#include "bvh_quant.hpp"
#include "device.hpp"

namespace quasi_hip {
namespace {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

template <bool BATCH>
__device__ __forceinline__ const SceneDesc& scene_at(const PassParams& P, uint32_t sid) {
    if (BATCH) return P.scenes[sid];
    return P.scene0;
}

// Scene owning path id p (BATCH): the last scene whose path_begin <= p.
__device__ __forceinline__ uint32_t scene_of_path(const PassParams& P, uint32_t p) {
    uint32_t lo = 0, hi = P.num_scenes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (P.scenes[mid].path_begin <= p) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t scene_of_pixel(const PassParams& P, uint32_t pix) {
    uint32_t lo = 0, hi = P.num_scenes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (P.scenes[mid].pix_begin <= pix) lo = mid; else hi = mid;
    }
    return lo;
}

// Pixel (row-major within the scene's band) and sample of the pass's path
// p_local = pixel * samples + sample.  A pixel's samples are neighbours in the
// camera queue, so a wave traces one pixel's camera rays together: on the 1M
// soup at 1080p, 64 spp, extend took 3.5 ms less per frame than with the
// sample-major numbering (sample * pixels + pixel), whose queue order puts 64
// pixels of a row in a wave.
__device__ __forceinline__ void path_pixel_sample(uint32_t p_local, uint32_t samples, uint32_t& pix,
                                                  uint32_t& s) {
    pix = p_local / samples;
    s = p_local - pix * samples;
}

__device__ __forceinline__ uint32_t pass_sample_begin(const PassParams& P) {
    return P.dyn_sample_begin ? *P.dyn_sample_begin : P.sample_begin;
}

__device__ __forceinline__ uint32_t wave_global() { return (blockIdx.x * blockDim.x + threadIdx.x) >> 6; }

// Adds the length of a segmented queue to a ray counter (first wave of the
// grid; one segment per lane).
static_assert(kQueueSegments == 64, "one segment per lane of a wave");
__device__ __forceinline__ void count_rays(const uint32_t* lengths, unsigned long long* counter) {
    if (blockIdx.x == 0 && threadIdx.x < 64u) {
        const uint32_t total = wave_sum(lengths[threadIdx.x * kCounterStride]);
        if (threadIdx.x == 0u && total) atomicAdd(counter, (unsigned long long)total);
    }
}

template <int WIDTH, bool BATCH>
__device__ __forceinline__ const float4* scene_nodes(const PassParams& P, uint32_t sid) {
    const SceneDesc& sc = scene_at<BATCH>(P, sid);
    return static_cast<const float4*>(WIDTH == 4 ? sc.nodes4 : sc.nodes);
}

__device__ __forceinline__ void add_shadow(float4* rad, const float4* sc_a, uint32_t k, uint32_t p) {
    const float4 c = sc_a[k];
    float4 l = rad[p];
    l.x += c.x; l.y += c.y; l.z += c.z;
    rad[p] = l;
}

constexpr uint32_t kRefillIdle = 32;   // measured best of 8..48 on the 1M soup (28-32 alike)

// The node loop of the while-while traversal hands over to the leaf tests
// once no more than this many lanes of the wave are still looking for a leaf
// and some lane holds one (0 would wait for every lane): the lanes holding
// leaves then test them instead of idling while a few lanes descend.  On the
// 1M soup at 1080p (quantised nodes), ms of extend + shadow per 64-spp frame:
// 106.5 at 0, 89.6 at 4, 86.3 at 8, 84.3-85.5 at 16, 86.7 at 24, 91.5 at 32,
// 103.0 at 40, 130.3 at 64 (leaves as soon as any lane has one).
constexpr uint32_t kLeafPhaseSearching = 16;
__device__ __forceinline__ bool leave_node_loop(bool searching, bool holds_leaf) {
    const uint32_t n = uint32_t(__popcll(__ballot(searching)));
    return n == 0u || (n <= kLeafPhaseSearching && __ballot(holds_leaf) != 0ull);
}
constexpr int kDone = int(0x80000000);   // no leaf reference has this value (fewer than 2^29 triangles)
constexpr float kMissKey = 3.0e38f;      // sort key of a child the ray misses

// The children of a BVH4 node in order of entry distance (keys; kMissKey for
// a miss, references c): returns the nearest as the next node and pushes the
// other hit ones, farthest first; with no hit, the top of the stack (kDone
// when empty).
template <int BLOCK, class Stack>
__device__ __forceinline__ int visit_children4(float k0, float k1, float k2, float k3, int c0, int c1, int c2, int c3,
                                               Stack& st) {
    cswap(k0, c0, k1, c1);
    cswap(k2, c2, k3, c3);
    cswap(k0, c0, k2, c2);
    cswap(k1, c1, k3, c3);
    cswap(k1, c1, k2, c2);
    if (k0 == kMissKey) return st.empty() ? kDone : st.pop();
    // push the hit children after the nearest, farthest first
    const int np = int(k1 != kMissKey) + int(k2 != kMissKey) + int(k3 != kMissKey);
    if (st.sp + 3 <= kShortStack) {
        // three unconditional LDS stores; slots above sp + np are overwritten
        // before they are read
        lds_int_ptr top = st.lds + st.sp * BLOCK;
        top[0] = np == 3 ? c3 : (np == 2 ? c2 : c1);
        top[BLOCK] = np == 3 ? c2 : c1;
        top[2 * BLOCK] = c1;
        st.sp += np;
    } else {
        if (k3 != kMissKey) st.push(c3);
        if (k2 != kMissKey) st.push(c2);
        if (k1 != kMissKey) st.push(c1);
    }
    return c0;
}

// Node step over a quantised BVH4 node (bvh_quant.hpp; q: its four float4):
// the next node, with the other hit children pushed.  Unused slots are
// rejected by their reference (slot 1 is unused only in the root of a
// one-triangle scene).
template <int BLOCK, class Stack>
__device__ __forceinline__ int quant_node_step(const float4& q0, const float4& q1, const float4& q2, const float4& q3,
                                               const Ray& r, const f3& inv, float tmin, float tmax, Stack& st) {
    float tn[4], tf[4];
    slab_quant4(q0, q1, q2, r.o.x, r.o.y, r.o.z, inv.x, inv.y, inv.z, tmin, tmax, tn, tf);
    const int c0 = __float_as_int(q3.x), c1 = __float_as_int(q3.y), c2 = __float_as_int(q3.z),
              c3 = __float_as_int(q3.w);
    return visit_children4<BLOCK>(tn[0] <= tf[0] ? tn[0] : kMissKey,
                                  tn[1] <= tf[1] && c1 != kEmptyChild ? tn[1] : kMissKey,
                                  tn[2] <= tf[2] && c2 != kEmptyChild ? tn[2] : kMissKey,
                                  tn[3] <= tf[3] && c3 != kEmptyChild ? tn[3] : kMissKey, c0, c1, c2, c3, st);
}

}  // namespace
}  // namespace quasi_hip
