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
//   nodes4 : BVH4 nodes, 8 float4 (128 B) per node, boxes stored per component
//            so the four slab tests vectorise:
//              [0] = lo.x of children 0..3   [1] = hi.x of children 0..3
//              [2] = lo.y                    [3] = hi.y
//              [4] = lo.z                    [5] = hi.z
//              [6] = child references (int bits, same encoding as BVH2;
//                    kEmptyChild marks an unused slot, whose box is
//                    lo = +inf, hi = -inf so that every slab test rejects it)
//              [7] = unused
//   qnodes4: the same BVH4 nodes (same indices) in 64 B each, child boxes
//            quantised to 8 bits against the node's box (bvh_quant.hpp)
//   lights : one LightEntry per emissive triangle
#pragma once

#include <cstdint>

namespace quasi_hip {

constexpr int kLeafMax = 4;
constexpr int kSahBins = 32;
// SAH cost of a binary node's traversal relative to one triangle test.  Split
// planes are ranked by 1 + (A_l n_l + A_r n_r) / A (the constant does not
// change the ranking); the constant decides only, for a node of at most
// kLeafMax triangles, between a leaf (cost n) and its best split (cost
// kSahNodeCost + (A_l n_l + A_r n_r) / A).  Measured on the 1M soup at 1080p,
// ms of extend + shadow per 64-spp frame: 93.9 at 2, 88.9 at 1.5, 84.8 at 1,
// 81.4-81.6 at 0.5, 81.9 at 0, 81.8 with every splittable node split.
constexpr float kSahNodeCost = 0.5f;
constexpr uint32_t kMissId = 0xFFFFFFFFu;
constexpr int32_t kEmptyChild = 0x7FFFFFFF;
constexpr float kEmptyLo = __builtin_huge_valf();    // box of an unused BVH4 slot
constexpr float kEmptyHi = -__builtin_huge_valf();

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

constexpr uint32_t kMaxDepth = 16;

// Wavefront queues are split into kQueueSegments segments of seg_cap slots:
// segment j of every queue holds slots [j * seg_cap, j * seg_cap + count_j),
// and each segment has its own length and work cursor, kCounterStride words
// (one 128-B line) apart.  A single counter per queue would serialise every
// wave's atomic on one address (about 11 ns each on MI355X); 64 addresses
// take them 50 times faster.  Paths keep their segment from bounce to bounce:
// the continuation and the shadow ray of a ray of segment j go to segment j of
// the next queues, so no segment can overflow.
constexpr uint32_t kQueueSegments = 64;
constexpr uint32_t kCounterStride = 32;

// Word index of the counter of segment `seg` in counter group `group`.
constexpr uint32_t queue_counter(uint32_t group, uint32_t seg) {
    return (group * kQueueSegments + seg) * kCounterStride;
}

// The camera bounce deals a pass's path ids to the queue segments in chunks of
// `chunk` consecutive ids (a multiple of 64), round robin: chunk k to segment
// k % kQueueSegments, after the segment's earlier chunks.  A path keeps its
// segment on later bounces.  Single-scene passes deal chunks of
// kQueueChunkScene paths, so that every segment holds a sample of the whole
// image (contiguous ranges made each segment an image band of its own cost:
// on a sphere on a ground plane, 512^2 at 32 spp, the forward pass took 5.3
// ms against 3.2, the waves of cheap bands all ending up on the cursors of
// the dear ones).  Batches keep contiguous ranges, one chunk per segment
// (queue_contiguous_chunk), so that a segment's rays come from one scene or
// a few: dealt in chunks of 64 to 1024 paths the 64-scene batch ran 12-13%
// slower (1024-path chunks: 5% slower once shade kept the queues' order).
constexpr uint32_t kQueueChunkScene = 1024;
constexpr uint32_t queue_contiguous_chunk(uint32_t paths) {
    return ((paths + kQueueSegments - 1) / kQueueSegments + 63u) / 64u * 64u;
}
// Slots per segment: the chunks a segment can receive, whole.
constexpr uint32_t queue_seg_cap(uint32_t paths, uint32_t chunk) {
    return ((paths + chunk - 1) / chunk + kQueueSegments - 1) / kQueueSegments * chunk;
}
// Slots of all segments: at most paths + kQueueSegments * chunk - 1.
constexpr uint64_t queue_slots(uint32_t paths, uint32_t chunk) {
    return uint64_t(queue_seg_cap(paths, chunk)) * kQueueSegments;
}
// Slots allocated for passes of up to `paths` paths, either layout.
constexpr uint64_t queue_slots_max(uint32_t paths) {
    const uint64_t a = queue_slots(paths, kQueueChunkScene), b = queue_slots(paths, queue_contiguous_chunk(paths));
    return a > b ? a : b;
}
// Camera-bounce slot of path p, and the path of slot i (i below its segment's length).
constexpr uint32_t queue_slot_of_path(uint32_t p, uint32_t seg_cap, uint32_t chunk) {
    return (p / chunk) % kQueueSegments * seg_cap + (p / chunk) / kQueueSegments * chunk + p % chunk;
}
constexpr uint32_t queue_path_of_slot(uint32_t i, uint32_t seg_cap, uint32_t chunk) {
    return ((i % seg_cap) / chunk * kQueueSegments + i / seg_cap) * chunk + i % seg_cap % chunk;
}
// Camera-bounce length of segment seg: its chunks seg, seg + kQueueSegments,
// ..., the pass's last chunk possibly short.
constexpr uint32_t queue_segment_paths(uint32_t paths, uint32_t seg, uint32_t chunk) {
    const uint32_t chunks = (paths + chunk - 1) / chunk;
    if (seg >= chunks) return 0u;
    const uint32_t n = ((chunks - 1 - seg) / kQueueSegments + 1) * chunk;
    return (chunks - 1) % kQueueSegments == seg ? n - (chunks * chunk - paths) : n;
}

// Wavefront path state.  Per-path arrays are indexed by the path id
// p = pixel * samples_in_pass + sample_in_pass; per-slot arrays by the position
// in the segmented queue of the current bounce (double buffered by bounce
// parity).
struct WaveState {
    void* ray_o[2];        // float4 per slot: origin.xyz, w = path id (bits)
    void* ray_d[2];        // float4 per slot: direction.xyz
    void* hit;             // uint2 per slot: (t bits, BVH triangle index or -1)
    void* thr;             // float4 per path: throughput
    void* radiance;        // float4 per path: radiance
    uint32_t* rng;         // per path
    void* aov0;            // float4 per path: first-hit albedo.xyz, w = first-hit t
    void* aov1;            // float4 per path: first-hit normal*0.5+0.5, w = first-hit input id (bits)
    void* sh_o;            // float4 per shadow slot: origin.xyz, w = t max
    void* sh_d;            // float4 per shadow slot: direction.xyz, w = path id (bits)
    void* sh_c;            // float4 per shadow slot: unoccluded contribution
    uint32_t* counts;      // extension queue lengths: group b (0..kMaxDepth), per segment
    uint32_t* shadow_counts;  // shadow queue lengths: group b (0..kMaxDepth-1), per segment
    uint32_t* fetch;       // work cursors: group 2b (extend) and 2b+1 (shadow), per segment
    uint32_t paths;        // paths in this pass
    uint32_t seg_cap;      // slots per queue segment: queue_seg_cap(paths, qchunk)
    uint32_t qchunk;       // paths per chunk the camera bounce deals to a segment
    // float4 per slot, double buffered like the rays: throughput.xyz, w = RNG
    // state (bits) of the path in that slot, written by the shade that queues
    // the ray and read by the next bounce's shade in the same (coalesced)
    // order.  Null: the per-path arrays thr and rng carry them instead.
    void* tq[2];
};

// One scene of a wavefront pass.  A single-scene pass carries it inline
// (PassParams::scene0); a batched pass reads a device table of them.
// Scene i owns paths [path_begin, path_begin + width*height*samples) of the
// pass and pixels [pix_begin, pix_begin + width*height) of the framebuffer
// arrays; path p_local = (y*width + x)*samples + s (pixel-major: a pixel's
// samples are neighbours in the camera queue, see path_pixel_sample).
// A pass may cover a band of rows [row_begin, row_begin + rows) of the image
// only (rows = 0: every row): its paths are then p_local = ((y - row_begin)*
// width + x)*samples + s, while the camera and the RNG streams stay those of
// the whole width x height image.
struct SceneDesc {
    const void* tris;
    const void* mats;
    const void* nodes;
    const void* nodes4;
    const void* qnodes4;   // the BVH4 as quantised 64-B nodes (bvh_quant.hpp)
    const LightEntry* lights;
    uint32_t num_nodes, num_lights;
    CameraBasis cam;
    uint32_t width, height;
    uint32_t path_begin, pix_begin;
    uint32_t seed;
    uint32_t num_tris;
    uint32_t row_begin, rows;
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
    uint32_t num_tris;     // BVH-order triangle records (a batch: of all scenes, whose arrays are concatenated)
    // 4-wide BVH over the same triangles, and the traversal-stack overflow
    // area of the persistent kernels (gstack_entries words per thread,
    // lane-interleaved; sized by the host for the deepest possible stack)
    const void* nodes4;
    const void* qnodes4;   // the same BVH4 as quantised 64-B nodes (bvh_quant.hpp)
    uint32_t num_nodes4;
    uint32_t gstack_entries;
    int* gstack;
    // wavefront scene(s): scene0 when num_scenes == 0, else scenes[0..num_scenes)
    SceneDesc scene0;
    const SceneDesc* scenes;
    uint32_t num_scenes;
    uint32_t total_pixels;             // framebuffer pixels over all scenes
    const uint32_t* dyn_sample_begin;  // when set, the pass's first sample index is read here
    // Path records of a wavefront pass run for the adjoint (null otherwise; the
    // pass then adds nothing to the framebuffer).  For bounce b and path p of
    // the pass, at [b * paths + p]: rec_hit = the bounce's hit as the traversal
    // returned it (t bits, BVH triangle or -1), rec_vis = 1 when the vertex's
    // shadow ray was unoccluded.
    void* rec_hit;      // uint2 per (bounce, path)
    uint8_t* rec_vis;
};

// One adjoint launch: path replay of pass.sample_begin .. + pass.sample_count - 1
// from the records of a wavefront pass over the same samples.
struct AdjointParams {
    PassParams pass;              // scene, camera, integrator settings, replayed samples
    const void* dimg;             // float4 per pixel: dL/dI (rgb; w unused)
    float scale;                  // 1 / samples averaged into the image
    uint32_t replicas;            // FP32 gradient copies (blockIdx % replicas)
    const uint32_t* indices;      // 3 vertex indices per input triangle
    float* grad_albedo;           // replicas x 3 x num_tris
    float* grad_vertices;         // replicas x 3 x num_vertices
    uint32_t num_tris, num_vertices;
    uint32_t rec_paths;           // paths of the recorded pass (pass.rec_hit / rec_vis stride)
    uint32_t row_begin, rows;     // the band of image rows replayed (the recorded pass's band)
};

// Launch shape of the persistent traversal kernels (autotuned).
struct TraceConfig {
    uint32_t width;   // BVH width: 2 or 4
    uint32_t block;   // threads per block: 64, 128 or 256
    uint32_t chunk;   // rays fetched per wave per queue access: 64 or 256; 1 = per-lane dynamic fetch;
                      // kChunkFlat = no persistence (one thread per queue slot)
};
constexpr uint32_t kChunkFlat = 0xFFFFFFFFu;
constexpr uint32_t kChunkLdsTop = 2;  // per-lane fetch, top BVH4 levels staged in LDS (single scenes, width 4)
constexpr uint32_t kChunkQuant = 3;   // per-lane fetch over the quantised 64-B BVH4 nodes (bvh_quant.hpp),
                                      // top levels staged in LDS for single scenes (width 4)
constexpr int kFlatStack = 48;        // LDS stack entries per thread of the non-persistent traversal

constexpr int kShortStack = 16;       // LDS entries per lane in the persistent kernels

// Widening of the far slab distance so that rounding in (lo - o) * (1/d) never
// culls a box that the ray touches (relative error of each slab distance is a
// few ulps).
constexpr float kSlabWiden = 1.000001f;

// Paths in flight per automatically sized wavefront pass: about 30 GB of path
// state (220 B per path), large enough that the short queues of the last
// bounces are a small part of a pass.
constexpr uint32_t kAutoPassPaths = 1u << 27;
constexpr uint32_t kMaxThreadsPerCU = 2048;

}  // namespace quasi_hip
