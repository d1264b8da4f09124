// kernels_wave.hip -- the wavefront engine, for one scene or a batch of scenes.
//
// One sample pass over N paths runs as
//   raygen -> (extend -> shade -> shadow) x max_depth -> accumulate
// with every queue length kept on the device, so the host enqueues a fixed
// launch sequence without synchronising, and the sequence can be captured in
// a hipGraph and replayed (the pass's first sample index can be read from
// device memory, and k_advance moves it on).  extend and shadow are
// persistent kernels: waves fetch chunks of rays from a per-bounce cursor until
// the queue is drained (or, in the dynamic variant, each lane fetches its next
// ray as soon as it is idle).  shade compacts the surviving paths and the
// requested shadow rays with one ballot and one atomic per wave.  Radiance
// lives in a per-path slot and is reduced per pixel in sample order by
// accumulate, so the result does not depend on the order in which queues were
// filled.
//
// BATCH = true: several scenes of different sizes share the queues of one
// pass; each ray carries its scene index and every kernel reads the scene's
// geometry, camera and framebuffer offsets from the scene table.
#include <algorithm>
#include <atomic>

#include "kernels.hpp"
#include "bvh_quant.hpp"
#include "wave_common.hpp"

namespace quasi_hip {

namespace {


template <bool BATCH>
__global__ __launch_bounds__(256) void k_raygen(PassParams P, WaveState S) {
    // queue lengths and cursors of the pass: the camera rays fill the
    // segments in chunks of consecutive paths (queue_slot_of_path)
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < 2 * kMaxDepth * kQueueSegments;
         e += gridDim.x * blockDim.x) {
        const uint32_t seg = e % kQueueSegments;
        if (e < (kMaxDepth + 1) * kQueueSegments)
            S.counts[e * kCounterStride] = e < kQueueSegments ? queue_segment_paths(S.paths, seg, S.qchunk) : 0u;
        if (e < kMaxDepth * kQueueSegments) S.shadow_counts[e * kCounterStride] = 0u;
        S.fetch[e * kCounterStride] = 0u;
    }
    // throughput, radiance and the first-hit records are not written here:
    // bounce 0's shade writes them for every path, hit or miss
    const uint32_t s0 = pass_sample_begin(P);
    float4* ro = static_cast<float4*>(S.ray_o[0]);
    float4* rd = static_cast<float4*>(S.ray_d[0]);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < S.paths; p += gridDim.x * blockDim.x) {
        const uint32_t sid = BATCH ? scene_of_path(P, p) : 0u;
        const SceneDesc& sc = scene_at<BATCH>(P, sid);
        uint32_t pix, s;
        path_pixel_sample(p - sc.path_begin, P.sample_count, pix, s);
        const uint32_t yl = pix / sc.width;
        const uint32_t x = pix - yl * sc.width;
        const uint32_t y = sc.row_begin + yl;
        uint32_t rng = path_seed(P.rng_mode, x, y, sc.width, sc.height, s0 + s, sc.seed);
        const Ray r = camera_ray(sc.cam, sc.width, sc.height, x, y, rng);
        const uint32_t slot = queue_slot_of_path(p, S.seg_cap, S.qchunk);
        ro[slot] = make_float4(r.o.x, r.o.y, r.o.z, __uint_as_float(p));
        rd[slot] = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(sid));
        S.rng[p] = rng;
    }
}



// Persistent closest-hit traversal: each wave reserves CHUNK rays of the
// bounce's queue at a time until the queue is drained.
template <int WIDTH, int BLOCK, int CHUNK, bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_extend(PassParams P, WaveState S, uint32_t b) {
    __shared__ int s_stack[kShortStack * BLOCK];
    ShortStack<kShortStack, BLOCK> st{s_stack + threadIdx.x, P.gstack + (blockIdx.x * BLOCK + threadIdx.x),
                                      gridDim.x * BLOCK, 0};
    count_rays(S.counts + queue_counter(b, 0), b == 0 ? &P.counters->rays_camera : &P.counters->rays_extension);
    const float4* __restrict__ ro = static_cast<const float4*>(S.ray_o[b & 1u]);
    const float4* __restrict__ rd = static_cast<const float4*>(S.ray_d[b & 1u]);
    uint2* hit = static_cast<uint2*>(S.hit);
    SegmentFetch f;
    f.init(S.counts + queue_counter(b, 0), S.fetch + queue_counter(2 * b, 0), S.seg_cap, wave_global());
    uint32_t base = 0, got;
    while ((got = f.take(uint32_t(CHUNK), base)) != 0u) {
        for (uint32_t j = 0; j < uint32_t(CHUNK); j += 64) {
            const uint32_t i = base + j + lane_id();
            if (j + lane_id() < got) {
                const float4 o = ro[i], d = rd[i];
                const uint32_t sid = BATCH ? __float_as_uint(d.w) : 0u;
                const SceneDesc& sc = scene_at<BATCH>(P, sid);
                TraceHit h{kTFar, -1, 0u};
                st.sp = 0;
                if (sc.num_nodes) {
                    const float4* nodes = scene_nodes<WIDTH, BATCH>(P, sid);
                    const float4* tris = static_cast<const float4*>(sc.tris);
                    const Ray ray{xyz(o), xyz(d)};
                    h = WIDTH == 4 ? traverse4<false>(nodes, tris, ray, kTMin, kTFar, st)
                                   : traverse2<false>(nodes, tris, ray, kTMin, kTFar, st);
                }
                hit[i] = make_uint2(__float_as_uint(h.t), uint32_t(h.tri));
            }
        }
    }
}

// One vertex per hit of the bounce's queue (shade_hit); compacts continuing
// paths and shadow rays into the next queues.  FIRST: bounce 0, which also
// records the first-hit AOVs (a separate instantiation keeps their registers
// out of the other bounces).
// With S.tq set, throughput and RNG state travel with the ray in queue order
// (read at the lane's slot, written beside the continuation ray) instead of
// being gathered and scattered by path id, and a next-event-estimation vertex
// after bounce 0 neither reads nor writes the path's radiance, which only the
// shadow rays change from then on (15.2 -> 11.3 ms of shade per 64-spp frame
// on the 1M soup at 1080p).
// At least 5 waves per SIMD: 96 VGPRs instead of the 100 the compiler picks
// on its own (4 waves), more gathers in flight (15.9 -> 15.3 ms of shade per
// 64-spp frame on the 1M soup at 1080p).  The single-scene bounce-0
// instantiation spills 8 bytes there; at 4 waves (no spill) it was slower.
template <int BLOCK, bool BATCH, bool FIRST>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_shade(PassParams P, WaveState S, uint32_t b) {
    // each wave serves one queue segment; the grid is a multiple of the
    // segment count (in waves), so every segment has the same number of waves
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    const uint32_t seg = wave_global() % kQueueSegments;
    const uint32_t n = S.counts[queue_counter(b, seg)];
    const uint32_t slot0 = seg * S.seg_cap;
    uint32_t* next_len = S.counts + queue_counter(b + 1, seg);
    uint32_t* shadow_len = S.shadow_counts + queue_counter(b, seg);
    const float4* __restrict__ ro = static_cast<const float4*>(S.ray_o[b & 1u]);
    const float4* __restrict__ rd = static_cast<const float4*>(S.ray_d[b & 1u]);
    float4* no = static_cast<float4*>(S.ray_o[(b + 1) & 1u]);
    float4* nd = static_cast<float4*>(S.ray_d[(b + 1) & 1u]);
    const uint2* hit = static_cast<const uint2*>(S.hit);
    float4* thr_a = static_cast<float4*>(S.thr);
    float4* rad_a = static_cast<float4*>(S.radiance);
    float4* a0 = static_cast<float4*>(S.aov0);
    float4* a1 = static_cast<float4*>(S.aov1);
    float4* so_a = static_cast<float4*>(S.sh_o);
    float4* sd_a = static_cast<float4*>(S.sh_d);
    float4* sc_a = static_cast<float4*>(S.sh_c);
    const bool qstate = S.tq[0] != nullptr;
    const float4* tq_in = static_cast<const float4*>(S.tq[b & 1u]);
    float4* tq_out = static_cast<float4*>(S.tq[(b + 1) & 1u]);
    // radiance changes at this vertex: always for the path-tracing integrator
    // (emission times throughput), for next-event estimation only at bounce 0
    const bool radiance = FIRST || P.integrator == 0u;
    // wave-uniform loop over the segment so that every lane reaches the ballots
    for (uint32_t wb = (wave_global() / kQueueSegments) * 64u; wb < n; wb += (waves / kQueueSegments) * 64u) {
        const uint32_t i = slot0 + wb + lane_id();
        bool cont = false, shadow = false;
        ShadeOut so;
        uint32_t p = 0, sid = 0;
        f3 thr_next = mk(0.f, 0.f, 0.f);
        uint32_t rng_next = 0u;
        if (wb + lane_id() < n) {
            const float4 o = ro[i], d = rd[i];
            p = __float_as_uint(o.w);
            sid = BATCH ? __float_as_uint(d.w) : 0u;
            const uint2 h = hit[i];
            const int tri = int(h.y);
            if (P.rec_hit) static_cast<uint2*>(P.rec_hit)[size_t(b) * S.paths + p] = h;
            if (tri >= 0) {
                const SceneDesc& sc = scene_at<BATCH>(P, sid);
                // a camera ray starts with throughput 1 and no radiance
                f3 thr = mk(1.f, 1.f, 1.f), L = mk(0.f, 0.f, 0.f);
                uint32_t rng;
                if (FIRST) {
                    rng = S.rng[p];
                } else if (qstate) {
                    const float4 tr = tq_in[i];
                    thr = xyz(tr);
                    rng = __float_as_uint(tr.w);
                } else {
                    thr = xyz(thr_a[p]);
                    rng = S.rng[p];
                }
                if (!FIRST && (radiance || !qstate)) L = xyz(rad_a[p]);
                FirstHit first;
                shade_hit(P, static_cast<const float4*>(sc.tris), static_cast<const float4*>(sc.mats), sc.lights,
                          sc.num_lights, b, Ray{xyz(o), xyz(d)}, __uint_as_float(h.x), tri, thr, L, rng,
                          FIRST ? &first : nullptr, so);
                if (FIRST) {
                    a0[p] = make_float4(first.albedo.x, first.albedo.y, first.albedo.z, first.t);
                    a1[p] = make_float4(first.normal.x, first.normal.y, first.normal.z, __uint_as_float(first.id));
                }
                if (radiance || !qstate) rad_a[p] = make_float4(L.x, L.y, L.z, 0.f);
                if (!qstate) {
                    thr_a[p] = make_float4(thr.x, thr.y, thr.z, 0.f);
                    S.rng[p] = rng;
                }
                thr_next = thr;
                rng_next = rng;
                cont = so.cont;
                shadow = so.shadow;
            } else if (FIRST) {
                // a camera ray that missed: no radiance, empty first-hit records
                rad_a[p] = make_float4(0.f, 0.f, 0.f, 0.f);
                a0[p] = make_float4(0.f, 0.f, 0.f, 0.f);
                a1[p] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kMissId));
            }
        }
        const uint32_t j = slot0 + wave_append(cont, next_len);
        if (cont) {
            no[j] = make_float4(so.next.o.x, so.next.o.y, so.next.o.z, __uint_as_float(p));
            nd[j] = make_float4(so.next.d.x, so.next.d.y, so.next.d.z, __uint_as_float(sid));
            if (qstate) tq_out[j] = make_float4(thr_next.x, thr_next.y, thr_next.z, __uint_as_float(rng_next));
        }
        const uint32_t k = slot0 + wave_append(shadow, shadow_len);
        if (shadow) {
            so_a[k] = make_float4(so.sray.o.x, so.sray.o.y, so.sray.o.z, so.stmax);
            sd_a[k] = make_float4(so.sray.d.x, so.sray.d.y, so.sray.d.z, __uint_as_float(p));
            sc_a[k] = make_float4(so.scontrib.x, so.scontrib.y, so.scontrib.z, __uint_as_float(sid));
        }
    }
}


// Persistent any-hit traversal of the bounce's shadow queue; an unoccluded
// ray adds its contribution to the path's radiance slot (one shadow ray per
// path and bounce, so no two lanes write the same slot).
template <int WIDTH, int BLOCK, int CHUNK, bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_shadow(PassParams P, WaveState S, uint32_t b) {
    __shared__ int s_stack[kShortStack * BLOCK];
    ShortStack<kShortStack, BLOCK> st{s_stack + threadIdx.x, P.gstack + (blockIdx.x * BLOCK + threadIdx.x),
                                      gridDim.x * BLOCK, 0};
    count_rays(S.shadow_counts + queue_counter(b, 0), &P.counters->rays_shadow);
    const float4* __restrict__ so_a = static_cast<const float4*>(S.sh_o);
    const float4* __restrict__ sd_a = static_cast<const float4*>(S.sh_d);
    const float4* __restrict__ sc_a = static_cast<const float4*>(S.sh_c);
    float4* rad = static_cast<float4*>(S.radiance);
    SegmentFetch f;
    f.init(S.shadow_counts + queue_counter(b, 0), S.fetch + queue_counter(2 * b + 1, 0), S.seg_cap, wave_global());
    uint32_t base = 0, got;
    while ((got = f.take(uint32_t(CHUNK), base)) != 0u) {
        for (uint32_t j = 0; j < uint32_t(CHUNK); j += 64) {
            const uint32_t k = base + j + lane_id();
            if (j + lane_id() < got) {
                const float4 o = so_a[k], d = sd_a[k];
                const uint32_t sid = BATCH ? __float_as_uint(sc_a[k].w) : 0u;
                const SceneDesc& sc = scene_at<BATCH>(P, sid);
                const float4* nodes = scene_nodes<WIDTH, BATCH>(P, sid);
                const float4* tris = static_cast<const float4*>(sc.tris);
                const Ray ray{xyz(o), xyz(d)};
                st.sp = 0;
                const bool occ = sc.num_nodes != 0 &&
                                 (WIDTH == 4 ? traverse4<true>(nodes, tris, ray, kTMin, o.w, st).tri >= 0
                                             : traverse2<true>(nodes, tris, ray, kTMin, o.w, st).tri >= 0);
                if (!occ) add_shadow(rad, sc_a, k, __float_as_uint(d.w));
                if (P.rec_vis) P.rec_vis[size_t(b) * S.paths + __float_as_uint(d.w)] = occ ? 0u : 1u;
            }
        }
    }
}

// Persistent traversal with per-lane dynamic ray fetch, in while-while form:
// every lane descends through internal nodes until it reaches a leaf (the wave
// loops while any lane is still descending), then the lanes at a leaf test its
// triangles together; a lane whose ray has finished takes the next ray from
// the queue as soon as kRefillIdle lanes of its wave are idle, so a wave never
// waits for its slowest ray.  Every fetch is one atomic per wave on the
// cursor of the wave's current queue segment.  SHADOW selects the any-hit
// shadow queue; otherwise the closest-hit queue.  Same visit order and tie
// rule as traverse2 / traverse4.
// OFF32 (single scenes whose node and triangle arrays are each under 4 GiB):
// nodes and triangles are addressed as the array's base pointer, uniform over
// the grid, plus a 32-bit byte offset, which the loads take as they are (with
// the record's field offsets as immediates); otherwise, and in every batch,
// each lane holds 64-bit pointers to its scene's arrays.  A batch lane's
// 32-bit scene offset added to the record's kept the compiler from folding the
// field offsets (four more adds per node step): the 1M soup as a batch of one
// ran at 4211 Mrays/s that way against 4398 with 64-bit pointers, the 64-scene
// batch at 4591 against 4696 (73 VGPRs and 6 waves per SIMD against 71 and 7).

// TOP (single-scene BVH4 kernels): the top levels of the tree (the root, its
// internal children, theirs, ...: four levels, at most 85 nodes, 9.5 KB, in
// 256-thread blocks; three levels, 21 nodes, in smaller ones) are copied into
// LDS when the block starts, without the nodes' unused last float4, their
// child references to staged nodes rewritten as kLdsRef + slot, and node steps
// on them read LDS instead of global memory.  Every ray starts at the root, so
// these few nodes take a share of all node requests that would otherwise
// concentrate on a handful of L2 lines; the LDS copy takes that load off the
// L2.  On the 1M soup at 1080p, extend + shadow time per 64-spp frame: 150.8 ms
// without staging, 142.3 ms with three levels (block 128), 139.4 ms with four
// (block 256; 5 resident blocks per CU, limited by LDS), 147.7 ms with four
// levels and an 8-entry short stack (8 blocks per CU, more stack spills).
//
// QUANT (BVH4, single scenes and batches): the nodes are the 64-byte quantised
// ones of bvh_quant.hpp (P.qnodes4, a scene's SceneDesc::qnodes4), read with
// four 16-byte loads per node step instead of seven; staged top levels take
// 4 float4 per node.  With 128-byte nodes the kernel was bound by the texture
// addresser's per-lane cache-line lookups (busy 86% of the extend launch on the
// 1M soup), so the loads, not the VALU work, set the node step's cost: extend +
// shadow per 64-spp frame 106.2 ms (three levels staged, block 128) against
// 136.1 ms (128-byte nodes, four levels, block 256), same images.
constexpr int kLdsRef = 0x40000000;   // node references >= this are LDS slots (fewer than 2^30 nodes)
constexpr int kTopFloat4 = 7;         // float4s of a BVH4 node that the traversal reads ([7] is unused)

template <int BLOCK>
constexpr int top_levels() { return BLOCK >= 256 ? 4 : 3; }
template <int BLOCK>
constexpr int top_nodes() { return top_levels<BLOCK>() == 4 ? 1 + 4 + 16 + 64 : 1 + 4 + 16; }

// TOP in a batch (quantised nodes): a block's rays belong to many scenes, so
// each wave stages the top three levels (21 nodes, 1.3 KB) of one scene for
// itself: slot 0 the root, slot 1 + k its child k, slot 5 + 4 k + j child j of
// that child, where internal; child references to staged nodes rewritten as
// kLdsRef + slot.  A wave restages, to the scene of the first ray it has just
// fetched, only when none of its lanes is still walking the staged copy;
// lanes of other scenes start at their scene's root in global memory.
constexpr int kWaveTopNodes = 1 + 4 + 16;
constexpr uint32_t kNoScene = 0xFFFFFFFFu;

__device__ __forceinline__ bool internal_ref(int c) { return c >= 0 && c != kEmptyChild; }

__device__ __forceinline__ int ref_of(const float4& v, int k) {
    return __float_as_int(k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)));
}

// The whole wave stages scene sid's top levels into wtop (kWaveTopNodes x 4 float4).
__device__ __forceinline__ void stage_wave_top(const PassParams& P, uint32_t sid, float4* wtop) {
    const float4* q = static_cast<const float4*>(P.scenes[sid].qnodes4);
    const uint32_t lane = lane_id();
    const float4 root_refs = q[3];
    int n = -1;   // the node this lane stages (slot = lane)
    if (lane == 0u) {
        n = 0;
    } else if (lane <= 4u) {
        const int c = ref_of(root_refs, int(lane) - 1);
        n = internal_ref(c) ? c : -1;
    } else if (lane < uint32_t(kWaveTopNodes)) {
        const int p = ref_of(root_refs, int(lane - 5u) >> 2);
        if (internal_ref(p)) {
            const int c = ref_of(q[size_t(p) * 4 + 3], int(lane - 5u) & 3);
            n = internal_ref(c) ? c : -1;
        }
    }
    if (n >= 0) {
        const float4* s = q + size_t(n) * 4;
        const float4 q0 = s[0], q1 = s[1], q2 = s[2];
        float4 q3 = s[3];
        if (lane <= 4u) {
            int c[4] = {__float_as_int(q3.x), __float_as_int(q3.y), __float_as_int(q3.z), __float_as_int(q3.w)};
            for (int j = 0; j < 4; ++j)
                if (internal_ref(c[j])) c[j] = kLdsRef + (lane == 0u ? 1 + j : 5 + 4 * int(lane - 1u) + j);
            q3 = make_float4(__int_as_float(c[0]), __int_as_float(c[1]), __int_as_float(c[2]), __int_as_float(c[3]));
        }
        float4* d = wtop + lane * 4u;
        d[0] = q0; d[1] = q1; d[2] = q2; d[3] = q3;
    }
    // the other lanes of the wave read these slots next
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int WIDTH, int BLOCK, bool SHADOW, bool BATCH, bool OFF32, bool TOP = false, bool QUANT = false>
__global__ __launch_bounds__(BLOCK) void k_trace_dynamic(PassParams P, WaveState S, uint32_t b) {
    static_assert(!(BATCH && OFF32), "a batch addresses its scenes through 64-bit pointers");
    constexpr bool kQuant = QUANT && WIDTH == 4;
    // 128-B BVH4 / 64-B quantised BVH4 / 64-B BVH2 nodes
    constexpr uint32_t kNodeShift = WIDTH == 4 && !kQuant ? 7u : 6u;
    constexpr bool kTop = TOP && WIDTH == 4 && !BATCH;
    constexpr bool kWaveTop = TOP && kQuant && BATCH;   // per-wave staging in a batch
    constexpr int kTopNodes = top_nodes<BLOCK>();
    constexpr int kTopF4 = kQuant ? 4 : kTopFloat4;   // float4 staged per node
    constexpr int kRefF4 = kQuant ? 3 : 6;            // the node's float4 of child references
    constexpr int kStrideF4 = kQuant ? 4 : 8;         // node size in float4
    __shared__ int s_stack[kShortStack * BLOCK];
    __shared__ float4 s_top[kTop ? kTopNodes * kTopF4 : 1];
    __shared__ int s_top_map[kTop ? kTopNodes : 1];
    __shared__ int s_top_n;
    __shared__ float4 s_wtop[kWaveTop ? (BLOCK / 64) * kWaveTopNodes * 4 : 1];
    float4* const wtop = s_wtop + (kWaveTop ? (threadIdx.x >> 6) * kWaveTopNodes * 4 : 0);
    uint32_t staged = kNoScene;   // wave-uniform: the scene in wtop
    if (kTop) {
        const float4* gn = static_cast<const float4*>(kQuant ? P.qnodes4 : P.nodes4);
        if (threadIdx.x == 0) {
            // breadth first over the top levels: slot -> global node index
            int n = 0;
            if (P.scene0.num_nodes != 0u) {
                s_top_map[n++] = 0;
                int lo = 0, hi = 1;
                for (int level = 1; level < top_levels<BLOCK>(); ++level) {
                    for (int k = lo; k < hi; ++k) {
                        const float4 chf = gn[size_t(s_top_map[k]) * kStrideF4 + kRefF4];
                        const int c[4] = {__float_as_int(chf.x), __float_as_int(chf.y), __float_as_int(chf.z),
                                          __float_as_int(chf.w)};
                        for (int j = 0; j < 4; ++j)
                            if (c[j] >= 0 && c[j] != kEmptyChild && n < kTopNodes) s_top_map[n++] = c[j];
                    }
                    lo = hi;
                    hi = n;
                }
            }
            s_top_n = n;
        }
        __syncthreads();
        const int n = s_top_n;
        for (int i = int(threadIdx.x); i < n * kTopF4; i += BLOCK)
            s_top[i] = gn[size_t(s_top_map[i / kTopF4]) * kStrideF4 + (i % kTopF4)];
        __syncthreads();
        // child references to staged nodes point at their slots
        for (int i = int(threadIdx.x); i < n * 4; i += BLOCK) {
            int* ref = reinterpret_cast<int*>(&s_top[(i >> 2) * kTopF4 + kRefF4]) + (i & 3);
            const int c = *ref;
            if (c < 0 || c == kEmptyChild) continue;
            for (int m = 1; m < n; ++m)
                if (s_top_map[m] == c) { *ref = kLdsRef + m; break; }
        }
        __syncthreads();
    }
    ShortStack<kShortStack, BLOCK> st{s_stack + threadIdx.x, P.gstack + (blockIdx.x * BLOCK + threadIdx.x),
                                      gridDim.x * BLOCK, 0};
    const uint32_t* lengths = SHADOW ? S.shadow_counts + queue_counter(b, 0) : S.counts + queue_counter(b, 0);
    count_rays(lengths, SHADOW ? &P.counters->rays_shadow
                               : (b == 0 ? &P.counters->rays_camera : &P.counters->rays_extension));
    const float4* __restrict__ ro = static_cast<const float4*>(SHADOW ? S.sh_o : S.ray_o[b & 1u]);
    const float4* __restrict__ rd = static_cast<const float4*>(SHADOW ? S.sh_d : S.ray_d[b & 1u]);
    const float4* __restrict__ sc_a = static_cast<const float4*>(S.sh_c);
    uint2* hit = static_cast<uint2*>(S.hit);
    float4* rad = static_cast<float4*>(S.radiance);
    SegmentFetch f;
    f.init(lengths, S.fetch + queue_counter(2 * b + (SHADOW ? 1 : 0), 0), S.seg_cap, wave_global());

    Ray r{mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 1.f)};
    f3 inv = mk(0.f, 0.f, 0.f);
    float tmin = kTMin, best_t = kTFar;
    int best_tri = -1;
    uint32_t best_id = 0u;
    int node = kDone;
    int leaf = 0;   // postponed leaf reference (leaf references are negative; 0 = none)
    bool busy = false;
    bool fresh = false;      // per-wave staging: the lane took its ray in this refill
    bool on_top = false;     // ... and walks the wave's staged copy
    uint32_t lane_sid = kNoScene;
    uint32_t idx = 0, path = 0;
    // byte offsets in a BVH4 node of each axis's entry plane (lo, or hi when
    // the direction component is negative); the exit plane is at offset ^ 16
    uint32_t sel_x = 0, sel_y = 32, sel_z = 64;
    // base of the node and triangle arrays (a batch: the concatenated ones)
    const char* const nbase = static_cast<const char*>(kQuant ? P.qnodes4 : (WIDTH == 4 ? P.nodes4 : P.nodes));
    const char* const tbase = static_cast<const char*>(P.tris);
    const char* nodes = BATCH ? nullptr : nbase;    // the lane's scene's arrays (64-bit addressing)
    const char* tris = BATCH ? nullptr : tbase;
    for (;;) {
        // refill idle lanes (wave-uniform: every lane reaches the ballots)
        const unsigned long long idle = __ballot(!busy);
        const unsigned long long act = __ballot(busy);
        unsigned long long want = f.live() && (act == 0ull || uint32_t(__popcll(idle)) >= kRefillIdle) ? idle : 0ull;
        fresh = false;
        while (want) {
            uint32_t first = 0;
            const uint32_t got = f.take(uint32_t(__popcll(want)), first);
            if (got == 0u) break;   // every segment exhausted
            const uint32_t lb = lanes_below(want);
            const bool mine = ((want >> (threadIdx.x & 63u)) & 1ull) && lb < got;
            want &= ~__ballot(mine);
            if (mine) {
                fresh = true;
                on_top = false;
                idx = first + lb;
                const float4 o = ro[idx];
                const float4 d = rd[idx];
                path = __float_as_uint(d.w);
                const uint32_t sid = BATCH ? __float_as_uint(SHADOW ? sc_a[idx].w : d.w) : 0u;
                const SceneDesc& sc = scene_at<BATCH>(P, sid);
                if (BATCH) {
                    nodes = static_cast<const char*>(kQuant ? sc.qnodes4 : scene_nodes<WIDTH, true>(P, sid));
                    tris = static_cast<const char*>(sc.tris);
                }
                r = Ray{xyz(o), xyz(d)};
                inv = mk(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
                sel_x = inv.x < 0.f ? 16u : 0u;
                sel_y = inv.y < 0.f ? 48u : 32u;
                sel_z = inv.z < 0.f ? 80u : 64u;
                best_t = SHADOW ? o.w : kTFar;
                best_tri = -1;
                best_id = 0u;
                st.sp = 0;
                node = sc.num_nodes ? (kTop ? kLdsRef : 0) : kDone;
                lane_sid = sc.num_nodes ? sid : kNoScene;
                leaf = 0;
                busy = true;
            }
        }
        if (kWaveTop) {
            // restage when no lane walks the staged copy any more
            if (__ballot(busy && on_top) == 0ull) {
                const unsigned long long cand = __ballot(fresh && lane_sid != kNoScene);
                if (cand) {
                    const uint32_t want_sid =
                        uint32_t(__builtin_amdgcn_readlane(int(lane_sid), int(__builtin_ctzll(cand))));
                    if (want_sid != staged) {
                        stage_wave_top(P, want_sid, wtop);
                        staged = want_sid;
                    }
                }
            }
            if (fresh && lane_sid == staged && node == 0) {
                node = kLdsRef;
                on_top = true;
            }
        }
        // no lane busy after a refill: every segment is exhausted
        if (__ballot(busy) == 0ull) break;
        // descend through internal nodes.  A lane that reaches a leaf keeps it
        // in `leaf` and goes on descending from its stack while lanes of the
        // wave without a leaf are still looking for one (postponed leaves:
        // lanes that would idle do node work instead), until only a few lanes
        // are still looking (leave_node_loop)
        for (;;) {
            const bool desc = busy && node >= 0;
            if (leave_node_loop(desc && leaf == 0, busy && leaf != 0)) break;
            if (!desc) continue;
            // the node's record: base + 32-bit offset (OFF32) or a 64-bit address
            const uint32_t off = uint32_t(node) << kNodeShift;
            const char* rec = nodes + (OFF32 ? size_t(0) : (size_t(uint32_t(node)) << kNodeShift));
            auto ld = [&](uint32_t o) -> float4 {
                return OFF32 ? *reinterpret_cast<const float4*>(nbase + (off + o))
                             : *reinterpret_cast<const float4*>(rec + o);
            };
            if (kQuant) {
                float4 q0, q1, q2, q3;
                if ((kTop || kWaveTop) && node >= kLdsRef) {
                    const float4* q = (kWaveTop ? wtop : s_top) + (node - kLdsRef) * 4;
                    q0 = q[0]; q1 = q[1]; q2 = q[2]; q3 = q[3];
                } else {
                    q0 = ld(0u); q1 = ld(16u); q2 = ld(32u); q3 = ld(48u);
                }
                node = quant_node_step<BLOCK>(q0, q1, q2, q3, r, inv, tmin, best_t, st);
            } else if (WIDTH == 4) {
                // entry and exit planes chosen by the direction's signs: the same
                // values slab() takes the min and max of, without the min/max
                float4 ex, fx, ey, fy, ez, fz, chf;
                if (kTop && node >= kLdsRef) {
                    const float4* q = s_top + (node - kLdsRef) * kTopFloat4;
                    ex = q[sel_x >> 4]; fx = q[(sel_x ^ 16u) >> 4];
                    ey = q[sel_y >> 4]; fy = q[(sel_y ^ 16u) >> 4];
                    ez = q[sel_z >> 4]; fz = q[(sel_z ^ 16u) >> 4];
                    chf = q[6];
                } else {
                    ex = ld(sel_x); fx = ld(sel_x ^ 16u);
                    ey = ld(sel_y); fy = ld(sel_y ^ 16u);
                    ez = ld(sel_z); fz = ld(sel_z ^ 16u);
                    chf = ld(96u);
                }
                const int4 ch = make_int4(__float_as_int(chf.x), __float_as_int(chf.y), __float_as_int(chf.z),
                                          __float_as_int(chf.w));
                float t0n, t0f, t1n, t1f, t2n, t2f, t3n, t3f;
                slab_sel(ex.x, fx.x, ey.x, fy.x, ez.x, fz.x, r, inv, tmin, best_t, t0n, t0f);
                slab_sel(ex.y, fx.y, ey.y, fy.y, ez.y, fz.y, r, inv, tmin, best_t, t1n, t1f);
                slab_sel(ex.z, fx.z, ey.z, fy.z, ez.z, fz.z, r, inv, tmin, best_t, t2n, t2f);
                slab_sel(ex.w, fx.w, ey.w, fy.w, ez.w, fz.w, r, inv, tmin, best_t, t3n, t3f);
                // unused slots hold inverted boxes (lo = +inf, hi = -inf), which
                // the sign-selected slab test always rejects
                node = visit_children4<BLOCK>(t0n <= t0f ? t0n : kMissKey, t1n <= t1f ? t1n : kMissKey,
                                              t2n <= t2f ? t2n : kMissKey, t3n <= t3f ? t3n : kMissKey, ch.x, ch.y,
                                              ch.z, ch.w, st);
            } else {
                const float4 a = ld(0u), bb = ld(16u), c = ld(32u), dd = ld(48u);
                const int cl = __float_as_int(dd.x);
                const int cr = __float_as_int(dd.y);
                float tnl, tfl, tnr, tfr;
                slab(a.x, a.y, a.z, a.w, c.x, c.y, r, inv, tmin, best_t, tnl, tfl);
                slab(bb.x, bb.y, bb.z, bb.w, c.z, c.w, r, inv, tmin, best_t, tnr, tfr);
                const bool hl = tnl <= tfl;
                const bool hr = tnr <= tfr;
                if (hl && hr) {
                    const bool left_first = tnl <= tnr;
                    st.push(left_first ? cr : cl);
                    node = left_first ? cl : cr;
                } else if (hl || hr) {
                    node = hl ? cl : cr;
                } else {
                    node = st.empty() ? kDone : st.pop();
                }
            }
            if (leaf == 0 && node < 0 && node != kDone) {
                leaf = node;
                node = st.empty() ? kDone : st.pop();
            }
        }
        // the lanes holding a leaf test its triangles; a lane whose descent
        // stopped at a second leaf takes that one next
        if (busy && leaf != 0) {
            const uint32_t v = ~uint32_t(leaf);
            const int first = int(v >> 2);
            const int count = int(v & 3u) + 1;
            bool found = false;
            uint32_t toff = uint32_t(first) * 48u;   // 48-B triangle records
            const char* trec = tris + (OFF32 ? size_t(0) : size_t(uint32_t(first)) * 48u);
            for (int i = first; i < first + count; ++i, toff += 48u, trec += OFF32 ? 0 : 48) {
                float t;
                uint32_t id;
                const float4* tr = reinterpret_cast<const float4*>(OFF32 ? tbase + toff : trec);
                if (!hit_triangle_rec(tr, r, tmin, best_t, t, id)) continue;
                if (SHADOW) { best_t = t; best_tri = i; best_id = id; found = true; break; }
                if (best_tri < 0 || t < best_t || id > best_id) {
                    best_t = t;
                    best_tri = i;
                    best_id = id;
                }
            }
            if (found) {
                node = kDone;
                leaf = 0;
            } else if (node < 0 && node != kDone) {
                leaf = node;
                node = st.empty() ? kDone : st.pop();
            } else {
                leaf = 0;
            }
        }
        if (busy && node == kDone && leaf == 0) {
            if (SHADOW) {
                if (best_tri < 0) add_shadow(rad, sc_a, idx, path);
                if (P.rec_vis) P.rec_vis[size_t(b) * S.paths + path] = best_tri < 0 ? 1u : 0u;
            } else {
                hit[idx] = make_uint2(__float_as_uint(best_t), uint32_t(best_tri));
            }
            busy = false;
        }
    }
}

// The plain wavefront traversal: no persistence and no work fetch, one thread
// per slot of the bounce's (segmented) queue; slots past their segment's
// length exit at once.  The whole stack lives in LDS (kFlatStack entries, which
// the host checks against the BVH depth).  SHADOW: any hit on the shadow queue.
template <int WIDTH, int BLOCK, bool SHADOW, bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_trace_flat(PassParams P, WaveState S, uint32_t b) {
    __shared__ int s_stack[kFlatStack * BLOCK];
    const uint32_t* lengths = SHADOW ? S.shadow_counts + queue_counter(b, 0) : S.counts + queue_counter(b, 0);
    count_rays(lengths, SHADOW ? &P.counters->rays_shadow
                               : (b == 0 ? &P.counters->rays_camera : &P.counters->rays_extension));
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t seg = i / S.seg_cap;
    if (seg >= kQueueSegments || i - seg * S.seg_cap >= lengths[seg * kCounterStride]) return;
    const float4* __restrict__ ro = static_cast<const float4*>(SHADOW ? S.sh_o : S.ray_o[b & 1u]);
    const float4* __restrict__ rd = static_cast<const float4*>(SHADOW ? S.sh_d : S.ray_d[b & 1u]);
    const float4* __restrict__ sc_a = static_cast<const float4*>(S.sh_c);
    const float4 o = ro[i], d = rd[i];
    const uint32_t sid = BATCH ? __float_as_uint(SHADOW ? sc_a[i].w : d.w) : 0u;
    const SceneDesc& sc = scene_at<BATCH>(P, sid);
    const float4* nodes = scene_nodes<WIDTH, BATCH>(P, sid);
    const float4* tris = static_cast<const float4*>(sc.tris);
    LdsStack<BLOCK> st{s_stack + threadIdx.x, 0};
    const Ray ray{xyz(o), xyz(d)};
    TraceHit h{SHADOW ? o.w : kTFar, -1, 0u};
    if (sc.num_nodes)
        h = WIDTH == 4 ? traverse4<SHADOW>(nodes, tris, ray, kTMin, SHADOW ? o.w : kTFar, st)
                       : traverse2<SHADOW>(nodes, tris, ray, kTMin, SHADOW ? o.w : kTFar, st);
    if (SHADOW) {
        const uint32_t path = __float_as_uint(d.w);
        if (h.tri < 0) add_shadow(static_cast<float4*>(S.radiance), sc_a, i, path);
        if (P.rec_vis) P.rec_vis[size_t(b) * S.paths + path] = h.tri < 0 ? 1u : 0u;
    } else {
        static_cast<uint2*>(S.hit)[i] = make_uint2(__float_as_uint(h.t), uint32_t(h.tri));
    }
}

// Per pixel: the pass's samples in sample order, added to the accumulation.
// A pixel's samples are neighbouring paths (p = gp * samples + s: every scene
// of a batch starts at path pix_begin * samples), so a block stages the
// records of a tile of up to 64 pixels in LDS with coalesced loads, in rounds
// of kAccRow samples per pixel (one row per pixel, padded by one float4
// against bank conflicts).  Each of the block's first three waves then adds
// one staged array (radiance; albedo and depth; normal and the id sample) for
// the tile's pixels, one pixel per lane, every component in sample order, so
// the sums are those of a sequential per-pixel loop.  Reading the records
// straight from HBM, one pixel per thread, strides the loads by samples x 16 B
// and took 6.6 ms per 64-spp frame at 1080p.
constexpr uint32_t kAccEntries = 1040;   // float4 per staged array (16.25 KB)
constexpr uint32_t kAccRow = 16;         // samples staged per pixel and round
constexpr uint32_t kAccTile = 64;        // pixels per tile at most (one wave's lanes)
template <bool BATCH>
__global__ __launch_bounds__(256) void k_accumulate(PassParams P, WaveState S) {
    __shared__ float4 s_rad[kAccEntries], s_a0[kAccEntries], s_a1[kAccEntries];
    const uint32_t spp = P.sample_count;
    const uint32_t s0 = pass_sample_begin(P);
    const float4* rad = static_cast<const float4*>(S.radiance);
    const float4* a0 = static_cast<const float4*>(S.aov0);
    const float4* a1 = static_cast<const float4*>(S.aov1);
    const uint32_t chunk = spp < kAccRow ? spp : kAccRow;      // samples per pixel and round
    const uint32_t stride = chunk + 1u;                          // padded row
    uint32_t tile = kAccEntries / stride;                        // pixels per tile
    if (tile > kAccTile) tile = kAccTile;
    const uint32_t role = threadIdx.x / kAccTile;                // 0 radiance, 1 albedo + depth, 2 normal
    const uint32_t j = threadIdx.x % kAccTile;                   // the lane's pixel in the tile
    const float4* src = role == 0 ? s_rad : (role == 1 ? s_a0 : s_a1);
    for (uint32_t t0 = blockIdx.x * tile; t0 < P.total_pixels; t0 += gridDim.x * tile) {
        const uint32_t np = P.total_pixels - t0 < tile ? P.total_pixels - t0 : tile;
        const bool mine = role < 3u && j < np;
        const uint32_t gp = t0 + j;
        float4 sum = {0, 0, 0, 0}, sd = {0, 0, 0, 0};
        for (uint32_t c0 = 0; c0 < spp; c0 += chunk) {
            const uint32_t cn = spp - c0 < chunk ? spp - c0 : chunk;
            __syncthreads();   // the previous round's rows are read
            for (uint32_t e = threadIdx.x; e < np * cn; e += blockDim.x) {
                const uint32_t jj = e / cn, k = e - jj * cn;
                const size_t p = size_t(t0 + jj) * spp + c0 + k;
                s_rad[jj * stride + k] = rad[p];
                s_a0[jj * stride + k] = a0[p];
                s_a1[jj * stride + k] = a1[p];
            }
            __syncthreads();
            if (mine) {
                for (uint32_t k = 0; k < cn; ++k) {
                    const float4 x = src[j * stride + k];
                    sum.x += x.x; sum.y += x.y; sum.z += x.z; sum.w += 1.0f;
                    if (role == 1u) {
                        const float dn = fminf(fmaxf((x.w - 1.0f) / 4.0f, 0.0f), 1.0f);   // plugin.mm:301
                        sd.x += dn; sd.y += dn; sd.z += dn; sd.w += 1.0f;
                    } else if (role == 2u && P.record_ids && s0 + c0 + k == P.id_sample) {
                        P.ids[gp] = __float_as_uint(x.w);
                    }
                }
            }
        }
        if (mine) {
            float4* acc = static_cast<float4*>(role == 0 ? P.accum_beauty : (role == 1 ? P.accum_albedo : P.accum_normal));
            float4 o = acc[gp];
            o.x += sum.x; o.y += sum.y; o.z += sum.z; o.w += sum.w;
            acc[gp] = o;
            if (role == 1u) {
                float4* ad = static_cast<float4*>(P.accum_depth);
                o = ad[gp]; o.x += sd.x; o.y += sd.y; o.z += sd.z; o.w += sd.w; ad[gp] = o;
            }
        }
    }
}

// Moves the device-side sample index on by one pass (last node of a
// captured pass graph).
__global__ void k_advance(uint32_t* sample_begin, uint32_t by) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *sample_begin += by;
}

uint32_t grid_for(size_t n, uint32_t block, uint32_t cap) {
    size_t g = (n + block - 1) / block;
    return uint32_t(g < 1 ? 1 : (g > cap ? cap : g));
}

constexpr int kShadeBlock = 256;

using TraceFn = hipError_t (*)(const PassParams&, const WaveState&, uint32_t, uint32_t, hipStream_t);
using OccFn = int (*)();

template <class K>
int occupancy_of(K kernel, int block) {
    int a = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, kernel, block, 0) != hipSuccess) a = 1;
    return a;
}

// CHUNK 64 / 256: static chunks per wave; CHUNK 1: per-lane dynamic fetch;
// CHUNK 2 (kChunkLdsTop): per-lane dynamic fetch with the top BVH4 levels in
// LDS; CHUNK 0 (kChunkFlat in a TraceConfig): the non-persistent traversal.
template <int WIDTH, int BLOCK, int CHUNK>
struct TraceVariant {
    static constexpr bool kDynamic = CHUNK == 1 || CHUNK == int(kChunkLdsTop) || CHUNK == int(kChunkQuant);
    static constexpr bool kTop = CHUNK == int(kChunkLdsTop) || CHUNK == int(kChunkQuant);
    static constexpr bool kQ = CHUNK == int(kChunkQuant);
    static constexpr int kStatic = kDynamic || CHUNK == 0 ? 64 : CHUNK;
    // one thread per queue slot; the stack must hold the deepest walk
    template <bool SHADOW, bool BATCH>
    static hipError_t flat(const PassParams& p, const WaveState& s, uint32_t b, hipStream_t st) {
        if (p.gstack_entries + uint32_t(kShortStack) > uint32_t(kFlatStack)) return hipErrorInvalidValue;
        const uint32_t blocks = uint32_t((uint64_t(s.seg_cap) * kQueueSegments + BLOCK - 1) / BLOCK);
        hipLaunchKernelGGL((k_trace_flat<WIDTH, BLOCK, SHADOW, BATCH>), dim3(blocks), dim3(BLOCK), 0, st, p, s, b);
        return hipGetLastError();
    }
    // node and triangle arrays each under 4 GiB (single scenes)
    static bool off32(const PassParams& p) {
        const uint64_t node_bytes = uint64_t(WIDTH == 4 ? p.num_nodes4 : p.num_nodes) * (WIDTH == 4 && !kQ ? 128u : 64u);
        return node_bytes < (1ull << 32) && uint64_t(p.num_tris) * 48u < (1ull << 32);
    }
    template <bool SHADOW, bool BATCH>
    static void dynamic(const PassParams& p, const WaveState& s, uint32_t b, uint32_t blocks, hipStream_t st) {
        if constexpr (BATCH)
            hipLaunchKernelGGL((k_trace_dynamic<WIDTH, BLOCK, SHADOW, true, false, kTop, kQ>), dim3(blocks), dim3(BLOCK), 0, st, p, s, b);
        else if (off32(p))
            hipLaunchKernelGGL((k_trace_dynamic<WIDTH, BLOCK, SHADOW, false, true, kTop, kQ>), dim3(blocks), dim3(BLOCK), 0, st, p, s, b);
        else
            hipLaunchKernelGGL((k_trace_dynamic<WIDTH, BLOCK, SHADOW, false, false, kTop, kQ>), dim3(blocks), dim3(BLOCK), 0, st, p, s, b);
    }
    template <bool BATCH>
    static hipError_t extend(const PassParams& p, const WaveState& s, uint32_t b, uint32_t blocks, hipStream_t st) {
        if constexpr (CHUNK == 0) return flat<false, BATCH>(p, s, b, st);
        if constexpr (kDynamic)
            dynamic<false, BATCH>(p, s, b, blocks, st);
        else
            hipLaunchKernelGGL((k_extend<WIDTH, BLOCK, kStatic, BATCH>), dim3(blocks), dim3(BLOCK), 0, st, p, s, b);
        return hipGetLastError();
    }
    template <bool BATCH>
    static hipError_t shadow(const PassParams& p, const WaveState& s, uint32_t b, uint32_t blocks, hipStream_t st) {
        if constexpr (CHUNK == 0) return flat<true, BATCH>(p, s, b, st);
        if constexpr (kDynamic)
            dynamic<true, BATCH>(p, s, b, blocks, st);
        else
            hipLaunchKernelGGL((k_shadow<WIDTH, BLOCK, kStatic, BATCH>), dim3(blocks), dim3(BLOCK), 0, st, p, s, b);
        return hipGetLastError();
    }
    // Resident blocks per CU of the single-scene kernels.  The grid is sized
    // by it; kernels with fewer resident blocks (a batch's) start the rest as
    // the first ones finish, which is harmless: no block waits for another.
    static int occupancy() {
        int a, c;
        if constexpr (CHUNK == 0) {
            a = occupancy_of(k_trace_flat<WIDTH, BLOCK, false, false>, BLOCK);
            c = occupancy_of(k_trace_flat<WIDTH, BLOCK, true, false>, BLOCK);
        } else if constexpr (kDynamic) {
            a = occupancy_of(k_trace_dynamic<WIDTH, BLOCK, false, false, true, kTop, kQ>, BLOCK);
            c = occupancy_of(k_trace_dynamic<WIDTH, BLOCK, true, false, true, kTop, kQ>, BLOCK);
        } else {
            a = occupancy_of(k_extend<WIDTH, BLOCK, kStatic, false>, BLOCK);
            c = occupancy_of(k_shadow<WIDTH, BLOCK, kStatic, false>, BLOCK);
        }
        return a < c ? a : c;
    }
};

struct VariantEntry {
    uint32_t width, block, chunk;
    TraceFn extend1, shadow1;   // one scene
    TraceFn extendB, shadowB;   // batch
    OccFn occupancy;
};

#define QH_VARIANT(W, B, C)                                                                                  \
    {W, B, C, &TraceVariant<W, B, C>::extend<false>, &TraceVariant<W, B, C>::shadow<false>,                  \
     &TraceVariant<W, B, C>::extend<true>, &TraceVariant<W, B, C>::shadow<true>, &TraceVariant<W, B, C>::occupancy}
const VariantEntry kVariants[] = {
    QH_VARIANT(2, 64, 64),  QH_VARIANT(2, 128, 64),  QH_VARIANT(2, 256, 64),
    QH_VARIANT(2, 64, 256), QH_VARIANT(2, 128, 256), QH_VARIANT(2, 256, 256),
    QH_VARIANT(4, 64, 64),  QH_VARIANT(4, 128, 64),  QH_VARIANT(4, 256, 64),
    QH_VARIANT(4, 64, 256), QH_VARIANT(4, 128, 256), QH_VARIANT(4, 256, 256),
    QH_VARIANT(2, 64, 1),   QH_VARIANT(2, 128, 1),   QH_VARIANT(2, 256, 1),
    QH_VARIANT(4, 64, 1),   QH_VARIANT(4, 128, 1),   QH_VARIANT(4, 256, 1),
    QH_VARIANT(2, 128, 0),  QH_VARIANT(4, 128, 0),
    QH_VARIANT(4, 128, 2),  QH_VARIANT(4, 256, 2),
    QH_VARIANT(4, 128, 3),  QH_VARIANT(4, 256, 3),
};
#undef QH_VARIANT

const VariantEntry* find_variant(const TraceConfig& c) {
    const uint32_t chunk = c.chunk == kChunkFlat ? 0u : c.chunk;
    for (const VariantEntry& v : kVariants)
        if (v.width == c.width && v.block == c.block && v.chunk == chunk) return &v;
    return nullptr;
}

}  // namespace

bool trace_config_supported(const TraceConfig& c) { return find_variant(c) != nullptr; }

uint32_t trace_config_count() { return uint32_t(sizeof(kVariants) / sizeof(kVariants[0])); }

TraceConfig trace_config_at(uint32_t i) {
    const VariantEntry& v = kVariants[i % trace_config_count()];
    return TraceConfig{v.width, v.block, v.chunk == 0 ? kChunkFlat : v.chunk};
}

namespace {

// Blocks of k_shade that device `dev` (the current device) holds resident at
// once: its CU count times the kernel's blocks per CU, cached per device
// (persistent_blocks, called when a plugin starts, fills the cache outside
// any stream capture).  The shade grid is capped there: each wave then walks
// its segment in consecutive groups of 64 slots beside the segment's other
// waves, so the rays it appends keep the order of the queue it reads (a
// pixel's samples stay neighbours bounce after bounce).  A grid of 8192 blocks, whose later waves start on groups far
// from the first ones', scattered that order: on the 1M soup at 1080p, 64
// spp, extend + shadow + shade per frame took 79.4-80.0 ms against 74.8-76.7
// with the resident grid (same images); half or three quarters of the
// resident grid left shade short of waves (14.5 and 12.7 ms against 10.9).
constexpr int kShadeCacheDevices = 64;
uint32_t shade_resident_blocks(int dev) {
    static std::atomic<uint32_t> cache[kShadeCacheDevices];
    if (dev < 0) dev = 0;
    const bool cached = dev < kShadeCacheDevices;
    if (cached) {
        const uint32_t c = cache[dev].load(std::memory_order_relaxed);
        if (c) return c;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    const int per_cu = std::max(1, occupancy_of(k_shade<kShadeBlock, false, false>, kShadeBlock));
    const uint32_t blocks = uint32_t(cus * per_cu);
    if (cached) cache[dev].store(blocks, std::memory_order_relaxed);
    return blocks;
}

}  // namespace

uint32_t persistent_blocks(int device, const TraceConfig& c) {
    (void)shade_resident_blocks(device);
    const VariantEntry* v = find_variant(c);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    int per_cu = v ? v->occupancy() : 1;
    if (per_cu <= 0) per_cu = 1;
    // never more threads than the stack-overflow area is sized for
    const int max_per_cu = int(kMaxThreadsPerCU / c.block);
    if (per_cu > max_per_cu) per_cu = max_per_cu;
    return uint32_t(cus * per_cu);
}

hipError_t launch_wavefront_pass(const PassParams& p, const WaveState& s, const TraceConfig& cfg,
                                 uint32_t trace_blocks, hipStream_t stream, WaveTimers* tm) {
    if (s.paths == 0) return hipSuccess;
    // without quantised nodes (a scene beyond their range) the same fetch runs
    // over the 128-byte nodes
    TraceConfig c = cfg;
    if (c.chunk == kChunkQuant && !p.qnodes4)
        c.chunk = find_variant(TraceConfig{c.width, c.block, kChunkLdsTop}) ? kChunkLdsTop : 1u;
    const VariantEntry* v = find_variant(c);
    if (!v || p.max_depth > kMaxDepth) return hipErrorInvalidValue;
    const bool batch = p.num_scenes > 0;
    const uint32_t gen_blocks = grid_for(s.paths, 256, 16384);
    // at most the resident grid, a multiple of the queue segments in waves
    // (one segment per wave)
    constexpr uint32_t kShadeGrain = kQueueSegments / (kShadeBlock / 64);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    const uint32_t shade_cap = std::max(kShadeGrain, shade_resident_blocks(dev) / kShadeGrain * kShadeGrain);
    const uint32_t shade_blocks = (grid_for(s.paths, kShadeBlock, shade_cap) + kShadeGrain - 1) / kShadeGrain * kShadeGrain;
    auto before = [&](uint32_t k) { if (tm && tm->before) tm->before(tm->user, k, stream); };
    auto after = [&](uint32_t k) { if (tm && tm->after) tm->after(tm->user, k, stream); };
    before(WAVE_K_RAYGEN);
    if (batch) hipLaunchKernelGGL(k_raygen<true>, dim3(gen_blocks), dim3(256), 0, stream, p, s);
    else hipLaunchKernelGGL(k_raygen<false>, dim3(gen_blocks), dim3(256), 0, stream, p, s);
    after(WAVE_K_RAYGEN);
    // shadow kernels run whenever next-event estimation is on (an empty
    // shadow queue makes them exit at once)
    const bool shadows = p.integrator != 0u;
    for (uint32_t b = 0; b < p.max_depth; ++b) {
        before(WAVE_K_EXTEND);
        hipError_t e = (batch ? v->extendB : v->extend1)(p, s, b, trace_blocks, stream);
        after(WAVE_K_EXTEND);
        if (e != hipSuccess) return e;
        before(WAVE_K_SHADE);
        if (b == 0) {
            if (batch) hipLaunchKernelGGL((k_shade<kShadeBlock, true, true>), dim3(shade_blocks), dim3(kShadeBlock), 0, stream, p, s, b);
            else hipLaunchKernelGGL((k_shade<kShadeBlock, false, true>), dim3(shade_blocks), dim3(kShadeBlock), 0, stream, p, s, b);
        } else {
            if (batch) hipLaunchKernelGGL((k_shade<kShadeBlock, true, false>), dim3(shade_blocks), dim3(kShadeBlock), 0, stream, p, s, b);
            else hipLaunchKernelGGL((k_shade<kShadeBlock, false, false>), dim3(shade_blocks), dim3(kShadeBlock), 0, stream, p, s, b);
        }
        after(WAVE_K_SHADE);
        if (shadows) {
            before(WAVE_K_SHADOW);
            e = (batch ? v->shadowB : v->shadow1)(p, s, b, trace_blocks, stream);
            after(WAVE_K_SHADOW);
            if (e != hipSuccess) return e;
        }
    }
    if (p.rec_hit) return hipGetLastError();   // a recorded pass for the adjoint: no accumulation
    before(WAVE_K_ACCUMULATE);
    const hipError_t e = launch_pass_accumulate(p, s, stream);
    after(WAVE_K_ACCUMULATE);
    return e;
}

hipError_t launch_pass_accumulate(const PassParams& p, const WaveState& s, hipStream_t stream) {
    // tiles of up to 64 pixels (fewer when a pixel's staged row is longer)
    const uint32_t chunk = p.sample_count < kAccRow ? p.sample_count : kAccRow;
    const uint32_t tile = std::min<uint32_t>(kAccTile, kAccEntries / (chunk + 1u));
    const uint32_t acc_blocks = grid_for((uint64_t(p.total_pixels) + tile - 1) / tile, 1, 16384);
    if (p.num_scenes > 0) hipLaunchKernelGGL(k_accumulate<true>, dim3(acc_blocks), dim3(256), 0, stream, p, s);
    else hipLaunchKernelGGL(k_accumulate<false>, dim3(acc_blocks), dim3(256), 0, stream, p, s);
    if (p.dyn_sample_begin)
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, stream, const_cast<uint32_t*>(p.dyn_sample_begin),
                           p.sample_count);
    return hipGetLastError();
}

}  // namespace quasi_hip
