// kernels_paths.hip -- the single-launch form of a sample pass.
//
// One persistent grid runs every path of the pass, of one scene or of a
// batch of scenes, from its camera ray to its end: closest hit, shading (the
// reference bounce of plugin.mm:223-276, or the NEE vertex), the NEE shadow
// ray and the next segment, with no drain between bounces.  Paths are taken
// from one shared path queue (the pass's path ids in kQueueSegments
// segments, a cursor per segment); a lane whose path has ended takes the next
// path as soon as kRefillIdle lanes of its wave are idle, so a wave's lanes
// are refilled with live rays of any bounce, of any scene, instead of waiting
// for the slowest ray of a bounce.  The traversal is the wavefront engine's
// (BVH4 or BVH2, while-while with postponed leaves, sign-selected slabs, the
// same visit order and tie rule); each lane switches between closest-hit
// segments and any-hit shadow rays.
//
// Path state lives in the wavefront engine's per-path arrays (throughput,
// radiance, RNG state, first-hit records; the continuation ray and the
// shadow contribution of a path whose shadow ray is in flight), so registers
// hold only the ray being traced and its traversal.  The pass ends with the
// wavefront engine's per-pixel accumulation in sample order, and every path
// computes what the wavefront kernels compute for it, in the same order, so
// the images equal the wavefront engine's bit for bit.
#include "kernels.hpp"
#include "wave_common.hpp"

namespace quasi_hip {

namespace {

// Segment lengths and cursors of the path queue: the path ids dealt to the
// segments in chunks, as the wavefront engine's camera bounce does
// (queue_slot_of_path, common.hpp).
__global__ void k_path_queue(WaveState S) {
    const uint32_t seg = threadIdx.x;
    if (seg >= kQueueSegments) return;
    S.counts[queue_counter(0, seg)] = queue_segment_paths(S.paths, seg, S.qchunk);
    S.fetch[queue_counter(0, seg)] = 0u;
}

template <int WIDTH, int BLOCK, bool BATCH, bool OFF32, bool QUANT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_paths(PassParams P, WaveState S) {
    constexpr bool kQuant = QUANT && WIDTH == 4;
    // 128-B BVH4 / 64-B quantised BVH4 / 64-B BVH2 nodes
    constexpr uint32_t kNodeShift = WIDTH == 4 && !kQuant ? 7u : 6u;
    __shared__ int s_stack[kShortStack * BLOCK];
    ShortStack<kShortStack, BLOCK> st{s_stack + threadIdx.x, P.gstack + (blockIdx.x * BLOCK + threadIdx.x),
                                      gridDim.x * BLOCK, 0};
    float4* thr_a = static_cast<float4*>(S.thr);
    float4* rad_a = static_cast<float4*>(S.radiance);
    float4* a0 = static_cast<float4*>(S.aov0);
    float4* a1 = static_cast<float4*>(S.aov1);
    // per path, while its shadow ray is traced: the continuation ray and the
    // shadow ray's contribution
    float4* pend_o = static_cast<float4*>(S.ray_o[0]);
    float4* pend_d = static_cast<float4*>(S.ray_d[0]);
    float4* pend_c = static_cast<float4*>(S.sh_c);
    SegmentFetch f;
    f.init(S.counts + queue_counter(0, 0), S.fetch + queue_counter(0, 0), S.seg_cap, wave_global());
    const uint32_t s0 = pass_sample_begin(P);

    Ray r{mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 1.f)};
    f3 inv = mk(0.f, 0.f, 0.f);
    const float tmin = kTMin;
    float best_t = kTFar;
    int best_tri = -1;
    uint32_t best_id = 0u;
    int node = kDone;
    int leaf = 0;   // postponed leaf reference (negative; 0 = none)
    bool busy = false;
    bool shadow = false;    // the lane traces its path's shadow ray
    bool pending = false;   // ... after which the path continues with its pending ray
    uint32_t path = 0, sid = 0, bounce = 0;
    uint32_t n_cam = 0, n_ext = 0, n_shadow = 0;
    uint32_t sel_x = 0, sel_y = 32, sel_z = 64;
    const char* const nbase = static_cast<const char*>(kQuant ? P.qnodes4 : (WIDTH == 4 ? P.nodes4 : P.nodes));
    const char* const tbase = static_cast<const char*>(P.tris);
    uint32_t nofs = 0, tofs = 0;
    const char* nodes = BATCH ? nullptr : nbase;
    const char* tris = BATCH ? nullptr : tbase;
    bool scene_empty = false;

    // a new segment from the lane's ray r (t in [kTMin, tmax])
    auto begin_segment = [&](float tmax) {
        inv = mk(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
        sel_x = inv.x < 0.f ? 16u : 0u;
        sel_y = inv.y < 0.f ? 48u : 32u;
        sel_z = inv.z < 0.f ? 80u : 64u;
        best_t = tmax;
        best_tri = -1;
        best_id = 0u;
        st.sp = 0;
        node = scene_empty ? kDone : 0;
        leaf = 0;
    };

    for (;;) {
        // refill idle lanes with new paths (wave-uniform: every lane reaches the ballots)
        const unsigned long long idle = __ballot(!busy);
        const unsigned long long act = __ballot(busy);
        unsigned long long want = f.live() && (act == 0ull || uint32_t(__popcll(idle)) >= kRefillIdle) ? idle : 0ull;
        while (want) {
            uint32_t first = 0;
            const uint32_t got = f.take(uint32_t(__popcll(want)), first);
            if (got == 0u) break;   // every path of the pass taken
            const uint32_t lb = lanes_below(want);
            const bool mine = ((want >> (threadIdx.x & 63u)) & 1ull) && lb < got;
            want &= ~__ballot(mine);
            if (mine) {
                path = queue_path_of_slot(first + lb, S.seg_cap, S.qchunk);
                sid = BATCH ? scene_of_path(P, path) : 0u;
                const SceneDesc& sc = scene_at<BATCH>(P, sid);
                if (BATCH) {
                    const char* sn = static_cast<const char*>(kQuant ? sc.qnodes4 : scene_nodes<WIDTH, true>(P, sid));
                    const char* stt = static_cast<const char*>(sc.tris);
                    if (OFF32) {
                        nofs = uint32_t(sn - nbase);
                        tofs = uint32_t(stt - tbase);
                    } else {
                        nodes = sn;
                        tris = stt;
                    }
                }
                scene_empty = sc.num_nodes == 0u;
                // the camera ray of path p_local = ((y - row_begin)*width + x)*samples + s
                uint32_t pix, s;
                path_pixel_sample(path - sc.path_begin, P.sample_count, pix, s);
                const uint32_t yl = pix / sc.width;
                const uint32_t x = pix - yl * sc.width;
                const uint32_t y = sc.row_begin + yl;
                uint32_t rng = path_seed(P.rng_mode, x, y, sc.width, sc.height, s0 + s, sc.seed);
                r = camera_ray(sc.cam, sc.width, sc.height, x, y, rng);
                S.rng[path] = rng;
                bounce = 0;
                shadow = false;
                pending = false;
                begin_segment(kTFar);
                busy = true;
                ++n_cam;
            }
        }
        if (__ballot(busy) == 0ull) break;   // the pass is done
        // descend through internal nodes (postponed leaves, as k_trace_dynamic)
        for (;;) {
            const bool desc = busy && node >= 0;
            if (leave_node_loop(desc && leaf == 0, busy && leaf != 0)) break;
            if (!desc) continue;
            const uint32_t off = nofs + (uint32_t(node) << kNodeShift);
            const char* rec = nodes + (OFF32 ? size_t(0) : (size_t(uint32_t(node)) << kNodeShift));
            auto ld = [&](uint32_t o) -> float4 {
                return OFF32 ? *reinterpret_cast<const float4*>(nbase + (off + o))
                             : *reinterpret_cast<const float4*>(rec + o);
            };
            if (kQuant) {
                node = quant_node_step<BLOCK>(ld(0u), ld(16u), ld(32u), ld(48u), r, inv, tmin, best_t, st);
            } else if (WIDTH == 4) {
                const float4 ex = ld(sel_x), fx = ld(sel_x ^ 16u);
                const float4 ey = ld(sel_y), fy = ld(sel_y ^ 16u);
                const float4 ez = ld(sel_z), fz = ld(sel_z ^ 16u);
                const float4 chf = ld(96u);
                float t0n, t0f, t1n, t1f, t2n, t2f, t3n, t3f;
                slab_sel(ex.x, fx.x, ey.x, fy.x, ez.x, fz.x, r, inv, tmin, best_t, t0n, t0f);
                slab_sel(ex.y, fx.y, ey.y, fy.y, ez.y, fz.y, r, inv, tmin, best_t, t1n, t1f);
                slab_sel(ex.z, fx.z, ey.z, fy.z, ez.z, fz.z, r, inv, tmin, best_t, t2n, t2f);
                slab_sel(ex.w, fx.w, ey.w, fy.w, ez.w, fz.w, r, inv, tmin, best_t, t3n, t3f);
                node = visit_children4<BLOCK>(t0n <= t0f ? t0n : kMissKey, t1n <= t1f ? t1n : kMissKey,
                                              t2n <= t2f ? t2n : kMissKey, t3n <= t3f ? t3n : kMissKey,
                                              __float_as_int(chf.x), __float_as_int(chf.y), __float_as_int(chf.z),
                                              __float_as_int(chf.w), st);
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
        // the lanes holding a leaf test its triangles
        if (busy && leaf != 0) {
            const uint32_t v = ~uint32_t(leaf);
            const int first = int(v >> 2);
            const int count = int(v & 3u) + 1;
            bool found = false;
            uint32_t toff = tofs + uint32_t(first) * 48u;
            const char* trec = tris + (OFF32 ? size_t(0) : size_t(uint32_t(first)) * 48u);
            for (int i = first; i < first + count; ++i, toff += 48u, trec += OFF32 ? 0 : 48) {
                float t;
                uint32_t id;
                const float4* tr = reinterpret_cast<const float4*>(OFF32 ? tbase + toff : trec);
                if (!hit_triangle_rec(tr, r, tmin, best_t, t, id)) continue;
                if (shadow) { best_t = t; best_tri = i; best_id = id; found = true; break; }
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
        // a finished segment: shade the hit, or end the shadow ray; then the
        // path's next segment, or the lane is idle
        if (busy && node == kDone && leaf == 0) {
            const SceneDesc& sc = scene_at<BATCH>(P, sid);
            if (shadow) {
                if (best_tri < 0) add_shadow(rad_a, pend_c, path, path);
                if (pending) {
                    const float4 o = pend_o[path], d = pend_d[path];
                    r = Ray{xyz(o), xyz(d)};
                    ++bounce;
                    shadow = false;
                    begin_segment(kTFar);
                    ++n_ext;
                } else {
                    busy = false;
                }
            } else if (best_tri >= 0) {
                f3 thr = mk(1.f, 1.f, 1.f), L = mk(0.f, 0.f, 0.f);
                if (bounce != 0u) {
                    thr = xyz(thr_a[path]);
                    L = xyz(rad_a[path]);
                }
                uint32_t rng = S.rng[path];
                FirstHit fh;
                ShadeOut so;
                // the first-hit record is formed at every vertex (a constant
                // pointer keeps it in registers) and stored at bounce 0 only
                shade_hit(P, static_cast<const float4*>(sc.tris), static_cast<const float4*>(sc.mats), sc.lights,
                          sc.num_lights, bounce, r, best_t, best_tri, thr, L, rng, &fh, so);
                if (bounce == 0u) {
                    a0[path] = make_float4(fh.albedo.x, fh.albedo.y, fh.albedo.z, fh.t);
                    a1[path] = make_float4(fh.normal.x, fh.normal.y, fh.normal.z, __uint_as_float(fh.id));
                }
                thr_a[path] = make_float4(thr.x, thr.y, thr.z, 0.f);
                rad_a[path] = make_float4(L.x, L.y, L.z, 0.f);
                S.rng[path] = rng;
                if (so.shadow) {
                    pend_c[path] = make_float4(so.scontrib.x, so.scontrib.y, so.scontrib.z, 0.f);
                    if (so.cont) {
                        pend_o[path] = make_float4(so.next.o.x, so.next.o.y, so.next.o.z, 0.f);
                        pend_d[path] = make_float4(so.next.d.x, so.next.d.y, so.next.d.z, 0.f);
                    }
                    pending = so.cont;
                    shadow = true;
                    r = so.sray;
                    begin_segment(so.stmax);
                    ++n_shadow;
                } else if (so.cont) {
                    r = so.next;
                    ++bounce;
                    begin_segment(kTFar);
                    ++n_ext;
                } else {
                    busy = false;
                }
            } else {
                if (bounce == 0u) {
                    // a camera ray that missed: no radiance, empty first-hit records
                    rad_a[path] = make_float4(0.f, 0.f, 0.f, 0.f);
                    a0[path] = make_float4(0.f, 0.f, 0.f, 0.f);
                    a1[path] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kMissId));
                }
                busy = false;
            }
        }
    }
    // one atomic per wave and counter
    n_cam = wave_sum(n_cam);
    n_ext = wave_sum(n_ext);
    n_shadow = wave_sum(n_shadow);
    if ((threadIdx.x & 63u) == 0u) {
        if (n_cam) atomicAdd(&P.counters->rays_camera, (unsigned long long)n_cam);
        if (n_ext) atomicAdd(&P.counters->rays_extension, (unsigned long long)n_ext);
        if (n_shadow) atomicAdd(&P.counters->rays_shadow, (unsigned long long)n_shadow);
    }
}

template <class K>
int occupancy_of(K kernel, int block) {
    int a = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, kernel, block, 0) != hipSuccess) a = 1;
    return a;
}

using PathsFn = void (*)(const PassParams&, const WaveState&, uint32_t, hipStream_t);

template <int WIDTH, int BLOCK>
struct PathsVariant {
    static bool off32(const PassParams& p) {
        const uint64_t node_bytes = uint64_t(WIDTH == 4 ? p.num_nodes4 : p.num_nodes) * (WIDTH == 4 ? 128u : 64u);
        return node_bytes < (1ull << 32) && uint64_t(p.num_tris) * 48u < (1ull << 32);
    }
    // the quantised BVH4 nodes whenever the pass has them (bvh_quant.hpp)
    template <bool BATCH, bool QUANT>
    static void launch_q(const PassParams& p, const WaveState& s, uint32_t blocks, hipStream_t st) {
        if (off32(p))
            hipLaunchKernelGGL((k_paths<WIDTH, BLOCK, BATCH, true, QUANT>), dim3(blocks), dim3(BLOCK), 0, st, p, s);
        else
            hipLaunchKernelGGL((k_paths<WIDTH, BLOCK, BATCH, false, QUANT>), dim3(blocks), dim3(BLOCK), 0, st, p, s);
    }
    template <bool BATCH>
    static void launch(const PassParams& p, const WaveState& s, uint32_t blocks, hipStream_t st) {
        if (WIDTH == 4 && p.qnodes4) launch_q<BATCH, true>(p, s, blocks, st);
        else launch_q<BATCH, false>(p, s, blocks, st);
    }
    static int occupancy() {
        const int a = occupancy_of(k_paths<WIDTH, BLOCK, false, true, WIDTH == 4>, BLOCK);
        const int b = occupancy_of(k_paths<WIDTH, BLOCK, true, true, WIDTH == 4>, BLOCK);
        return a < b ? a : b;
    }
};

struct PathsEntry {
    uint32_t width, block;
    PathsFn one, batch;
    int (*occupancy)();
};

#define QH_PATHS(W, B) \
    {W, B, &PathsVariant<W, B>::launch<false>, &PathsVariant<W, B>::launch<true>, &PathsVariant<W, B>::occupancy}
const PathsEntry kPaths[] = {QH_PATHS(2, 128), QH_PATHS(2, 256), QH_PATHS(4, 128), QH_PATHS(4, 256)};
#undef QH_PATHS

const PathsEntry* find_paths(const TraceConfig& c) {
    for (const PathsEntry& e : kPaths)
        if (e.width == c.width && e.block == c.block) return &e;
    return nullptr;
}

}  // namespace

bool paths_config_supported(const TraceConfig& c) { return find_paths(c) != nullptr; }

uint32_t paths_pass_blocks(int device, const TraceConfig& c) {
    const PathsEntry* e = find_paths(c);
    if (!e) return 0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    int per_cu = e->occupancy();
    if (per_cu <= 0) per_cu = 1;
    // never more threads than the stack-overflow area is sized for
    const int max_per_cu = int(kMaxThreadsPerCU / c.block);
    if (per_cu > max_per_cu) per_cu = max_per_cu;
    return uint32_t(cus * per_cu);
}

hipError_t launch_paths_pass(const PassParams& p, const WaveState& s, const TraceConfig& cfg, uint32_t blocks,
                             hipStream_t stream) {
    if (s.paths == 0) return hipSuccess;
    const PathsEntry* e = find_paths(cfg);
    if (!e || blocks == 0 || p.max_depth == 0 || p.max_depth > kMaxDepth || p.rec_hit) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_path_queue, dim3(1), dim3(kQueueSegments), 0, stream, s);
    (p.num_scenes > 0 ? e->batch : e->one)(p, s, blocks, stream);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    return launch_pass_accumulate(p, s, stream);
}

}  // namespace quasi_hip
