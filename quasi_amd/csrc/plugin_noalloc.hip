// plugin_noalloc.hip -- Q_plugin_render_noalloc: the wavefront render of
// Q_plugin_render into caller-owned device sums, with every piece of working
// memory (path state, queues, queue counters, traversal-stack overflow, ray
// counters) carved out of one fixed device array, so that the call allocates
// nothing.  For callers that cannot allocate device memory: code under HIP
// graph capture, serving loops, harnesses that forbid it.
//
// The array lives in this file's code object.  Its host-side name is not a
// device address: the address of each device's copy is taken once per device
// with hipGetSymbolAddress (at Q_plugin_create, with that device current) and
// only that address is ever handed to kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "plugin_state.hpp"

#define Q_EXPORT __attribute__((visibility("default")))

namespace quasi_hip {
namespace host {

namespace {

// One device-wide working area, shared by every call on that device: calls
// that overlap on one device race.
__device__ __align__(256) unsigned char scratch[Q_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Bytes of the queue counters (extension lengths, shadow lengths, work
// cursors): kCounterStride words per segment and group.
constexpr size_t kCounterGroup = size_t(kQueueSegments) * kCounterStride * 4;
constexpr size_t kCountersBytes = kCounterGroup * ((kMaxDepth + 1) + kMaxDepth + 2 * kMaxDepth);
// Path state per path (throughput, radiance, two first-hit records, RNG
// state) and per queue slot (two ray buffers of origin + direction, the
// hit, the shadow ray's origin, direction and contribution).  Throughput and
// RNG stay in the per-path arrays (WaveState::tq null), which holds 32 B per
// path less than the queue-ordered copies the allocating entry points use.
constexpr size_t kPerPath = 16 * 4 + 4;
constexpr size_t kPerSlot = 2 * 32 + 8 + 3 * 16;

// Where each part of a call's working memory sits in the area.
struct Layout {
    size_t counters, rays, gstack, path_state;
    uint32_t trace_blocks;
    uint32_t max_paths;       // paths per pass the path-state part holds
};

Layout layout_for(const plugin_state* st, uint32_t gstack_entries) {
    Layout L{};
    L.counters = 0;
    L.rays = align_up(kCountersBytes);
    L.gstack = align_up(L.rays + sizeof(Counters));
    // the overflow stacks: at most a quarter of the area; a grid that would
    // need more runs with fewer persistent blocks
    const size_t per_block = size_t(st->trace_cfg.block) * gstack_entries * 4;
    uint32_t blocks = st->trace_blocks;
    if (per_block) {
        const size_t cap = Q_NOALLOC_SCRATCH_BYTES / 4 / per_block;
        blocks = uint32_t(std::min<size_t>(blocks, cap));
    }
    L.trace_blocks = blocks;
    L.path_state = align_up(L.gstack + per_block * blocks);
    const size_t left = Q_NOALLOC_SCRATCH_BYTES > L.path_state ? Q_NOALLOC_SCRATCH_BYTES - L.path_state : 0;
    // 5 per-path arrays and 8 per-slot arrays (queue_slots(paths, kQueueChunkScene):
    // at most paths + kQueueSegments * kQueueChunkScene slots), each starting on a kAlign boundary
    const size_t fixed = 13 * kAlign + kPerSlot * size_t(kQueueSegments) * kQueueChunkScene;
    L.max_paths = left > fixed ? uint32_t(std::min<size_t>((left - fixed) / (kPerPath + kPerSlot), 0x7FFFFFFFu)) : 0u;
    return L;
}

// The path state of `paths` paths in the area; `end` = the offset one past
// its last byte.
WaveState carve(unsigned char* base, const Layout& L, uint32_t paths, size_t& end) {
    WaveState s{};
    const size_t n = paths, q = size_t(queue_slots(paths, kQueueChunkScene));
    size_t off = L.path_state;
    auto take = [&](size_t bytes) {
        void* p = base + off;
        off = align_up(off + bytes);
        return p;
    };
    for (int k = 0; k < 2; ++k) {
        s.ray_o[k] = take(16 * q);
        s.ray_d[k] = take(16 * q);
    }
    s.hit = take(8 * q);
    s.thr = take(16 * n);
    s.radiance = take(16 * n);
    s.rng = static_cast<uint32_t*>(take(4 * n));
    s.aov0 = take(16 * n);
    s.aov1 = take(16 * n);
    s.sh_o = take(16 * q);
    s.sh_d = take(16 * q);
    s.sh_c = take(16 * q);
    end = off;
    unsigned char* c = base + L.counters;
    s.counts = reinterpret_cast<uint32_t*>(c);
    s.shadow_counts = reinterpret_cast<uint32_t*>(c + kCounterGroup * (kMaxDepth + 1));
    s.fetch = reinterpret_cast<uint32_t*>(c + kCounterGroup * (2 * kMaxDepth + 1));
    s.paths = paths;
    s.qchunk = kQueueChunkScene;
    s.seg_cap = queue_seg_cap(paths, s.qchunk);
    return s;
}

}  // namespace

bool noalloc_init(plugin_state* st) {
    if (st->device < 0 || st->device >= kMaxDevices) return false;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[st->device]) {
        void* p = nullptr;
        // the caller (Q_plugin_create) has made st->device current
        if (!hip_ok(st, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "no-alloc scratch address")) return false;
        g_scratch_ptr[st->device] = p;
    }
    st->noalloc_scratch = g_scratch_ptr[st->device];
    return true;
}

}  // namespace host
}  // namespace quasi_hip

using namespace quasi_hip;
using namespace quasi_hip::host;

extern "C" {

Q_EXPORT int32_t Q_plugin_render_noalloc(Q_plugin_handle* handle, const Q_camera* camera, uint32_t width,
                                         uint32_t height, uint32_t sample_begin, uint32_t spp, float* sums,
                                         void* stream) {
    if (!handle) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = reinterpret_cast<plugin_state*>(handle);
    try {
        if (!camera || !sums || width == 0 || height == 0 || spp == 0) {
            log_msg(st, "render_noalloc: camera, sums, a non-empty image and spp >= 1 are required");
            return Q_EXT_ERR_ARGUMENT;
        }
        if (!st->scene_ok) {
            log_msg(st, "render_noalloc: no scene on the device");
            return Q_EXT_ERR_STATE;
        }
        if (!st->noalloc_scratch) {
            log_msg(st, "render_noalloc: the device's working area was not set up");
            return Q_EXT_ERR_STATE;
        }
        const uint64_t npix = uint64_t(width) * height;
        if (npix > 0xFFFFFFFFull / 2) {
            log_msg(st, "render_noalloc: image of %llu pixels is too large", (unsigned long long)npix);
            return Q_EXT_ERR_ARGUMENT;
        }
        const Q_render_params& rp = st->params;
        const Layout L = layout_for(st, st->gstack_entries);
        if (L.trace_blocks == 0 || L.max_paths < npix) {
            log_msg(st, "render_noalloc: an image of %llu pixels needs %llu paths per pass; the %llu-byte working "
                        "area holds %u",
                    (unsigned long long)npix, (unsigned long long)npix, (unsigned long long)Q_NOALLOC_SCRATCH_BYTES,
                    L.max_paths);
            return Q_EXT_ERR_ARGUMENT;
        }
        uint32_t per_pass = rp.samples_per_pass ? rp.samples_per_pass : uint32_t(L.max_paths / npix);
        per_pass = std::min(per_pass, spp);
        if (uint64_t(per_pass) * npix > L.max_paths) {
            log_msg(st, "render_noalloc: a pass of %u samples x %llu pixels exceeds the %u paths the working area "
                        "holds",
                    per_pass, (unsigned long long)npix, L.max_paths);
            return Q_EXT_ERR_ARGUMENT;
        }
        if (rp.max_depth > kMaxDepth) return Q_EXT_ERR_ARGUMENT;
        (void)hipSetDevice(st->device);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : st->stream;
        unsigned char* base = static_cast<unsigned char*>(st->noalloc_scratch);
        Counters* counters = reinterpret_cast<Counters*>(base + L.rays);
        if (!hip_ok(st, launch_fill_u32(counters, sizeof(Counters) / 4, 0u, s), "render_noalloc counter clear"))
            return Q_EXT_ERR_DEVICE;

        PassParams p{};
        p.tris = st->tris.ptr;
        p.mats = st->mats.ptr;
        p.nodes = st->nodes.ptr;
        p.lights = static_cast<const LightEntry*>(st->lights.ptr);
        float4* planes = reinterpret_cast<float4*>(sums);
        p.accum_beauty = planes;
        p.accum_albedo = planes + npix;
        p.accum_normal = planes + 2 * npix;
        p.accum_depth = planes + 3 * npix;
        p.ids = nullptr;
        p.record_ids = 0;
        p.counters = counters;
        p.cam = make_camera(*camera, width, height);
        p.width = width;
        p.height = height;
        p.max_depth = rp.max_depth;
        p.rr_start = rp.rr_start;
        p.seed = rp.seed;
        p.rng_mode = rp.rng_mode;
        p.integrator = rp.integrator;
        p.num_lights = st->num_lights;
        p.num_nodes = st->num_nodes;
        p.num_tris = st->num_prims;
        p.nodes4 = st->nodes4.ptr;
        p.qnodes4 = st->quant_ok ? st->qnodes4.ptr : nullptr;
        p.num_nodes4 = st->num_nodes4;
        p.gstack = reinterpret_cast<int*>(base + L.gstack);
        p.gstack_entries = st->gstack_entries;
        SceneDesc& sd = p.scene0;
        sd.tris = p.tris;
        sd.mats = p.mats;
        sd.nodes = p.nodes;
        sd.nodes4 = p.nodes4;
        sd.qnodes4 = p.qnodes4;
        sd.lights = p.lights;
        sd.num_nodes = p.num_nodes;
        sd.num_lights = p.num_lights;
        sd.cam = p.cam;
        sd.width = width;
        sd.height = height;
        sd.seed = rp.seed;
        sd.num_tris = st->num_prims;
        p.num_scenes = 0;
        p.total_pixels = uint32_t(npix);
        for (uint32_t done = 0; done < spp; done += per_pass) {
            const uint32_t n = std::min(per_pass, spp - done);
            p.sample_begin = sample_begin + done;
            p.sample_count = n;
            size_t end = 0;
            const WaveState ws = carve(base, L, uint32_t(n * npix), end);
            if (end > Q_NOALLOC_SCRATCH_BYTES) {
                log_msg(st, "render_noalloc: path state of %llu paths overruns the working area",
                        (unsigned long long)(n * npix));
                return Q_EXT_ERR_ARGUMENT;
            }
            if (!hip_ok(st, launch_wavefront_pass(p, ws, st->trace_cfg, L.trace_blocks, s, nullptr),
                        "render_noalloc launch"))
                return Q_EXT_ERR_DEVICE;
        }
        return Q_EXT_OK;
    } catch (...) {
        log_msg(st, "render_noalloc: exception");
        return Q_EXT_ERR_STATE;
    }
}

}  // extern "C"
