// plugin_batch.cpp -- batched scenes, captured pass graphs and the autotuner
// (extension exports of include/quasi/quasi_hip_ext.h; no reference
// counterpart: the reference renders its one hard-coded scene,
// backends/metal/plugin.mm:584).
//
// A batch concatenates every scene's triangles, materials, BVH nodes and
// lights into one set of device arrays and keeps a device table of
// SceneDesc entries pointing into them.  One sample pass covers all scenes
// with shared ray queues (kernels_wave.hip, BATCH = true); the pass's launch
// sequence is captured once into a hipGraph and replayed per pass, with the
// first sample index read from device memory and advanced by the graph's last
// node, so the host enqueues one graph launch per pass.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "plugin_state.hpp"

#define Q_EXPORT __attribute__((visibility("default")))

namespace quasi_hip {
namespace host {

void free_batch(plugin_state* st) {
    BatchState& b = st->batch;
    if (b.graph_exec) (void)hipGraphExecDestroy(b.graph_exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.graph_exec = nullptr;
    b.graph = nullptr;
    b.graph_key = 0;
    for (auto* x : {&b.tris, &b.mats, &b.nodes, &b.nodes4, &b.qnodes4, &b.lights, &b.table, &b.ids, &b.resolved,
                    &b.sample_begin})
        free_buffer(*x);
    b.quant_ok = false;
    for (auto& a : b.accum) free_buffer(a);
    b.descs.clear();
    b.active = false;
    b.num_scenes = b.total_pixels = b.total_tris = b.total_prims = b.sample_count = 0;
}

namespace {

plugin_state* as_state(Q_plugin_handle* h) { return reinterpret_cast<plugin_state*>(h); }

bool clear_batch(plugin_state* st, hipStream_t s) {
    BatchState& b = st->batch;
    for (auto& a : b.accum)
        if (!hip_ok(st, launch_fill_u32(a.ptr, 4 * size_t(b.total_pixels), 0u, s), "batch clear")) return false;
    if (!hip_ok(st, launch_fill_u32(b.ids.ptr, b.total_pixels, kMissId, s), "batch id clear")) return false;
    b.sample_count = 0;
    return true;
}

// Samples per batched pass: the explicit setting, or (0) the largest power of
// two that divides spp and keeps a pass within kAutoPassPaths paths.
uint32_t batch_samples_per_pass(const plugin_state* st, uint32_t spp) {
    if (st->params.samples_per_pass) return st->params.samples_per_pass;
    const uint64_t cap = std::max<uint64_t>(1, kAutoPassPaths / std::max<uint32_t>(1u, st->batch.total_pixels));
    uint32_t s = 1;
    while (uint64_t(s) * 2 <= cap && spp % (s * 2) == 0) s *= 2;
    return s;
}

// FNV-1a over the values a captured pass depends on.
struct KeyHash {
    uint64_t h = 1469598103934665603ull;
    void add(uint64_t v) {
        for (int i = 0; i < 8; ++i) { h ^= (v >> (8 * i)) & 0xFF; h *= 1099511628211ull; }
    }
};

// The pass parameters of the current batch for passes of `spp_pass` samples.
PassParams batch_params(plugin_state* st, uint32_t spp_pass) {
    BatchState& b = st->batch;
    const Q_render_params& rp = st->params;
    PassParams p{};
    p.accum_beauty = b.accum[Q_AOV_BEAUTY].ptr;
    p.accum_albedo = b.accum[Q_AOV_ALBEDO].ptr;
    p.accum_normal = b.accum[Q_AOV_NORMAL].ptr;
    p.accum_depth = b.accum[Q_AOV_DEPTH].ptr;
    p.ids = static_cast<uint32_t*>(b.ids.ptr);
    p.counters = static_cast<Counters*>(st->counters.ptr);
    p.sample_begin = 0;
    p.sample_count = spp_pass;
    p.max_depth = rp.max_depth;
    p.rr_start = rp.rr_start;
    p.seed = rp.seed;
    p.rng_mode = rp.rng_mode;
    p.integrator = rp.integrator;
    p.record_ids = (rp.flags & Q_FLAG_RECORD_IDS) ? 1u : 0u;
    p.id_sample = rp.id_sample;
    p.gstack = static_cast<int*>(st->gstack.ptr);
    p.gstack_entries = b.gstack_entries;
    // the concatenated arrays: every scene's pointers in the table lie inside them
    p.tris = b.tris.ptr;
    p.mats = b.mats.ptr;
    p.nodes = b.nodes.ptr;
    p.nodes4 = b.nodes4.ptr;
    p.qnodes4 = b.quant_ok ? b.qnodes4.ptr : nullptr;
    p.lights = static_cast<const LightEntry*>(b.lights.ptr);
    p.num_tris = b.total_prims;
    p.num_nodes = b.total_nodes;
    p.num_nodes4 = b.total_nodes4;
    p.scenes = static_cast<const SceneDesc*>(b.table.ptr);
    p.num_scenes = b.num_scenes;
    p.total_pixels = b.total_pixels;
    p.dyn_sample_begin = static_cast<const uint32_t*>(b.sample_begin.ptr);
    return p;
}

// Uploads the scene table with path offsets for passes of `spp_pass` samples
// and (re)captures the pass graph when anything it depends on changed.
bool prepare_batch_pass(plugin_state* st, uint32_t spp_pass) {
    BatchState& b = st->batch;
    const uint64_t paths = uint64_t(b.total_pixels) * spp_pass;
    if (paths == 0 || paths > 0x7FFFFFFFull) {
        log_msg(st, "render_batch: %llu paths per pass is out of range", (unsigned long long)paths);
        return false;
    }
    if (!ensure_wave(st, uint32_t(paths))) return false;
    KeyHash k;
    k.add(spp_pass);
    k.add(st->trace_cfg.width); k.add(st->trace_cfg.block); k.add(st->trace_cfg.chunk); k.add(st->trace_blocks);
    const Q_render_params& rp = st->params;
    const bool single_launch = rp.engine == Q_ENGINE_PERSISTENT;
    k.add(single_launch); k.add(st->paths_blocks);
    k.add(rp.max_depth); k.add(rp.rr_start); k.add(rp.rng_mode); k.add(rp.integrator);
    k.add(rp.flags & Q_FLAG_RECORD_IDS); k.add(rp.id_sample);
    k.add(reinterpret_cast<uintptr_t>(st->wave.ray_o[0].ptr)); k.add(st->wave.capacity);
    k.add(reinterpret_cast<uintptr_t>(st->gstack.ptr));
    k.add(reinterpret_cast<uintptr_t>(b.table.ptr));
    k.add(reinterpret_cast<uintptr_t>(b.qnodes4.ptr)); k.add(b.quant_ok);
    if (k.h == b.graph_key && b.graph_exec) return true;

    // scene table for this pass size
    for (uint32_t i = 0; i < b.num_scenes; ++i) b.descs[i].path_begin = b.descs[i].pix_begin * spp_pass;
    if (!hip_ok(st, hipMemcpy(b.table.ptr, b.descs.data(), b.descs.size() * sizeof(SceneDesc), hipMemcpyHostToDevice),
                "scene table upload"))
        return false;
    if (b.graph_exec) { (void)hipGraphExecDestroy(b.graph_exec); b.graph_exec = nullptr; }
    if (b.graph) { (void)hipGraphDestroy(b.graph); b.graph = nullptr; }
    if (!st->capture_stream &&
        !hip_ok(st, hipStreamCreateWithFlags(&st->capture_stream, hipStreamNonBlocking), "capture stream"))
        return false;
    const PassParams p = batch_params(st, spp_pass);
    const WaveState ws = wave_state(st->wave, uint32_t(paths), true);
    if (!hip_ok(st, hipStreamBeginCapture(st->capture_stream, hipStreamCaptureModeThreadLocal), "begin capture"))
        return false;
    // the batch renders with the single-launch pass under Q_ENGINE_PERSISTENT,
    // else with the wavefront launch sequence
    const hipError_t le =
        single_launch ? launch_paths_pass(p, ws, st->trace_cfg, st->paths_blocks, st->capture_stream)
                      : launch_wavefront_pass(p, ws, st->trace_cfg, st->trace_blocks, st->capture_stream, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(st->capture_stream, &g);
    if (!hip_ok(st, le, "captured pass launch") || !hip_ok(st, ce, "end capture")) {
        if (g) (void)hipGraphDestroy(g);
        return false;
    }
    hipGraphExec_t ex = nullptr;
    if (!hip_ok(st, hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "graph instantiate")) {
        (void)hipGraphDestroy(g);
        return false;
    }
    b.graph = g;
    b.graph_exec = ex;
    b.graph_key = k.h;
    return true;
}

int32_t render_batch(plugin_state* st, uint32_t spp, hipStream_t s) {
    BatchState& b = st->batch;
    if (!b.active) {
        log_msg(st, "render_batch: no batch set");
        return Q_EXT_ERR_STATE;
    }
    if (spp == 0) {
        log_msg(st, "render_batch: no samples");
        return Q_EXT_ERR_ARGUMENT;
    }
    const uint32_t per_pass = batch_samples_per_pass(st, spp);
    if (spp % per_pass) {
        log_msg(st, "render_batch: %u samples is not a multiple of %u samples per pass", spp, per_pass);
        return Q_EXT_ERR_ARGUMENT;
    }
    if (st->last_stream && st->last_stream != s) (void)hipStreamSynchronize(st->last_stream);
    st->last_stream = s;
    if (!prepare_batch_pass(st, per_pass)) return Q_EXT_ERR_DEVICE;
    if (!hip_ok(st, launch_fill_u32(b.sample_begin.ptr, 1, b.sample_count, s), "sample index")) return Q_EXT_ERR_DEVICE;
    for (uint32_t done = 0; done < spp; done += per_pass) {
        LaunchTimer lt(st, s, K_BATCH_PASS);
        if (!hip_ok(st, hipGraphLaunch(b.graph_exec, s), "graph launch")) return Q_EXT_ERR_DEVICE;
    }
    b.sample_count += spp;
    return Q_EXT_OK;
}

int32_t set_batch(plugin_state* st, const Q_batch_desc* desc) {
    const uint32_t n = desc->num_scenes;
    if (n == 0 || !desc->scenes) return Q_EXT_ERR_ARGUMENT;
    uint64_t total_pix = 0, total_tris = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const Q_batch_scene& sc = desc->scenes[i];
        const Q_scene_desc& m = sc.mesh;
        if (sc.width == 0 || sc.height == 0 || m.struct_size < sizeof(Q_scene_desc) ||
            (m.num_triangles && (!m.indices || !m.albedo || !m.emission)) || (m.num_vertices && !m.vertices)) {
            log_msg(st, "set_batch: scene %u is malformed", i);
            return Q_EXT_ERR_ARGUMENT;
        }
        total_pix += uint64_t(sc.width) * sc.height;
        total_tris += m.num_triangles;
    }
    if (total_pix > 0x3FFFFFFFull || total_tris >= (1ull << 29)) {
        log_msg(st, "set_batch: batch too large (%llu pixels, %llu triangles)", (unsigned long long)total_pix,
                (unsigned long long)total_tris);
        return Q_EXT_ERR_UNSUPPORTED;
    }
    (void)hipSetDevice(st->device);
    if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
    free_batch(st);
    auto t0 = std::chrono::steady_clock::now();

    // BVH builds: scenes the builder setting sends to the GPU are built on the
    // device one after another by this thread, the others meanwhile on host
    // threads, one scene per thread
    std::vector<SceneHost> hosts(n);
    struct DeviceGuard {   // device arrays of GPU-built scenes are freed once copied (or on failure)
        std::vector<SceneHost>& h;
        ~DeviceGuard() { for (SceneHost& x : h) x.device.release(); }
    } device_guard{hosts};
    const uint32_t builder = st->params.bvh_builder;
    auto on_gpu = [&](uint32_t i) {
        const uint32_t nt = desc->scenes[i].mesh.num_triangles;
        return builder == Q_BVH_BUILD_GPU || (builder == Q_BVH_BUILD_AUTO && nt >= kGpuBuildMinTriangles);
    };
    std::vector<uint32_t> host_list, gpu_list;
    for (uint32_t i = 0; i < n; ++i) (on_gpu(i) ? gpu_list : host_list).push_back(i);
    std::vector<int32_t> rcs(n, Q_EXT_OK);
    std::atomic<uint32_t> next{0};
    auto worker = [&] {
        for (;;) {
            const uint32_t k = next.fetch_add(1);
            if (k >= host_list.size()) break;
            const uint32_t i = host_list[k];
            const Q_scene_desc& m = desc->scenes[i].mesh;
            rcs[i] = prepare_scene(st, m.vertices, m.num_vertices, m.indices, m.num_triangles, m.albedo, m.emission,
                                   1, hosts[i]);
        }
    };
    const unsigned nth = std::max(1u, std::min({16u, std::thread::hardware_concurrency(), uint32_t(host_list.size())}));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth && !host_list.empty(); ++t) pool.emplace_back(worker);
    for (uint32_t i : gpu_list) {
        const Q_scene_desc& m = desc->scenes[i].mesh;
        rcs[i] = prepare_scene(st, m.vertices, m.num_vertices, m.indices, m.num_triangles, m.albedo, m.emission, 0,
                               hosts[i], builder);
    }
    for (auto& t : pool) t.join();
    for (uint32_t i = 0; i < n; ++i)
        if (rcs[i] != Q_EXT_OK) return rcs[i];
    // floats of each array, for host- and GPU-built scenes alike
    auto tri_f = [](const SceneHost& h) { return size_t(12) * h.num_prims; };
    auto mat_f = [](const SceneHost& h) { return size_t(8) * h.num_prims; };
    auto node_f = [](const SceneHost& h) { return size_t(16) * h.num_nodes; };
    auto node4_f = [](const SceneHost& h) { return size_t(32) * h.num_nodes4; };

    // concatenated arrays
    size_t tri_total = 0, mat_total = 0, node_total = 0, node4_total = 0, light_n = 0;
    uint32_t gentries = 0;
    for (const SceneHost& h : hosts) {
        tri_total += tri_f(h);
        mat_total += mat_f(h);
        node_total += node_f(h);
        node4_total += node4_f(h);
        light_n += h.lights.size();
        gentries = std::max(gentries, stack_overflow_entries(h.depth, h.depth4));
    }
    BatchState& b = st->batch;
    const size_t npix = size_t(total_pix);
    if (!ensure_buffer(st, b.tris, std::max<size_t>(16, tri_total * 4), "batch triangles") ||
        !ensure_buffer(st, b.mats, std::max<size_t>(16, mat_total * 4), "batch materials") ||
        !ensure_buffer(st, b.nodes, std::max<size_t>(64, node_total * 4), "batch BVH nodes") ||
        !ensure_buffer(st, b.nodes4, std::max<size_t>(128, node4_total * 4), "batch BVH4 nodes") ||
        !ensure_buffer(st, b.qnodes4, std::max<size_t>(64, node4_total * 2), "batch quantised BVH4 nodes") ||
        !ensure_buffer(st, b.lights, std::max<size_t>(16, light_n * sizeof(LightEntry)), "batch lights") ||
        !ensure_buffer(st, b.table, n * sizeof(SceneDesc), "scene table") ||
        !ensure_buffer(st, b.ids, 4 * npix, "batch ids") ||
        !ensure_buffer(st, b.sample_begin, 4, "batch sample index") ||
        !ensure_gstack(st, std::max(gentries, st->gstack_entries))) {
        free_batch(st);
        return Q_EXT_ERR_OUT_OF_MEMORY;
    }
    for (auto& a : b.accum)
        if (!ensure_buffer(st, a, 16 * npix, "batch accumulation")) { free_batch(st); return Q_EXT_ERR_OUT_OF_MEMORY; }
    b.descs.resize(n);
    size_t to = 0, mo = 0, no = 0, n4o = 0, lo = 0;
    uint32_t pix = 0;
    char* tb = static_cast<char*>(b.tris.ptr);
    char* mb = static_cast<char*>(b.mats.ptr);
    char* nb = static_cast<char*>(b.nodes.ptr);
    char* n4b = static_cast<char*>(b.nodes4.ptr);
    char* q4b = static_cast<char*>(b.qnodes4.ptr);   // 64 B per node: 2 bytes per float of the 128-B nodes
    LightEntry* lb = static_cast<LightEntry*>(b.lights.ptr);
    for (uint32_t i = 0; i < n; ++i) {
        const SceneHost& h = hosts[i];
        const Q_batch_scene& sc = desc->scenes[i];
        // a GPU-built scene's arrays are already on the device; the others come from host memory
        const bool dev = h.builder == Q_BVH_BUILD_GPU;
        auto put = [&](char* dst, const void* host_src, const void* dev_src, size_t bytes) {
            if (bytes == 0) return true;
            return hip_ok(st, hipMemcpy(dst, dev ? dev_src : host_src, bytes,
                                        dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice), "batch upload");
        };
        if (!put(tb + 4 * to, h.tri4.data(), h.device.tris, 4 * tri_f(h)) ||
            !put(mb + 4 * mo, h.mat4.data(), h.device.mats, 4 * mat_f(h)) ||
            !put(nb + 4 * no, h.nodes.data(), h.device.nodes, 4 * node_f(h)) ||
            !put(n4b + 4 * n4o, h.nodes4.data(), h.device.nodes4, 4 * node4_f(h)) ||
            (!h.lights.empty() && !hip_ok(st, hipMemcpy(lb + lo, h.lights.data(), sizeof(LightEntry) * h.lights.size(), hipMemcpyHostToDevice), "batch upload"))) {
            free_batch(st);
            return Q_EXT_ERR_DEVICE;
        }
        SceneDesc& d = b.descs[i];
        d = SceneDesc{};
        d.tris = tb + 4 * to;
        d.mats = mb + 4 * mo;
        d.nodes = nb + 4 * no;
        d.nodes4 = n4b + 4 * n4o;
        d.qnodes4 = q4b + 2 * n4o;
        d.lights = lb + lo;
        d.num_nodes = h.num_nodes;
        d.num_lights = uint32_t(h.lights.size());
        d.cam = make_camera(sc.camera, sc.width, sc.height);
        d.width = sc.width;
        d.height = sc.height;
        d.pix_begin = pix;
        d.path_begin = pix;
        d.seed = sc.seed;
        d.num_tris = h.num_prims;
        to += tri_f(h);
        mo += mat_f(h);
        no += node_f(h);
        n4o += node4_f(h);
        lo += h.lights.size();
        pix += sc.width * sc.height;
    }
    b.num_scenes = n;
    b.total_pixels = uint32_t(total_pix);
    b.total_tris = uint32_t(total_tris);
    b.total_prims = uint32_t(tri_total / 12);
    b.total_nodes = uint32_t(node_total / 16);     // 16 floats per BVH2 pair node
    b.total_nodes4 = uint32_t(node4_total / 32);   // 32 floats per BVH4 node
    // every scene's nodes keep scene-local child references, so the
    // concatenated array is quantised node by node like a single scene's
    if (!quantize_bvh4(st, b.nodes4.ptr, b.total_nodes4, b.qnodes4, b.quant_ok)) {
        free_batch(st);
        return Q_EXT_ERR_DEVICE;
    }
    b.gstack_entries = std::max(gentries, st->gstack_entries);
    st->gstack_entries = b.gstack_entries;
    b.active = true;
    if (!clear_batch(st, st->stream) || !hip_ok(st, hipStreamSynchronize(st->stream), "batch clear")) {
        free_batch(st);
        return Q_EXT_ERR_DEVICE;
    }
    b.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    log_msg(st, "batch: %u scenes, %u triangles, %u pixels, prepared in %.0f ms", n, b.total_tris, b.total_pixels,
            b.build_ms);
    return Q_EXT_OK;
}

// ---------------------------------------------------------------- autotuner

std::string tune_cache_path() {
    if (const char* e = std::getenv("QUASI_HIP_TUNE_CACHE")) return e;
    const char* home = std::getenv("HOME");
    return std::string(home ? home : "/tmp") + "/.cache/quasi_hip/autotune.jsonl";
}

// A fingerprint of the scene's content: FNV-1a over up to 4096 evenly spaced
// vertex coordinates and indices of the single scene, or over every batch
// scene's triangle count, resolution and seed.
uint64_t scene_fingerprint(const plugin_state* st, uint32_t target) {
    KeyHash k;
    if (target == Q_AUTOTUNE_BATCH) {
        for (const SceneDesc& d : st->batch.descs) {
            k.add(d.num_tris);
            k.add(uint64_t(d.width) << 32 | d.height);
            k.add(d.seed);
        }
        return k.h;
    }
    auto sample = [&](const auto& v) {
        const size_t n = v.size(), step = std::max<size_t>(1, n / 4096);
        k.add(n);
        for (size_t i = 0; i < n; i += step) {
            uint32_t bits;
            std::memcpy(&bits, &v[i], 4);
            k.add(bits);
        }
    };
    sample(st->verts);
    sample(st->indices);
    return k.h;
}

std::string tune_key(const plugin_state* st, uint32_t target) {
    char buf[320];
    const unsigned long long fp = (unsigned long long)scene_fingerprint(st, target);
    if (target == Q_AUTOTUNE_BATCH)
        std::snprintf(buf, sizeof buf, "%s|batch|%u|%u|%u|d%u|i%u|%016llx", st->arch, st->batch.num_scenes,
                      st->batch.total_tris, st->batch.total_pixels, st->params.max_depth, st->params.integrator, fp);
    else
        std::snprintf(buf, sizeof buf, "%s|single|1|%u|%u|d%u|i%u|%016llx", st->arch, st->num_tris,
                      st->width * st->height, st->params.max_depth, st->params.integrator, fp);
    return buf;
}

// A cached choice: traversal shape and samples per pass.
struct TuneChoice {
    TraceConfig cfg;
    uint32_t samples_per_pass;
};

bool tune_cache_lookup(const std::string& key, TuneChoice& choice, double& ms) {
    std::ifstream in(tune_cache_path());
    std::string line;
    bool found = false;
    while (std::getline(in, line)) {
        char k[320] = {0};
        unsigned w = 0, bl = 0, ch = 0, spp = 0;
        double t = 0;
        if (std::sscanf(line.c_str(),
                        "{\"key\": \"%319[^\"]\", \"bvh_width\": %u, \"block_size\": %u, \"queue_chunk\": %u, "
                        "\"samples_per_pass\": %u, \"ms\": %lf}",
                        k, &w, &bl, &ch, &spp, &t) == 6 &&
            key == k) {
            choice = TuneChoice{TraceConfig{w, bl, ch}, spp};
            ms = t;
            found = trace_config_supported(choice.cfg);   // the last matching line wins
        }
    }
    return found;
}

void tune_cache_store(plugin_state* st, const std::string& key, const TraceConfig& c, uint32_t spp_pass, double ms) {
    const std::string path = tune_cache_path();
    std::error_code ec;
    std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
    std::ofstream out(path, std::ios::app);
    if (!out) {
        log_msg(st, "autotune: cannot write %s", path.c_str());
        return;
    }
    char buf[640];
    std::snprintf(buf, sizeof buf,
                  "{\"key\": \"%s\", \"bvh_width\": %u, \"block_size\": %u, \"queue_chunk\": %u, "
                  "\"samples_per_pass\": %u, \"ms\": %.4f}\n",
                  key.c_str(), c.width, c.block, c.chunk, spp_pass, ms);
    out << buf;
}

void apply_config(plugin_state* st, const TraceConfig& c) {
    st->trace_cfg = c;
    st->trace_blocks = persistent_blocks(st->device, c);
    st->paths_blocks = paths_pass_blocks(st->device, c);
    st->params.bvh_width = c.width;
    st->params.block_size = c.block;
    st->params.queue_chunk = c.chunk;
}

// One timed probe with the current trace configuration.
bool time_probe(plugin_state* st, uint32_t target, uint32_t spp, float& ms) {
    hipStream_t s = st->stream;
    hipEvent_t a = take_event(st), b = take_event(st);
    if (!a || !b) return false;
    bool ok = true;
    if (target == Q_AUTOTUNE_BATCH) {
        ok = clear_batch(st, s) && hip_ok(st, hipEventRecord(a, s), "event") && render_batch(st, spp, s) == Q_EXT_OK;
    } else {
        Q_render_frame f{};
        f.width = st->width;
        f.height = st->height;
        f.camera = st->last_camera;
        f.camera_dirty = 1;
        const uint32_t saved = st->params.spp_per_render;
        st->params.spp_per_render = spp;
        ok = hip_ok(st, hipEventRecord(a, s), "event") && render_frame(st, &f);
        st->params.spp_per_render = saved;
    }
    ok = ok && hip_ok(st, hipEventRecord(b, s), "event") && hip_ok(st, hipEventSynchronize(b), "probe");
    if (ok) ok = hip_ok(st, hipEventElapsedTime(&ms, a, b), "probe time");
    st->event_pool.push_back(a);
    st->event_pool.push_back(b);
    return ok;
}

// Samples per pass the probes of `probe_spp` samples use now.
uint32_t probe_pass_size(const plugin_state* st, uint32_t target, uint32_t probe_spp) {
    if (target == Q_AUTOTUNE_BATCH) return batch_samples_per_pass(st, probe_spp);
    const uint32_t npix = std::max<uint32_t>(1u, st->width * st->height);
    uint32_t s = st->params.samples_per_pass ? st->params.samples_per_pass
                                             : std::max<uint32_t>(1u, kAutoPassPaths / npix);
    return std::max(1u, std::min(s, probe_spp));
}

// Candidate pass sizes (samples per pass, hence queue lengths) for probes of
// probe_spp samples: powers of two up to the automatic size of about 2^27
// paths, the six largest; a batch's must divide probe_spp.
std::vector<uint32_t> pass_candidates(const plugin_state* st, uint32_t target, uint32_t probe_spp) {
    const uint32_t npix = std::max<uint32_t>(1u, target == Q_AUTOTUNE_BATCH ? st->batch.total_pixels
                                                                             : st->width * st->height);
    const uint32_t cap = std::min<uint32_t>(probe_spp, std::max<uint32_t>(1u, kAutoPassPaths / npix));
    std::vector<uint32_t> out;
    for (uint32_t s = 1; s <= cap; s *= 2)
        if (target != Q_AUTOTUNE_BATCH || probe_spp % s == 0) out.push_back(s);
    if (out.size() > 6) out.erase(out.begin(), out.end() - 6);
    return out;
}

int32_t autotune(plugin_state* st, uint32_t target, uint32_t probe_spp, uint32_t flags, Q_autotune_result* out) {
    if (target == Q_AUTOTUNE_BATCH && !st->batch.active) {
        log_msg(st, "autotune: no batch set");
        return Q_EXT_ERR_STATE;
    }
    if (target != Q_AUTOTUNE_BATCH && st->params.engine == Q_ENGINE_MEGAKERNEL) {
        log_msg(st, "autotune: the single-scene target needs the wavefront or single-launch engine");
        return Q_EXT_ERR_STATE;
    }
    if (probe_spp == 0) probe_spp = 1;
    if (target == Q_AUTOTUNE_BATCH) {
        const uint32_t pp = batch_samples_per_pass(st, probe_spp);
        probe_spp = std::max(pp, probe_spp - probe_spp % pp);
    }
    const std::string key = tune_key(st, target);
    TuneChoice cached{};
    double cached_ms = 0;
    const bool from_cache = (flags & Q_AUTOTUNE_USE_CACHE) && tune_cache_lookup(key, cached, cached_ms);
    // probes render the last frame's camera; a choice served from the cache
    // renders nothing
    if (!from_cache && target != Q_AUTOTUNE_BATCH && (!st->have_camera || st->width == 0 || !st->scene_ok)) {
        log_msg(st, "autotune: the single-scene probes render the last frame's camera: render a frame first");
        return Q_EXT_ERR_STATE;
    }
    (void)hipSetDevice(st->device);
    if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
    // the probes' rays, launches and kernel times are not the caller's: the
    // statistics are put back as they were once tuning ends
    harvest_events(st);
    Counters counters{};
    if (!hip_ok(st, hipMemcpy(&counters, st->counters.ptr, sizeof counters, hipMemcpyDeviceToHost), "autotune"))
        return Q_EXT_ERR_DEVICE;
    uint64_t launches[K_COUNT];
    double kernel_ms[K_COUNT];
    std::copy(std::begin(st->launches), std::end(st->launches), launches);
    std::copy(std::begin(st->kernel_ms), std::end(st->kernel_ms), kernel_ms);
    auto restore_stats = [&]() {
        harvest_events(st);
        std::copy(launches, launches + K_COUNT, std::begin(st->launches));
        std::copy(kernel_ms, kernel_ms + K_COUNT, std::begin(st->kernel_ms));
        return hip_ok(st, hipMemcpy(st->counters.ptr, &counters, sizeof counters, hipMemcpyHostToDevice), "autotune");
    };
    Q_autotune_result res{};
    res.struct_size = sizeof(Q_autotune_result);
    if (from_cache) {
        apply_config(st, cached.cfg);
        st->params.samples_per_pass = cached.samples_per_pass;
        res.num_entries = 1;
        res.entries[0] = Q_autotune_entry{cached.cfg.width, cached.cfg.block, cached.cfg.chunk,
                                          cached.samples_per_pass, cached_ms};
        res.best = 0;
        res.from_cache = 1;
    } else {
        const TraceConfig saved = st->trace_cfg;
        const uint32_t saved_pass = st->params.samples_per_pass;
        auto fail = [&]() {
            apply_config(st, saved);
            st->params.samples_per_pass = saved_pass;
            (void)sync_render(st);
            (void)restore_stats();
            return Q_EXT_ERR_DEVICE;
        };
        // first run warms code and caches, the second is kept
        auto probe = [&](float& ms) {
            float warm = 0.f;
            return time_probe(st, target, probe_spp, warm) && time_probe(st, target, probe_spp, ms);
        };
        // stage 1: every compiled traversal shape at the current pass size
        // (the single-launch engine: each BVH width and block size it has once)
        const bool single_launch = st->params.engine == Q_ENGINE_PERSISTENT;
        const uint32_t pass0 = probe_pass_size(st, target, probe_spp);
        double best = 1e300;
        for (uint32_t i = 0; i < trace_config_count() && res.num_entries < Q_AUTOTUNE_MAX_ENTRIES; ++i) {
            const TraceConfig c = trace_config_at(i);
            if (single_launch && (c.chunk != 1u || !paths_config_supported(c))) continue;
            apply_config(st, c);
            float ms = 0.f;
            if (!probe(ms)) return fail();
            const uint32_t k = res.num_entries++;
            res.entries[k] = Q_autotune_entry{c.width, c.block, c.chunk, pass0, double(ms)};
            if (ms < best) { best = ms; res.best = k; }
        }
        {
            const Q_autotune_entry& e = res.entries[res.best];
            apply_config(st, TraceConfig{e.bvh_width, e.block_size, e.queue_chunk});
        }
        // stage 2: pass sizes (queue lengths) with the chosen shape, when the
        // probe leaves a choice
        const std::vector<uint32_t> passes = pass_candidates(st, target, probe_spp);
        if (passes.size() > 1) {
            const Q_autotune_entry shape = res.entries[res.best];
            for (uint32_t sp : passes) {
                if (res.num_entries >= Q_AUTOTUNE_MAX_ENTRIES) break;
                st->params.samples_per_pass = sp;
                float ms = 0.f;
                if (!probe(ms)) return fail();
                const uint32_t i = res.num_entries++;
                res.entries[i] = Q_autotune_entry{shape.bvh_width, shape.block_size, shape.queue_chunk, sp, double(ms)};
                if (ms < best) { best = ms; res.best = i; }
            }
        }
        const Q_autotune_entry& e = res.entries[res.best];
        st->params.samples_per_pass = passes.size() > 1 ? e.samples_per_pass : saved_pass;
        if (!(flags & Q_AUTOTUNE_NO_STORE))
            tune_cache_store(st, key, TraceConfig{e.bvh_width, e.block_size, e.queue_chunk},
                             st->params.samples_per_pass, e.ms);
        log_msg(st, "autotune (%s): best width %u, block %u, chunk %u, %u samples per pass: %.3f ms",
                target == Q_AUTOTUNE_BATCH ? "batch" : "single", e.bvh_width, e.block_size, e.queue_chunk,
                e.samples_per_pass, e.ms);
    }
    // leave the pass graph of the chosen configuration captured, the
    // accumulations empty and the statistics as they were before tuning
    if (target == Q_AUTOTUNE_BATCH && !prepare_batch_pass(st, batch_samples_per_pass(st, probe_spp)))
        return Q_EXT_ERR_DEVICE;
    if (st->batch.active && !clear_batch(st, st->stream)) return Q_EXT_ERR_DEVICE;
    if (!clear_single(st, st->stream)) return Q_EXT_ERR_DEVICE;
    if (!sync_render(st) || !restore_stats()) return Q_EXT_ERR_DEVICE;
    if (out) *out = res;
    return Q_EXT_OK;
}

}  // namespace
}  // namespace host
}  // namespace quasi_hip

using namespace quasi_hip;
using namespace quasi_hip::host;

extern "C" {

Q_EXPORT int32_t Q_plugin_set_batch(Q_plugin_handle* handle, const Q_batch_desc* batch) {
    if (!handle || !batch || batch->struct_size < sizeof(Q_batch_desc)) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        return set_batch(st, batch);
    } catch (const std::bad_alloc&) {
        log_msg(st, "set_batch: host allocation failed");
        free_batch(st);
        return Q_EXT_ERR_OUT_OF_MEMORY;
    } catch (...) {
        log_msg(st, "set_batch: exception");
        free_batch(st);
        return Q_EXT_ERR_STATE;
    }
}

Q_EXPORT int32_t Q_plugin_render_batch(Q_plugin_handle* handle, uint32_t spp, void* stream) {
    if (!handle) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        (void)hipSetDevice(st->device);
        return render_batch(st, spp, stream ? static_cast<hipStream_t>(stream) : st->stream);
    } catch (...) {
        log_msg(st, "render_batch: exception");
        return Q_EXT_ERR_STATE;
    }
}

Q_EXPORT int32_t Q_plugin_readback_batch(Q_plugin_handle* handle, uint32_t scene, Q_readback_aov_result* out) {
    if (!handle || !out) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    *out = Q_readback_aov_result{};
    try {
        BatchState& b = st->batch;
        if (!b.active || scene >= b.num_scenes) return Q_EXT_ERR_ARGUMENT;
        if (b.sample_count == 0) return Q_EXT_ERR_STATE;
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
        const SceneDesc& d = b.descs[scene];
        for (uint32_t a = 0; a < Q_AOV_COUNT; ++a) {
            const char* base = static_cast<const char*>(b.accum[a].ptr) + 16 * size_t(d.pix_begin);
            out->buffers[a] = read_mean_of(st, base, d.width, d.height, b.sample_count);
            if (!out->buffers[a].data) return Q_EXT_ERR_DEVICE;
        }
        return Q_EXT_OK;
    } catch (...) {
        return Q_EXT_ERR_STATE;
    }
}

Q_EXPORT int32_t Q_plugin_readback_batch_ids(Q_plugin_handle* handle, uint32_t scene, uint32_t* out, uint64_t count) {
    if (!handle || !out) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    BatchState& b = st->batch;
    if (!b.active || scene >= b.num_scenes) return Q_EXT_ERR_ARGUMENT;
    const SceneDesc& d = b.descs[scene];
    const uint64_t npix = uint64_t(d.width) * d.height;
    if (count < npix) return Q_EXT_ERR_ARGUMENT;
    (void)hipSetDevice(st->device);
    if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
    const char* base = static_cast<const char*>(b.ids.ptr) + 4 * size_t(d.pix_begin);
    return hip_ok(st, hipMemcpy(out, base, 4 * npix, hipMemcpyDeviceToHost), "batch id readback") ? Q_EXT_OK
                                                                                                 : Q_EXT_ERR_DEVICE;
}

Q_EXPORT int32_t Q_plugin_autotune(Q_plugin_handle* handle, uint32_t target, uint32_t probe_spp, uint32_t flags,
                                   Q_autotune_result* out) {
    if (!handle || target > Q_AUTOTUNE_BATCH) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        return autotune(st, target, probe_spp, flags, out);
    } catch (...) {
        log_msg(st, "autotune: exception");
        return Q_EXT_ERR_STATE;
    }
}

}  // extern "C"
