// plugin.cpp -- libquasi_hip.so: the Quasi plugin C ABI on HIP (gfx950).
//
// Replaces the host side of the reference backend, backends/metal/plugin.mm:
// the 10 exported symbols (plugin.mm:543-902, declared in
// src/quasi/plugin/plugin_interface.hpp:108-148) plus the single-scene
// extension exports declared in include/quasi/quasi_hip_ext.h (the batch and
// autotuner exports live in plugin_batch.cpp).  No exception and no HIP error
// crosses the ABI: failures are logged through Q_plugin_context.log and
// reported as NULL handles, NULL readback data or negative return codes.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "bvh_build.hpp"
#include "bvh_split.hpp"
#include "plugin_state.hpp"
#include "scene_cornell.hpp"

#define Q_EXPORT __attribute__((visibility("default")))

namespace quasi_hip {
namespace host {

const char* const kKernelNames[K_COUNT] = {"megakernel", "resolve",  "fill",         "wf_raygen", "wf_extend",
                                           "wf_shade",   "wf_shadow", "wf_accumulate", "batch_pass", "paths_pass"};
const bool kRenderKernel[K_COUNT] = {true, false, false, true, true, true, true, true, true, true};

// Scene builds of a batch run on host threads and may log at once: the
// host's callback is entered by one thread at a time.
void log_msg(plugin_state* st, const char* fmt, ...) {
    if (!st || !st->context || !st->context->log) return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    static std::mutex log_mutex;
    std::lock_guard<std::mutex> lock(log_mutex);
    st->context->log(st->context->host_data, buf);
}

bool hip_ok(plugin_state* st, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    log_msg(st, "HIP error in %s: %s", what, hipGetErrorString(e));
    return false;
}

void free_buffer(DeviceBuffer& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

bool ensure_buffer(plugin_state* st, DeviceBuffer& b, size_t bytes, const char* what) {
    if (b.bytes >= bytes && b.ptr) return true;
    free_buffer(b);
    if (bytes == 0) return true;
    if (!hip_ok(st, hipMalloc(&b.ptr, bytes), what)) {
        b.ptr = nullptr;
        return false;
    }
    b.bytes = bytes;
    return true;
}

bool quantize_bvh4(plugin_state* st, const void* nodes4, uint32_t n, DeviceBuffer& q, bool& in_range) {
    in_range = false;
    if (!ensure_buffer(st, q, std::max<size_t>(64, 64 * size_t(n)), "quantised BVH4 nodes") ||
        !ensure_buffer(st, st->quant_flag, 4, "quantisation flag"))
        return false;
    uint32_t failed = 0;
    if (!hip_ok(st, launch_fill_u32(st->quant_flag.ptr, 1, 0u, st->stream), "quantisation flag clear") ||
        !hip_ok(st, launch_quantize4(nodes4, n, q.ptr, static_cast<uint32_t*>(st->quant_flag.ptr), st->stream),
                "BVH4 quantisation") ||
        !hip_ok(st, hipMemcpyAsync(&failed, st->quant_flag.ptr, 4, hipMemcpyDeviceToHost, st->stream),
                "quantisation flag read") ||
        !hip_ok(st, hipStreamSynchronize(st->stream), "BVH4 quantisation"))
        return false;
    in_range = failed == 0;
    if (!in_range) log_msg(st, "scene: BVH extents beyond the quantised nodes' range; tracing the 128-byte nodes");
    return true;
}

hipEvent_t take_event(plugin_state* st) {
    if (!st->event_pool.empty()) {
        hipEvent_t e = st->event_pool.back();
        st->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

LaunchTimer::LaunchTimer(plugin_state* st_, hipStream_t s_, uint32_t slot_) : st(st_), s(s_), slot(slot_) {
    st->launches[slot]++;
    if (st->params.flags & Q_FLAG_TIME_KERNELS) {
        a = take_event(st);
        b = take_event(st);
        if (a && b) (void)hipEventRecord(a, s);
    }
}

LaunchTimer::~LaunchTimer() {
    if (a && b) {
        (void)hipEventRecord(b, s);
        st->timed.push_back({slot, a, b});
    } else {
        if (a) st->event_pool.push_back(a);
        if (b) st->event_pool.push_back(b);
    }
}

// Sums finished event pairs into the per-kernel totals (synchronises on them).
void harvest_events(plugin_state* st) {
    for (const TimedLaunch& t : st->timed) {
        float ms = 0.f;
        if (hipEventSynchronize(t.stop) == hipSuccess && hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess)
            st->kernel_ms[t.slot] += ms;
        st->event_pool.push_back(t.start);
        st->event_pool.push_back(t.stop);
    }
    st->timed.clear();
}

bool sync_render(plugin_state* st) {
    if (st->last_stream && !hip_ok(st, hipStreamSynchronize(st->last_stream), "render synchronise")) return false;
    return hip_ok(st, hipStreamSynchronize(st->stream), "stream synchronise");
}

namespace {

// f(begin, end) over [0, n) split across up to `threads` host threads.
template <class F>
void parallel_ranges(size_t n, unsigned threads, F&& f) {
    if (threads <= 1 || n < (size_t(1) << 16)) {
        f(size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (n + threads - 1) / threads;
    for (unsigned i = 1; i < threads; ++i) {
        const size_t b = i * chunk, e = std::min(n, b + chunk);
        if (b < e) pool.emplace_back([&f, b, e] { f(b, e); });
    }
    f(size_t(0), std::min(n, chunk));
    for (auto& t : pool) t.join();
}

// Emissive triangles (max emission > 0.1, the light table's) and triangles
// whose emission has any nonzero bit (what a sparse upload must carry), in
// input order.
struct EmissionScan {
    std::vector<uint32_t> lights, nonzero;
};
EmissionScan scan_emission(const float* emission, uint32_t nt, unsigned threads) {
    const unsigned parts = std::max(1u, std::min(threads, 16u));
    std::vector<EmissionScan> found(parts);
    const size_t chunk = (size_t(nt) + parts - 1) / parts;
    parallel_ranges(parts, nt >= (1u << 16) ? parts : 1u, [&](size_t pb, size_t pe) {
        for (size_t part = pb; part < pe; ++part)
            for (size_t t = part * chunk; t < std::min<size_t>(nt, (part + 1) * chunk); ++t) {
                const float* e = &emission[3 * t];
                if (std::max(e[0], std::max(e[1], e[2])) > 0.1f) found[part].lights.push_back(uint32_t(t));
                uint32_t bits[3];
                std::memcpy(bits, e, sizeof bits);
                if (bits[0] | bits[1] | bits[2]) found[part].nonzero.push_back(uint32_t(t));
            }
    });
    EmissionScan all;
    for (auto& f : found) {
        all.lights.insert(all.lights.end(), f.lights.begin(), f.lights.end());
        all.nonzero.insert(all.nonzero.end(), f.nonzero.begin(), f.nonzero.end());
    }
    return all;
}

// Light table in input-triangle order: weight = area * (Le.r + Le.g + Le.b),
// prefix sums in double, cdf stored as float with the last entry set to 1.
// pos[i] is the BVH-order position of triangle lights[i].
void light_table(const float* verts, const uint32_t* idx, const float* emission, const std::vector<uint32_t>& tris,
                 const uint32_t* pos, std::vector<LightEntry>& out) {
    out.clear();
    std::vector<double> wts;
    for (size_t i = 0; i < tris.size(); ++i) {
        const uint32_t t = tris[i];
        const float* v0 = &verts[3 * size_t(idx[3 * size_t(t)])];
        const float* v1 = &verts[3 * size_t(idx[3 * size_t(t) + 1])];
        const float* v2 = &verts[3 * size_t(idx[3 * size_t(t) + 2])];
        double a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = double(v1[k]) - double(v0[k]);
            b[k] = double(v2[k]) - double(v0[k]);
        }
        const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
        const double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        const float* e = &emission[3 * size_t(t)];
        out.push_back({pos[i], 0.f, 0.f, float(area)});
        wts.push_back(area * (double(e[0]) + double(e[1]) + double(e[2])));
    }
    double total = 0;
    for (double w : wts) total += w;
    double run = 0;
    for (size_t i = 0; i < out.size(); ++i) {
        run += wts[i];
        out[i].cdf = total > 0 ? float(run / total) : 0.f;
        out[i].psel = total > 0 ? float(wts[i] / total) : 0.f;
    }
    if (!out.empty()) out.back().cdf = 1.0f;
}

}  // namespace

int32_t prepare_scene(plugin_state* st, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const float* albedo, const float* emission, unsigned threads, SceneHost& out,
                      uint32_t builder) {
    const unsigned host_threads = threads ? threads : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    using clock = std::chrono::steady_clock;
    auto ms_since = [](clock::time_point t) { return std::chrono::duration<double, std::milli>(clock::now() - t).count(); };
    out.builder = Q_BVH_BUILD_HOST;
    out.device_ms = 0;
    const EmissionScan scan = scan_emission(emission, nt, host_threads);
    const std::vector<uint32_t>& lights = scan.lights;
    const bool gpu = threads == 0 && (builder == Q_BVH_BUILD_GPU ||
                                      (builder == Q_BVH_BUILD_AUTO && nt >= kGpuBuildMinTriangles));
    if (gpu) {
        // the whole scene on the device: gather, tree, layouts, BVH-order arrays
        std::vector<uint32_t> lpos(lights.size());
        const char* why = "";
        if (build_scene_gpu(verts, nv, idx, nt, albedo, emission, scan.nonzero.data(), uint32_t(scan.nonzero.size()),
                            lights.data(), uint32_t(lights.size()), lpos.data(), st->stream, out.device,
                            &out.gpu_times, &why)) {
            const DeviceScene& d = out.device;
            if (d.depth > megakernel_max_depth()) {
                log_msg(st, "scene: BVH depth %u exceeds the traversal stack (%u)", d.depth, megakernel_max_depth());
                out.device.release();
                return Q_EXT_ERR_UNSUPPORTED;
            }
            out.builder = Q_BVH_BUILD_GPU;
            out.device_ms = out.gpu_times.device_ms;
            out.sah_cost = d.sah_cost;
            out.num_tris = nt;
            out.num_prims = d.num_prims;
            out.num_nodes = d.num_nodes;
            out.num_nodes4 = d.num_nodes4;
            out.depth = d.depth;
            out.depth4 = d.depth4;
            light_table(verts, idx, emission, lights, lpos.data(), out.lights);
            return Q_EXT_OK;
        }
        log_msg(st, "scene: GPU build failed (%s); building on the host", why);
    }

    for (size_t i = 0; i < 3 * size_t(nt); ++i)
        if (idx[i] >= nv) {
            log_msg(st, "scene: vertex index %u out of range (%u vertices)", idx[i], nv);
            return Q_EXT_ERR_ARGUMENT;
        }
    clock::time_point t0 = clock::now();
    std::vector<float> p0(3 * size_t(nt)), p1(3 * size_t(nt)), p2(3 * size_t(nt));
    parallel_ranges(nt, host_threads, [&](size_t tb, size_t te) {
        for (size_t t = tb; t < te; ++t)
            for (int k = 0; k < 3; ++k) {
                p0[3 * t + k] = verts[3 * size_t(idx[3 * t + 0]) + k];
                p1[3 * t + k] = verts[3 * size_t(idx[3 * t + 1]) + k];
                p2[3 * t + k] = verts[3 * size_t(idx[3 * t + 2]) + k];
            }
    });
    // references (bvh_split.hpp): the threshold from the box of the finite
    // corner coordinates, the pieces of each triangle in triangle order
    float slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const std::vector<float>* pv : {&p0, &p1, &p2})
        for (size_t i = 0; i < 3 * size_t(nt); ++i) {
            const float v = (*pv)[i];
            if (!std::isfinite(v)) continue;
            slo[i % 3] = std::min(slo[i % 3], v);
            shi[i % 3] = std::max(shi[i % 3], v);
        }
    const bool finite_box = slo[0] <= shi[0] && slo[1] <= shi[1] && slo[2] <= shi[2];
    float T = finite_box ? split_threshold(slo, shi, nt) : 3.0e38f;
    auto emits = [&](size_t t) {
        uint32_t b[3];
        std::memcpy(b, &emission[3 * t], 12);
        return (b[0] | b[1] | b[2]) != 0u;
    };
    std::vector<uint64_t> tri_off(size_t(nt) + 1, 0);
    // references beyond the builder's limit: build the triangles unsplit
    for (int attempt = 0; attempt < 2; ++attempt) {
        parallel_ranges(nt, host_threads, [&](size_t tb, size_t te) {
            PieceBox pc[kMaxPieces];
            for (size_t t = tb; t < te; ++t)
                tri_off[t + 1] = split_triangle(&p0[3 * t], &p1[3 * t], &p2[3 * t], T, !emits(t), pc);
        });
        for (size_t t = 0; t < nt; ++t) tri_off[t + 1] += tri_off[t];
        if (tri_off[nt] <= (1u << 29)) break;
        T = 3.0e38f;
    }
    const uint32_t nr = uint32_t(tri_off[nt]);
    std::vector<float> rlo(3 * size_t(nr)), rhi(3 * size_t(nr));
    std::vector<uint32_t> ref_tri(nr);
    parallel_ranges(nt, host_threads, [&](size_t tb, size_t te) {
        PieceBox pc[kMaxPieces];
        for (size_t t = tb; t < te; ++t) {
            const uint32_t c = split_triangle(&p0[3 * t], &p1[3 * t], &p2[3 * t], T, !emits(t), pc);
            for (uint32_t j = 0; j < c; ++j) {
                const size_t r = tri_off[t] + j;
                std::memcpy(&rlo[3 * r], pc[j].lo, 12);
                std::memcpy(&rhi[3 * r], pc[j].hi, 12);
                ref_tri[r] = uint32_t(t);
            }
        }
    });
    out.gather_ms = ms_since(t0);
    t0 = clock::now();
    // the tree over the reference boxes (each given as the triangle lo, hi, lo)
    BvhBuildResult bvh = build_bvh(rlo.data(), rhi.data(), rlo.data(), nr, threads);
    out.sah_cost = bvh.sah_cost;
    out.bvh_ms = ms_since(t0);
    t0 = clock::now();
    if (bvh.depth > megakernel_max_depth()) {
        log_msg(st, "scene: BVH depth %u exceeds the traversal stack (%u)", bvh.depth, megakernel_max_depth());
        return Q_EXT_ERR_UNSUPPORTED;
    }
    // position of each reference; a triangle's record at every one of them
    std::vector<uint32_t> ref_pos(nr);
    out.tri4.assign(12 * size_t(nr), 0.f);
    out.mat4.assign(8 * size_t(nr), 0.f);
    parallel_ranges(nr, host_threads, [&](size_t ib, size_t ie) {
        for (size_t i = ib; i < ie; ++i) {
            const uint32_t r = bvh.order[i];
            const uint32_t t = ref_tri[r];
            ref_pos[r] = uint32_t(i);
            float* d = &out.tri4[12 * i];
            for (int k = 0; k < 3; ++k) {
                d[k] = p0[3 * size_t(t) + k];
                d[4 + k] = p1[3 * size_t(t) + k] - p0[3 * size_t(t) + k];
                d[8 + k] = p2[3 * size_t(t) + k] - p0[3 * size_t(t) + k];
            }
            std::memcpy(&d[3], &t, 4);
            float* m = &out.mat4[8 * i];
            const float* a = &albedo[3 * size_t(t)];
            const float* e = &emission[3 * size_t(t)];
            m[0] = a[0]; m[1] = a[1]; m[2] = a[2]; m[3] = std::max(e[0], std::max(e[1], e[2]));
            m[4] = e[0]; m[5] = e[1]; m[6] = e[2]; m[7] = 0.f;
        }
    });
    // an emitter has one reference (never split)
    std::vector<uint32_t> lpos(lights.size());
    for (size_t i = 0; i < lights.size(); ++i) lpos[i] = ref_pos[tri_off[lights[i]]];
    light_table(verts, idx, emission, lights, lpos.data(), out.lights);
    out.nodes = std::move(bvh.nodes);
    out.nodes4 = std::move(bvh.nodes4);
    out.layout_ms = ms_since(t0);
    out.num_tris = nt;
    out.num_prims = nr;
    out.num_nodes = bvh.num_nodes;
    out.num_nodes4 = bvh.num_nodes4;
    out.depth = bvh.depth;
    out.depth4 = bvh.depth4;
    return Q_EXT_OK;
}

uint32_t stack_overflow_entries(uint32_t depth, uint32_t depth4) {
    // the BVH2 walk keeps at most one entry per level, the BVH4 walk at most three
    const uint32_t need = std::max(depth, 3 * depth4);
    return need > uint32_t(kShortStack) ? need - uint32_t(kShortStack) : 0u;
}

bool ensure_gstack(plugin_state* st, uint32_t entries) {
    const size_t threads = size_t(st->num_cus) * kMaxThreadsPerCU;
    return ensure_buffer(st, st->gstack, std::max<size_t>(4, size_t(entries) * threads * 4), "traversal stacks");
}

namespace {

plugin_state* as_state(Q_plugin_handle* h) { return reinterpret_cast<plugin_state*>(h); }

void log_ctx(Q_plugin_context* ctx, const char* msg) {
    if (ctx && ctx->log) ctx->log(ctx->host_data, msg);
}

constexpr const char* kName = "HIP Path Tracer";
constexpr const char* kDescription = "Cornell Box path tracer using HIP on AMD Instinct (gfx950)";
constexpr const char* kAuthor = "Quasi";

Q_render_params default_params() {
    Q_render_params p{};
    p.struct_size = sizeof(Q_render_params);
    p.spp_per_render = 1;          // plugin.mm:23 (one sample per frame)
    p.max_depth = 5;               // plugin.mm:22
    p.rr_start = 3;                // plugin.mm:264 ("bounce > 2")
    p.seed = 0;
    p.rng_mode = Q_RNG_REFERENCE;  // plugin.mm:295
    p.integrator = Q_INTEGRATOR_REFERENCE;
    p.engine = Q_ENGINE_MEGAKERNEL;
    p.bvh_width = 4;
    p.samples_per_pass = 0;
    p.block_size = 0;
    p.flags = 0;
    p.id_sample = 0;
    p.queue_chunk = 0;
    return p;
}

bool clear_accumulation(plugin_state* st, hipStream_t s) {
    const size_t npix = size_t(st->width) * st->height;
    for (auto& b : st->accum) {
        LaunchTimer lt(st, s, K_FILL);
        if (!hip_ok(st, launch_fill_u32(b.ptr, 4 * npix, 0u, s), "clear accumulation")) return false;
    }
    {
        LaunchTimer lt(st, s, K_FILL);
        if (!hip_ok(st, launch_fill_u32(st->ids.ptr, npix, kMissId, s), "clear ids")) return false;
    }
    st->sample_count = 0;
    return true;
}

bool resize_framebuffer(plugin_state* st, uint32_t w, uint32_t h, hipStream_t s) {
    const size_t npix = size_t(w) * h;
    bool ok = true;
    for (auto& b : st->accum) ok = ok && ensure_buffer(st, b, std::max<size_t>(16, 16 * npix), "accumulation buffer");
    ok = ok && ensure_buffer(st, st->ids, std::max<size_t>(4, 4 * npix), "id buffer");
    if (!ok) {
        // release the framebuffer altogether: as 0x0, the next frame
        // allocates again whatever its size, and readbacks return nothing
        for (auto& b : st->accum) free_buffer(b);
        free_buffer(st->ids);
        st->width = st->height = 0;
        st->sample_count = 0;
        return false;
    }
    st->width = w;
    st->height = h;
    return clear_accumulation(st, s);
}

// Builds and uploads the single scene held in st->verts / indices / albedo / emission.
int32_t upload_scene(plugin_state* st) {
    const uint32_t nt = uint32_t(st->indices.size() / 3);
    const uint32_t nv = uint32_t(st->verts.size() / 3);
    auto t0 = std::chrono::steady_clock::now();
    SceneHost sh;
    int32_t rc = prepare_scene(st, st->verts.data(), nv, st->indices.data(), nt, st->albedo.data(),
                               st->emission.data(), 0, sh, st->params.bvh_builder);
    if (rc != Q_EXT_OK) return rc;
    // From here on the device arrays change: until every one of them holds
    // the new scene, nothing may be traced against them.
    st->scene_ok = false;
    if (sh.builder == Q_BVH_BUILD_GPU) {
        // adopt the device arrays of the GPU build (the stream is idle: the build synchronised it)
        DeviceScene& d = sh.device;
        for (DeviceBuffer* b : {&st->tris, &st->mats, &st->nodes, &st->nodes4}) free_buffer(*b);
        st->tris = {d.tris, 48 * size_t(d.num_prims)};
        st->mats = {d.mats, 32 * size_t(d.num_prims)};
        st->nodes = {d.nodes, 64 * size_t(d.num_nodes)};
        st->nodes4 = {d.nodes4, 128 * size_t(d.num_nodes4)};
        d.tris = d.mats = nullptr;
        d.nodes = d.nodes4 = nullptr;
    }
    const uint32_t gentries = std::max(stack_overflow_entries(sh.depth, sh.depth4), st->batch.gstack_entries);
    if (!ensure_buffer(st, st->tris, std::max<size_t>(16, sh.tri4.size() * 4), "triangles") ||
        !ensure_buffer(st, st->mats, std::max<size_t>(16, sh.mat4.size() * 4), "materials") ||
        !ensure_buffer(st, st->nodes, std::max<size_t>(64, sh.nodes.size() * 4), "BVH nodes") ||
        !ensure_buffer(st, st->nodes4, std::max<size_t>(128, sh.nodes4.size() * 4), "BVH4 nodes") ||
        !ensure_gstack(st, gentries) ||
        !ensure_buffer(st, st->lights, std::max<size_t>(16, sh.lights.size() * sizeof(LightEntry)), "lights"))
        return Q_EXT_ERR_OUT_OF_MEMORY;
    if (!hip_ok(st, hipStreamSynchronize(st->stream), "scene upload") ||
        (!sh.tri4.empty() && !hip_ok(st, hipMemcpy(st->tris.ptr, sh.tri4.data(), sh.tri4.size() * 4, hipMemcpyHostToDevice), "triangle upload")) ||
        (!sh.mat4.empty() && !hip_ok(st, hipMemcpy(st->mats.ptr, sh.mat4.data(), sh.mat4.size() * 4, hipMemcpyHostToDevice), "material upload")) ||
        (!sh.nodes.empty() && !hip_ok(st, hipMemcpy(st->nodes.ptr, sh.nodes.data(), sh.nodes.size() * 4, hipMemcpyHostToDevice), "BVH upload")) ||
        (!sh.nodes4.empty() && !hip_ok(st, hipMemcpy(st->nodes4.ptr, sh.nodes4.data(), sh.nodes4.size() * 4, hipMemcpyHostToDevice), "BVH4 upload")) ||
        (!sh.lights.empty() && !hip_ok(st, hipMemcpy(st->lights.ptr, sh.lights.data(), sh.lights.size() * sizeof(LightEntry), hipMemcpyHostToDevice), "light upload")))
        return Q_EXT_ERR_DEVICE;
    st->num_tris = nt;
    st->num_prims = sh.num_prims;
    st->num_nodes = sh.num_nodes;
    st->bvh_depth = sh.depth;
    st->num_lights = uint32_t(sh.lights.size());
    st->num_nodes4 = sh.num_nodes4;
    st->bvh_depth4 = sh.depth4;
    if (!quantize_bvh4(st, st->nodes4.ptr, sh.num_nodes4, st->qnodes4, st->quant_ok)) return Q_EXT_ERR_DEVICE;
    st->gstack_entries = gentries;
    ++st->scene_generation;
    st->bvh_builder = sh.builder;
    st->bvh_sah_cost = sh.sah_cost;
    st->bvh_device_ms = sh.device_ms;
    st->bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    st->scene_ok = true;
    if (nt >= (1u << 20) && sh.builder == Q_BVH_BUILD_GPU) {
        const BvhGpuTimes& g = sh.gpu_times;
        log_msg(st, "scene: %u triangles (%u references) built on the GPU in %.0f ms: allocate %.0f, mesh upload "
                "and corners %.0f, split %.0f, tree %.0f (kernels %.0f), layouts and triangle arrays %.0f (kernels "
                "%.0f)", nt, sh.num_prims, st->bvh_build_ms, g.alloc_ms, g.upload_ms, g.split_ms, g.build_ms,
                g.device_ms, g.layout_ms, g.layout_device_ms);
    } else if (nt >= (1u << 20)) {
        log_msg(st, "scene: %u triangles (%u references) built on the host in %.0f ms: corners and split %.0f, "
                "BVH %.0f, layout %.0f, upload %.0f", nt, sh.num_prims, st->bvh_build_ms, sh.gather_ms, sh.bvh_ms,
                sh.layout_ms,
                st->bvh_build_ms - sh.gather_ms - sh.bvh_ms - sh.layout_ms);
    }
    return Q_EXT_OK;
}

void free_wave(WaveBuffers& w) {
    for (auto* b : {&w.ray_o[0], &w.ray_o[1], &w.ray_d[0], &w.ray_d[1], &w.tq[0], &w.tq[1], &w.hit, &w.thr,
                    &w.radiance, &w.rng, &w.aov0, &w.aov1, &w.sh_o, &w.sh_d, &w.sh_c, &w.counts, &w.shadow_counts, &w.fetch})
        free_buffer(*b);
    w.capacity = 0;
}

// Samples per wavefront pass: about 16M paths in flight unless the caller set it.
uint32_t wave_samples_per_pass(const plugin_state* st) {
    const uint32_t npix = st->width * st->height;
    uint32_t s = st->params.samples_per_pass;
    if (s == 0) s = std::max<uint32_t>(1u, kAutoPassPaths / std::max<uint32_t>(1u, npix));
    s = std::min(s, st->params.spp_per_render);
    while (s > 1 && uint64_t(s) * npix > 0xFFFFFFFFull / 2) s /= 2;
    return std::max(1u, s);
}

void destroy_state(plugin_state* st) {
    if (!st) return;
    (void)hipSetDevice(st->device);
    if (st->last_stream) (void)hipStreamSynchronize(st->last_stream);
    if (st->stream) (void)hipStreamSynchronize(st->stream);
    harvest_events(st);
    for (hipEvent_t e : st->event_pool) (void)hipEventDestroy(e);
    st->event_pool.clear();
    free_batch(st);
    free_adjoint(st);
    free_buffer(st->tris);
    free_buffer(st->mats);
    free_buffer(st->nodes);
    free_buffer(st->nodes4);
    free_buffer(st->qnodes4);
    free_buffer(st->quant_flag);
    free_buffer(st->lights);
    free_buffer(st->gstack);
    for (auto& b : st->accum) free_buffer(b);
    free_buffer(st->ids);
    free_buffer(st->resolved);
    free_buffer(st->counters);
    free_wave(st->wave);
    if (st->capture_stream) (void)hipStreamDestroy(st->capture_stream);
    if (st->own_stream && st->stream) (void)hipStreamDestroy(st->stream);
    delete st;
}

}  // namespace

bool ensure_wave(plugin_state* st, uint32_t paths) {
    WaveBuffers& w = st->wave;
    if (w.capacity >= paths) return true;
    const size_t n = paths;
    // queue slots: enough for either layout (single scene or batch)
    const size_t q = size_t(queue_slots_max(paths));
    const size_t group = size_t(kQueueSegments) * kCounterStride * 4;   // bytes per counter group
    bool ok = true;
    for (int k = 0; k < 2; ++k) {
        ok = ok && ensure_buffer(st, w.ray_o[k], 16 * q, "wavefront rays");
        ok = ok && ensure_buffer(st, w.ray_d[k], 16 * q, "wavefront rays");
        ok = ok && ensure_buffer(st, w.tq[k], 16 * q, "queue-ordered path state");
    }
    ok = ok && ensure_buffer(st, w.hit, 8 * q, "wavefront hits") && ensure_buffer(st, w.thr, 16 * n, "throughput") &&
         ensure_buffer(st, w.radiance, 16 * n, "radiance") && ensure_buffer(st, w.rng, 4 * n, "rng states") &&
         ensure_buffer(st, w.aov0, 16 * n, "first-hit data") && ensure_buffer(st, w.aov1, 16 * n, "first-hit data") &&
         ensure_buffer(st, w.sh_o, 16 * q, "shadow rays") && ensure_buffer(st, w.sh_d, 16 * q, "shadow rays") &&
         ensure_buffer(st, w.sh_c, 16 * q, "shadow contributions") &&
         ensure_buffer(st, w.counts, group * (kMaxDepth + 1), "queue lengths") &&
         ensure_buffer(st, w.shadow_counts, group * kMaxDepth, "queue lengths") &&
         ensure_buffer(st, w.fetch, group * 2 * kMaxDepth, "work cursors");
    if (!ok) { free_wave(w); return false; }
    w.capacity = paths;
    return true;
}

WaveState wave_state(const WaveBuffers& w, uint32_t paths, bool batch) {
    WaveState s{};
    for (int k = 0; k < 2; ++k) {
        s.ray_o[k] = w.ray_o[k].ptr;
        s.ray_d[k] = w.ray_d[k].ptr;
        s.tq[k] = w.tq[k].ptr;
    }
    s.hit = w.hit.ptr;
    s.thr = w.thr.ptr;
    s.radiance = w.radiance.ptr;
    s.rng = static_cast<uint32_t*>(w.rng.ptr);
    s.aov0 = w.aov0.ptr;
    s.aov1 = w.aov1.ptr;
    s.sh_o = w.sh_o.ptr;
    s.sh_d = w.sh_d.ptr;
    s.sh_c = w.sh_c.ptr;
    s.counts = static_cast<uint32_t*>(w.counts.ptr);
    s.shadow_counts = static_cast<uint32_t*>(w.shadow_counts.ptr);
    s.fetch = static_cast<uint32_t*>(w.fetch.ptr);
    s.paths = paths;
    s.qchunk = batch ? queue_contiguous_chunk(paths) : kQueueChunkScene;
    s.seg_cap = queue_seg_cap(paths, s.qchunk);
    return s;
}

void wave_timer_before(void* user, uint32_t kernel, hipStream_t s) {
    auto* st = static_cast<plugin_state*>(user);
    st->launches[K_RAYGEN + kernel]++;
    if (!(st->params.flags & Q_FLAG_TIME_KERNELS)) return;
    st->pending_start = take_event(st);
    if (st->pending_start) (void)hipEventRecord(st->pending_start, s);
}

void wave_timer_after(void* user, uint32_t kernel, hipStream_t s) {
    auto* st = static_cast<plugin_state*>(user);
    if (!st->pending_start) return;
    hipEvent_t stop = take_event(st);
    if (!stop) { st->event_pool.push_back(st->pending_start); st->pending_start = nullptr; return; }
    (void)hipEventRecord(stop, s);
    st->timed.push_back({K_RAYGEN + kernel, st->pending_start, stop});
    st->pending_start = nullptr;
}

// Host camera packing (plugin.mm:649-674) and the basis of get_camera_ray
// (plugin.mm:134-145), evaluated once per frame in double.
CameraBasis make_camera(const Q_camera& c, uint32_t w, uint32_t h) {
    float dx = c.target[0] - c.position[0], dy = c.target[1] - c.position[1], dz = c.target[2] - c.position[2];
    const float len = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (len > 0.0f) { dx /= len; dy /= len; dz /= len; }
    const double aspect = double(w) / double(h);
    const double theta = double(c.fov) * M_PI / 180.0;
    const double hh = std::tan(theta / 2.0);
    const double vh = 2.0 * hh, vw = aspect * vh;
    double wv[3] = {-double(dx), -double(dy), -double(dz)};
    double wl = std::sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
    for (double& x : wv) x /= wl;
    const double up[3] = {c.up[0], c.up[1], c.up[2]};
    double right[3] = {up[1] * wv[2] - up[2] * wv[1], up[2] * wv[0] - up[0] * wv[2], up[0] * wv[1] - up[1] * wv[0]};
    double rl = std::sqrt(right[0] * right[0] + right[1] * right[1] + right[2] * right[2]);
    for (double& x : right) x /= rl;
    const double cup[3] = {wv[1] * right[2] - wv[2] * right[1], wv[2] * right[0] - wv[0] * right[2],
                           wv[0] * right[1] - wv[1] * right[0]};
    CameraBasis cb{};
    for (int k = 0; k < 3; ++k) {
        const double hor = right[k] * vw, ver = cup[k] * vh;
        cb.origin[k] = c.position[k];
        cb.horizontal[k] = float(hor);
        cb.vertical[k] = float(ver);
        cb.llo[k] = float(-hor * 0.5 - ver * 0.5 - wv[k]);
    }
    return cb;
}

Q_aov_buffer read_mean_of(plugin_state* st, const void* sums, uint32_t width, uint32_t height, uint32_t n) {
    Q_aov_buffer out{};
    const size_t npix = size_t(width) * height;
    if (n == 0 || npix == 0) return out;
    if (!ensure_buffer(st, st->resolved, 16 * npix, "readback staging")) return out;
    {
        LaunchTimer lt(st, st->stream, K_RESOLVE);
        if (!hip_ok(st, launch_resolve(static_cast<const float4*>(sums), static_cast<float4*>(st->resolved.ptr), npix,
                                       n, st->stream),
                    "resolve"))
            return out;
    }
    float* data = static_cast<float*>(std::malloc(16 * npix));
    if (!data) return out;
    if (!hip_ok(st, hipMemcpyAsync(data, st->resolved.ptr, 16 * npix, hipMemcpyDeviceToHost, st->stream), "readback copy") ||
        !hip_ok(st, hipStreamSynchronize(st->stream), "readback synchronise")) {
        std::free(data);
        return out;
    }
    out.data = data;
    out.width = width;
    out.height = height;
    out.channels = 4;
    return out;
}

bool clear_single(plugin_state* st, hipStream_t s) { return clear_accumulation(st, s); }

bool render_frame(plugin_state* st, const Q_render_frame* frame) {
    if (frame->width == 0 || frame->height == 0) return false;
    if (!st->scene_ok) {
        log_msg(st, "render: no scene on the device (the last scene upload failed)");
        return false;
    }
    (void)hipSetDevice(st->device);
    hipStream_t s = frame->command_buffer ? static_cast<hipStream_t>(frame->command_buffer) : st->stream;
    if (st->last_stream && st->last_stream != s) (void)hipStreamSynchronize(st->last_stream);
    st->last_stream = s;
    if (frame->width != st->width || frame->height != st->height) {      // plugin.mm:638-640
        if (!resize_framebuffer(st, frame->width, frame->height, s)) return false;
    }
    if (frame->camera_dirty) {                                            // plugin.mm:643-646
        if (!clear_accumulation(st, s)) return false;
    }
    st->last_camera = frame->camera;
    st->have_camera = true;
    const Q_render_params& rp = st->params;
    PassParams p{};
    p.tris = st->tris.ptr;
    p.mats = st->mats.ptr;
    p.nodes = st->nodes.ptr;
    p.lights = static_cast<const LightEntry*>(st->lights.ptr);
    p.accum_beauty = st->accum[Q_AOV_BEAUTY].ptr;
    p.accum_albedo = st->accum[Q_AOV_ALBEDO].ptr;
    p.accum_normal = st->accum[Q_AOV_NORMAL].ptr;
    p.accum_depth = st->accum[Q_AOV_DEPTH].ptr;
    p.ids = static_cast<uint32_t*>(st->ids.ptr);
    p.counters = static_cast<Counters*>(st->counters.ptr);
    p.cam = make_camera(frame->camera, st->width, st->height);
    p.width = st->width;
    p.height = st->height;
    p.sample_begin = st->sample_count;
    p.sample_count = rp.spp_per_render;
    p.max_depth = rp.max_depth;
    p.rr_start = rp.rr_start;
    p.seed = rp.seed;
    p.rng_mode = rp.rng_mode;
    p.integrator = rp.integrator;
    p.num_lights = st->num_lights;
    p.num_nodes = st->num_nodes;
    p.num_tris = st->num_prims;
    p.record_ids = (rp.flags & Q_FLAG_RECORD_IDS) ? 1u : 0u;
    p.id_sample = rp.id_sample;
    p.nodes4 = st->nodes4.ptr;
    p.qnodes4 = st->quant_ok ? st->qnodes4.ptr : nullptr;
    p.num_nodes4 = st->num_nodes4;
    p.gstack = static_cast<int*>(st->gstack.ptr);
    p.gstack_entries = st->gstack_entries;
    if (rp.engine == Q_ENGINE_WAVEFRONT || rp.engine == Q_ENGINE_PERSISTENT) {
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
        sd.width = p.width;
        sd.height = p.height;
        sd.path_begin = 0;
        sd.pix_begin = 0;
        sd.seed = rp.seed;
        sd.num_tris = st->num_prims;
        p.num_scenes = 0;
        p.total_pixels = st->width * st->height;
        const uint32_t per_pass = wave_samples_per_pass(st);
        const uint32_t npix = st->width * st->height;
        if (!ensure_wave(st, per_pass * npix)) return false;
        WaveTimers tm{st, wave_timer_before, wave_timer_after};
        for (uint32_t done = 0; done < rp.spp_per_render; done += per_pass) {
            const uint32_t n = std::min(per_pass, rp.spp_per_render - done);
            p.sample_begin = st->sample_count + done;
            p.sample_count = n;
            const WaveState ws = wave_state(st->wave, n * npix, false);
            if (rp.engine == Q_ENGINE_PERSISTENT) {
                LaunchTimer lt(st, s, K_PATHS);
                if (!hip_ok(st, launch_paths_pass(p, ws, st->trace_cfg, st->paths_blocks, s), "single-launch pass"))
                    return false;
            } else if (!hip_ok(st, launch_wavefront_pass(p, ws, st->trace_cfg, st->trace_blocks, s, &tm),
                               "wavefront launch")) {
                return false;
            }
        }
    } else {
        LaunchTimer lt(st, s, K_MEGAKERNEL);
        if (!hip_ok(st, launch_megakernel(p, st->bvh_depth, s), "megakernel launch")) return false;
    }
    st->sample_count += rp.spp_per_render;                                // plugin.mm:791
    return true;
}

}  // namespace host
}  // namespace quasi_hip

using namespace quasi_hip;
using namespace quasi_hip::host;

extern "C" {

Q_EXPORT uint32_t Q_plugin_abi_version(void) { return Q_PLUGIN_ABI_VERSION; }

Q_EXPORT Q_plugin_info Q_plugin_get_info(void) {
    Q_plugin_info info{};
    info.name = kName;
    info.version = {1, 0, 0};
    info.description = kDescription;
    info.author = kAuthor;
    return info;
}

Q_EXPORT Q_plugin_handle* Q_plugin_create(Q_plugin_context* ctx) {
    if (!ctx || !ctx->gpu) return nullptr;
    if (ctx->gpu->backend != Q_GPU_BACKEND_HIP) {
        log_ctx(ctx, "HIP backend requires HIP GPU context");
        return nullptr;
    }
    plugin_state* st = nullptr;
    try {
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count <= 0) {
            log_ctx(ctx, "HIP backend: no HIP device available");
            return nullptr;
        }
        const int dev = ctx->gpu->device ? *static_cast<const int*>(ctx->gpu->device) : 0;
        if (dev < 0 || dev >= count) {
            log_ctx(ctx, "HIP backend: device ordinal out of range");
            return nullptr;
        }
        st = new plugin_state{};
        st->context = ctx;
        st->device = dev;
        st->params = default_params();
        if (!hip_ok(st, hipSetDevice(dev), "hipSetDevice")) { delete st; return nullptr; }
        if (hipDeviceGetAttribute(&st->num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            st->num_cus <= 0)
            st->num_cus = 256;
        hipDeviceProp_t prop{};
        const bool have_prop = hipGetDeviceProperties(&prop, dev) == hipSuccess;
        std::snprintf(st->arch, sizeof st->arch, "%s", have_prop ? prop.gcnArchName : "unknown");
        st->trace_cfg = TraceConfig{4, 128, kChunkQuant};
        st->trace_blocks = persistent_blocks(dev, st->trace_cfg);
        st->paths_blocks = paths_pass_blocks(dev, st->trace_cfg);
        if (ctx->gpu->queue) {
            st->stream = static_cast<hipStream_t>(ctx->gpu->queue);
        } else {
            if (!hip_ok(st, hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking), "stream create")) {
                delete st;
                return nullptr;
            }
            st->own_stream = true;
        }
        if (!noalloc_init(st)) log_msg(st, "Q_plugin_render_noalloc is unavailable on device %d", dev);
        if (!ensure_buffer(st, st->counters, sizeof(Counters), "counters") ||
            !hip_ok(st, launch_fill_u32(st->counters.ptr, sizeof(Counters) / 4, 0u, st->stream), "counter clear")) {
            destroy_state(st);
            return nullptr;
        }
        // Default scene, as plugin.mm:581-584.
        const float aspect = ctx->viewport_height ? float(ctx->viewport_width) / float(ctx->viewport_height) : 1.0f;
        make_cornell_box(aspect, st->verts, st->indices, st->albedo, st->emission);
        if (upload_scene(st) != Q_EXT_OK) { destroy_state(st); return nullptr; }
        if (!resize_framebuffer(st, std::max(1u, ctx->viewport_width), std::max(1u, ctx->viewport_height), st->stream)) {
            destroy_state(st);
            return nullptr;
        }
        if (have_prop)
            log_msg(st, "Cornell Box path tracer initialized (HIP device %d: %s, %s)", dev, prop.name, prop.gcnArchName);
        else
            log_msg(st, "Cornell Box path tracer initialized (HIP device %d)", dev);
        return reinterpret_cast<Q_plugin_handle*>(st);
    } catch (...) {
        log_ctx(ctx, "HIP backend: exception during create");
        if (st) destroy_state(st);
        return nullptr;
    }
}

Q_EXPORT void Q_plugin_destroy(Q_plugin_handle* handle) {
    if (!handle) return;
    plugin_state* st = as_state(handle);
    try {
        log_msg(st, "Path tracer destroyed");
        destroy_state(st);
    } catch (...) {
    }
}

Q_EXPORT void Q_plugin_update(Q_plugin_handle* handle, float delta_time) {
    (void)handle;
    (void)delta_time;
}

Q_EXPORT void Q_plugin_render(Q_plugin_handle* handle, Q_render_frame* frame) {
    if (!handle || !frame) return;
    plugin_state* st = as_state(handle);
    try {
        render_frame(st, frame);
    } catch (...) {
        log_msg(st, "HIP backend: exception during render");
    }
}

Q_EXPORT Q_readback_result Q_plugin_readback(Q_plugin_handle* handle) {
    Q_readback_result result{};
    if (!handle) return result;
    plugin_state* st = as_state(handle);
    try {
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return result;
        Q_aov_buffer b = read_mean_of(st, st->accum[Q_AOV_BEAUTY].ptr, st->width, st->height, st->sample_count);
        result.data = b.data;
        result.width = b.width;
        result.height = b.height;
        result.channels = b.channels;
        if (b.data) log_msg(st, "HDR readback complete");
    } catch (...) {
        log_msg(st, "HIP backend: exception during readback");
    }
    return result;
}

Q_EXPORT void Q_plugin_readback_free(Q_readback_result* result) {
    if (result && result->data) {
        std::free(result->data);
        result->data = nullptr;
    }
}

Q_EXPORT Q_readback_aov_result Q_plugin_readback_aov(Q_plugin_handle* handle) {
    Q_readback_aov_result result{};
    if (!handle) return result;
    plugin_state* st = as_state(handle);
    try {
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return result;
        for (uint32_t a = 0; a < Q_AOV_COUNT; ++a)
            result.buffers[a] = read_mean_of(st, st->accum[a].ptr, st->width, st->height, st->sample_count);
        log_msg(st, "AOV readback complete");
    } catch (...) {
        log_msg(st, "HIP backend: exception during AOV readback");
    }
    return result;
}

Q_EXPORT void Q_plugin_readback_aov_free(Q_readback_aov_result* result) {
    if (!result) return;
    for (int i = 0; i < Q_AOV_COUNT; ++i) {
        if (result->buffers[i].data) {
            std::free(result->buffers[i].data);
            result->buffers[i].data = nullptr;
        }
    }
}

// ---------------------------------------------------------------- extensions

Q_EXPORT int32_t Q_plugin_set_scene(Q_plugin_handle* handle, const Q_scene_desc* scene) {
    if (!handle || !scene || scene->struct_size < sizeof(Q_scene_desc)) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        if ((scene->num_triangles && (!scene->indices || !scene->albedo || !scene->emission)) ||
            (scene->num_vertices && !scene->vertices)) {
            log_msg(st, "set_scene: missing array");
            return Q_EXT_ERR_ARGUMENT;
        }
        if (scene->num_triangles >= (1u << 29)) {
            log_msg(st, "set_scene: too many triangles (%u)", scene->num_triangles);
            return Q_EXT_ERR_UNSUPPORTED;
        }
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
        const size_t nv = scene->num_vertices, nt = scene->num_triangles;
        std::vector<float> verts(scene->vertices, scene->vertices + 3 * nv);
        std::vector<uint32_t> indices(scene->indices, scene->indices + 3 * nt);
        std::vector<float> albedo(scene->albedo, scene->albedo + 3 * nt);
        std::vector<float> emission(scene->emission, scene->emission + 3 * nt);
        std::swap(st->verts, verts);
        std::swap(st->indices, indices);
        std::swap(st->albedo, albedo);
        std::swap(st->emission, emission);
        int32_t rc = upload_scene(st);
        if (rc != Q_EXT_OK) {
            if (st->scene_ok) {
                // rejected before the device arrays changed: the previous scene stays, host copy included
                std::swap(st->verts, verts);
                std::swap(st->indices, indices);
                std::swap(st->albedo, albedo);
                std::swap(st->emission, emission);
            }
            return rc;
        }
        if (!clear_accumulation(st, st->stream)) return Q_EXT_ERR_DEVICE;
        return Q_EXT_OK;
    } catch (const std::bad_alloc&) {
        log_msg(st, "set_scene: host allocation failed");
        return Q_EXT_ERR_OUT_OF_MEMORY;
    } catch (...) {
        log_msg(st, "set_scene: exception");
        return Q_EXT_ERR_STATE;
    }
}

// Structs grow at the end; a caller built against an older header passes its
// smaller struct_size and reads / writes only the fields it knows.
constexpr size_t kParamsMinSize = offsetof(Q_render_params, bvh_builder);
constexpr size_t kStatsMinSize = offsetof(Q_render_stats, bvh_builder);

Q_EXPORT int32_t Q_plugin_get_params(Q_plugin_handle* handle, Q_render_params* out) {
    if (!handle || !out || out->struct_size < kParamsMinSize) return Q_EXT_ERR_ARGUMENT;
    const uint32_t size = out->struct_size;
    Q_render_params p = as_state(handle)->params;
    p.struct_size = size;
    std::memcpy(out, &p, std::min<size_t>(size, sizeof p));
    return Q_EXT_OK;
}

Q_EXPORT int32_t Q_plugin_set_params(Q_plugin_handle* handle, const Q_render_params* in) {
    if (!handle || !in || in->struct_size < kParamsMinSize) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        Q_render_params p{};
        std::memcpy(&p, in, std::min<size_t>(in->struct_size, sizeof p));
        p.struct_size = sizeof(Q_render_params);
        if (p.spp_per_render == 0 || p.spp_per_render > 4096 || p.max_depth == 0 || p.max_depth > kMaxDepth ||
            p.rng_mode > Q_RNG_HASHED || p.integrator > Q_INTEGRATOR_NEE || p.engine > Q_ENGINE_PERSISTENT ||
            p.bvh_builder > Q_BVH_BUILD_GPU) {
            log_msg(st, "set_params: unsupported parameter combination");
            return Q_EXT_ERR_UNSUPPORTED;
        }
        if (p.bvh_width == 0) p.bvh_width = 4;
        // the default traversal (measured fastest on the 1M soup): BVH4 in
        // 128-thread blocks over the quantised nodes with the top levels in
        // LDS; per-lane fetch where no such kernel exists (BVH2: 256 threads)
        const uint32_t block = p.block_size ? p.block_size : (p.bvh_width == 4 ? 128u : 256u);
        TraceConfig cfg{p.bvh_width, block, p.queue_chunk ? p.queue_chunk : kChunkQuant};
        if (!p.queue_chunk && !trace_config_supported(cfg)) cfg.chunk = kChunkLdsTop;
        if (!p.queue_chunk && !trace_config_supported(cfg)) cfg.chunk = 1u;
        if (!trace_config_supported(cfg)) {
            log_msg(st, "set_params: no traversal kernel for width %u, block %u, chunk %u", cfg.width, cfg.block,
                    cfg.chunk);
            return Q_EXT_ERR_UNSUPPORTED;
        }
        if (p.engine == Q_ENGINE_PERSISTENT && !paths_config_supported(cfg)) {
            log_msg(st, "set_params: the single-launch engine has no kernel for width %u, block %u (blocks 128, 256)",
                    cfg.width, cfg.block);
            return Q_EXT_ERR_UNSUPPORTED;
        }
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
        st->params = p;
        st->trace_cfg = cfg;
        st->trace_blocks = persistent_blocks(st->device, cfg);
        st->paths_blocks = paths_pass_blocks(st->device, cfg);
        if (!clear_accumulation(st, st->stream)) return Q_EXT_ERR_DEVICE;
        return Q_EXT_OK;
    } catch (...) {
        return Q_EXT_ERR_STATE;
    }
}

Q_EXPORT int32_t Q_plugin_readback_ids(Q_plugin_handle* handle, uint32_t* out, uint64_t count) {
    if (!handle || !out) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        const uint64_t npix = uint64_t(st->width) * st->height;
        if (count < npix) {
            log_msg(st, "readback_ids: buffer holds %llu ids, %llu needed", (unsigned long long)count,
                    (unsigned long long)npix);
            return Q_EXT_ERR_ARGUMENT;
        }
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
        if (npix == 0) return Q_EXT_OK;   // no framebuffer (a resize failed)
        if (!hip_ok(st, hipMemcpy(out, st->ids.ptr, 4 * npix, hipMemcpyDeviceToHost), "id readback")) return Q_EXT_ERR_DEVICE;
        return Q_EXT_OK;
    } catch (...) {
        log_msg(st, "%s: exception", "readback_ids");
        return Q_EXT_ERR_STATE;
    }
}

Q_EXPORT int32_t Q_plugin_readback_display(Q_plugin_handle* handle, uint8_t* out, uint64_t bytes) {
    if (!handle || !out) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        const uint64_t npix = uint64_t(st->width) * st->height;
        if (bytes < 4 * npix) {
            log_msg(st, "readback_display: buffer holds %llu bytes, %llu needed", (unsigned long long)bytes,
                    (unsigned long long)(4 * npix));
            return Q_EXT_ERR_ARGUMENT;
        }
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
        if (st->sample_count == 0 || npix == 0) {
            log_msg(st, "readback_display: nothing rendered yet");
            return Q_EXT_ERR_STATE;
        }
        if (!ensure_buffer(st, st->resolved, 16 * npix, "readback staging")) return Q_EXT_ERR_OUT_OF_MEMORY;
        {
            LaunchTimer lt(st, st->stream, K_RESOLVE);
            if (!hip_ok(st, launch_tonemap(static_cast<const float4*>(st->accum[Q_AOV_BEAUTY].ptr), st->resolved.ptr, npix,
                                           st->sample_count, st->stream),
                        "tonemap"))
                return Q_EXT_ERR_DEVICE;
        }
        if (!hip_ok(st, hipMemcpyAsync(out, st->resolved.ptr, 4 * npix, hipMemcpyDeviceToHost, st->stream),
                    "display readback") ||
            !hip_ok(st, hipStreamSynchronize(st->stream), "display readback"))
            return Q_EXT_ERR_DEVICE;
        return Q_EXT_OK;
    } catch (...) {
        log_msg(st, "%s: exception", "readback_display");
        return Q_EXT_ERR_STATE;
    }
}

Q_EXPORT int32_t Q_plugin_get_stats(Q_plugin_handle* handle, Q_render_stats* out) {
    if (!handle || !out || out->struct_size < kStatsMinSize) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
        harvest_events(st);
        const uint32_t out_size = out->struct_size;
        Counters c{};
        if (!hip_ok(st, hipMemcpy(&c, st->counters.ptr, sizeof c, hipMemcpyDeviceToHost), "counter readback"))
            return Q_EXT_ERR_DEVICE;
        Q_render_stats s{};
        s.struct_size = sizeof(Q_render_stats);
        s.width = st->width;
        s.height = st->height;
        s.sample_count = st->sample_count;
        s.num_triangles = st->num_tris;
        s.num_lights = st->num_lights;
        s.bvh_nodes = st->num_nodes;
        s.bvh_depth = st->bvh_depth;
        s.rays_camera = c.rays_camera;
        s.rays_extension = c.rays_extension;
        s.rays_shadow = c.rays_shadow;
        s.bvh_build_ms = st->bvh_build_ms;
        s.bvh_builder = st->bvh_builder;
        s.bvh_sah_cost = st->bvh_sah_cost;
        s.bvh_device_ms = st->bvh_device_ms;
        s.bvh_references = st->num_prims;
        s.render_ms = 0;
        for (uint32_t k = 0; k < K_COUNT; ++k)
            if (kRenderKernel[k] && k != K_BATCH_PASS) s.render_ms += st->kernel_ms[k];
        s.num_kernels = K_COUNT;
        for (uint32_t k = 0; k < K_COUNT; ++k) {
            std::snprintf(s.kernels[k].name, sizeof s.kernels[k].name, "%s", kKernelNames[k]);
            s.kernels[k].launches = st->launches[k];
            s.kernels[k].total_ms = st->kernel_ms[k];
        }
        s.struct_size = out_size;
        std::memcpy(out, &s, std::min<size_t>(out_size, sizeof s));
        return Q_EXT_OK;
    } catch (...) {
        log_msg(st, "%s: exception", "get_stats");
        return Q_EXT_ERR_STATE;
    }
}

Q_EXPORT int32_t Q_plugin_reset_stats(Q_plugin_handle* handle) {
    if (!handle) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        (void)hipSetDevice(st->device);
        if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
        harvest_events(st);
        for (uint32_t k = 0; k < K_COUNT; ++k) { st->launches[k] = 0; st->kernel_ms[k] = 0; }
        if (!hip_ok(st, launch_fill_u32(st->counters.ptr, sizeof(Counters) / 4, 0u, st->stream), "counter clear") ||
            !hip_ok(st, hipStreamSynchronize(st->stream), "counter clear"))
            return Q_EXT_ERR_DEVICE;
        return Q_EXT_OK;
    } catch (...) {
        log_msg(st, "%s: exception", "reset_stats");
        return Q_EXT_ERR_STATE;
    }
}

Q_EXPORT int32_t Q_plugin_synchronize(Q_plugin_handle* handle) {
    if (!handle) return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        (void)hipSetDevice(st->device);
        return sync_render(st) ? Q_EXT_OK : Q_EXT_ERR_DEVICE;
    } catch (...) {
        log_msg(st, "%s: exception", "synchronize");
        return Q_EXT_ERR_STATE;
    }
}

}  // extern "C"
