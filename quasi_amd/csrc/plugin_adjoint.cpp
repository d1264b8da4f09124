// plugin_adjoint.cpp -- Q_plugin_adjoint: gradients of an image loss with
// respect to the current scene's per-triangle albedo and vertex positions
// (extension export of include/quasi/quasi_hip_ext.h; the reference renders a
// fixed scene and has no derivatives, backends/metal/plugin.mm:584).
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>

#include "plugin_state.hpp"

#define Q_EXPORT __attribute__((visibility("default")))

namespace quasi_hip {
namespace host {

constexpr uint32_t kAdjointReplicas = 8;   // FP32 gradient copies, one per XCD of an MI355X

void free_adjoint(plugin_state* st) {
    AdjointState& a = st->adjoint;
    for (auto* b : {&a.indices, &a.image, &a.dimg, &a.grads, &a.out, &a.partials, &a.rec_hit, &a.rec_vis,
                    &a.counters, &a.param_map, &a.param_out})
        free_buffer(*b);
    a.indices_generation = 0;
}

namespace {

plugin_state* as_state(Q_plugin_handle* h) { return reinterpret_cast<plugin_state*>(h); }

// How the samples of one call are split into recorded passes: blocks of
// `per_pass` samples of whole images, or single samples of bands of
// `band_rows` full rows when one sample of the image exceeds the pass bound.
struct PassPlan {
    uint32_t per_pass, band_rows, bands;
    size_t max_paths;    // paths of the largest pass
};

PassPlan plan_passes(uint32_t width, uint32_t height, uint32_t spp, uint32_t max_pass_paths) {
    PassPlan pl{};
    const size_t npix = size_t(width) * height;
    if (npix <= max_pass_paths) {
        pl.per_pass = uint32_t(std::min<size_t>(spp, max_pass_paths / npix));
        pl.band_rows = height;
    } else {
        pl.per_pass = 1;
        pl.band_rows = std::max<uint32_t>(1u, uint32_t(max_pass_paths / width));
    }
    pl.bands = (height + pl.band_rows - 1) / pl.band_rows;
    pl.max_paths = size_t(pl.per_pass) * width * pl.band_rows;
    return pl;
}

int32_t adjoint(plugin_state* st, const Q_adjoint_desc* d, Q_adjoint_result* res, bool res_ext) {
    if (d->mode > Q_ADJOINT_L2_TARGET || d->spp == 0 || !d->image) {
        log_msg(st, "adjoint: bad descriptor (mode %u, spp %u)", d->mode, d->spp);
        return Q_EXT_ERR_ARGUMENT;
    }
    const Q_render_params& rp = st->params;
    if (st->num_tris == 0 || !st->scene_ok || !st->have_camera || st->width == 0 || st->height == 0) {
        log_msg(st, "adjoint: needs a scene and a rendered frame (camera and resolution)");
        return Q_EXT_ERR_STATE;
    }
    if (rp.max_depth > adjoint_max_depth()) {
        log_msg(st, "adjoint: depth %u is not supported (at most %u)", rp.max_depth, adjoint_max_depth());
        return Q_EXT_ERR_UNSUPPORTED;
    }
    if (d->mode == Q_ADJOINT_L2_TARGET && st->sample_count == 0) {
        log_msg(st, "adjoint: the L2 target mode needs an accumulated image");
        return Q_EXT_ERR_STATE;
    }
    if (uint64_t(d->sample_begin) + d->spp > 0xFFFFFFFFull) return Q_EXT_ERR_ARGUMENT;
    (void)hipSetDevice(st->device);
    if (!sync_render(st)) return Q_EXT_ERR_DEVICE;
    AdjointState& a = st->adjoint;
    const size_t npix = size_t(st->width) * st->height;
    const size_t nt = st->num_tris;
    const size_t nv = st->verts.size() / 3;
    const uint32_t nblk = l2_image_blocks(npix);
    const uint32_t np = d->albedo_param ? d->num_albedo_params : 0u;
    if (d->albedo_param) {
        if (np == 0) {
            log_msg(st, "adjoint: an albedo parameter map needs num_albedo_params >= 1");
            return Q_EXT_ERR_ARGUMENT;
        }
        for (size_t t = 0; t < nt; ++t)
            if (d->albedo_param[t] >= np && d->albedo_param[t] != Q_ADJOINT_NO_PARAM) {
                log_msg(st, "adjoint: triangle %zu maps to albedo parameter %u of %u", t, d->albedo_param[t], np);
                return Q_EXT_ERR_ARGUMENT;
            }
    }
    // recorded passes: at most max_pass_paths paths each (samples of the whole
    // image, or one sample of a band of rows)
    const uint32_t max_pass = d->max_pass_paths ? d->max_pass_paths : Q_ADJOINT_PASS_PATHS;
    const PassPlan plan = plan_passes(st->width, st->height, d->spp, max_pass);
    const size_t rec_paths = plan.max_paths;
    if (!ensure_buffer(st, a.rec_hit, 8 * size_t(rp.max_depth) * rec_paths, "adjoint path records") ||
        !ensure_buffer(st, a.rec_vis, size_t(rp.max_depth) * rec_paths, "adjoint path records") ||
        !ensure_buffer(st, a.counters, sizeof(Counters), "adjoint ray counters") || !ensure_wave(st, uint32_t(rec_paths)))
        return Q_EXT_ERR_OUT_OF_MEMORY;
    if (!ensure_buffer(st, a.indices, 4 * 3 * nt, "adjoint indices") ||
        !ensure_buffer(st, a.image, 16 * npix, "adjoint image") ||
        !ensure_buffer(st, a.dimg, 16 * npix, "adjoint image gradient") ||
        !ensure_buffer(st, a.grads, 4 * size_t(kAdjointReplicas) * 3 * (nt + nv), "adjoint gradients") ||
        !ensure_buffer(st, a.out, 8 * 3 * (nt + nv), "adjoint FP64 gradients") ||
        !ensure_buffer(st, a.partials, 4 * size_t(nblk), "adjoint loss partials"))
        return Q_EXT_ERR_OUT_OF_MEMORY;
    if (np && (!ensure_buffer(st, a.param_map, 4 * nt, "albedo parameter map") ||
               !ensure_buffer(st, a.param_out, 8 * 3 * size_t(np), "albedo parameter gradients")))
        return Q_EXT_ERR_OUT_OF_MEMORY;
    hipStream_t s = st->stream;
    if (a.indices_generation != st->scene_generation) {
        if (!hip_ok(st, hipMemcpy(a.indices.ptr, st->indices.data(), 4 * 3 * nt, hipMemcpyHostToDevice),
                    "adjoint index upload"))
            return Q_EXT_ERR_DEVICE;
        a.indices_generation = st->scene_generation;
    }
    if (!hip_ok(st, hipMemcpy(a.image.ptr, d->image, 16 * npix, hipMemcpyHostToDevice), "adjoint image upload") ||
        !hip_ok(st, launch_fill_u32(a.grads.ptr, size_t(kAdjointReplicas) * 3 * (nt + nv), 0u, s), "gradient clear"))
        return Q_EXT_ERR_DEVICE;
    if (np && (!hip_ok(st, hipMemcpy(a.param_map.ptr, d->albedo_param, 4 * nt, hipMemcpyHostToDevice),
                       "albedo parameter map upload") ||
               !hip_ok(st, launch_fill_u32(a.param_out.ptr, 2 * 3 * size_t(np), 0u, s), "parameter gradient clear")))
        return Q_EXT_ERR_DEVICE;

    hipEvent_t e0 = take_event(st), e1 = take_event(st);
    if (!e0 || !e1) return Q_EXT_ERR_DEVICE;
    bool ok = hip_ok(st, hipEventRecord(e0, s), "event");
    const void* dimg = a.image.ptr;
    if (ok && d->mode == Q_ADJOINT_L2_TARGET) {
        ok = hip_ok(st, launch_l2_image(static_cast<const float4*>(st->accum[Q_AOV_BEAUTY].ptr), st->sample_count,
                                        static_cast<const float4*>(a.image.ptr), static_cast<float4*>(a.dimg.ptr),
                                        npix, static_cast<float*>(a.partials.ptr), s),
                    "L2 image gradient");
        dimg = a.dimg.ptr;
    }
    AdjointParams A{};
    PassParams& p = A.pass;
    p.tris = st->tris.ptr;
    p.mats = st->mats.ptr;
    p.nodes = st->nodes.ptr;
    p.lights = static_cast<const LightEntry*>(st->lights.ptr);
    p.cam = make_camera(st->last_camera, st->width, st->height);
    p.width = st->width;
    p.height = st->height;
    p.sample_begin = d->sample_begin;
    p.sample_count = d->spp;
    p.max_depth = rp.max_depth;
    p.rr_start = rp.rr_start;
    p.seed = (d->flags & Q_ADJOINT_SEED) ? d->seed : rp.seed;
    p.rng_mode = rp.rng_mode;
    p.integrator = rp.integrator;
    p.num_lights = st->num_lights;
    p.num_nodes = st->num_nodes;
    p.num_tris = st->num_prims;
    p.nodes4 = st->nodes4.ptr;
    p.qnodes4 = st->quant_ok ? st->qnodes4.ptr : nullptr;
    p.num_nodes4 = st->num_nodes4;
    p.gstack = static_cast<int*>(st->gstack.ptr);
    p.gstack_entries = st->gstack_entries;
    p.counters = static_cast<Counters*>(a.counters.ptr);
    p.rec_hit = a.rec_hit.ptr;
    p.rec_vis = static_cast<uint8_t*>(a.rec_vis.ptr);
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
    sd.seed = p.seed;
    sd.num_tris = st->num_prims;
    p.total_pixels = uint32_t(npix);
    A.dimg = dimg;
    A.scale = 1.0f / float(d->spp);
    A.replicas = kAdjointReplicas;
    A.indices = static_cast<const uint32_t*>(a.indices.ptr);
    A.grad_albedo = static_cast<float*>(a.grads.ptr);
    A.grad_vertices = static_cast<float*>(a.grads.ptr) + size_t(kAdjointReplicas) * 3 * nt;
    A.num_tris = uint32_t(nt);
    A.num_vertices = uint32_t(nv);
    // the persistent traversal stacks hold kMaxThreadsPerCU threads per CU
    const uint32_t tiles = ((st->width + 15) / 16) * ((st->height + 7) / 8);
    const uint32_t cap = uint32_t(st->num_cus) * (kMaxThreadsPerCU / adjoint_block_size());
    const uint32_t blocks = std::max(1u, std::min(tiles, cap));
    double* out = static_cast<double*>(a.out.ptr);
    // per pass: the forward's wavefront kernels record the paths' hits and
    // shadow visibility (of a band of rows when the image is split), then
    // k_adjoint replays and differentiates them
    uint32_t passes = 0;
    for (uint32_t done = 0; ok && done < d->spp; done += plan.per_pass) {
        const uint32_t n = std::min(plan.per_pass, d->spp - done);
        p.sample_begin = d->sample_begin + done;
        p.sample_count = n;
        for (uint32_t band = 0; ok && band < plan.bands; ++band) {
            const uint32_t row0 = band * plan.band_rows;
            const uint32_t rows = std::min(plan.band_rows, st->height - row0);
            sd.row_begin = plan.bands > 1 ? row0 : 0u;
            sd.rows = plan.bands > 1 ? rows : 0u;
            A.row_begin = sd.row_begin;
            A.rows = sd.rows;
            A.rec_paths = uint32_t(size_t(n) * st->width * rows);
            const WaveState ws = wave_state(st->wave, A.rec_paths, false);
            ok = hip_ok(st, launch_wavefront_pass(p, ws, st->trace_cfg, st->trace_blocks, s, nullptr),
                        "adjoint record pass") &&
                 hip_ok(st, launch_adjoint(A, blocks, s), "adjoint");
            ++passes;
        }
    }
    ok = ok &&
         hip_ok(st, launch_replica_reduce(A.grad_albedo, 3 * nt, kAdjointReplicas, out, s), "albedo reduction") &&
         hip_ok(st, launch_replica_reduce(A.grad_vertices, 3 * nv, kAdjointReplicas, out + 3 * nt, s),
                "vertex reduction");
    if (ok && np)
        ok = hip_ok(st, launch_param_scatter(out, static_cast<const uint32_t*>(a.param_map.ptr), nt, np,
                                             static_cast<double*>(a.param_out.ptr), s),
                    "albedo parameter reduction");
    ok = ok && hip_ok(st, hipEventRecord(e1, s), "event") && hip_ok(st, hipEventSynchronize(e1), "adjoint");
    float ms = 0.f;
    if (ok) ok = hip_ok(st, hipEventElapsedTime(&ms, e0, e1), "adjoint time");
    st->event_pool.push_back(e0);
    st->event_pool.push_back(e1);
    if (!ok) return Q_EXT_ERR_DEVICE;
    if (d->grad_albedo &&
        !hip_ok(st, hipMemcpy(d->grad_albedo, np ? a.param_out.ptr : static_cast<const void*>(out),
                              8 * 3 * (np ? size_t(np) : nt), hipMemcpyDeviceToHost),
                "albedo gradient readback"))
        return Q_EXT_ERR_DEVICE;
    if (d->grad_vertices &&
        !hip_ok(st, hipMemcpy(d->grad_vertices, out + 3 * nt, 8 * 3 * nv, hipMemcpyDeviceToHost),
                "vertex gradient readback"))
        return Q_EXT_ERR_DEVICE;
    double loss = 0.0;
    if (d->mode == Q_ADJOINT_L2_TARGET) {
        std::vector<float> part(nblk);
        if (!hip_ok(st, hipMemcpy(part.data(), a.partials.ptr, 4 * size_t(nblk), hipMemcpyDeviceToHost),
                    "loss readback"))
            return Q_EXT_ERR_DEVICE;
        for (float v : part) loss += double(v);
    }
    if (res) {
        res->paths = uint64_t(npix) * d->spp;
        res->loss = loss;
        res->kernel_ms = ms;
        if (res_ext) {
            // what this call's passes need: path records and wavefront path
            // state of the largest pass, gradient replicas and FP64 sums,
            // images, vertex indices and the parameter map
            res->working_bytes = uint64_t(rec_paths) * (9u * rp.max_depth + 188u) +
                                 4ull * kAdjointReplicas * 3 * (nt + nv) + 8ull * 3 * (nt + nv + np) +
                                 3ull * 16 * npix + 12ull * nt + (np ? 4ull * nt : 0ull) + 4ull * nblk +
                                 sizeof(Counters);
            res->passes = passes;
            res->bands = plan.bands;
        }
    }
    return Q_EXT_OK;
}

}  // namespace
}  // namespace host
}  // namespace quasi_hip

using namespace quasi_hip::host;

extern "C" {

Q_EXPORT int32_t Q_plugin_adjoint(Q_plugin_handle* handle, const Q_adjoint_desc* desc, Q_adjoint_result* result) {
    // callers built against the header before max_pass_paths / working_bytes
    // pass the smaller sizes; the missing fields read as 0
    constexpr size_t kDescBase = offsetof(Q_adjoint_desc, max_pass_paths);
    constexpr size_t kResultBase = offsetof(Q_adjoint_result, working_bytes);
    if (!handle || !desc || desc->struct_size < kDescBase || (result && result->struct_size < kResultBase))
        return Q_EXT_ERR_ARGUMENT;
    plugin_state* st = as_state(handle);
    try {
        Q_adjoint_desc d{};
        std::memcpy(&d, desc, std::min<size_t>(desc->struct_size, sizeof d));
        return adjoint(st, &d, result, result && result->struct_size >= sizeof(Q_adjoint_result));
    } catch (...) {
        log_msg(st, "adjoint: exception");
        return Q_EXT_ERR_STATE;
    }
}

}  // extern "C"
