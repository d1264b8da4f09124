// oracle_capi.cpp -- C entry points of the CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the CPU
// baseline leg of bench.py through ctypes, never by the product.
#include "oracle.hpp"
#include "oracle_adjoint.hpp"

#include <chrono>
#include <cstdio>

using namespace qo;

struct qo_scene {
    Scene scene;
    std::vector<uint32_t> idx;   // 3 vertex indices per triangle
    uint32_t nv = 0;
    Bvh bvh;
    bool has_bvh = false;
    double bvh_build_ms = 0;
};

// Mirrors RenderParams field for field (all u32, then the camera).
struct qo_params {
    uint32_t width, height, spp, sample_offset, max_depth, rr_start, seed, rng_mode, integrator,
        use_bvh, threads, pixel_stride, id_sample, precision;
    float cam_position[3], cam_target[3], cam_up[3], cam_fov;
};

struct qo_counts {
    uint64_t camera_rays, camera_aabb, camera_tri;
    uint64_t extension_rays, extension_aabb, extension_tri;
    uint64_t shadow_rays, shadow_aabb, shadow_tri;
    double seconds;
    uint32_t threads;
    uint32_t pad;
};

static RenderParams to_rp(const qo_params* q) {
    RenderParams p{};
    p.width = q->width; p.height = q->height; p.spp = q->spp; p.sample_offset = q->sample_offset;
    p.max_depth = q->max_depth; p.rr_start = q->rr_start; p.seed = q->seed; p.rng_mode = q->rng_mode;
    p.integrator = q->integrator; p.use_bvh = q->use_bvh; p.threads = q->threads;
    p.pixel_stride = q->pixel_stride; p.id_sample = q->id_sample;
    for (int k = 0; k < 3; ++k) {
        p.camera.position[k] = q->cam_position[k];
        p.camera.target[k] = q->cam_target[k];
        p.camera.up[k] = q->cam_up[k];
    }
    p.camera.fov = q->cam_fov;
    return p;
}

extern "C" {

uint32_t qo_pcg_hash(uint32_t x) { return pcg_hash(x); }

void qo_random_floats(uint32_t state, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = random_float(state);
}

uint32_t qo_seed(uint32_t mode, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample, uint32_t seed) {
    return mode == RNG_REFERENCE ? seed_reference(x, y, sample) : seed_hashed(x, y, w, h, sample, seed);
}

// Camera ray direction (FP64) for an explicit uv (no jitter).
void qo_camera_direction(const float* pos, const float* target, const float* up, float fov,
                         uint32_t w, uint32_t h, double u, double v, double* out) {
    CameraDesc c{};
    for (int k = 0; k < 3; ++k) { c.position[k] = pos[k]; c.target[k] = target[k]; c.up[k] = up[k]; }
    c.fov = fov;
    Camera<double> cam = Camera<double>::make(c, w, h);
    V3<double> d = cam.direction(u, v);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

// Primary ray (FP64) of pixel (x, y), sample s, including the jitter draws.
void qo_primary_ray(const qo_params* q, uint32_t x, uint32_t y, uint32_t s, double* origin, double* dir) {
    RenderParams p = to_rp(q);
    Camera<double> cam = Camera<double>::make(p.camera, p.width, p.height);
    uint32_t rng = p.rng_mode == RNG_REFERENCE ? seed_reference(x, y, s) : seed_hashed(x, y, p.width, p.height, s, p.seed);
    double u, v;
    pixel_uv<double>(x, y, p.width, p.height, rng, u, v);
    V3<double> d = cam.direction(u, v);
    origin[0] = cam.origin.x; origin[1] = cam.origin.y; origin[2] = cam.origin.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

// Cosine-weighted direction about n from the stream `state` (two draws).
void qo_cosine_sample(const double* n, uint32_t state, double* out) {
    V3<double> d = cosine_sample(V3<double>{n[0], n[1], n[2]}, state);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

int qo_intersect_tri(const double* o, const double* d, const double* p0, const double* p1, const double* p2,
                     double tmin, double tmax, double* t) {
    V3<double> O{o[0], o[1], o[2]}, D{d[0], d[1], d[2]};
    V3<double> A{p0[0], p0[1], p0[2]}, B{p1[0], p1[1], p1[2]}, C{p2[0], p2[1], p2[2]};
    return intersect_tri(O, D, A, B, C, tmin, tmax, *t) ? 1 : 0;
}

qo_scene* qo_scene_create(uint32_t nv, uint32_t nt, const float* verts, const uint32_t* idx,
                          const float* albedo, const float* emission, int build_bvh) {
    auto* s = new qo_scene;
    s->scene.set(nv, nt, verts, idx, albedo, emission);
    s->idx.assign(idx, idx + 3 * size_t(nt));
    s->nv = nv;
    if (build_bvh) {
        auto t0 = std::chrono::steady_clock::now();
        s->bvh.build(s->scene);
        s->bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        s->has_bvh = true;
    }
    return s;
}

void qo_scene_destroy(qo_scene* s) { delete s; }

// info: [nodes, leaves, depth, max_leaf, nlights, build_ms*1000]
void qo_scene_info(const qo_scene* s, double* info) {
    uint32_t leaves = 0, max_leaf = 0;
    for (const auto& n : s->bvh.nodes) if (n.count) { leaves++; max_leaf = std::max(max_leaf, n.count); }
    info[0] = double(s->bvh.nodes.size());
    info[1] = leaves;
    info[2] = s->bvh.depth;
    info[3] = max_leaf;
    info[4] = double(s->scene.light_tri.size());
    info[5] = s->bvh_build_ms;
}

// Copies the BVH out: nodes as 8 u32 each (lo[3], hi[3] as float bits, first, count) and prims.
void qo_scene_bvh(const qo_scene* s, uint32_t* nodes, uint32_t* prims) {
    for (size_t i = 0; i < s->bvh.nodes.size(); ++i) std::memcpy(nodes + 8 * i, &s->bvh.nodes[i], 32);
    std::memcpy(prims, s->bvh.prims.data(), 4 * s->bvh.prims.size());
}

// Light table: [tri, cdf, psel, area] per light.
uint32_t qo_scene_lights(const qo_scene* s, uint32_t* tri, float* cdf, float* psel, float* area) {
    const Scene& sc = s->scene;
    for (size_t i = 0; i < sc.light_tri.size(); ++i) {
        if (tri) tri[i] = sc.light_tri[i];
        if (cdf) cdf[i] = sc.light_cdf[i];
        if (psel) psel[i] = sc.light_psel[i];
        if (area) area[i] = sc.light_area[i];
    }
    return uint32_t(sc.light_tri.size());
}

// Closest hit of n rays (FP64), with or without the BVH.
void qo_closest_hits(const qo_scene* s, uint32_t n, const double* o, const double* d, int use_bvh,
                     uint32_t* tri, double* t) {
    for (uint32_t i = 0; i < n; ++i) {
        V3<double> O{o[3 * i], o[3 * i + 1], o[3 * i + 2]}, D{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        Hit<double> h = closest_hit(s->scene, use_bvh && s->has_bvh ? &s->bvh : nullptr, O, D,
                                    double(kTMin), double(kTFar), nullptr);
        tri[i] = h.tri;
        t[i] = h.t;
    }
}

int qo_render(const qo_scene* s, const qo_params* q, double* beauty, double* albedo, double* normal,
              double* depth, uint32_t* ids, double* tprim, qo_counts* counts) {
    RenderParams p = to_rp(q);
    const Bvh* bvh = (p.use_bvh && s->has_bvh) ? &s->bvh : nullptr;
    auto t0 = std::chrono::steady_clock::now();
    PathCounts pc;
    if (q->precision == 64)
        pc = render<double>(s->scene, bvh, p, beauty, albedo, normal, depth, ids, tprim);
    else
        pc = render<float>(s->scene, bvh, p, beauty, albedo, normal, depth, ids, tprim);
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (counts) {
        counts->camera_rays = pc.camera.rays; counts->camera_aabb = pc.camera.aabb; counts->camera_tri = pc.camera.tri;
        counts->extension_rays = pc.extension.rays; counts->extension_aabb = pc.extension.aabb; counts->extension_tri = pc.extension.tri;
        counts->shadow_rays = pc.shadow.rays; counts->shadow_aabb = pc.shadow.aabb; counts->shadow_tri = pc.shadow.tri;
        counts->seconds = sec;
        counts->threads = p.threads ? p.threads : std::max(1u, std::thread::hardware_concurrency());
    }
    return 0;
}

uint32_t qo_hardware_concurrency(void) { return std::thread::hardware_concurrency(); }

// ---------------------------------------------------------------- adjoint
// Reverse-mode gradients of l = sum_pixels dimg . (1/spp) sum_s C_s (dimg:
// RGBA per pixel) with respect to albedo (3 per triangle) and vertex
// positions (3 per vertex), in q->precision (32 or 64).
int qo_adjoint(const qo_scene* s, const qo_params* q, const float* dimg, double* grad_albedo,
               double* grad_vertices, double* seconds) {
    RenderParams p = to_rp(q);
    const Bvh* bvh = (p.use_bvh && s->has_bvh) ? &s->bvh : nullptr;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> ga, gc;
    if (q->precision == 64)
        adjoint_render<double>(s->scene, bvh, p, dimg, ga, gc);
    else
        adjoint_render<float>(s->scene, bvh, p, dimg, ga, gc);
    if (grad_albedo) std::copy(ga.begin(), ga.end(), grad_albedo);
    if (grad_vertices) {
        std::fill(grad_vertices, grad_vertices + 3 * size_t(s->nv), 0.0);
        for (size_t t = 0; t < s->scene.ntri; ++t)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) grad_vertices[3 * size_t(s->idx[3 * t + k]) + c] += gc[9 * t + 3 * k + c];
    }
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

struct qo_tape {
    const qo_scene* scene;
    Tape tape;
};

// Records (in FP64) the decisions of every path of the pixels with non-zero
// dimg, for qo_tape_eval.
qo_tape* qo_tape_create(const qo_scene* s, const qo_params* q, const float* dimg) {
    auto* t = new qo_tape;
    t->scene = s;
    RenderParams p = to_rp(q);
    t->tape.record(s->scene, (p.use_bvh && s->has_bvh) ? &s->bvh : nullptr, p, dimg);
    return t;
}

void qo_tape_destroy(qo_tape* t) { delete t; }

uint64_t qo_tape_paths(const qo_tape* t) { return t->tape.path_pixel.size(); }
uint64_t qo_tape_vertices(const qo_tape* t) { return t->tape.vx.size(); }

// l over the recorded paths for vertex positions `verts` (3 per vertex) and
// albedos `albedo` (3 per triangle), FP64, decisions frozen.  Only the paths
// that visit one of the `ntris` triangles listed are evaluated (all when
// ntris is 0): the others do not depend on those triangles' parameters.
double qo_tape_eval(const qo_tape* t, const double* verts, const double* albedo, const uint32_t* tris,
                    uint32_t ntris, uint32_t threads) {
    const qo_scene* s = t->scene;
    const size_t nt = s->scene.ntri;
    std::vector<double> corners(9 * nt);
    for (size_t i = 0; i < nt; ++i)
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) corners[9 * i + 3 * k + c] = verts[3 * size_t(s->idx[3 * i + k]) + c];
    ParamView<double> pv{&s->scene, corners.data(), albedo};
    std::vector<char> rel;
    if (ntris) {
        rel.assign(nt, 0);
        for (uint32_t i = 0; i < ntris; ++i) rel[tris[i]] = 1;
    }
    unsigned nth = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    return t->tape.eval(pv, ntris ? &rel : nullptr, nth);
}

}  // extern "C"
