// oracle.hpp -- scalar CPU oracle for the Quasi render path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (libquasi_hip.so, the
// quasi_amd package) includes, links or calls this code; only tests/, the
// smoke check and bench.py's CPU-baseline leg load it, as the checker.
//
// It restates the reference's Metal shader (backends/metal/plugin.mm:98-330)
// on the CPU, templated on the arithmetic type Real (float or double).  Random
// numbers are always produced in u32/float exactly as the shader does and then
// widened to Real, so the FP32 and FP64 instantiations consume the same samples.
//
// Parity status of each function (how it is pinned) is stated next to it:
//   pcg_hash / random_float   known-answer values + bit-exact float conversion
//   camera_ray                closed forms + the reference's own camera.hpp
//                             (built from its sources into oracle/_ref)
//   intersect_tri             brute-force/closed-form tests, quad <=> 2 triangles
//   cosine_sample             closed forms (r1=r2=0 -> n) + moments
//   Bvh                       invariants + closest-hit equality with brute force
//   trace_path_reference      furnace closed form, energy bounds
//   trace_path_nee            equal expectation with trace_path_reference
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace qo {

// ---------------------------------------------------------------- RNG
// plugin.mm:98-102
inline uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// plugin.mm:104-107.  float(0xFFFFFFFF) rounds to 2^32, so the quotient is
// float(state) * 2^-32 exactly, in [0, 1] (1.0 is reachable).
inline float random_float(uint32_t& state) {
    state = pcg_hash(state);
    return static_cast<float>(state) / static_cast<float>(0xFFFFFFFFu);
}

// plugin.mm:295 -- the reference seeds with a 1920-wide stride whatever the
// framebuffer size is.
inline uint32_t seed_reference(uint32_t x, uint32_t y, uint32_t sample) {
    return pcg_hash(x + y * 1920u + sample * 1920u * 1080u);
}

// Hashed seeding for the synthetic workloads: a distinct linear path index per
// (pixel, sample) while W*H*spp < 2^32, mixed with a hashed global seed.
inline uint32_t seed_hashed(uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                            uint32_t sample, uint32_t seed) {
    uint32_t path = sample * (w * h) + y * w + x;
    return pcg_hash(path ^ pcg_hash(seed));
}

// ---------------------------------------------------------------- vec3
template <class R>
struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline V3<R> normalize(V3<R> a) { return a / std::sqrt(dot(a, a)); }
template <class R> inline R max3(V3<R> a) { return std::max(a.x, std::max(a.y, a.z)); }

template <class R>
inline R pi() { return static_cast<R>(3.14159265358979323846); }

// ---------------------------------------------------------------- camera
struct CameraDesc {           // Q_camera (types.hpp:45-50)
    float position[3];
    float target[3];
    float up[3];
    float fov;
};

template <class R>
struct Camera {
    V3<R> origin, lower_left, horizontal, vertical;

    // Host packing plugin.mm:649-674 (direction = normalize(target - position),
    // aspect = W/H) followed by the basis of get_camera_ray, plugin.mm:134-145.
    static Camera make(const CameraDesc& c, uint32_t width, uint32_t height) {
        V3<R> pos{R(c.position[0]), R(c.position[1]), R(c.position[2])};
        V3<R> dir{R(c.target[0]) - pos.x, R(c.target[1]) - pos.y, R(c.target[2]) - pos.z};
        R len = std::sqrt(dot(dir, dir));
        if (len > R(0)) dir = dir / len;
        V3<R> up{R(c.up[0]), R(c.up[1]), R(c.up[2])};
        R aspect = R(width) / R(height);
        R theta = R(c.fov) * pi<R>() / R(180);
        R h = std::tan(theta / R(2));
        R vh = R(2) * h;
        R vw = aspect * vh;
        V3<R> w = normalize(-dir);
        V3<R> right = normalize(cross(up, w));
        V3<R> cup = cross(w, right);
        Camera cam;
        cam.origin = pos;
        cam.horizontal = right * vw;
        cam.vertical = cup * vh;
        cam.lower_left = pos - cam.horizontal * R(0.5) - cam.vertical * R(0.5) - w;
        return cam;
    }

    // plugin.mm:147-151 for a uv already jittered.
    V3<R> direction(R u, R v) const {
        V3<R> target = lower_left + horizontal * u + vertical * v;
        return normalize(target - origin);
    }
};

// uv of pixel (x, y), y counted from the top row: the fullscreen-triangle uv of
// plugin.mm:39-44 interpolated at the pixel centre (plugin.mm:291).  The jitter
// of plugin.mm:131-132 is drawn x first, then y.
template <class R>
inline void pixel_uv(uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t& rng, R& u, R& v) {
    u = (R(x) + R(0.5)) / R(w);
    v = R(1) - (R(y) + R(0.5)) / R(h);
    float jx = random_float(rng);
    float jy = random_float(rng);
    u += (R(jx) - R(0.5)) * R(0.001);
    v += (R(jy) - R(0.5)) * R(0.001);
}

// ---------------------------------------------------------------- sampling
// plugin.mm:111-125 (r1 drawn before r2).
template <class R>
inline V3<R> cosine_sample(V3<R> normal, uint32_t& rng) {
    R r1 = R(random_float(rng));
    R r2 = R(random_float(rng));
    R phi = R(2) * pi<R>() * r1;
    R cos_t = std::sqrt(R(1) - r2);
    R sin_t = std::sqrt(r2);
    V3<R> w = normalize(normal);
    V3<R> a = (std::fabs(w.x) > R(0.9)) ? V3<R>{0, 1, 0} : V3<R>{1, 0, 0};
    V3<R> v = normalize(cross(w, a));
    V3<R> u = cross(w, v);
    return normalize(u * (std::cos(phi) * sin_t) + v * (std::sin(phi) * sin_t) + w * cos_t);
}

// ---------------------------------------------------------------- scene
struct Scene {
    uint32_t ntri = 0;
    std::vector<float> p0, p1, p2;      // 3 floats per triangle each (vertex positions)
    std::vector<float> albedo, emission;  // 3 per triangle
    // lights: triangles whose max emission channel exceeds 0.1 (plugin.mm:243)
    std::vector<uint32_t> light_tri;
    std::vector<float> light_cdf, light_psel, light_area;
    double total_light_weight = 0;

    void set(uint32_t nv, uint32_t nt, const float* verts, const uint32_t* idx,
             const float* alb, const float* emi) {
        (void)nv;
        ntri = nt;
        p0.resize(3 * size_t(nt));
        p1.resize(3 * size_t(nt));
        p2.resize(3 * size_t(nt));
        for (size_t t = 0; t < nt; ++t) {
            for (int k = 0; k < 3; ++k) {
                p0[3 * t + k] = verts[3 * size_t(idx[3 * t + 0]) + k];
                p1[3 * t + k] = verts[3 * size_t(idx[3 * t + 1]) + k];
                p2[3 * t + k] = verts[3 * size_t(idx[3 * t + 2]) + k];
            }
        }
        albedo.assign(alb, alb + 3 * size_t(nt));
        emission.assign(emi, emi + 3 * size_t(nt));
        build_lights();
    }

    // Light selection table: weight = area * (Le.r + Le.g + Le.b), prefix sums in
    // double in triangle order, cdf stored as float with the last entry forced
    // to 1.
    void build_lights() {
        light_tri.clear();
        std::vector<double> wts;
        std::vector<double> areas;
        for (uint32_t t = 0; t < ntri; ++t) {
            const float* e = &emission[3 * size_t(t)];
            float m = std::max(e[0], std::max(e[1], e[2]));
            if (!(m > 0.1f)) continue;
            double a[3], b[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = double(p1[3 * t + k]) - double(p0[3 * t + k]);
                b[k] = double(p2[3 * t + k]) - double(p0[3 * t + k]);
            }
            double cx = a[1] * b[2] - a[2] * b[1];
            double cy = a[2] * b[0] - a[0] * b[2];
            double cz = a[0] * b[1] - a[1] * b[0];
            double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            light_tri.push_back(t);
            areas.push_back(area);
            wts.push_back(area * (double(e[0]) + double(e[1]) + double(e[2])));
        }
        double total = 0;
        for (double w : wts) total += w;
        total_light_weight = total;
        light_cdf.resize(wts.size());
        light_psel.resize(wts.size());
        light_area.resize(wts.size());
        double run = 0;
        for (size_t i = 0; i < wts.size(); ++i) {
            run += wts[i];
            light_cdf[i] = total > 0 ? float(run / total) : 0.f;
            light_psel[i] = total > 0 ? float(wts[i] / total) : 0.f;
            light_area[i] = float(areas[i]);
        }
        if (!light_cdf.empty()) light_cdf.back() = 1.0f;
    }

    // first i with r < cdf[i]; the last light when none (r == 1).
    uint32_t select_light(float r) const {
        size_t lo = 0, hi = light_cdf.size();
        while (lo < hi) {
            size_t mid = (lo + hi) / 2;
            if (r < light_cdf[mid]) hi = mid; else lo = mid + 1;
        }
        if (lo >= light_cdf.size()) lo = light_cdf.size() - 1;
        return uint32_t(lo);
    }

    template <class R> V3<R> v0(uint32_t t) const { return {R(p0[3 * t]), R(p0[3 * t + 1]), R(p0[3 * t + 2])}; }
    template <class R> V3<R> v1(uint32_t t) const { return {R(p1[3 * t]), R(p1[3 * t + 1]), R(p1[3 * t + 2])}; }
    template <class R> V3<R> v2(uint32_t t) const { return {R(p2[3 * t]), R(p2[3 * t + 1]), R(p2[3 * t + 2])}; }
};

// ---------------------------------------------------------------- intersection
// Moller-Trumbore with inclusive edges (u >= 0, v >= 0, u + v <= 1) and an
// inclusive t range, the triangle analogue of plugin.mm:157-195 (which accepts
// t in [t_min, t_max] and alpha, beta in [0, 1]).  Two-sided.
template <class R>
inline bool intersect_tri(V3<R> o, V3<R> d, V3<R> p0, V3<R> p1, V3<R> p2,
                          R tmin, R tmax, R& t_out) {
    V3<R> e1 = p1 - p0;
    V3<R> e2 = p2 - p0;
    V3<R> pv = cross(d, e2);
    R det = dot(e1, pv);
    if (det == R(0)) return false;
    R inv = R(1) / det;
    V3<R> s = o - p0;
    R u = dot(s, pv) * inv;
    if (u < R(0) || u > R(1)) return false;
    V3<R> q = cross(s, e1);
    R v = dot(d, q) * inv;
    if (v < R(0) || u + v > R(1)) return false;
    R t = dot(e2, q) * inv;
    if (t < tmin || t > tmax) return false;
    t_out = t;
    return true;
}

template <class R>
inline V3<R> tri_normal(const Scene& sc, uint32_t t) {
    V3<R> p0 = sc.v0<R>(t);
    return normalize(cross(sc.v1<R>(t) - p0, sc.v2<R>(t) - p0));
}

// ---------------------------------------------------------------- BVH (oracle's own)
// Binned SAH, 32 bins on centroids, at most 4 triangles per leaf.
struct BvhNode {
    float lo[3], hi[3];
    uint32_t left_or_first;  // internal: left child index (right = left + 1); leaf: first prim
    uint32_t count;          // 0 = internal
};

struct Bvh {
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> prims;
    uint32_t depth = 0;

    void build(const Scene& sc) {
        nodes.clear();
        prims.resize(sc.ntri);
        for (uint32_t i = 0; i < sc.ntri; ++i) prims[i] = i;
        depth = 0;
        if (sc.ntri == 0) return;
        std::vector<float> bl(3 * size_t(sc.ntri)), bh(3 * size_t(sc.ntri)), cen(3 * size_t(sc.ntri));
        for (size_t t = 0; t < sc.ntri; ++t) {
            for (int k = 0; k < 3; ++k) {
                float a = sc.p0[3 * t + k], b = sc.p1[3 * t + k], c = sc.p2[3 * t + k];
                bl[3 * t + k] = std::min(a, std::min(b, c));
                bh[3 * t + k] = std::max(a, std::max(b, c));
                cen[3 * t + k] = 0.5f * (bl[3 * t + k] + bh[3 * t + k]);
            }
        }
        nodes.reserve(2 * size_t(sc.ntri));
        nodes.push_back({});
        struct Item { uint32_t node, begin, end, depth; };
        std::vector<Item> work{{0, 0, sc.ntri, 1}};
        while (!work.empty()) {
            Item it = work.back();
            work.pop_back();
            depth = std::max(depth, it.depth);
            BvhNode& nd = nodes[it.node];
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t i = it.begin; i < it.end; ++i) {
                uint32_t p = prims[i];
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], bl[3 * size_t(p) + k]);
                    hi[k] = std::max(hi[k], bh[3 * size_t(p) + k]);
                    clo[k] = std::min(clo[k], cen[3 * size_t(p) + k]);
                    chi[k] = std::max(chi[k], cen[3 * size_t(p) + k]);
                }
            }
            for (int k = 0; k < 3; ++k) { nd.lo[k] = lo[k]; nd.hi[k] = hi[k]; }
            uint32_t n = it.end - it.begin;
            auto area = [](const float* a, const float* b) {
                double dx = double(b[0]) - a[0], dy = double(b[1]) - a[1], dz = double(b[2]) - a[2];
                if (dx < 0 || dy < 0 || dz < 0) return 0.0;
                return 2.0 * (dx * dy + dy * dz + dz * dx);
            };
            // SAH split search
            constexpr int B = 32;
            double best_cost = INFINITY;
            int best_axis = -1, best_split = -1;
            double parent_area = area(lo, hi);
            if (n > 1) {
                for (int ax = 0; ax < 3; ++ax) {
                    float ext = chi[ax] - clo[ax];
                    if (!(ext > 0.f)) continue;
                    uint32_t cnt[B] = {};
                    float blo[B][3], bhi[B][3];
                    for (int b = 0; b < B; ++b)
                        for (int k = 0; k < 3; ++k) { blo[b][k] = INFINITY; bhi[b][k] = -INFINITY; }
                    float scale = float(B) / ext;
                    for (uint32_t i = it.begin; i < it.end; ++i) {
                        uint32_t p = prims[i];
                        int b = int((cen[3 * size_t(p) + ax] - clo[ax]) * scale);
                        b = std::min(std::max(b, 0), B - 1);
                        cnt[b]++;
                        for (int k = 0; k < 3; ++k) {
                            blo[b][k] = std::min(blo[b][k], bl[3 * size_t(p) + k]);
                            bhi[b][k] = std::max(bhi[b][k], bh[3 * size_t(p) + k]);
                        }
                    }
                    double ra[B];
                    uint32_t rc[B];
                    float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                    uint32_t run = 0;
                    for (int b = B - 1; b >= 1; --b) {
                        run += cnt[b];
                        for (int k = 0; k < 3; ++k) { rlo[k] = std::min(rlo[k], blo[b][k]); rhi[k] = std::max(rhi[k], bhi[b][k]); }
                        ra[b] = area(rlo, rhi);
                        rc[b] = run;
                    }
                    float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                    uint32_t lrun = 0;
                    for (int s = 1; s < B; ++s) {   // split between bin s-1 and s
                        lrun += cnt[s - 1];
                        for (int k = 0; k < 3; ++k) { llo[k] = std::min(llo[k], blo[s - 1][k]); lhi[k] = std::max(lhi[k], bhi[s - 1][k]); }
                        if (lrun == 0 || rc[s] == 0) continue;
                        double cost = 1.0 + (area(llo, lhi) * lrun + ra[s] * rc[s]) / std::max(parent_area, 1e-30);
                        if (cost < best_cost) { best_cost = cost; best_axis = ax; best_split = s; }
                    }
                }
            }
            bool make_leaf = (n <= 1) || (n <= 4 && (best_axis < 0 || double(n) <= best_cost));
            if (make_leaf) {
                nd.left_or_first = it.begin;
                nd.count = n;
                continue;
            }
            uint32_t mid;
            if (best_axis < 0) {
                mid = it.begin + n / 2;   // all centroids coincide: split the range
            } else {
                float ext = chi[best_axis] - clo[best_axis];
                float scale = float(B) / ext;
                auto* first = prims.data() + it.begin;
                auto* last = prims.data() + it.end;
                auto* m = std::partition(first, last, [&](uint32_t p) {
                    int b = int((cen[3 * size_t(p) + best_axis] - clo[best_axis]) * scale);
                    b = std::min(std::max(b, 0), B - 1);
                    return b < best_split;
                });
                mid = uint32_t(m - prims.data());
                if (mid == it.begin || mid == it.end) mid = it.begin + n / 2;
            }
            uint32_t left = uint32_t(nodes.size());
            nodes[it.node].left_or_first = left;
            nodes[it.node].count = 0;
            nodes.push_back({});
            nodes.push_back({});
            work.push_back({left + 1, mid, it.end, it.depth + 1});
            work.push_back({left, it.begin, mid, it.depth + 1});
        }
    }
};

// ---------------------------------------------------------------- traversal
struct TraversalCounts {
    uint64_t rays = 0, aabb = 0, tri = 0;
    void add(const TraversalCounts& o) { rays += o.rays; aabb += o.aabb; tri += o.tri; }
};

template <class R>
struct Hit {
    R t;
    uint32_t tri;  // UINT32_MAX = miss
};

// Conservative slab test (the far distance is widened by a few ulps so that
// rounding never culls a box that the ray touches).
template <class R>
inline bool box_hit(const BvhNode& nd, V3<R> o, V3<R> inv, R tmin, R tmax, R& tnear) {
    R t0x = (R(nd.lo[0]) - o.x) * inv.x, t1x = (R(nd.hi[0]) - o.x) * inv.x;
    R t0y = (R(nd.lo[1]) - o.y) * inv.y, t1y = (R(nd.hi[1]) - o.y) * inv.y;
    R t0z = (R(nd.lo[2]) - o.z) * inv.z, t1z = (R(nd.hi[2]) - o.z) * inv.z;
    R tn = std::max(std::max(std::min(t0x, t1x), std::min(t0y, t1y)), std::max(std::min(t0z, t1z), tmin));
    R tf = std::min(std::min(std::max(t0x, t1x), std::max(t0y, t1y)), std::min(std::max(t0z, t1z), tmax));
    tf *= R(1) + R(4) * std::numeric_limits<R>::epsilon();
    tnear = tn;
    return tn <= tf;
}

template <class R>
inline V3<R> safe_inverse(V3<R> d) {
    const R tiny = R(1e-20);
    auto f = [&](R c) { return std::fabs(c) < tiny ? std::copysign(tiny, c) : c; };
    return {R(1) / f(d.x), R(1) / f(d.y), R(1) / f(d.z)};
}

// Closest hit over [tmin, tmax]; on equal t the larger triangle index wins,
// which is what the reference's linear loop does (plugin.mm:204-209 with the
// `t > t_max` rejection of plugin.mm:177 accepting equality).
template <class R>
inline Hit<R> closest_hit(const Scene& sc, const Bvh* bvh, V3<R> o, V3<R> d, R tmin, R tmax,
                          TraversalCounts* cnt) {
    Hit<R> best{tmax, UINT32_MAX};
    if (cnt) cnt->rays++;
    auto test = [&](uint32_t t) {
        R th;
        if (cnt) cnt->tri++;
        if (intersect_tri(o, d, sc.v0<R>(t), sc.v1<R>(t), sc.v2<R>(t), tmin, best.t, th)) {
            if (best.tri == UINT32_MAX || th < best.t || (th == best.t && t > best.tri)) {
                best.t = th;
                best.tri = t;
            }
        }
    };
    if (!bvh) {
        for (uint32_t t = 0; t < sc.ntri; ++t) test(t);
        return best;
    }
    if (bvh->nodes.empty()) return best;
    V3<R> inv = safe_inverse(d);
    uint32_t stack[128];
    int sp = 0;
    R tn;
    if (cnt) cnt->aabb++;
    if (!box_hit(bvh->nodes[0], o, inv, tmin, best.t, tn)) return best;
    stack[sp++] = 0;
    while (sp > 0) {
        const BvhNode& nd = bvh->nodes[stack[--sp]];
        if (nd.count > 0) {
            for (uint32_t i = 0; i < nd.count; ++i) test(bvh->prims[nd.left_or_first + i]);
            continue;
        }
        uint32_t l = nd.left_or_first, r = l + 1;
        R tl, tr;
        if (cnt) cnt->aabb += 2;
        bool hl = box_hit(bvh->nodes[l], o, inv, tmin, best.t, tl);
        bool hr = box_hit(bvh->nodes[r], o, inv, tmin, best.t, tr);
        if (hl && hr) {
            if (tl <= tr) { stack[sp++] = r; stack[sp++] = l; }
            else { stack[sp++] = l; stack[sp++] = r; }
        } else if (hl) {
            stack[sp++] = l;
        } else if (hr) {
            stack[sp++] = r;
        }
    }
    return best;
}

// Any hit over [tmin, tmax] (shadow rays); stops at the first intersection.
template <class R>
inline bool any_hit(const Scene& sc, const Bvh* bvh, V3<R> o, V3<R> d, R tmin, R tmax,
                    TraversalCounts* cnt) {
    if (cnt) cnt->rays++;
    R th;
    if (!bvh) {
        for (uint32_t t = 0; t < sc.ntri; ++t) {
            if (cnt) cnt->tri++;
            if (intersect_tri(o, d, sc.v0<R>(t), sc.v1<R>(t), sc.v2<R>(t), tmin, tmax, th)) return true;
        }
        return false;
    }
    if (bvh->nodes.empty()) return false;
    V3<R> inv = safe_inverse(d);
    uint32_t stack[128];
    int sp = 0;
    R tn;
    if (cnt) cnt->aabb++;
    if (!box_hit(bvh->nodes[0], o, inv, tmin, tmax, tn)) return false;
    stack[sp++] = 0;
    while (sp > 0) {
        const BvhNode& nd = bvh->nodes[stack[--sp]];
        if (nd.count > 0) {
            for (uint32_t i = 0; i < nd.count; ++i) {
                uint32_t t = bvh->prims[nd.left_or_first + i];
                if (cnt) cnt->tri++;
                if (intersect_tri(o, d, sc.v0<R>(t), sc.v1<R>(t), sc.v2<R>(t), tmin, tmax, th)) return true;
            }
            continue;
        }
        uint32_t l = nd.left_or_first, r = l + 1;
        R tl, tr;
        if (cnt) cnt->aabb += 2;
        bool hl = box_hit(bvh->nodes[l], o, inv, tmin, tmax, tl);
        bool hr = box_hit(bvh->nodes[r], o, inv, tmin, tmax, tr);
        if (hl && hr) {
            if (tl <= tr) { stack[sp++] = r; stack[sp++] = l; }
            else { stack[sp++] = l; stack[sp++] = r; }
        } else if (hl) {
            stack[sp++] = l;
        } else if (hr) {
            stack[sp++] = r;
        }
    }
    return false;
}

// ---------------------------------------------------------------- integrators
enum { RNG_REFERENCE = 0, RNG_HASHED = 1 };
enum { INTEGRATOR_REFERENCE = 0, INTEGRATOR_NEE = 1 };

struct RenderParams {
    uint32_t width, height;
    uint32_t spp, sample_offset;
    uint32_t max_depth, rr_start;
    uint32_t seed, rng_mode, integrator;
    uint32_t use_bvh;
    uint32_t threads;
    uint32_t pixel_stride;   // render pixels whose linear index % stride == 0
    uint32_t id_sample;      // sample whose primary ids / t are recorded
    CameraDesc camera;
};

struct PathCounts {
    TraversalCounts camera, extension, shadow;
    void add(const PathCounts& o) { camera.add(o.camera); extension.add(o.extension); shadow.add(o.shadow); }
};

template <class R>
struct SampleResult {
    V3<R> color{0, 0, 0};
    V3<R> albedo{0, 0, 0};
    V3<R> normal{0, 0, 0};
    R depth = 0;
    uint32_t primary_tri = UINT32_MAX;
    R primary_t = 0;
};

constexpr float kTMin = 0.001f;     // plugin.mm:205
constexpr float kTFar = 1e30f;      // plugin.mm:202
constexpr float kOffset = 0.001f;   // plugin.mm:271

// The reference bounce loop, plugin.mm:223-276.
template <class R>
inline void trace_path_reference(const Scene& sc, const Bvh* bvh, V3<R> o, V3<R> d, uint32_t& rng,
                                 const RenderParams& p, SampleResult<R>& res, PathCounts* pc) {
    V3<R> thr{1, 1, 1};
    for (uint32_t b = 0; b < p.max_depth; ++b) {
        Hit<R> h = closest_hit(sc, bvh, o, d, R(kTMin), R(kTFar), pc ? (b == 0 ? &pc->camera : &pc->extension) : nullptr);
        if (h.tri == UINT32_MAX) break;
        V3<R> n = tri_normal<R>(sc, h.tri);
        R denom = dot(n, d);
        V3<R> nf = denom < R(0) ? n : -n;            // plugin.mm:192
        V3<R> pt = o + d * h.t;                       // plugin.mm:179
        const float* e = &sc.emission[3 * size_t(h.tri)];
        const float* a = &sc.albedo[3 * size_t(h.tri)];
        V3<R> Le{R(e[0]), R(e[1]), R(e[2])};
        V3<R> alb{R(a[0]), R(a[1]), R(a[2])};
        float emit_str = std::max(e[0], std::max(e[1], e[2]));
        if (b == 0) {                                 // plugin.mm:242-249
            res.albedo = emit_str > 0.1f ? Le : alb;
            res.normal = nf * R(0.5) + V3<R>{R(0.5), R(0.5), R(0.5)};
            res.depth = h.t;
            res.primary_tri = h.tri;
            res.primary_t = h.t;
        }
        res.color = res.color + thr * Le;             // plugin.mm:252
        if (emit_str > 0.1f) break;                   // plugin.mm:255-258
        thr = thr * alb;                              // plugin.mm:261
        if (b >= p.rr_start) {                        // plugin.mm:264-268 ("bounce > 2")
            R q = std::max(R(0.05), max3(thr));
            if (R(random_float(rng)) > q) break;
            thr = thr / q;
        }
        o = pt + nf * R(kOffset);                     // plugin.mm:271-272
        d = cosine_sample(nf, rng);
    }
}

// Next-event estimation variant (no reference counterpart): emission is seen
// directly only by camera rays; each non-emissive hit before the last segment
// samples one light triangle and traces a shadow ray, so path lengths stay
// <= max_depth and the expectation equals trace_path_reference's.
// RNG order per vertex: light select, r1, r2 (only when a next segment exists),
// then the Russian-roulette draw, then the two direction draws.
template <class R>
inline void trace_path_nee(const Scene& sc, const Bvh* bvh, V3<R> o, V3<R> d, uint32_t& rng,
                           const RenderParams& p, SampleResult<R>& res, PathCounts* pc) {
    V3<R> thr{1, 1, 1};
    const bool have_lights = !sc.light_tri.empty();
    for (uint32_t b = 0; b < p.max_depth; ++b) {
        Hit<R> h = closest_hit(sc, bvh, o, d, R(kTMin), R(kTFar), pc ? (b == 0 ? &pc->camera : &pc->extension) : nullptr);
        if (h.tri == UINT32_MAX) break;
        V3<R> n = tri_normal<R>(sc, h.tri);
        R denom = dot(n, d);
        V3<R> nf = denom < R(0) ? n : -n;
        V3<R> pt = o + d * h.t;
        const float* e = &sc.emission[3 * size_t(h.tri)];
        const float* a = &sc.albedo[3 * size_t(h.tri)];
        V3<R> Le{R(e[0]), R(e[1]), R(e[2])};
        V3<R> alb{R(a[0]), R(a[1]), R(a[2])};
        float emit_str = std::max(e[0], std::max(e[1], e[2]));
        if (b == 0) {
            res.albedo = emit_str > 0.1f ? Le : alb;
            res.normal = nf * R(0.5) + V3<R>{R(0.5), R(0.5), R(0.5)};
            res.depth = h.t;
            res.primary_tri = h.tri;
            res.primary_t = h.t;
            res.color = res.color + Le;
        }
        if (emit_str > 0.1f) break;
        if (b + 1 >= p.max_depth) break;
        V3<R> origin = pt + nf * R(kOffset);
        if (have_lights) {
            uint32_t li = sc.select_light(random_float(rng));
            R r1 = R(random_float(rng));
            R r2 = R(random_float(rng));
            uint32_t lt = sc.light_tri[li];
            V3<R> l0 = sc.v0<R>(lt);
            V3<R> le1 = sc.v1<R>(lt) - l0;
            V3<R> le2 = sc.v2<R>(lt) - l0;
            R su = std::sqrt(r1);
            V3<R> xl = l0 + le1 * (su * (R(1) - r2)) + le2 * (su * r2);
            // geometry term at the hit point itself; the offset origin is
            // used only for the visibility ray
            V3<R> w = xl - pt;
            R dist2 = dot(w, w);
            R dist = std::sqrt(dist2);
            V3<R> wi = w / dist;
            V3<R> nl = normalize(cross(le1, le2));
            R cos_x = dot(nf, wi);
            R cos_l = std::fabs(dot(nl, wi));
            if (cos_x > R(0) && cos_l > R(0)) {
                V3<R> sv = xl - origin;
                R sl = std::sqrt(dot(sv, sv));
                V3<R> ds = sv / sl;
                bool occ = any_hit(sc, bvh, origin, ds, R(kTMin), sl * R(1.0 - 0.001), pc ? &pc->shadow : nullptr);
                if (!occ) {
                    const float* el = &sc.emission[3 * size_t(lt)];
                    V3<R> Ll{R(el[0]), R(el[1]), R(el[2])};
                    R g = cos_x * cos_l * R(sc.light_area[li]) / (pi<R>() * dist2 * R(sc.light_psel[li]));
                    res.color = res.color + thr * alb * Ll * g;
                }
            }
        }
        thr = thr * alb;
        if (b >= p.rr_start) {
            R q = std::max(R(0.05), max3(thr));
            if (R(random_float(rng)) > q) break;
            thr = thr / q;
        }
        o = origin;
        d = cosine_sample(nf, rng);
    }
}

// ---------------------------------------------------------------- render
// Per-pixel sums over samples sample_offset .. sample_offset+spp-1 (RGBA, the
// alpha of every AOV is 1 per sample, plugin.mm:303-307), primary ids and t of
// sample id_sample.  Pixels not selected by pixel_stride are left untouched.
template <class R>
inline PathCounts render(const Scene& sc, const Bvh* bvh, const RenderParams& p,
                         double* beauty, double* albedo, double* normal, double* depth,
                         uint32_t* ids, double* tprim) {
    Camera<R> cam = Camera<R>::make(p.camera, p.width, p.height);
    const uint32_t W = p.width, H = p.height;
    const uint32_t stride = std::max(1u, p.pixel_stride);
    std::atomic<uint32_t> next_row{0};
    unsigned nthreads = p.threads ? p.threads : std::max(1u, std::thread::hardware_concurrency());
    std::vector<PathCounts> counts(nthreads);
    auto worker = [&](unsigned tid) {
        PathCounts& pc = counts[tid];
        for (;;) {
            uint32_t y = next_row.fetch_add(1);
            if (y >= H) break;
            for (uint32_t x = 0; x < W; ++x) {
                size_t pix = size_t(y) * W + x;
                if (pix % stride) continue;
                V3<R> sb{0, 0, 0}, sa{0, 0, 0}, sn{0, 0, 0};
                R sd = 0;
                for (uint32_t s = p.sample_offset; s < p.sample_offset + p.spp; ++s) {
                    uint32_t rng = p.rng_mode == RNG_REFERENCE ? seed_reference(x, y, s)
                                                               : seed_hashed(x, y, W, H, s, p.seed);
                    R u, v;
                    pixel_uv<R>(x, y, W, H, rng, u, v);
                    V3<R> d = cam.direction(u, v);
                    SampleResult<R> res;
                    if (p.integrator == INTEGRATOR_NEE)
                        trace_path_nee(sc, bvh, cam.origin, d, rng, p, res, &pc);
                    else
                        trace_path_reference(sc, bvh, cam.origin, d, rng, p, res, &pc);
                    // plugin.mm:301: depth normalised for display
                    R nd = std::min(std::max((res.depth - R(1)) / R(4), R(0)), R(1));
                    sb = sb + res.color;
                    sa = sa + res.albedo;
                    sn = sn + res.normal;
                    sd += nd;
                    if (s == p.id_sample) {
                        if (ids) ids[pix] = res.primary_tri;
                        if (tprim) tprim[pix] = double(res.primary_t);
                    }
                }
                double n = double(p.spp);
                if (beauty) { beauty[4 * pix] = double(sb.x); beauty[4 * pix + 1] = double(sb.y); beauty[4 * pix + 2] = double(sb.z); beauty[4 * pix + 3] = n; }
                if (albedo) { albedo[4 * pix] = double(sa.x); albedo[4 * pix + 1] = double(sa.y); albedo[4 * pix + 2] = double(sa.z); albedo[4 * pix + 3] = n; }
                if (normal) { normal[4 * pix] = double(sn.x); normal[4 * pix + 1] = double(sn.y); normal[4 * pix + 2] = double(sn.z); normal[4 * pix + 3] = n; }
                if (depth) { depth[4 * pix] = double(sd); depth[4 * pix + 1] = double(sd); depth[4 * pix + 2] = double(sd); depth[4 * pix + 3] = n; }
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
    PathCounts total;
    for (auto& c : counts) total.add(c);
    return total;
}

}  // namespace qo
