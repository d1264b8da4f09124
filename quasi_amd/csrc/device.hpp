// device.hpp -- device-side math, RNG, camera, sampling and BVH traversal for gfx950.
//
// Everything here is __device__ only; the CPU has its own, separately written
// checker.  Shader functions restated from the reference cite
// backends/metal/plugin.mm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace quasi_hip {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ float max3(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTMin = 0.001f;     // plugin.mm:205
constexpr float kTFar = 1e30f;      // plugin.mm:202
constexpr float kOffset = 0.001f;   // plugin.mm:271

// ---------------------------------------------------------------- RNG (plugin.mm:98-107)
__device__ __forceinline__ uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// float(0xFFFFFFFF) == 2^32, so the quotient is an exact scaling of the
// round-to-nearest u32 -> f32 conversion.
__device__ __forceinline__ float random_float(uint32_t& state) {
    state = pcg_hash(state);
    return static_cast<float>(state) * 2.3283064365386963e-10f;
}

__device__ __forceinline__ uint32_t path_seed(uint32_t mode, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                                              uint32_t sample, uint32_t seed) {
    if (mode == 0u) return pcg_hash(x + y * 1920u + sample * 1920u * 1080u);  // plugin.mm:295
    return pcg_hash((sample * (w * h) + y * w + x) ^ pcg_hash(seed));
}

// ---------------------------------------------------------------- sampling (plugin.mm:111-125)
__device__ __forceinline__ f3 cosine_sample(f3 n, uint32_t& rng) {
    float r1 = random_float(rng);
    float r2 = random_float(rng);
    float phi = 2.0f * kPi * r1;
    float cos_t = sqrtf(1.0f - r2);
    float sin_t = sqrtf(r2);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    f3 w = normalize(n);
    f3 a = fabsf(w.x) > 0.9f ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
    f3 v = normalize(cross(w, a));
    f3 u = cross(w, v);
    return normalize(u * (cp * sin_t) + v * (sp * sin_t) + w * cos_t);
}

// ---------------------------------------------------------------- traversal
__device__ __forceinline__ float safe_rcp(float c) {
    const float tiny = 1e-20f;
    float s = fabsf(c) < tiny ? copysignf(tiny, c) : c;
    return 1.0f / s;
}

struct Ray {
    f3 o, d;
};

struct TraceHit {
    float t;
    int tri;        // BVH-order index, -1 = miss
    uint32_t id;    // input-order index
};


// Traversal stacks.  Entry i of a lane lives at base[i * stride], so the 64
// lanes of a wave touch 64 consecutive words (distinct LDS banks, coalesced
// global lines).
//
// LDS pointer type: accesses through it are ds_* instructions (a generic
// int* that may point to either space becomes a flat access).
using lds_int_ptr = __attribute__((address_space(3))) int*;

// LdsStack: the whole stack in LDS (bounded by the BVH depth).
template <int STRIDE>
struct LdsStack {
    lds_int_ptr lds;
    int sp;
    __device__ __forceinline__ LdsStack(int* base, int sp0) : lds((lds_int_ptr)base), sp(sp0) {}
    __device__ __forceinline__ void push(int v) { lds[sp * STRIDE] = v; ++sp; }
    __device__ __forceinline__ int pop() { --sp; return lds[sp * STRIDE]; }
    __device__ __forceinline__ bool empty() const { return sp == 0; }
};

// ShortStack: the top LDS_N entries in LDS, deeper entries in a per-thread
// global overflow area (lane-interleaved with stride gstride).  Keeps the LDS
// footprint small enough for full occupancy; the overflow is rarely touched.
template <int LDS_N, int STRIDE>
struct ShortStack {
    lds_int_ptr lds;
    int* glob;
    uint32_t gstride;
    int sp;
    __device__ __forceinline__ ShortStack(int* lds_base, int* glob_base, uint32_t stride, int sp0)
        : lds((lds_int_ptr)lds_base), glob(glob_base), gstride(stride), sp(sp0) {}
    __device__ __forceinline__ void push(int v) {
        if (sp < LDS_N) lds[sp * STRIDE] = v;
        else glob[size_t(sp - LDS_N) * gstride] = v;
        ++sp;
    }
    __device__ __forceinline__ int pop() {
        --sp;
        return sp < LDS_N ? lds[sp * STRIDE] : glob[size_t(sp - LDS_N) * gstride];
    }
    __device__ __forceinline__ bool empty() const { return sp == 0; }
};

// Moller-Trumbore against BVH-order triangle i; true with t when the hit lies
// in [tmin, tmax].  Inclusive edges.
__device__ __forceinline__ bool hit_triangle(const float4* __restrict__ tris, int i, const Ray& r, float tmin,
                                             float tmax, float& t_out, uint32_t& id);

// The same test on the triangle record at tri (its three float4s).  Written
// without early exits: a wave's lanes test different triangles, so every lane
// runs the whole test anyway, and the three loads issue together.  The
// predicate is the early-exit form's, NaN cases included: a value is rejected
// only when one of its comparisons is true.
// Every multiply-add is written out (fmaf) and nothing else is contracted, so
// the test computes the same bits wherever it is inlined: the traversal
// variants, the megakernel and the batch kernels all agree on every hit
// whatever code the compiler places around it.
__device__ __forceinline__ float dot_fma(const f3& a, const f3& b) {
#pragma clang fp contract(off)
    return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z));
}
__device__ __forceinline__ f3 cross_fma(const f3& a, const f3& b) {
#pragma clang fp contract(off)
    return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ bool hit_triangle_rec(const float4* __restrict__ tri, const Ray& r, float tmin,
                                                 float tmax, float& t_out, uint32_t& id) {
#pragma clang fp contract(off)
    const float4 t0 = tri[0];
    const float4 t1 = tri[1];
    const float4 t2 = tri[2];
    const f3 e1 = xyz(t1), e2 = xyz(t2);
    const f3 pv = cross_fma(r.d, e2);
    const float det = dot_fma(e1, pv);
    const float idet = 1.0f / det;
    const f3 s = mk(r.o.x - t0.x, r.o.y - t0.y, r.o.z - t0.z);
    const float u = dot_fma(s, pv) * idet;
    const f3 q = cross_fma(s, e1);
    const float vv = dot_fma(r.d, q) * idet;
    const float t = dot_fma(e2, q) * idet;
    const bool reject = det == 0.0f || u < 0.0f || u > 1.0f || vv < 0.0f || u + vv > 1.0f || t < tmin || t > tmax;
    if (reject) return false;
    t_out = t;
    id = __float_as_uint(t0.w);
    return true;
}

__device__ __forceinline__ bool hit_triangle(const float4* __restrict__ tris, int i, const Ray& r, float tmin,
                                             float tmax, float& t_out, uint32_t& id) {
    return hit_triangle_rec(tris + 3 * i, r, tmin, tmax, t_out, id);
}

// Slab distances of one box: returns (t_near, widened t_far) clipped to
// [tmin, tmax].
__device__ __forceinline__ void slab(float lx, float hx, float ly, float hy, float lz, float hz, const Ray& r,
                                     const f3& inv, float tmin, float tmax, float& tn, float& tf) {
    const float x0 = (lx - r.o.x) * inv.x, x1 = (hx - r.o.x) * inv.x;
    const float y0 = (ly - r.o.y) * inv.y, y1 = (hy - r.o.y) * inv.y;
    const float z0 = (lz - r.o.z) * inv.z, z1 = (hz - r.o.z) * inv.z;
    tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax)) * kSlabWiden;
}

// slab() with the entry (e*) and exit (f*) planes of each axis already chosen
// from the sign of the direction: (e - o) * inv <= (f - o) * inv, so the
// values are those slab() takes the min and the max of, bit for bit.
__device__ __forceinline__ void slab_sel(float ex, float fx, float ey, float fy, float ez, float fz, const Ray& r,
                                         const f3& inv, float tmin, float tmax, float& tn, float& tf) {
    const float x0 = (ex - r.o.x) * inv.x, x1 = (fx - r.o.x) * inv.x;
    const float y0 = (ey - r.o.y) * inv.y, y1 = (fy - r.o.y) * inv.y;
    const float z0 = (ez - r.o.z) * inv.z, z1 = (fz - r.o.z) * inv.z;
    tn = fmaxf(fmaxf(x0, y0), fmaxf(z0, tmin));
    tf = fminf(fminf(x1, y1), fminf(z1, tmax)) * kSlabWiden;
}

// Closest hit over the pair-node BVH2.  Children are visited near first.  On
// equal t the larger input index wins (the reference's linear loop,
// plugin.mm:204-209 with the inclusive `t > t_max` test of plugin.mm:177).
// ANY: return at the first hit in [tmin, tmax] (shadow rays).
template <bool ANY, class Stack>
__device__ __forceinline__ TraceHit traverse2(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                              const Ray& r, float tmin, float tmax, Stack& st) {
    TraceHit best{tmax, -1, 0u};
    const f3 inv = mk(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
    int node = 0;
    for (;;) {
        while (node >= 0) {
            const float4 a = nodes[4 * node + 0];
            const float4 b = nodes[4 * node + 1];
            const float4 c = nodes[4 * node + 2];
            const float4 dd = nodes[4 * node + 3];
            const int cl = __float_as_int(dd.x);
            const int cr = __float_as_int(dd.y);
            float tnl, tfl, tnr, tfr;
            slab(a.x, a.y, a.z, a.w, c.x, c.y, r, inv, tmin, best.t, tnl, tfl);
            slab(b.x, b.y, b.z, b.w, c.z, c.w, r, inv, tmin, best.t, tnr, tfr);
            const bool hl = tnl <= tfl;
            const bool hr = tnr <= tfr;
            if (hl && hr) {
                const bool left_first = tnl <= tnr;
                st.push(left_first ? cr : cl);
                node = left_first ? cl : cr;
            } else if (hl) {
                node = cl;
            } else if (hr) {
                node = cr;
            } else {
                if (st.empty()) return best;
                node = st.pop();
            }
        }
        const uint32_t v = ~uint32_t(node);
        const int first = int(v >> 2);
        const int count = int(v & 3u) + 1;
        for (int i = first; i < first + count; ++i) {
            float t;
            uint32_t id;
            if (!hit_triangle(tris, i, r, tmin, best.t, t, id)) continue;
            if (ANY) { best.t = t; best.tri = i; best.id = id; return best; }
            if (best.tri < 0 || t < best.t || id > best.id) {
                best.t = t;
                best.tri = i;
                best.id = id;
            }
        }
        if (st.empty()) return best;
        node = st.pop();
    }
}

// Compare-exchange on (key, child) pairs, smaller key first.
__device__ __forceinline__ void cswap(float& ka, int& ca, float& kb, int& cb) {
    const bool sw = kb < ka;
    const float k = sw ? kb : ka;
    const int c = sw ? cb : ca;
    kb = sw ? ka : kb;
    cb = sw ? ca : cb;
    ka = k;
    ca = c;
}

// Closest (or any) hit over the 4-wide BVH.  The four slab tests are
// independent; hit children are ordered by entry distance with a 5-comparator
// sorting network, the nearest is visited next and the others are pushed far
// to near.  Same tie rule as traverse2.
template <bool ANY, class Stack>
__device__ __forceinline__ TraceHit traverse4(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                              const Ray& r, float tmin, float tmax, Stack& st) {
    TraceHit best{tmax, -1, 0u};
    const f3 inv = mk(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
    // entry-plane float4 of each axis in a node (lo, or hi for a negative
    // direction component); the exit plane is the other one of the pair
    const uint32_t sx = inv.x < 0.f ? 1u : 0u, sy = inv.y < 0.f ? 3u : 2u, sz = inv.z < 0.f ? 5u : 4u;
    constexpr float kMissKey = 3.0e38f;
    int node = 0;
    for (;;) {
        while (node >= 0) {
            const float4* n = nodes + 8 * node;
            const float4 ex = n[sx], fx = n[sx ^ 1u], ey = n[sy], fy = n[sy ^ 1u], ez = n[sz], fz = n[sz ^ 1u];
            const int4 ch = *reinterpret_cast<const int4*>(n + 6);
            float t0n, t0f, t1n, t1f, t2n, t2f, t3n, t3f;
            slab_sel(ex.x, fx.x, ey.x, fy.x, ez.x, fz.x, r, inv, tmin, best.t, t0n, t0f);
            slab_sel(ex.y, fx.y, ey.y, fy.y, ez.y, fz.y, r, inv, tmin, best.t, t1n, t1f);
            slab_sel(ex.z, fx.z, ey.z, fy.z, ez.z, fz.z, r, inv, tmin, best.t, t2n, t2f);
            slab_sel(ex.w, fx.w, ey.w, fy.w, ez.w, fz.w, r, inv, tmin, best.t, t3n, t3f);
            // unused slots hold inverted boxes, which slab_sel always rejects
            float k0 = t0n <= t0f ? t0n : kMissKey;
            float k1 = t1n <= t1f ? t1n : kMissKey;
            float k2 = t2n <= t2f ? t2n : kMissKey;
            float k3 = t3n <= t3f ? t3n : kMissKey;
            int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
            cswap(k0, c0, k1, c1);
            cswap(k2, c2, k3, c3);
            cswap(k0, c0, k2, c2);
            cswap(k1, c1, k3, c3);
            cswap(k1, c1, k2, c2);
            if (k0 == kMissKey) {
                if (st.empty()) return best;
                node = st.pop();
                continue;
            }
            if (k3 != kMissKey) st.push(c3);
            if (k2 != kMissKey) st.push(c2);
            if (k1 != kMissKey) st.push(c1);
            node = c0;
        }
        const uint32_t v = ~uint32_t(node);
        const int first = int(v >> 2);
        const int count = int(v & 3u) + 1;
        for (int i = first; i < first + count; ++i) {
            float t;
            uint32_t id;
            if (!hit_triangle(tris, i, r, tmin, best.t, t, id)) continue;
            if (ANY) { best.t = t; best.tri = i; best.id = id; return best; }
            if (best.tri < 0 || t < best.t || id > best.id) {
                best.t = t;
                best.tri = i;
                best.id = id;
            }
        }
        if (st.empty()) return best;
        node = st.pop();
    }
}

// Convenience wrappers with an LDS stack (megakernel).
template <int STRIDE>
__device__ __forceinline__ TraceHit trace_closest(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                                  Ray r, float tmin, float tmax, int* __restrict__ stk) {
    LdsStack<STRIDE> st{stk, 0};
    return traverse2<false>(nodes, tris, r, tmin, tmax, st);
}

template <int STRIDE>
__device__ __forceinline__ bool trace_any(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                          Ray r, float tmin, float tmax, int* __restrict__ stk) {
    LdsStack<STRIDE> st{stk, 0};
    return traverse2<true>(nodes, tris, r, tmin, tmax, st).tri >= 0;
}

// ---------------------------------------------------------------- shading
// What one path vertex produces.  The shadow ray (NEE) is returned instead of
// traced so that the megakernel can trace it in place and the wavefront
// engine can queue it.
struct ShadeOut {
    bool cont;       // the path continues with `next`
    bool shadow;     // a shadow ray towards a light sample is requested
    Ray next;
    Ray sray;
    float stmax;
    f3 scontrib;     // added to the path radiance when the shadow ray is unoccluded
};

struct FirstHit {
    f3 albedo, normal;
    float t;
    uint32_t id;
};

// Light choice: the first entry with rs < cdf (the last when none).
__device__ __forceinline__ uint32_t select_light(const LightEntry* __restrict__ lights, uint32_t num_lights,
                                                 float rs) {
    uint32_t lo = 0, hi = num_lights;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (rs < lights[mid].cdf) hi = mid; else lo = mid + 1;
    }
    return lo >= num_lights ? num_lights - 1 : lo;
}

// One vertex of either integrator at bounce b for a ray that hit BVH triangle
// `tri` at distance t.  Updates throughput, radiance and the random stream.
__device__ __forceinline__ void shade_hit(const PassParams& P, const float4* __restrict__ tris,
                                          const float4* __restrict__ mats, const LightEntry* __restrict__ lights,
                                          uint32_t num_lights, uint32_t b, const Ray& ray,
                                          float t, int tri, f3& thr, f3& L, uint32_t& rng, FirstHit* first,
                                          ShadeOut& out) {
    out.cont = false;
    out.shadow = false;
    const float4 t0 = tris[3 * tri + 0];
    const float4 t1 = tris[3 * tri + 1];
    const float4 t2 = tris[3 * tri + 2];
    const f3 n = normalize(cross(xyz(t1), xyz(t2)));
    const f3 nf = dot(n, ray.d) < 0.f ? n : -n;           // plugin.mm:192
    const f3 pt = ray.o + ray.d * t;                        // plugin.mm:179
    const float4 m0 = mats[2 * tri + 0];
    const float4 m1 = mats[2 * tri + 1];
    const f3 alb = xyz(m0), Le = xyz(m1);
    const float emit_str = m0.w;
    if (first) {                                            // plugin.mm:242-249
        first->albedo = emit_str > 0.1f ? Le : alb;
        first->normal = nf * 0.5f + mk(0.5f, 0.5f, 0.5f);
        first->t = t;
        first->id = __float_as_uint(t0.w);
    }
    if (P.integrator == 0u) {
        L = L + thr * Le;                                   // plugin.mm:252
        if (emit_str > 0.1f) return;                        // plugin.mm:255-258
        thr = thr * alb;                                    // plugin.mm:261
        if (b + 1 >= P.max_depth) return;                   // no further trace
        if (b >= P.rr_start) {                              // plugin.mm:264-268
            const float q = fmaxf(0.05f, max3(thr));
            if (random_float(rng) > q) return;
            thr = thr * (1.0f / q);
        }
        out.next.o = pt + nf * kOffset;                     // plugin.mm:271-272
        out.next.d = cosine_sample(nf, rng);
        out.cont = true;
        return;
    }
    if (b == 0) L = L + Le;
    if (emit_str > 0.1f) return;
    if (b + 1 >= P.max_depth) return;
    const f3 org = pt + nf * kOffset;
    if (num_lights) {
        const LightEntry le = lights[select_light(lights, num_lights, random_float(rng))];
        const float r1 = random_float(rng);
        const float r2 = random_float(rng);
        const float4 l0 = tris[3 * le.tri + 0];
        const float4 l1 = tris[3 * le.tri + 1];
        const float4 l2 = tris[3 * le.tri + 2];
        const float su = sqrtf(r1);
        const f3 xl = xyz(l0) + xyz(l1) * (su * (1.0f - r2)) + xyz(l2) * (su * r2);
        // geometry term at the hit point; the offset origin is used only for
        // the visibility ray
        const f3 w = xl - pt;
        const float dist2 = dot(w, w);
        const f3 wi = w * (1.0f / sqrtf(dist2));
        const f3 nl = normalize(cross(xyz(l1), xyz(l2)));
        const float cos_x = dot(nf, wi);
        const float cos_l = fabsf(dot(nl, wi));
        if (cos_x > 0.f && cos_l > 0.f) {
            const f3 sv = xl - org;
            const float sl = sqrtf(dot(sv, sv));
            out.shadow = true;
            out.sray.o = org;
            out.sray.d = sv * (1.0f / sl);
            out.stmax = sl * 0.999f;
            const f3 Ll = xyz(mats[2 * le.tri + 1]);
            const float g = cos_x * cos_l * le.area / (kPi * dist2 * le.psel);
            out.scontrib = thr * alb * Ll * g;
        }
    }
    thr = thr * alb;
    if (b >= P.rr_start) {
        const float q = fmaxf(0.05f, max3(thr));
        if (random_float(rng) > q) return;
        thr = thr * (1.0f / q);
    }
    out.next.o = org;
    out.next.d = cosine_sample(nf, rng);
    out.cont = true;
}

// Primary ray of pixel (x, y) for one sample (plugin.mm:129-153 with the
// camera basis evaluated on the host; x jitter drawn first, plugin.mm:131).
__device__ __forceinline__ Ray camera_ray(const CameraBasis& cam, uint32_t width, uint32_t height, uint32_t x,
                                          uint32_t y, uint32_t& rng) {
    const f3 llo = mk(cam.llo[0], cam.llo[1], cam.llo[2]);
    const f3 hor = mk(cam.horizontal[0], cam.horizontal[1], cam.horizontal[2]);
    const f3 ver = mk(cam.vertical[0], cam.vertical[1], cam.vertical[2]);
    const float u0 = (float(x) + 0.5f) / float(width);
    const float v0 = 1.0f - (float(y) + 0.5f) / float(height);
    const float jx = random_float(rng);
    const float jy = random_float(rng);
    const float u = u0 + (jx - 0.5f) * 0.001f;
    const float v = v0 + (jy - 0.5f) * 0.001f;
    return Ray{mk(cam.origin[0], cam.origin[1], cam.origin[2]), normalize(llo + hor * u + ver * v)};
}

// Sum over the 64 lanes of a wave (all lanes must be present).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Lanes below this one in `mask` (wave64).
__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Wave-wide stream compaction: every lane of the wave calls this with its
// predicate; lanes with pred get consecutive slots reserved by one atomic per
// wave on *counter.  Returns the lane's slot (meaningful only when pred).
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t total = uint32_t(__popcll(mask));
    uint32_t base = 0;
    if ((threadIdx.x & 63u) == 0u && total) base = atomicAdd(counter, total);
    base = __shfl(base, 0, 64);
    return base + lanes_below(mask);
}

// Wave-uniform work fetch over a segmented queue (kQueueSegments segments of
// `cap` slots, each with its length and cursor kCounterStride words apart).
// A wave starts at its own segment and moves on to the next one when that is
// exhausted; every lane of the wave calls take() together.
struct SegmentFetch {
    const uint32_t* counts;   // segment lengths of the queue
    uint32_t* cursors;        // segment cursors of the queue
    uint32_t cap, seg, tried, seg_n;

    __device__ __forceinline__ void init(const uint32_t* c, uint32_t* cur, uint32_t cap_, uint32_t start) {
        counts = c;
        cursors = cur;
        cap = cap_;
        seg = start % kQueueSegments;
        tried = 0;
        seg_n = counts[seg * kCounterStride];
    }
    __device__ __forceinline__ bool live() const { return tried < kQueueSegments; }
    __device__ __forceinline__ void advance() {
        seg = seg + 1 == kQueueSegments ? 0u : seg + 1;
        if (++tried < kQueueSegments) seg_n = counts[seg * kCounterStride];
    }
    // Reserves up to `want` (1..64 or a chunk) consecutive slots; returns how
    // many (0 once every segment is exhausted) and the first slot's index.
    __device__ __forceinline__ uint32_t take(uint32_t want, uint32_t& first) {
        while (tried < kQueueSegments) {
            if (seg_n != 0u) {
                uint32_t base = 0;
                if ((threadIdx.x & 63u) == 0u) base = atomicAdd(&cursors[seg * kCounterStride], want);
                base = __builtin_amdgcn_readlane(base, 0);
                if (base < seg_n) {
                    const uint32_t got = seg_n - base < want ? seg_n - base : want;
                    first = seg * cap + base;
                    if (got < want) advance();
                    return got;
                }
            }
            advance();
        }
        return 0u;
    }
};

}  // namespace quasi_hip
