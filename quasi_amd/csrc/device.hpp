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

// Widening of the far slab distance so that rounding in (lo - o) * (1/d) never
// culls a box that the ray touches (relative error of each slab distance is a
// few ulps).
constexpr float kSlabWiden = 1.000001f;

// Stack-based closest-hit traversal of the pair-node BVH2.  The per-lane stack
// lives in LDS, interleaved by lane (entry i of lane l at stk[i * STRIDE]), so a
// wave's pushes and pops hit 64 distinct banks.  Children are visited near
// first.  On equal t the larger input index wins (the reference's linear loop,
// plugin.mm:204-209 with the inclusive `t > t_max` test of plugin.mm:177).
template <int STRIDE>
__device__ __forceinline__ TraceHit trace_closest(const float4* __restrict__ nodes,
                                                  const float4* __restrict__ tris, Ray r, float tmin,
                                                  float tmax, int* __restrict__ stk) {
    TraceHit best{tmax, -1, 0u};
    const f3 inv = mk(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
    int sp = 0;
    int node = 0;
    for (;;) {
        while (node >= 0) {
            const float4 a = nodes[4 * node + 0];
            const float4 b = nodes[4 * node + 1];
            const float4 c = nodes[4 * node + 2];
            const float4 dd = nodes[4 * node + 3];
            const int cl = __float_as_int(dd.x);
            const int cr = __float_as_int(dd.y);
            // left child
            float lx0 = (a.x - r.o.x) * inv.x, lx1 = (a.y - r.o.x) * inv.x;
            float ly0 = (a.z - r.o.y) * inv.y, ly1 = (a.w - r.o.y) * inv.y;
            float lz0 = (c.x - r.o.z) * inv.z, lz1 = (c.y - r.o.z) * inv.z;
            float rx0 = (b.x - r.o.x) * inv.x, rx1 = (b.y - r.o.x) * inv.x;
            float ry0 = (b.z - r.o.y) * inv.y, ry1 = (b.w - r.o.y) * inv.y;
            float rz0 = (c.z - r.o.z) * inv.z, rz1 = (c.w - r.o.z) * inv.z;
            float tnl = fmaxf(fmaxf(fminf(lx0, lx1), fminf(ly0, ly1)), fmaxf(fminf(lz0, lz1), tmin));
            float tfl = fminf(fminf(fmaxf(lx0, lx1), fmaxf(ly0, ly1)), fminf(fmaxf(lz0, lz1), best.t)) * kSlabWiden;
            float tnr = fmaxf(fmaxf(fminf(rx0, rx1), fminf(ry0, ry1)), fmaxf(fminf(rz0, rz1), tmin));
            float tfr = fminf(fminf(fmaxf(rx0, rx1), fmaxf(ry0, ry1)), fminf(fmaxf(rz0, rz1), best.t)) * kSlabWiden;
            const bool hl = tnl <= tfl;
            const bool hr = tnr <= tfr;
            if (hl && hr) {
                const bool left_first = tnl <= tnr;
                stk[sp * STRIDE] = left_first ? cr : cl;
                ++sp;
                node = left_first ? cl : cr;
            } else if (hl) {
                node = cl;
            } else if (hr) {
                node = cr;
            } else {
                if (sp == 0) return best;
                --sp;
                node = stk[sp * STRIDE];
            }
        }
        // leaf
        const uint32_t v = ~uint32_t(node);
        const int first = int(v >> 2);
        const int count = int(v & 3u) + 1;
        for (int i = first; i < first + count; ++i) {
            const float4 t0 = tris[3 * i + 0];
            const float4 t1 = tris[3 * i + 1];
            const float4 t2 = tris[3 * i + 2];
            const f3 e1 = xyz(t1), e2 = xyz(t2);
            const f3 pv = cross(r.d, e2);
            const float det = dot(e1, pv);
            if (det == 0.0f) continue;
            const float idet = 1.0f / det;
            const f3 s = r.o - xyz(t0);
            const float u = dot(s, pv) * idet;
            if (u < 0.0f || u > 1.0f) continue;
            const f3 q = cross(s, e1);
            const float vv = dot(r.d, q) * idet;
            if (vv < 0.0f || u + vv > 1.0f) continue;
            const float t = dot(e2, q) * idet;
            if (t < tmin || t > best.t) continue;
            const uint32_t id = __float_as_uint(t0.w);
            if (best.tri < 0 || t < best.t || id > best.id) {
                best.t = t;
                best.tri = i;
                best.id = id;
            }
        }
        if (sp == 0) return best;
        --sp;
        node = stk[sp * STRIDE];
    }
}

// Any-hit traversal for shadow rays over [tmin, tmax]: returns at the first
// intersection found.
template <int STRIDE>
__device__ __forceinline__ bool trace_any(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                          Ray r, float tmin, float tmax, int* __restrict__ stk) {
    const f3 inv = mk(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
    int sp = 0;
    int node = 0;
    for (;;) {
        while (node >= 0) {
            const float4 a = nodes[4 * node + 0];
            const float4 b = nodes[4 * node + 1];
            const float4 c = nodes[4 * node + 2];
            const float4 dd = nodes[4 * node + 3];
            const int cl = __float_as_int(dd.x);
            const int cr = __float_as_int(dd.y);
            float lx0 = (a.x - r.o.x) * inv.x, lx1 = (a.y - r.o.x) * inv.x;
            float ly0 = (a.z - r.o.y) * inv.y, ly1 = (a.w - r.o.y) * inv.y;
            float lz0 = (c.x - r.o.z) * inv.z, lz1 = (c.y - r.o.z) * inv.z;
            float rx0 = (b.x - r.o.x) * inv.x, rx1 = (b.y - r.o.x) * inv.x;
            float ry0 = (b.z - r.o.y) * inv.y, ry1 = (b.w - r.o.y) * inv.y;
            float rz0 = (c.z - r.o.z) * inv.z, rz1 = (c.w - r.o.z) * inv.z;
            float tnl = fmaxf(fmaxf(fminf(lx0, lx1), fminf(ly0, ly1)), fmaxf(fminf(lz0, lz1), tmin));
            float tfl = fminf(fminf(fmaxf(lx0, lx1), fmaxf(ly0, ly1)), fminf(fmaxf(lz0, lz1), tmax)) * kSlabWiden;
            float tnr = fmaxf(fmaxf(fminf(rx0, rx1), fminf(ry0, ry1)), fmaxf(fminf(rz0, rz1), tmin));
            float tfr = fminf(fminf(fmaxf(rx0, rx1), fmaxf(ry0, ry1)), fminf(fmaxf(rz0, rz1), tmax)) * kSlabWiden;
            const bool hl = tnl <= tfl;
            const bool hr = tnr <= tfr;
            if (hl && hr) {
                const bool left_first = tnl <= tnr;
                stk[sp * STRIDE] = left_first ? cr : cl;
                ++sp;
                node = left_first ? cl : cr;
            } else if (hl) {
                node = cl;
            } else if (hr) {
                node = cr;
            } else {
                if (sp == 0) return false;
                --sp;
                node = stk[sp * STRIDE];
            }
        }
        const uint32_t v = ~uint32_t(node);
        const int first = int(v >> 2);
        const int count = int(v & 3u) + 1;
        for (int i = first; i < first + count; ++i) {
            const float4 t0 = tris[3 * i + 0];
            const float4 t1 = tris[3 * i + 1];
            const float4 t2 = tris[3 * i + 2];
            const f3 e1 = xyz(t1), e2 = xyz(t2);
            const f3 pv = cross(r.d, e2);
            const float det = dot(e1, pv);
            if (det == 0.0f) continue;
            const float idet = 1.0f / det;
            const f3 s = r.o - xyz(t0);
            const float u = dot(s, pv) * idet;
            if (u < 0.0f || u > 1.0f) continue;
            const f3 q = cross(s, e1);
            const float vv = dot(r.d, q) * idet;
            if (vv < 0.0f || u + vv > 1.0f) continue;
            const float t = dot(e2, q) * idet;
            if (t < tmin || t > tmax) continue;
            return true;
        }
        if (sp == 0) return false;
        --sp;
        node = stk[sp * STRIDE];
    }
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

// One vertex of either integrator at bounce b for a ray that hit BVH triangle
// `tri` at distance t.  Updates throughput, radiance and the random stream.
__device__ __forceinline__ void shade_hit(const PassParams& P, const float4* __restrict__ tris,
                                          const float4* __restrict__ mats, uint32_t b, const Ray& ray,
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
    if (P.num_lights) {
        // light choice: first entry with r < cdf (the last when none)
        const float rs = random_float(rng);
        uint32_t lo = 0, hi = P.num_lights;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (rs < P.lights[mid].cdf) hi = mid; else lo = mid + 1;
        }
        if (lo >= P.num_lights) lo = P.num_lights - 1;
        const LightEntry le = P.lights[lo];
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
__device__ __forceinline__ Ray camera_ray(const PassParams& P, uint32_t x, uint32_t y, uint32_t& rng) {
    const f3 llo = mk(P.cam.llo[0], P.cam.llo[1], P.cam.llo[2]);
    const f3 hor = mk(P.cam.horizontal[0], P.cam.horizontal[1], P.cam.horizontal[2]);
    const f3 ver = mk(P.cam.vertical[0], P.cam.vertical[1], P.cam.vertical[2]);
    const float u0 = (float(x) + 0.5f) / float(P.width);
    const float v0 = 1.0f - (float(y) + 0.5f) / float(P.height);
    const float jx = random_float(rng);
    const float jy = random_float(rng);
    const float u = u0 + (jx - 0.5f) * 0.001f;
    const float v = v0 + (jy - 0.5f) * 0.001f;
    return Ray{mk(P.cam.origin[0], P.cam.origin[1], P.cam.origin[2]), normalize(llo + hor * u + ver * v)};
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

// Wave-uniform work fetch: lane 0 reserves the next 64 items of a queue of n.
__device__ __forceinline__ bool fetch_chunk(uint32_t* cursor, uint32_t n, uint32_t& base) {
    uint32_t b = 0;
    if ((threadIdx.x & 63u) == 0u) b = atomicAdd(cursor, 64u);
    base = __shfl(b, 0, 64);
    return base < n;
}

}  // namespace quasi_hip
