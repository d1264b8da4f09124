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

// Sum over the 64 lanes of a wave (all lanes must be present).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace quasi_hip
