// bvh_quant.hpp -- 64-byte BVH4 nodes with child boxes quantised against the
// node's own box, and the conservative slab test the traversal runs on them.
//
// The traversal kernels are bound by vector-memory address processing: a
// node step of the 128-byte BVH4 node issues seven 16-byte loads per lane,
// each one a cache-line lookup in the texture addresser.  The quantised node
// holds the same four children in 64 bytes, read with four loads:
//
//   q[0] = origin.x, origin.y, origin.z, scale.x
//   q[1] = scale.y, scale.z, lo.x bytes, hi.x bytes
//   q[2] = lo.y bytes, hi.y bytes, lo.z bytes, hi.z bytes
//   q[3] = child references (int bits, as in the 128-byte node)
//
// origin = the minimum corner of the union of the used children's boxes;
// scale_a = 2^e_a, the smallest power of two (e_a in [kQuantMinExp,
// kQuantMaxExp]) with 255 * scale_a >= the union's extent along a (with a
// margin of 2^-20); byte k of a lo (hi) word is child k's plane, rounded down
// (up) onto that grid, so origin + q * scale never cuts into a child's box.
// Unused slots (reference kEmptyChild) hold lo = 255, hi = 0, and the
// traversal rejects them by their reference.  Node indices are those of the
// 128-byte layout; a node lives at byte 64 * index.
//
// Both functions are shared by the device kernels and a host test: the
// quantiser computes in double, where every step it takes is exact or
// covered by the margins, and the slab test is written with contraction off,
// so host and device compute the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.hpp"

namespace quasi_hip {

constexpr int kQuantMinExp = -100;
constexpr int kQuantMaxExp = 60;
constexpr float kQuantSlack = 2.3841858e-7f;   // 2^-22

// Quantises one 128-byte BVH4 node (n: 8 float4) into q (4 float4).  False
// when an extent needs a scale above 2^kQuantMaxExp (the node is then
// written anyway and must not be used).
__host__ __device__ inline bool quantize_node4(const float4* n, float4* q) {
    float pl[6][4];
    for (int j = 0; j < 6; ++j) {
        pl[j][0] = n[j].x;
        pl[j][1] = n[j].y;
        pl[j][2] = n[j].z;
        pl[j][3] = n[j].w;
    }
    const float4 chf = n[6];
    int c[4];
    {
        union { float f; int i; } u;
        u.f = chf.x; c[0] = u.i;
        u.f = chf.y; c[1] = u.i;
        u.f = chf.z; c[2] = u.i;
        u.f = chf.w; c[3] = u.i;
    }
    bool ok = true;
    float org[3], scl[3];
    uint32_t wl[3], wh[3];
    for (int a = 0; a < 3; ++a) {
        float mn = __builtin_huge_valf(), mx = -__builtin_huge_valf();
        for (int k = 0; k < 4; ++k)
            if (c[k] != kEmptyChild) {
                mn = pl[2 * a][k] < mn ? pl[2 * a][k] : mn;
                mx = pl[2 * a + 1][k] > mx ? pl[2 * a + 1][k] : mx;
            }
        if (!(mn <= mx)) { mn = 0.f; mx = 0.f; }
        const double need = (double(mx) - double(mn)) * (1.0 + 0x1p-20);
        int e = kQuantMinExp;
        if (need > 0.0) {
            e = ilogb(need / 255.0);   // 255 * 2^e <= need (to rounding)
            e = e < kQuantMinExp ? kQuantMinExp : (e > kQuantMaxExp ? kQuantMaxExp : e);
            while (e < kQuantMaxExp && ldexp(255.0, e) < need) ++e;
        }
        if (ldexp(255.0, e) < need) ok = false;
        const double sc = ldexp(1.0, e);
        org[a] = mn;
        scl[a] = float(sc);
        wl[a] = wh[a] = 0u;
        for (int k = 0; k < 4; ++k) {
            uint32_t ql = 255u, qh = 0u;
            if (c[k] != kEmptyChild) {
                const double lo = floor((double(pl[2 * a][k]) - double(mn)) / sc - 0x1p-20);
                const double hi = ceil((double(pl[2 * a + 1][k]) - double(mn)) / sc + 0x1p-20);
                ql = uint32_t(lo < 0.0 ? 0.0 : (lo > 255.0 ? 255.0 : lo));
                qh = uint32_t(hi < 0.0 ? 0.0 : (hi > 255.0 ? 255.0 : hi));
            }
            wl[a] |= ql << (8 * k);
            wh[a] |= qh << (8 * k);
        }
    }
    union { uint32_t u; float f; } w[6];
    w[0].u = wl[0]; w[1].u = wh[0]; w[2].u = wl[1]; w[3].u = wh[1]; w[4].u = wl[2]; w[5].u = wh[2];
    q[0] = make_float4(org[0], org[1], org[2], scl[0]);
    q[1] = make_float4(scl[1], scl[2], w[0].f, w[1].f);
    q[2] = make_float4(w[2].f, w[3].f, w[4].f, w[5].f);
    q[3] = chf;
    return ok;
}

// Slab test of the four children of a quantised node for a ray with origin
// o and inverse direction inv: tn[k] <= tf[k] unless child k's box (as stored
// in the 128-byte node) is missed on [tmin, tmax].  Plane q of axis a lies at
// ray distance q * (scale_a * inv_a) + (origin_a - o_a) * inv_a; the first
// product is exact (a power-of-two scale), the second term's two roundings
// are covered by a slack of 2^-22 of its magnitude (subtracted for entry
// planes, added for exit planes), and the fused multiply-add's own rounding
// by the relative widening of tf, as in the 128-byte test.  The entry planes
// of an axis are the lo bytes when inv_a >= 0, else the hi bytes.
__host__ __device__ inline void slab_quant4(const float4& q0, const float4& q1, const float4& q2, float ox, float oy,
                                            float oz, float ix, float iy, float iz, float tmin, float tmax,
                                            float (&tn)[4], float (&tf)[4]) {
#pragma clang fp contract(off)
    const float ax = q0.w * ix, ay = q1.x * iy, az = q1.y * iz;
    const float bx = (q0.x - ox) * ix, by = (q0.y - oy) * iy, bz = (q0.z - oz) * iz;
    // b -/+ |b| 2^-22 in one rounding each
    const float bxl = fmaf(-fabsf(bx), kQuantSlack, bx), bxh = fmaf(fabsf(bx), kQuantSlack, bx);
    const float byl = fmaf(-fabsf(by), kQuantSlack, by), byh = fmaf(fabsf(by), kQuantSlack, by);
    const float bzl = fmaf(-fabsf(bz), kQuantSlack, bz), bzh = fmaf(fabsf(bz), kQuantSlack, bz);
    union { float f; uint32_t u; } lx{q1.z}, hx{q1.w}, ly{q2.x}, hy{q2.y}, lz{q2.z}, hz{q2.w};
    const uint32_t wex = ix < 0.f ? hx.u : lx.u, wfx = ix < 0.f ? lx.u : hx.u;
    const uint32_t wey = iy < 0.f ? hy.u : ly.u, wfy = iy < 0.f ? ly.u : hy.u;
    const uint32_t wez = iz < 0.f ? hz.u : lz.u, wfz = iz < 0.f ? lz.u : hz.u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float x0 = fmaf(float((wex >> (8 * k)) & 0xFFu), ax, bxl);
        const float x1 = fmaf(float((wfx >> (8 * k)) & 0xFFu), ax, bxh);
        const float y0 = fmaf(float((wey >> (8 * k)) & 0xFFu), ay, byl);
        const float y1 = fmaf(float((wfy >> (8 * k)) & 0xFFu), ay, byh);
        const float z0 = fmaf(float((wez >> (8 * k)) & 0xFFu), az, bzl);
        const float z1 = fmaf(float((wfz >> (8 * k)) & 0xFFu), az, bzh);
        tn[k] = fmaxf(fmaxf(x0, y0), fmaxf(z0, tmin));
        tf[k] = fminf(fminf(x1, y1), fminf(z1, tmax)) * kSlabWiden;
    }
}

}  // namespace quasi_hip
