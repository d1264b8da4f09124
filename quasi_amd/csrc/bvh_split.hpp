// bvh_split.hpp -- spatial pre-split of large triangles into BVH references.
//
// No reference counterpart: the reference intersects at most 32 quads by brute
// force (backends/metal/plugin.mm:199-212).  Before the binned-SAH build, a
// triangle whose bounding box has a half surface area (dx dy + dy dz + dz dx)
// above the threshold T is cut at the middle of its box's longest axis, and
// each half's box is shrunk to (a conservative box of) the part of the
// triangle inside it; a half whose box is still above T is cut once more, so a
// triangle gives at most four references.  Each reference is a primitive of the build with its own box, and
// every leaf position holds a full copy of its triangle's record, so a ray
// meets the same triangle (same record, same intersection) through whichever
// reference it reaches and the closest hit does not change.  Emitting
// triangles are never cut (the light table names one position per light), nor
// are axis-aligned flat ones.
// T is the half area of the average leaf cell, half_area(scene box) / N^(2/3)
// (split_threshold).  The host and the GPU builders run the same functions
// with contraction off, so both get the same references bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace quasi_hip {

constexpr uint32_t kMaxPieces = 4;

struct PieceBox {
    float lo[3], hi[3];
};

__host__ __device__ inline float half_area_box(const PieceBox& b) {
#pragma clang fp contract(off)
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

__host__ __device__ inline PieceBox triangle_box(const float* a, const float* b, const float* c) {
    PieceBox r;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = fminf(a[k], fminf(b[k], c[k]));
        r.hi[k] = fmaxf(a[k], fmaxf(b[k], c[k]));
    }
    return r;
}

// Box of the part of triangle (a, b, c) on one side of the plane x_k = w
// (upper: x_k >= w, else x_k <= w): its vertices on that side and the
// crossings of its edges with the plane, grown on each axis by 2^-20 of the
// triangle's largest coordinate magnitude there (a crossing is computed from
// the exact vertices, so it is off by a few ulps at most) and held to the
// plane on axis k.  False when no part is on that side.
__host__ __device__ inline bool clip_halfspace(const float* a, const float* b, const float* c, int k, float w,
                                               bool upper, PieceBox& out) {
#pragma clang fp contract(off)
    const float* v[3] = {a, b, c};
    float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
    float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
    bool any = false;
    for (int i = 0; i < 3; ++i) {
        const float* p = v[i];
        const float* q = v[i < 2 ? i + 1 : 0];
        const bool pin = upper ? p[k] >= w : p[k] <= w;
        const bool qin = upper ? q[k] >= w : q[k] <= w;
        if (pin) {
            any = true;
            for (int j = 0; j < 3; ++j) { lo[j] = fminf(lo[j], p[j]); hi[j] = fmaxf(hi[j], p[j]); }
        }
        if (pin != qin) {
            const float u = (w - p[k]) / (q[k] - p[k]);
            for (int j = 0; j < 3; ++j) {
                const float x = j == k ? w : p[j] + u * (q[j] - p[j]);
                lo[j] = fminf(lo[j], x);
                hi[j] = fmaxf(hi[j], x);
            }
        }
    }
    if (!any) return false;
    for (int j = 0; j < 3; ++j) {
        const float e = fmaxf(fabsf(a[j]), fmaxf(fabsf(b[j]), fabsf(c[j]))) * 0x1p-20f;
        out.lo[j] = lo[j] - e;
        out.hi[j] = hi[j] + e;
    }
    if (upper) out.lo[k] = fmaxf(out.lo[k], w);
    else out.hi[k] = fminf(out.hi[k], w);
    return true;
}

// The two halves of piece box `box` of triangle (a, b, c), cut at the middle
// of the box's longest axis (the lowest such axis on ties): each is the box of
// the triangle's part on that side of the cut intersected with `box`, so it
// holds the triangle's part inside the half whenever `box` holds the part
// inside the whole.  Returns how many halves are nonempty (in order lower,
// upper); 0 when the box cannot be cut (its middle rounds onto an end).
__host__ __device__ inline int cut_piece(const float* a, const float* b, const float* c, const PieceBox& box,
                                         PieceBox (&half)[2]) {
#pragma clang fp contract(off)
    const float d[3] = {box.hi[0] - box.lo[0], box.hi[1] - box.lo[1], box.hi[2] - box.lo[2]};
    int ax = d[1] > d[0] ? 1 : 0;
    if (d[2] > d[ax]) ax = 2;
    const float mid = 0.5f * (box.lo[ax] + box.hi[ax]);
    if (!(mid > box.lo[ax] && mid < box.hi[ax])) return 0;
    int n = 0;
    for (int side = 0; side < 2; ++side) {
        PieceBox h;
        if (!clip_halfspace(a, b, c, ax, mid, side == 1, h)) continue;
        bool empty = false;
        for (int j = 0; j < 3; ++j) {
            h.lo[j] = fmaxf(h.lo[j], box.lo[j]);
            h.hi[j] = fminf(h.hi[j], box.hi[j]);
            empty = empty || h.lo[j] > h.hi[j];
        }
        if (!empty) half[n++] = h;
    }
    return n;
}

// The references of triangle (a, b, c): 1 to kMaxPieces boxes in a fixed order
// (lower half before upper, depth first).  may_split false (an emitter), a box
// at most T, or a box of zero extent on some axis (an axis-aligned flat
// triangle such as a wall, whose box is already tight across it; cutting the
// walls of the batch's rooms cost 3% of the 64-scene batch) gives the
// triangle's own box.
__host__ __device__ inline uint32_t split_triangle(const float* a, const float* b, const float* c, float T,
                                                   bool may_split, PieceBox (&out)[kMaxPieces]) {
    const PieceBox box = triangle_box(a, b, c);
    const bool flat = box.lo[0] == box.hi[0] || box.lo[1] == box.hi[1] || box.lo[2] == box.hi[2];
    if (!may_split || flat || !(half_area_box(box) > T)) {
        out[0] = box;
        return 1;
    }
    PieceBox half[2];
    const int nh = cut_piece(a, b, c, box, half);
    if (nh == 0) {
        out[0] = box;
        return 1;
    }
    uint32_t n = 0;
    for (int i = 0; i < nh; ++i) {
        PieceBox q[2];
        const int nq = half_area_box(half[i]) > T ? cut_piece(a, b, c, half[i], q) : 0;
        if (nq == 0) {
            out[n++] = half[i];
        } else {
            for (int j = 0; j < nq; ++j) out[n++] = q[j];
        }
    }
    return n;
}

// The threshold T for N triangles inside the scene box [lo, hi]: the half
// area of one of N equal cubic cells of the box's volume's shape,
// half_area(box) / N^(2/3).  Host only (both builders take it from the host).
inline float split_threshold(const float* lo, const float* hi, uint32_t n) {
    const double dx = double(hi[0]) - double(lo[0]), dy = double(hi[1]) - double(lo[1]),
                 dz = double(hi[2]) - double(lo[2]);
    const double a = dx * dy + dy * dz + dz * dx;
    const double c = std::cbrt(double(n > 0 ? n : 1u));
    const double t = a / (c * c);
    return t > 0 && t < 3.0e38 ? float(t) : 3.0e38f;
}

}  // namespace quasi_hip
