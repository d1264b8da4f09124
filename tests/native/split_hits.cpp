// split_hits.cpp -- test-only check that splitting triangles into references
// (quasi_amd/csrc/bvh_split.hpp) leaves every closest hit unchanged.
//
// Reads triangles from a raw float file (9 floats per triangle), builds the
// host BVH (bvh_build.cpp) once over the triangles and once over their
// references (the split threshold of the scene box, every triangle allowed to
// split), and traces the same random rays through both with a scalar walk: a
// slab test in double with a relative margin of 1e-6 and the Moller-Trumbore
// test in float of the traversal kernels, ties to the larger triangle index.
// Prints "triangles N references R rays K differing D".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "bvh_build.hpp"
#include "bvh_split.hpp"

using namespace quasi_hip;

namespace {

std::vector<float> g_tri;   // 9 floats per input triangle

bool hit_triangle(size_t t, const float* o, const float* d, float& best) {
    const float* p = &g_tri[9 * t];
    float e1[3], e2[3], s[3], pv[3], q[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = p[3 + k] - p[k];
        e2[k] = p[6 + k] - p[k];
        s[k] = o[k] - p[k];
    }
    pv[0] = d[1] * e2[2] - d[2] * e2[1];
    pv[1] = d[2] * e2[0] - d[0] * e2[2];
    pv[2] = d[0] * e2[1] - d[1] * e2[0];
    const float det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det == 0.f) return false;
    const float inv = 1.f / det;
    const float u = (s[0] * pv[0] + s[1] * pv[1] + s[2] * pv[2]) * inv;
    if (u < 0.f || u > 1.f) return false;
    q[0] = s[1] * e1[2] - s[2] * e1[1];
    q[1] = s[2] * e1[0] - s[0] * e1[2];
    q[2] = s[0] * e1[1] - s[1] * e1[0];
    const float v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    const float t_hit = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    if (t_hit < 1e-3f || t_hit > best) return false;
    best = t_hit;
    return true;
}

struct Tree {
    BvhBuildResult bvh;
    std::vector<uint32_t> ref_tri;   // empty: leaves hold triangles
};

bool box_hit(const float* lo, const float* hi, const float* o, const double* inv, double tmax) {
    double tn = 0.0, tf = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = (double(lo[k]) - o[k]) * inv[k], b = (double(hi[k]) - o[k]) * inv[k];
        if (a > b) std::swap(a, b);
        tn = std::max(tn, a);
        tf = std::min(tf, b);
    }
    return tn <= tf * (1.0 + 1e-6) + 1e-9;
}

void trace(const Tree& tr, const float* o, const float* d, float& best, long& hit) {
    double inv[3];
    for (int k = 0; k < 3; ++k) inv[k] = 1.0 / d[k];
    const std::vector<float>& nodes = tr.bvh.nodes;   // BVH2 pair nodes, 16 floats each (common.hpp)
    std::vector<int> stack{0};
    best = 1e30f;
    hit = -1;
    while (!stack.empty()) {
        const float* nd = &nodes[16 * size_t(stack.back())];
        stack.pop_back();
        for (int side = 0; side < 2; ++side) {
            const float lo[3] = {nd[4 * side], nd[4 * side + 2], nd[8 + 2 * side]};
            const float hi[3] = {nd[4 * side + 1], nd[4 * side + 3], nd[9 + 2 * side]};
            int ref;
            std::memcpy(&ref, &nd[12 + side], 4);
            if (!box_hit(lo, hi, o, inv, best)) continue;
            if (ref >= 0) {
                stack.push_back(ref);
                continue;
            }
            const unsigned v = ~unsigned(ref), first = v >> 2, count = (v & 3u) + 1u;
            for (unsigned i = first; i < first + count; ++i) {
                const size_t t = tr.ref_tri.empty() ? tr.bvh.order[i] : tr.ref_tri[tr.bvh.order[i]];
                float tt = best;
                if (hit_triangle(t, o, d, tt) && (tt < best || long(t) > hit)) {
                    best = tt;
                    hit = long(t);
                }
            }
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::fseek(f, 0, SEEK_END);
    const size_t nt = size_t(std::ftell(f)) / 36;
    std::fseek(f, 0, SEEK_SET);
    g_tri.resize(9 * nt);
    if (std::fread(g_tri.data(), 4, g_tri.size(), f) != g_tri.size()) return 2;
    std::fclose(f);
    std::vector<float> p0(3 * nt), p1(3 * nt), p2(3 * nt);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            p0[3 * t + k] = g_tri[9 * t + k];
            p1[3 * t + k] = g_tri[9 * t + 3 + k];
            p2[3 * t + k] = g_tri[9 * t + 6 + k];
            for (int c = 0; c < 3; ++c) {
                lo[k] = std::min(lo[k], g_tri[9 * t + 3 * c + k]);
                hi[k] = std::max(hi[k], g_tri[9 * t + 3 * c + k]);
            }
        }
    Tree plain, split;
    plain.bvh = build_bvh(p0.data(), p1.data(), p2.data(), uint32_t(nt), 8);
    const float T = split_threshold(lo, hi, uint32_t(nt));
    std::vector<float> rlo, rhi;
    for (size_t t = 0; t < nt; ++t) {
        PieceBox pc[kMaxPieces];
        const uint32_t c = split_triangle(&p0[3 * t], &p1[3 * t], &p2[3 * t], T, true, pc);
        for (uint32_t j = 0; j < c; ++j) {
            rlo.insert(rlo.end(), pc[j].lo, pc[j].lo + 3);
            rhi.insert(rhi.end(), pc[j].hi, pc[j].hi + 3);
            split.ref_tri.push_back(uint32_t(t));
        }
    }
    split.bvh = build_bvh(rlo.data(), rhi.data(), rlo.data(), uint32_t(split.ref_tri.size()), 8);
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> U(-1.2f, 1.2f), D(-1.f, 1.f);
    const long n = std::atol(argv[2]);
    long differing = 0;
    for (long i = 0; i < n; ++i) {
        const float o[3] = {U(rng), U(rng), U(rng)};
        float d[3] = {D(rng), D(rng), D(rng)};
        const float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (float& x : d) x /= l;
        float ta, tb;
        long ha, hb;
        trace(plain, o, d, ta, ha);
        trace(split, o, d, tb, hb);
        if (ha != hb || ta != tb) ++differing;
    }
    std::printf("triangles %zu references %zu rays %ld differing %ld\n", nt, split.ref_tri.size(), n, differing);
    return 0;
}
