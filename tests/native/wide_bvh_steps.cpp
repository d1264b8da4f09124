// wide_bvh_steps.cpp -- test-only measurement of what a wider BVH saves.
//
// The plugin's host build gives a BVH2; it is collapsed to 2-, 4- and 8-wide
// trees the way the BVH4 layout is (repeatedly opening the child of largest
// surface area), and a scalar front-to-back closest-hit traversal counts node
// steps and child-box tests per ray, for camera rays of the soup's camera and
// for one cosine-sampled secondary ray from each camera hit.  Reads triangles
// from a raw float file (9 floats per triangle); prints one line per width:
//   width W nodes N cam_steps C cam_boxes B sec_steps S sec_boxes X sec_tris T
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "bvh_build.hpp"

using namespace quasi_hip;

struct Box { float lo[3], hi[3]; };
struct WNode { std::vector<Box> box; std::vector<int> ref; };   // ref >= 0 internal wnode, < 0 leaf ~(first<<2|cnt-1)

const std::vector<float>* g_nodes;
static Box child_box(int node, int side) {
    const float* n = g_nodes->data() + 16 * size_t(node);
    Box b;
    if (side == 0) { b.lo[0] = n[0]; b.hi[0] = n[1]; b.lo[1] = n[2]; b.hi[1] = n[3]; b.lo[2] = n[8]; b.hi[2] = n[9]; }
    else { b.lo[0] = n[4]; b.hi[0] = n[5]; b.lo[1] = n[6]; b.hi[1] = n[7]; b.lo[2] = n[10]; b.hi[2] = n[11]; }
    return b;
}
static int child_ref(int node, int side) { int r; std::memcpy(&r, g_nodes->data() + 16 * size_t(node) + 12 + side, 4); return r; }
static float area(const Box& b) {
    float d[3]; for (int k = 0; k < 3; ++k) d[k] = b.hi[k] - b.lo[k];
    return d[0] * d[1] + d[1] * d[2] + d[0] * d[2];
}

std::vector<WNode> wn;
int collapse(int node, int width) {   // node: BVH2 internal index -> wide node index
    struct C { Box b; int ref; };
    std::vector<C> ch = {{child_box(node, 0), child_ref(node, 0)}, {child_box(node, 1), child_ref(node, 1)}};
    while (int(ch.size()) < width) {
        int best = -1; float ba = -1;
        for (int i = 0; i < int(ch.size()); ++i) if (ch[i].ref >= 0 && area(ch[i].b) > ba) { ba = area(ch[i].b); best = i; }
        if (best < 0) break;
        int n = ch[best].ref;
        ch[best] = {child_box(n, 0), child_ref(n, 0)};
        ch.push_back({child_box(n, 1), child_ref(n, 1)});
    }
    int id = int(wn.size());
    wn.emplace_back();
    std::vector<Box> bx; std::vector<int> rf;
    for (auto& c : ch) { bx.push_back(c.b); rf.push_back(c.ref >= 0 ? collapse(c.ref, width) : c.ref); }
    wn[id].box = bx; wn[id].ref = rf;
    return id;
}

std::vector<float> tri;   // 9 floats per BVH-order triangle
bool hit_tri(int i, const float* o, const float* d, float tmin, float& t) {
    const float* p = &tri[9 * size_t(i)];
    float e1[3], e2[3], s[3], pv[3], q[3];
    for (int k = 0; k < 3; ++k) { e1[k] = p[3 + k] - p[k]; e2[k] = p[6 + k] - p[k]; s[k] = o[k] - p[k]; }
    pv[0] = d[1] * e2[2] - d[2] * e2[1]; pv[1] = d[2] * e2[0] - d[0] * e2[2]; pv[2] = d[0] * e2[1] - d[1] * e2[0];
    float det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det == 0) return false;
    float inv = 1 / det;
    float u = (s[0] * pv[0] + s[1] * pv[1] + s[2] * pv[2]) * inv;
    if (u < 0 || u > 1) return false;
    q[0] = s[1] * e1[2] - s[2] * e1[1]; q[1] = s[2] * e1[0] - s[0] * e1[2]; q[2] = s[0] * e1[1] - s[1] * e1[0];
    float v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (v < 0 || u + v > 1) return false;
    float tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    if (tt < tmin || tt > t) return false;
    t = tt;
    return true;
}

struct Stats { double steps = 0, boxes = 0, tris = 0, rays = 0; };
float trace(const float* o, const float* d, Stats& st, int& hit) {
    float inv[3]; for (int k = 0; k < 3; ++k) inv[k] = 1.0f / d[k];
    float best = 1e30f; hit = -1;
    std::vector<int> stack; stack.push_back(0);
    st.rays++;
    while (!stack.empty()) {
        int r = stack.back(); stack.pop_back();
        if (r < 0) {
            unsigned v = ~unsigned(r); int first = v >> 2, cnt = (v & 3) + 1;
            for (int i = first; i < first + cnt; ++i) { st.tris++; if (hit_tri(i, o, d, 1e-3f, best)) hit = i; }
            continue;
        }
        st.steps++;
        const WNode& n = wn[r];
        std::vector<std::pair<float, int>> h;
        for (size_t c = 0; c < n.ref.size(); ++c) {
            st.boxes++;
            float tn = 1e-3f, tf = best;
            for (int k = 0; k < 3; ++k) {
                float a = (n.box[c].lo[k] - o[k]) * inv[k], b = (n.box[c].hi[k] - o[k]) * inv[k];
                tn = std::max(tn, std::min(a, b)); tf = std::min(tf, std::max(a, b));
            }
            if (tn <= tf) h.push_back({tn, n.ref[c]});
        }
        std::sort(h.begin(), h.end(), [](auto& a, auto& b) { return a.first > b.first; });
        for (auto& x : h) stack.push_back(x.second);
    }
    return best;
}

int main(int argc, char** argv) {
    FILE* f = std::fopen(argv[1], "rb");
    std::fseek(f, 0, SEEK_END); long sz = std::ftell(f); std::fseek(f, 0, SEEK_SET);
    size_t nt = sz / 36;
    std::vector<float> raw(9 * nt);
    std::fread(raw.data(), 4, raw.size(), f); std::fclose(f);
    std::vector<float> p0(3 * nt), p1(3 * nt), p2(3 * nt);
    for (size_t t = 0; t < nt; ++t) for (int k = 0; k < 3; ++k) { p0[3*t+k] = raw[9*t+k]; p1[3*t+k] = raw[9*t+3+k]; p2[3*t+k] = raw[9*t+6+k]; }
    BvhBuildResult b = build_bvh(p0.data(), p1.data(), p2.data(), uint32_t(nt), 8);
    g_nodes = &b.nodes;
    tri.resize(9 * nt);
    for (size_t i = 0; i < nt; ++i) std::memcpy(&tri[9 * i], &raw[9 * size_t(b.order[i])], 36);
    for (int width : {2, 4, 8}) {
        wn.clear();
        collapse(0, width);
        Stats cam, sec;
        std::mt19937 rng(1);
        std::uniform_real_distribution<float> U(0, 1);
        const float eye[3] = {0, 0, 3.5f};
        const float th = std::tan(20.0f * 3.14159265f / 180.0f);
        const int n = argc > 2 ? std::atoi(argv[2]) : 20000;
        for (int i = 0; i < n; ++i) {
            float u = U(rng), v = U(rng);
            float d[3] = {(2 * u - 1) * th * 16.f / 9.f, (2 * v - 1) * th, -1};
            float l = std::sqrt(d[0]*d[0]+d[1]*d[1]+d[2]*d[2]); for (float& x : d) x /= l;
            int h; float t = trace(eye, d, cam, h);
            if (h < 0) continue;
            // a secondary ray from the hit point, cosine about the facing normal
            const float* p = &tri[9 * size_t(h)];
            float e1[3], e2[3], nn[3];
            for (int k = 0; k < 3; ++k) { e1[k] = p[3+k]-p[k]; e2[k] = p[6+k]-p[k]; }
            nn[0] = e1[1]*e2[2]-e1[2]*e2[1]; nn[1] = e1[2]*e2[0]-e1[0]*e2[2]; nn[2] = e1[0]*e2[1]-e1[1]*e2[0];
            float nl = std::sqrt(nn[0]*nn[0]+nn[1]*nn[1]+nn[2]*nn[2]); for (float& x : nn) x /= nl;
            if (nn[0]*d[0]+nn[1]*d[1]+nn[2]*d[2] > 0) for (float& x : nn) x = -x;
            float o2[3]; for (int k = 0; k < 3; ++k) o2[k] = eye[k] + d[k] * t + nn[k] * 1e-3f;
            float r1 = U(rng), r2 = U(rng), phi = 6.2831853f * r1, ct = std::sqrt(1 - r2), stt = std::sqrt(r2);
            float a[3] = {std::fabs(nn[0]) > 0.9f ? 0.f : 1.f, std::fabs(nn[0]) > 0.9f ? 1.f : 0.f, 0};
            float vv[3] = {nn[1]*a[2]-nn[2]*a[1], nn[2]*a[0]-nn[0]*a[2], nn[0]*a[1]-nn[1]*a[0]};
            float vl = std::sqrt(vv[0]*vv[0]+vv[1]*vv[1]+vv[2]*vv[2]); for (float& x : vv) x /= vl;
            float uu[3] = {nn[1]*vv[2]-nn[2]*vv[1], nn[2]*vv[0]-nn[0]*vv[2], nn[0]*vv[1]-nn[1]*vv[0]};
            float d2[3]; for (int k = 0; k < 3; ++k) d2[k] = uu[k]*std::cos(phi)*stt + vv[k]*std::sin(phi)*stt + nn[k]*ct;
            int h2; trace(o2, d2, sec, h2);
        }
        std::printf("width %d nodes %zu cam_steps %.3f cam_boxes %.3f sec_steps %.3f sec_boxes %.3f sec_tris %.3f\n",
                    width, wn.size(), cam.steps / cam.rays, cam.boxes / cam.rays, sec.steps / sec.rays,
                    sec.boxes / sec.rays, sec.tris / sec.rays);
    }
    return 0;
}
