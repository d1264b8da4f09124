// scene_cornell.cpp -- see scene_cornell.hpp.
#include "scene_cornell.hpp"

#include <cmath>

namespace quasi_hip {
namespace {

struct V {
    float x, y, z;
};
V add(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

struct Mesh {
    std::vector<float>& verts;
    std::vector<uint32_t>& indices;
    std::vector<float>& albedo;
    std::vector<float>& emission;

    void quad(V o, V u, V v, V alb, V emi) {
        const uint32_t base = uint32_t(verts.size() / 3);
        const V corners[4] = {o, add(o, u), add(add(o, u), v), add(o, v)};
        for (const V& c : corners) { verts.push_back(c.x); verts.push_back(c.y); verts.push_back(c.z); }
        const uint32_t tri[6] = {0, 1, 2, 0, 2, 3};
        for (uint32_t k : tri) indices.push_back(base + k);
        for (int t = 0; t < 2; ++t) {
            albedo.push_back(alb.x); albedo.push_back(alb.y); albedo.push_back(alb.z);
            emission.push_back(emi.x); emission.push_back(emi.y); emission.push_back(emi.z);
        }
    }
};

// Rotation about +y by `deg` degrees, evaluated in float as cornell_box.hpp:32-37.
V rotate_y(V p, float deg) {
    const float rad = deg * 3.14159265359f / 180.0f;
    const float c = std::cos(rad), s = std::sin(rad);
    return {p.x * c + p.z * s, p.y, -p.x * s + p.z * c};
}

// Five faces of a box standing on y = 0 (no bottom face), cornell_box.hpp:45-90.
void box(Mesh& m, V center, V size, float deg, V alb) {
    const float hw = size.x * 0.5f, h = size.y, hd = size.z * 0.5f;
    const V local[8] = {{-hw, 0, -hd}, {hw, 0, -hd}, {-hw, 0, hd}, {hw, 0, hd},
                        {-hw, h, -hd}, {hw, h, -hd}, {-hw, h, hd}, {hw, h, hd}};
    V p[8];
    for (int i = 0; i < 8; ++i) p[i] = add(rotate_y(local[i], deg), center);
    const V black{0, 0, 0};
    m.quad(p[4], sub(p[5], p[4]), sub(p[6], p[4]), alb, black);  // top
    m.quad(p[0], sub(p[1], p[0]), sub(p[4], p[0]), alb, black);  // -z side
    m.quad(p[3], sub(p[2], p[3]), sub(p[7], p[3]), alb, black);  // +z side
    m.quad(p[2], sub(p[0], p[2]), sub(p[6], p[2]), alb, black);  // -x side
    m.quad(p[1], sub(p[3], p[1]), sub(p[5], p[1]), alb, black);  // +x side
}

}  // namespace

void make_cornell_box(float aspect, std::vector<float>& verts, std::vector<uint32_t>& indices,
                      std::vector<float>& albedo, std::vector<float>& emission) {
    (void)aspect;
    verts.clear(); indices.clear(); albedo.clear(); emission.clear();
    Mesh m{verts, indices, albedo, emission};
    const V white{0.73f, 0.73f, 0.73f}, red{0.65f, 0.05f, 0.05f}, green{0.12f, 0.45f, 0.15f};
    const V black{0, 0, 0};
    m.quad({-1, 0, -1}, {2, 0, 0}, {0, 0, 2}, white, black);    // floor
    m.quad({-1, 2, 1}, {2, 0, 0}, {0, 0, -2}, white, black);    // ceiling
    m.quad({-1, 0, -1}, {2, 0, 0}, {0, 2, 0}, white, black);    // back wall
    m.quad({-1, 0, 1}, {0, 0, -2}, {0, 2, 0}, red, black);      // left wall
    m.quad({1, 0, -1}, {0, 0, 2}, {0, 2, 0}, green, black);     // right wall
    m.quad({-0.25f, 1.99f, -0.25f}, {0.5f, 0, 0}, {0, 0, 0.5f}, black, {15.f, 15.f, 15.f});  // light
    box(m, {-0.35f, 0.0f, 0.3f}, {0.5f, 1.2f, 0.5f}, 15.0f, white);
    box(m, {0.35f, 0.0f, -0.3f}, {0.55f, 0.55f, 0.55f}, -18.0f, white);
}

}  // namespace quasi_hip
