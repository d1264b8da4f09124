// ref_dump.cpp -- prints what the reference's own scene and camera code
// computes, so the product's scene builder and the oracle's camera can be
// checked against it.  Compiled by oracle/Makefile (target `ref`) against the
// reference headers where they lie (src/quasi/scene/cornell_box.hpp,
// camera.hpp, quad.hpp, math/vec.hpp); output goes to oracle/_ref/.
// Test infrastructure only.
#include <quasi/scene/cornell_box.hpp>

#include <cstdio>

int main() {
    auto s = Q::scene::make_cornell_box(1.0f);
    std::printf("{\n  \"light_index\": %zu,\n  \"quads\": [\n", s.light_index);
    for (size_t i = 0; i < s.quads.size(); ++i) {
        const auto& q = s.quads[i];
        std::printf("    {\"origin\": [%.9g, %.9g, %.9g], \"u\": [%.9g, %.9g, %.9g], \"v\": [%.9g, %.9g, %.9g], "
                    "\"albedo\": [%.9g, %.9g, %.9g], \"emission\": [%.9g, %.9g, %.9g]}%s\n",
                    q.geometry.origin.x, q.geometry.origin.y, q.geometry.origin.z,
                    q.geometry.u.x, q.geometry.u.y, q.geometry.u.z,
                    q.geometry.v.x, q.geometry.v.y, q.geometry.v.z,
                    q.mat.albedo.x, q.mat.albedo.y, q.mat.albedo.z,
                    q.mat.emission.x, q.mat.emission.y, q.mat.emission.z,
                    i + 1 < s.quads.size() ? "," : "");
    }
    const auto& c = s.cam;
    std::printf("  ],\n  \"camera\": {\"position\": [%.9g, %.9g, %.9g], \"direction\": [%.9g, %.9g, %.9g], "
                "\"up\": [%.9g, %.9g, %.9g], \"fov\": %.9g, \"aspect\": %.9g},\n",
                c.position.x, c.position.y, c.position.z, c.direction.x, c.direction.y, c.direction.z,
                c.up.x, c.up.y, c.up.z, c.fov, c.aspect);
    // camera.hpp:36-52 (the CPU twin of the shader's get_camera_ray, no jitter)
    std::printf("  \"rays\": [\n");
    const float uvs[][2] = {{0.5f, 0.5f}, {0.f, 0.f}, {1.f, 1.f}, {0.25f, 0.75f}, {0.9f, 0.1f}, {0.3f, 0.6f}};
    const int n = sizeof(uvs) / sizeof(uvs[0]);
    for (int i = 0; i < n; ++i) {
        auto r = c.get_ray(uvs[i][0], uvs[i][1]);
        std::printf("    {\"uv\": [%.9g, %.9g], \"origin\": [%.9g, %.9g, %.9g], \"direction\": [%.9g, %.9g, %.9g]}%s\n",
                    uvs[i][0], uvs[i][1], r.origin.x, r.origin.y, r.origin.z,
                    r.direction.x, r.direction.y, r.direction.z, i + 1 < n ? "," : "");
    }
    std::printf("  ]\n}\n");
    return 0;
}
