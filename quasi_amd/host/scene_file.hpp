// scene_file.hpp -- the host's scene file: an indexed triangle mesh with
// per-triangle albedo and emission and a camera, the input of
// Q_plugin_set_scene (the reference hard-codes its scene, plugin.mm:584).
//
// Little-endian binary:
//   char[4] "QSCN", uint32 version (1), uint32 num_vertices, uint32 num_triangles,
//   float camera[10] (position xyz, target xyz, up xyz, vertical fov in degrees),
//   float vertices[3 * V], uint32 indices[3 * T], float albedo[3 * T],
//   float emission[3 * T]
#pragma once

#include <cstdint>
#include <filesystem>
#include <vector>

#include "expected.hpp"

namespace Q::io {

enum class scene_error {
    open_failed,
    bad_format,
    bad_index,
};

constexpr const char* to_string(scene_error e) noexcept {
    switch (e) {
        case scene_error::open_failed: return "cannot open scene file";
        case scene_error::bad_format: return "not a scene file (or truncated)";
        case scene_error::bad_index: return "scene file has a vertex index out of range";
    }
    return "Unknown error";
}

struct scene_data {
    float camera[10];
    std::vector<float> vertices, albedo, emission;
    std::vector<uint32_t> indices;
};

Q::expected<scene_data, scene_error> read_scene(const std::filesystem::path& path);

}  // namespace Q::io
