// io_capi.cpp -- C entry points of libquasi_io.so: the host's EXR writer and
// scene-file reader, for the Python tests (ctypes).  Not used by the plugin.
#include <cstring>

#include "exr_writer.hpp"
#include "scene_file.hpp"

extern "C" {

__attribute__((visibility("default"))) uint16_t q_io_float_to_half(float f) { return Q::io::float_to_half(f); }

// Writes a multi-layer AOV file from RGBA32F buffers (albedo, normal, depth
// may be NULL); compression 0 = none, 3 = ZIP.  0 on success, else the
// exr_error value + 1.
__attribute__((visibility("default"))) int q_io_write_exr_aov(const char* path, uint32_t w, uint32_t h,
                                                              const float* beauty, const float* albedo,
                                                              const float* normal, const float* depth,
                                                              int compression) {
    Q_readback_aov_result r{};
    const float* bufs[4] = {beauty, albedo, normal, depth};
    for (int a = 0; a < 4; ++a) {
        r.buffers[a].data = const_cast<float*>(bufs[a]);
        r.buffers[a].width = bufs[a] ? w : 0;
        r.buffers[a].height = bufs[a] ? h : 0;
        r.buffers[a].channels = bufs[a] ? 4 : 0;
    }
    auto res = Q::io::write_exr(path, r, compression == 3 ? Q::io::exr_compression::zip : Q::io::exr_compression::none);
    return res ? 0 : int(res.error()) + 1;
}

// Single-layer RGBA file (the reference's write_exr(Q_readback_result)).
__attribute__((visibility("default"))) int q_io_write_exr_rgba(const char* path, uint32_t w, uint32_t h,
                                                               const float* rgba, int compression) {
    Q_readback_result r{const_cast<float*>(rgba), w, h, 4};
    auto res = Q::io::write_exr(path, r, compression == 3 ? Q::io::exr_compression::zip : Q::io::exr_compression::none);
    return res ? 0 : int(res.error()) + 1;
}

// Reads a scene file's counts (out[0] = vertices, out[1] = triangles) and
// camera (10 floats); 0 on success.
__attribute__((visibility("default"))) int q_io_scene_info(const char* path, uint32_t* counts, float* camera) {
    auto s = Q::io::read_scene(path);
    if (!s) return int(s.error()) + 1;
    counts[0] = uint32_t(s->vertices.size() / 3);
    counts[1] = uint32_t(s->indices.size() / 3);
    std::memcpy(camera, s->camera, sizeof s->camera);
    return 0;
}

}  // extern "C"
