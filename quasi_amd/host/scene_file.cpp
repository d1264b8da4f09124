// scene_file.cpp -- see scene_file.hpp.
#include "scene_file.hpp"

#include <cstring>
#include <fstream>

namespace Q::io {

namespace {
template <class T>
bool read_array(std::ifstream& f, std::vector<T>& v, size_t n) {
    v.resize(n);
    f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(T)));
    return bool(f);
}
}  // namespace

Q::expected<scene_data, scene_error> read_scene(const std::filesystem::path& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return Q::unexpected(scene_error::open_failed);
    char magic[4];
    uint32_t hdr[3];
    f.read(magic, 4);
    f.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    if (!f || std::memcmp(magic, "QSCN", 4) != 0 || hdr[0] != 1u) return Q::unexpected(scene_error::bad_format);
    const size_t nv = hdr[1], nt = hdr[2];
    if (nv > (size_t(1) << 31) || nt > (size_t(1) << 29)) return Q::unexpected(scene_error::bad_format);
    scene_data s;
    f.read(reinterpret_cast<char*>(s.camera), sizeof s.camera);
    if (!f || !read_array(f, s.vertices, 3 * nv) || !read_array(f, s.indices, 3 * nt) ||
        !read_array(f, s.albedo, 3 * nt) || !read_array(f, s.emission, 3 * nt))
        return Q::unexpected(scene_error::bad_format);
    for (uint32_t i : s.indices)
        if (i >= nv) return Q::unexpected(scene_error::bad_index);
    return s;
}

}  // namespace Q::io
