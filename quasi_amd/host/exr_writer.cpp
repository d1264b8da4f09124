// exr_writer.cpp -- see exr_writer.hpp.
//
// File layout (OpenEXR 2.0 single-part scanline image):
//   magic 76 2f 31 01, version 2 (int32)
//   header: attributes (name\0 type\0 int32 size, value), then \0
//   offset table: one uint64 file offset per scanline block
//   blocks: int32 first y, int32 data size, data
// A block holds, for each of its scanlines, each channel's row (channels in
// name order).  ZIP blocks are 16 scanlines: the bytes are split into even and
// odd positions, delta-coded (d = b[i] - b[i-1] + 128, mod 256) and deflated;
// a block that does not shrink is stored raw.
#include "exr_writer.hpp"

#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

namespace Q::io {

uint16_t float_to_half(float f) noexcept {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u)                                   // inf, NaN
        return uint16_t(sign | 0x7C00u | (ax > 0x7F800000u ? (0x200u | ((ax >> 13) & 0x3FFu)) : 0u));
    if (ax >= 0x477FF000u) return uint16_t(sign | 0x7C00u);  // >= 65520 rounds to infinity
    if (ax < 0x38800000u) {                                  // below 2^-14: subnormal or zero
        if (ax <= 0x33000000u) return uint16_t(sign);        // <= 2^-25: rounds to zero
        const uint32_t e = ax >> 23;
        const uint32_t mant = (ax & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - e;                     // units of 2^-24
        uint32_t m = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (m & 1u))) ++m;
        return uint16_t(sign | m);
    }
    const uint32_t r = ax - 0x38000000u;                     // rebias 127 -> 15
    uint32_t h = r >> 13;
    const uint32_t rem = r & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return uint16_t(sign | h);
}

namespace {

enum : int32_t { kUint = 0, kHalf = 1, kFloat = 2 };

struct Channel {
    std::string name;
    int32_t type;
    const float* rgba;   // source RGBA32F buffer
    int comp;            // component of it
};

struct Bytes {
    std::vector<uint8_t> b;
    void raw(const void* p, size_t n) {
        const uint8_t* c = static_cast<const uint8_t*>(p);
        b.insert(b.end(), c, c + n);
    }
    void i32(int32_t v) { raw(&v, 4); }
    void u8(uint8_t v) { b.push_back(v); }
    void f32(float v) { raw(&v, 4); }
    void str(const char* s) { raw(s, std::strlen(s) + 1); }
    void attr(const char* name, const char* type, const std::vector<uint8_t>& value) {
        str(name);
        str(type);
        i32(int32_t(value.size()));
        raw(value.data(), value.size());
    }
};

std::vector<uint8_t> chlist(const std::vector<Channel>& chans) {
    Bytes v;
    for (const Channel& c : chans) {
        v.str(c.name.c_str());
        v.i32(c.type);
        v.u8(0);            // pLinear
        v.u8(0); v.u8(0); v.u8(0);
        v.i32(1);           // x sampling
        v.i32(1);           // y sampling
    }
    v.u8(0);
    return v.b;
}

std::vector<uint8_t> box2i(int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
    Bytes v;
    v.i32(x0); v.i32(y0); v.i32(x1); v.i32(y1);
    return v.b;
}

// The uncompressed bytes of scanlines [y0, y1).
std::vector<uint8_t> block_bytes(const std::vector<Channel>& chans, uint32_t w, uint32_t y0, uint32_t y1) {
    std::vector<uint8_t> out;
    size_t per_line = 0;
    for (const Channel& c : chans) per_line += size_t(w) * (c.type == kHalf ? 2 : 4);
    out.resize(per_line * (y1 - y0));
    uint8_t* p = out.data();
    for (uint32_t y = y0; y < y1; ++y)
        for (const Channel& c : chans) {
            const float* row = c.rgba + size_t(y) * w * 4 + c.comp;
            if (c.type == kHalf) {
                for (uint32_t x = 0; x < w; ++x) {
                    const uint16_t h = float_to_half(row[4 * size_t(x)]);
                    *p++ = uint8_t(h & 0xFF);
                    *p++ = uint8_t(h >> 8);
                }
            } else {
                for (uint32_t x = 0; x < w; ++x) {
                    std::memcpy(p, &row[4 * size_t(x)], 4);
                    p += 4;
                }
            }
        }
    return out;
}

bool zip_block(const std::vector<uint8_t>& raw, std::vector<uint8_t>& out) {
    const size_t n = raw.size();
    std::vector<uint8_t> t(n);
    size_t i1 = 0, i2 = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) t[(i & 1) ? i2++ : i1++] = raw[i];
    for (size_t i = n; i-- > 1;) t[i] = uint8_t(int(t[i]) - int(t[i - 1]) + (128 + 256));
    uLongf cap = compressBound(uLong(n));
    out.resize(cap);
    if (compress2(out.data(), &cap, t.data(), uLong(n), 4) != Z_OK) return false;
    out.resize(cap);
    return true;
}

Q::expected<void, exr_error> write_channels(const std::filesystem::path& path, uint32_t w, uint32_t h,
                                            std::vector<Channel> chans, exr_compression comp) {
    std::sort(chans.begin(), chans.end(), [](const Channel& a, const Channel& b) { return a.name < b.name; });
    const uint32_t lines = comp == exr_compression::zip ? 16u : 1u;
    const uint32_t nblocks = (h + lines - 1) / lines;
    // blocks, compressed on the host threads
    std::vector<std::vector<uint8_t>> blocks(nblocks);
    std::atomic<uint32_t> next{0};
    std::atomic<bool> ok{true};
    auto worker = [&] {
        for (;;) {
            const uint32_t b = next.fetch_add(1);
            if (b >= nblocks) break;
            const uint32_t y0 = b * lines, y1 = std::min(h, y0 + lines);
            std::vector<uint8_t> raw = block_bytes(chans, w, y0, y1);
            if (comp == exr_compression::zip) {
                std::vector<uint8_t> z;
                if (!zip_block(raw, z)) {
                    ok = false;
                    break;
                }
                blocks[b] = z.size() < raw.size() ? std::move(z) : std::move(raw);
            } else {
                blocks[b] = std::move(raw);
            }
        }
    };
    const unsigned nth = std::max(1u, std::min<unsigned>(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nth && t < nblocks; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    if (!ok) return Q::unexpected(exr_error::write_failed);

    Bytes hd;
    const uint8_t magic[4] = {0x76, 0x2F, 0x31, 0x01};
    hd.raw(magic, 4);
    hd.i32(2);
    hd.attr("channels", "chlist", chlist(chans));
    hd.attr("compression", "compression", {uint8_t(comp)});
    hd.attr("dataWindow", "box2i", box2i(0, 0, int32_t(w) - 1, int32_t(h) - 1));
    hd.attr("displayWindow", "box2i", box2i(0, 0, int32_t(w) - 1, int32_t(h) - 1));
    hd.attr("lineOrder", "lineOrder", {0});
    {
        Bytes v;
        v.f32(1.0f);
        hd.attr("pixelAspectRatio", "float", v.b);
    }
    {
        Bytes v;
        v.f32(0.0f);
        v.f32(0.0f);
        hd.attr("screenWindowCenter", "v2f", v.b);
    }
    {
        Bytes v;
        v.f32(1.0f);
        hd.attr("screenWindowWidth", "float", v.b);
    }
    hd.u8(0);
    uint64_t off = hd.b.size() + 8ull * nblocks;
    std::vector<uint64_t> offsets(nblocks);
    for (uint32_t b = 0; b < nblocks; ++b) {
        offsets[b] = off;
        off += 8 + blocks[b].size();
    }
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return Q::unexpected(exr_error::write_failed);
    f.write(reinterpret_cast<const char*>(hd.b.data()), std::streamsize(hd.b.size()));
    f.write(reinterpret_cast<const char*>(offsets.data()), std::streamsize(8 * offsets.size()));
    for (uint32_t b = 0; b < nblocks; ++b) {
        const int32_t y = int32_t(b * lines);
        const int32_t size = int32_t(blocks[b].size());
        f.write(reinterpret_cast<const char*>(&y), 4);
        f.write(reinterpret_cast<const char*>(&size), 4);
        f.write(reinterpret_cast<const char*>(blocks[b].data()), size);
    }
    if (!f) return Q::unexpected(exr_error::write_failed);
    return {};
}

bool make_parent(const std::filesystem::path& path) {
    const auto parent = path.parent_path();
    if (parent.empty()) return true;
    std::error_code ec;
    std::filesystem::create_directories(parent, ec);
    return !ec;
}

}  // namespace

Q::expected<void, exr_error> write_exr(const std::filesystem::path& path, const Q_readback_result& r,
                                       exr_compression compression) {
    if (!r.data || r.width == 0 || r.height == 0 || r.channels != 4) return Q::unexpected(exr_error::invalid_data);
    if (!make_parent(path)) return Q::unexpected(exr_error::directory_error);
    std::vector<Channel> ch = {{"R", kHalf, r.data, 0}, {"G", kHalf, r.data, 1}, {"B", kHalf, r.data, 2},
                               {"A", kHalf, r.data, 3}};
    return write_channels(path, r.width, r.height, std::move(ch), compression);
}

Q::expected<void, exr_error> write_exr(const std::filesystem::path& path, const Q_readback_aov_result& r,
                                       exr_compression compression) {
    const Q_aov_buffer& beauty = r.buffers[Q_AOV_BEAUTY];
    if (!beauty.data || beauty.width == 0 || beauty.height == 0) return Q::unexpected(exr_error::invalid_data);
    for (int a = 1; a < Q_AOV_COUNT; ++a) {
        const Q_aov_buffer& b = r.buffers[a];
        if (b.data && (b.width != beauty.width || b.height != beauty.height)) return Q::unexpected(exr_error::invalid_data);
    }
    if (!make_parent(path)) return Q::unexpected(exr_error::directory_error);
    std::vector<Channel> ch = {{"R", kHalf, beauty.data, 0}, {"G", kHalf, beauty.data, 1},
                               {"B", kHalf, beauty.data, 2}, {"A", kHalf, beauty.data, 3}};
    if (const float* a = r.buffers[Q_AOV_ALBEDO].data) {
        ch.push_back({"albedo.R", kHalf, a, 0});
        ch.push_back({"albedo.G", kHalf, a, 1});
        ch.push_back({"albedo.B", kHalf, a, 2});
    }
    if (const float* n = r.buffers[Q_AOV_NORMAL].data) {
        ch.push_back({"normal.X", kHalf, n, 0});
        ch.push_back({"normal.Y", kHalf, n, 1});
        ch.push_back({"normal.Z", kHalf, n, 2});
    }
    if (const float* d = r.buffers[Q_AOV_DEPTH].data) ch.push_back({"depth.Z", kFloat, d, 0});
    return write_channels(path, beauty.width, beauty.height, std::move(ch), compression);
}

std::filesystem::path make_timestamped_path(const std::filesystem::path& directory) {
    const std::time_t t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    std::tm local{};
    localtime_r(&t, &local);
    char name[64];
    std::snprintf(name, sizeof name, "quasi_%04d%02d%02d_%02d%02d%02d.exr", local.tm_year + 1900, local.tm_mon + 1,
                  local.tm_mday, local.tm_hour, local.tm_min, local.tm_sec);
    return directory / name;
}

}  // namespace Q::io
