// exr_writer.hpp -- OpenEXR output of the plugin's float framebuffer.
//
// Same entry points, channel layout and error codes as the reference's
// src/quasi/io/exr_writer.{hpp,cpp}, written against the published OpenEXR
// file format instead of the OpenEXR library:
//   write_exr(path, Q_readback_result)      R, G, B, A as HALF (exr_writer.cpp:20-62)
//   write_exr(path, Q_readback_aov_result)  R, G, B, A HALF; albedo.R/G/B HALF;
//       normal.X/Y/Z HALF; depth.Z FLOAT from the depth buffer's first channel
//       (exr_writer.cpp:64-172)
//   make_timestamped_path(dir)              quasi_YYYYMMDD_HHMMSS.exr (:174-185)
// Single-part scanline files, INCREASING_Y, ZIP compression (zlib, 16 lines
// per block: the OpenEXR library's default) or none.  Float to half rounds to
// nearest even.
#pragma once

#include <cstdint>
#include <filesystem>

#include "expected.hpp"
#include "quasi/plugin_interface.h"

namespace Q::io {

enum class exr_error {
    invalid_data,
    write_failed,
    directory_error,
};

constexpr const char* to_string(exr_error e) noexcept {
    switch (e) {
        case exr_error::invalid_data: return "Invalid readback data";
        case exr_error::write_failed: return "EXR write failed";
        case exr_error::directory_error: return "Could not create output directory";
    }
    return "Unknown error";
}

enum class exr_compression : uint8_t {
    none = 0,
    zip = 3,
};

Q::expected<void, exr_error> write_exr(const std::filesystem::path& path, const Q_readback_result& result,
                                       exr_compression compression = exr_compression::zip);

Q::expected<void, exr_error> write_exr(const std::filesystem::path& path, const Q_readback_aov_result& result,
                                       exr_compression compression = exr_compression::zip);

std::filesystem::path make_timestamped_path(const std::filesystem::path& directory = ".");

// IEEE binary16 bits of f, round to nearest even (overflow to infinity, NaN
// kept NaN with the top mantissa bits).
uint16_t float_to_half(float f) noexcept;

}  // namespace Q::io
