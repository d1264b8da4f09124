// dynamic_library.hpp -- RAII handle of a shared library opened with dlopen.
//
// Same interface and error strings as the reference's
// src/quasi/plugin/dynamic_library.hpp:22-170 (library_error and to_string
// :22-38, shared_library_extension / ensure_library_extension :41-58,
// dynamic_library::open / get_symbol / close :85-150), so a host written
// against it links unchanged.
#pragma once

#include <dlfcn.h>

#include <filesystem>
#include <string>
#include <string_view>
#include <utility>

#include "expected.hpp"

namespace Q::plugin {

enum class library_error {
    file_not_found,
    load_failed,
    symbol_not_found,
    not_loaded,
};

constexpr std::string_view to_string(library_error e) noexcept {
    switch (e) {
        case library_error::file_not_found: return "file not found";
        case library_error::load_failed: return "failed to load library";
        case library_error::symbol_not_found: return "symbol not found";
        case library_error::not_loaded: return "library not loaded";
    }
    return "unknown error";
}

constexpr std::string_view shared_library_extension() noexcept { return ".so"; }

inline std::filesystem::path ensure_library_extension(const std::filesystem::path& p) {
    const std::string ext = p.extension().string();
    if (ext == ".so" || ext == ".dylib" || ext == ".dll") return p;
    std::filesystem::path out = p;
    out += std::string(shared_library_extension());
    return out;
}

class dynamic_library {
public:
    template <class T>
    using result = Q::expected<T, library_error>;

    dynamic_library() noexcept = default;
    ~dynamic_library() { close(); }
    dynamic_library(const dynamic_library&) = delete;
    dynamic_library& operator=(const dynamic_library&) = delete;
    dynamic_library(dynamic_library&& o) noexcept : handle_(std::exchange(o.handle_, nullptr)), path_(std::move(o.path_)) {}
    dynamic_library& operator=(dynamic_library&& o) noexcept {
        if (this != &o) {
            close();
            handle_ = std::exchange(o.handle_, nullptr);
            path_ = std::move(o.path_);
        }
        return *this;
    }

    // RTLD_NOW | RTLD_LOCAL: every symbol resolved at load, none leaked into
    // the global namespace (the plugin can be reopened after a rebuild).
    static result<dynamic_library> open(const std::filesystem::path& path) {
        std::error_code ec;
        if (!std::filesystem::exists(path, ec)) return Q::unexpected(library_error::file_not_found);
        (void)dlerror();
        void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) return Q::unexpected(library_error::load_failed);
        dynamic_library lib;
        lib.handle_ = h;
        lib.path_ = path;
        return lib;
    }

    template <class Fn>
    result<Fn> get_symbol(std::string_view name) const {
        if (!handle_) return Q::unexpected(library_error::not_loaded);
        (void)dlerror();
        const std::string n(name);
        void* sym = dlsym(handle_, n.c_str());
        if (dlerror() != nullptr) return Q::unexpected(library_error::symbol_not_found);
        return reinterpret_cast<Fn>(sym);
    }

    void close() noexcept {
        if (handle_) {
            dlclose(handle_);
            handle_ = nullptr;
            path_.clear();
        }
    }

    bool is_loaded() const noexcept { return handle_ != nullptr; }
    explicit operator bool() const noexcept { return is_loaded(); }
    const std::filesystem::path& path() const noexcept { return path_; }
    void* native_handle() const noexcept { return handle_; }
    static const char* last_error() noexcept {
        const char* e = dlerror();
        return e ? e : "no error";
    }

private:
    void* handle_ = nullptr;
    std::filesystem::path path_;
};

}  // namespace Q::plugin
