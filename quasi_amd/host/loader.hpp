// loader.hpp -- a loaded plugin: resolved entry points, ABI check, lifetime.
//
// Same behaviour as the reference's src/quasi/plugin/loader.hpp:65-234: the
// six required symbols must resolve (else symbol_not_found), the ABI version
// must be 1..Q_PLUGIN_ABI_VERSION (else abi_mismatch), the optional readback
// (v2) and AOV readback (v3) pairs are looked up, and Q_plugin_create
// returning NULL is create_failed.  Destruction calls Q_plugin_destroy.
#pragma once

#include <utility>

#include "dynamic_library.hpp"
#include "quasi/plugin_interface.h"

namespace Q::plugin {

class loader {
public:
    enum class error {
        abi_mismatch,
        symbol_not_found,
        create_failed,
    };
    template <class T>
    using result = Q::expected<T, error>;

    using abi_version_fn = uint32_t (*)();
    using get_info_fn = Q_plugin_info (*)();
    using create_fn = Q_plugin_handle* (*)(Q_plugin_context*);
    using destroy_fn = void (*)(Q_plugin_handle*);
    using update_fn = void (*)(Q_plugin_handle*, float);
    using render_fn = void (*)(Q_plugin_handle*, Q_render_frame*);
    using readback_fn = Q_readback_result (*)(Q_plugin_handle*);
    using readback_free_fn = void (*)(Q_readback_result*);
    using readback_aov_fn = Q_readback_aov_result (*)(Q_plugin_handle*);
    using readback_aov_free_fn = void (*)(Q_readback_aov_result*);

    static result<loader> load(dynamic_library& lib, Q_plugin_context* ctx) {
        loader p;
        if (!resolve(lib, "Q_plugin_abi_version", p.abi_version_) || !resolve(lib, "Q_plugin_get_info", p.get_info_) ||
            !resolve(lib, "Q_plugin_create", p.create_) || !resolve(lib, "Q_plugin_destroy", p.destroy_) ||
            !resolve(lib, "Q_plugin_update", p.update_) || !resolve(lib, "Q_plugin_render", p.render_))
            return Q::unexpected(error::symbol_not_found);
        const uint32_t v = p.abi_version_();
        if (v < 1 || v > Q_PLUGIN_ABI_VERSION) return Q::unexpected(error::abi_mismatch);
        resolve(lib, "Q_plugin_readback", p.readback_);
        resolve(lib, "Q_plugin_readback_free", p.readback_free_);
        resolve(lib, "Q_plugin_readback_aov", p.readback_aov_);
        resolve(lib, "Q_plugin_readback_aov_free", p.readback_aov_free_);
        p.handle_ = p.create_(ctx);
        if (!p.handle_) return Q::unexpected(error::create_failed);
        return p;
    }

    loader() = default;
    ~loader() { destroy(); }
    loader(const loader&) = delete;
    loader& operator=(const loader&) = delete;
    loader(loader&& o) noexcept { *this = std::move(o); }
    loader& operator=(loader&& o) noexcept {
        if (this != &o) {
            destroy();
            handle_ = std::exchange(o.handle_, nullptr);
            abi_version_ = std::exchange(o.abi_version_, nullptr);
            get_info_ = std::exchange(o.get_info_, nullptr);
            create_ = std::exchange(o.create_, nullptr);
            destroy_ = std::exchange(o.destroy_, nullptr);
            update_ = std::exchange(o.update_, nullptr);
            render_ = std::exchange(o.render_, nullptr);
            readback_ = std::exchange(o.readback_, nullptr);
            readback_free_ = std::exchange(o.readback_free_, nullptr);
            readback_aov_ = std::exchange(o.readback_aov_, nullptr);
            readback_aov_free_ = std::exchange(o.readback_aov_free_, nullptr);
        }
        return *this;
    }

    void update(float dt) {
        if (handle_ && update_) update_(handle_, dt);
    }
    void render(Q_render_frame* frame) {
        if (handle_ && render_) render_(handle_, frame);
    }
    bool supports_readback() const noexcept { return readback_ && readback_free_; }
    Q_readback_result readback() { return handle_ && readback_ ? readback_(handle_) : Q_readback_result{}; }
    void readback_free(Q_readback_result* r) {
        if (readback_free_ && r) readback_free_(r);
    }
    bool supports_readback_aov() const noexcept { return readback_aov_ && readback_aov_free_; }
    Q_readback_aov_result readback_aov() {
        return handle_ && readback_aov_ ? readback_aov_(handle_) : Q_readback_aov_result{};
    }
    void readback_aov_free(Q_readback_aov_result* r) {
        if (readback_aov_free_ && r) readback_aov_free_(r);
    }
    Q_plugin_info info() const { return get_info_ ? get_info_() : Q_plugin_info{}; }
    uint32_t abi_version() const { return abi_version_ ? abi_version_() : 0u; }
    bool is_valid() const noexcept { return handle_ != nullptr; }
    Q_plugin_handle* handle() const noexcept { return handle_; }

    void destroy() {
        if (handle_ && destroy_) destroy_(handle_);
        handle_ = nullptr;
    }

private:
    template <class Fn>
    static bool resolve(dynamic_library& lib, const char* name, Fn& out) {
        auto s = lib.get_symbol<Fn>(name);
        if (!s) return false;
        out = *s;
        return true;
    }

    Q_plugin_handle* handle_ = nullptr;
    abi_version_fn abi_version_ = nullptr;
    get_info_fn get_info_ = nullptr;
    create_fn create_ = nullptr;
    destroy_fn destroy_ = nullptr;
    update_fn update_ = nullptr;
    render_fn render_ = nullptr;
    readback_fn readback_ = nullptr;
    readback_free_fn readback_free_ = nullptr;
    readback_aov_fn readback_aov_ = nullptr;
    readback_aov_free_fn readback_aov_free_ = nullptr;
};

constexpr const char* to_string(loader::error e) noexcept {
    switch (e) {
        case loader::error::abi_mismatch: return "ABI version mismatch";
        case loader::error::symbol_not_found: return "Required symbol not found";
        case loader::error::create_failed: return "Plugin creation failed";
    }
    return "Unknown error";
}

}  // namespace Q::plugin
