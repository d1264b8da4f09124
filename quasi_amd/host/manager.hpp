// manager.hpp -- a plugin that can be replaced while the host runs (hot reload).
//
// Behaviour of the reference's src/quasi/plugin/manager.hpp:78-330, driven
// synchronously by the host loop instead of a coroutine scheduler:
//  - the library is opened from a private copy in the temporary directory
//    (quasi_plugin_<n>_<file name>), so a rebuild may overwrite the original
//    while the copy is loaded; the copy is removed when the next one is made
//    and when the manager goes away;
//  - poll() reloads when the original's modification time changes (the host
//    calls it between frames; the reference polls every 100 ms);
//  - reload() = pre_unload hook, destroy the handle and close the library,
//    wait for the file system to settle, open a fresh copy, post_load hook.
//    A failed load calls on_error and leaves no plugin loaded; the next
//    change of the file tries again;
//  - stats() counts reload attempts, successes and failures and times the
//    last reload.
// The manager owns the Q_plugin_context that the plugin keeps a pointer to,
// so it is neither copied nor moved.
#pragma once

#include <atomic>
#include <chrono>
#include <cstdint>
#include <filesystem>
#include <functional>
#include <optional>
#include <string>
#include <string_view>
#include <system_error>
#include <thread>
#include <utility>

#include "dynamic_library.hpp"
#include "expected.hpp"
#include "file_watcher.hpp"
#include "loader.hpp"
#include "quasi/plugin_interface.h"

namespace Q::plugin {

struct reload_hooks {
    std::function<void()> pre_unload = [] {};            // e.g. wait for the GPU to go idle
    std::function<void()> post_load = [] {};             // e.g. hand the new plugin its scene again
    std::function<void(const std::string&)> on_error = [](const std::string&) {};
};

struct reload_stats {
    uint64_t reload_count = 0;      // attempts
    uint64_t success_count = 0;
    uint64_t failure_count = 0;
    float last_reload_time = 0.f;   // seconds, of the last successful reload
};

class manager {
public:
    enum class error {
        file_not_found,
        load_failed,
        abi_mismatch,
        symbol_missing,
        create_failed,
    };
    template <class T>
    using result = Q::expected<T, error>;

    manager(std::filesystem::path library_path, const Q_plugin_context& context, reload_hooks hooks = {},
            std::chrono::milliseconds settle = std::chrono::milliseconds(300))
        : path_(std::move(library_path)), watcher_(path_), hooks_(std::move(hooks)), context_(context),
          settle_(settle) {}
    ~manager() {
        unload();
        remove_copy();
    }
    manager(const manager&) = delete;
    manager& operator=(const manager&) = delete;

    // The first load.
    result<void> load() { return open_copy(); }

    result<void> reload() {
        const auto t0 = std::chrono::steady_clock::now();
        ++stats_.reload_count;
        hooks_.pre_unload();
        unload();
        if (settle_.count() > 0) std::this_thread::sleep_for(settle_);
        result<void> r = open_copy();
        if (!r) {
            ++stats_.failure_count;
            hooks_.on_error(std::string("reload failed: ") + to_string(r.error()));
            watcher_.refresh();   // wait for the next change of the file
            return r;
        }
        hooks_.post_load();
        ++stats_.success_count;
        stats_.last_reload_time = std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
        return r;
    }

    // Reloads when the library file changed since the last load; true when a
    // reload was attempted.
    bool poll() {
        if (!watcher_.has_changed()) return false;
        (void)reload();
        return true;
    }

    void update(float dt) {
        if (plugin_) plugin_->update(dt);
    }
    void render(Q_render_frame* frame) {
        if (plugin_) plugin_->render(frame);
    }

    bool is_loaded() const noexcept { return plugin_.has_value() && plugin_->is_valid(); }
    loader* plugin() noexcept { return plugin_ ? &*plugin_ : nullptr; }
    const dynamic_library& library() const noexcept { return library_; }
    std::optional<Q_plugin_info> info() const {
        if (plugin_) return plugin_->info();
        return std::nullopt;
    }
    const reload_stats& stats() const noexcept { return stats_; }
    const std::filesystem::path& library_path() const noexcept { return path_; }
    const std::filesystem::path& loaded_copy() const noexcept { return copy_; }
    Q_plugin_context& context() noexcept { return context_; }

    static constexpr const char* to_string(error e) noexcept {
        switch (e) {
            case error::file_not_found: return "file not found";
            case error::load_failed: return "load failed";
            case error::abi_mismatch: return "ABI version mismatch";
            case error::symbol_missing: return "missing symbols";
            case error::create_failed: return "plugin creation failed";
        }
        return "unknown error";
    }

private:
    result<void> open_copy() {
        std::error_code ec;
        if (!std::filesystem::exists(path_, ec)) return Q::unexpected(error::file_not_found);
        remove_copy();
        static std::atomic<uint64_t> counter{0};
        const std::filesystem::path copy = std::filesystem::temp_directory_path(ec) /
            ("quasi_plugin_" + std::to_string(counter++) + "_" + path_.filename().string());
        if (ec) return Q::unexpected(error::load_failed);
        watcher_.refresh();   // the time of the file as copied
        if (!std::filesystem::copy_file(path_, copy, std::filesystem::copy_options::overwrite_existing, ec))
            return Q::unexpected(error::load_failed);
        copy_ = copy;
        auto lib = dynamic_library::open(copy_);
        if (!lib) return Q::unexpected(error::load_failed);
        library_ = std::move(*lib);
        auto p = loader::load(library_, &context_);
        if (!p) {
            library_.close();
            switch (p.error()) {
                case loader::error::abi_mismatch: return Q::unexpected(error::abi_mismatch);
                case loader::error::symbol_not_found: return Q::unexpected(error::symbol_missing);
                case loader::error::create_failed: return Q::unexpected(error::create_failed);
            }
            return Q::unexpected(error::load_failed);
        }
        plugin_ = std::move(*p);
        return {};
    }

    void unload() {
        if (plugin_) {
            plugin_->destroy();
            plugin_.reset();
        }
        library_.close();
    }

    void remove_copy() {
        if (copy_.empty()) return;
        std::error_code ec;
        std::filesystem::remove(copy_, ec);
        copy_.clear();
    }

    std::filesystem::path path_, copy_;
    Q::io::file_watcher watcher_;
    reload_hooks hooks_;
    Q_plugin_context context_;
    std::chrono::milliseconds settle_;
    reload_stats stats_;
    dynamic_library library_;
    std::optional<loader> plugin_;
};

}  // namespace Q::plugin
