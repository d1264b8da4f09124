// file_watcher.hpp -- change detection for one file by its modification time.
//
// Behaviour of the reference's src/quasi/async/file_watcher.hpp:28-133 without
// the coroutine awaitables: the watcher records the file's last write time;
// has_changed() is true once the file exists with a different time than the
// recorded one (a missing file, or no recorded time, is "no change"), and
// refresh() records the current time.  The host polls it between frames.
#pragma once

#include <filesystem>
#include <optional>
#include <system_error>
#include <utility>

namespace Q::io {

class file_watcher {
public:
    using path_type = std::filesystem::path;
    using time_type = std::filesystem::file_time_type;

    explicit file_watcher(path_type path) : path_(std::move(path)) { refresh(); }

    bool has_changed() const {
        std::error_code ec;
        if (!last_ || !std::filesystem::exists(path_, ec)) return false;
        const time_type now = std::filesystem::last_write_time(path_, ec);
        return !ec && now != *last_;
    }

    void refresh() {
        std::error_code ec;
        last_.reset();
        if (!std::filesystem::exists(path_, ec)) return;
        const time_type now = std::filesystem::last_write_time(path_, ec);
        if (!ec) last_ = now;
    }

    const path_type& path() const noexcept { return path_; }
    std::optional<time_type> last_write_time() const noexcept { return last_; }

private:
    path_type path_;
    std::optional<time_type> last_;
};

}  // namespace Q::io
