// manager_probe.cpp -- test-only C shim over the host's plugin manager
// (quasi_amd/host/manager.hpp), so tests can load, poll and reload plugins
// from Python.  Everything the plugin logs and every hook call is appended to
// one event log ("pre_unload", "post_load", "error: ...", plugin messages).
// Built by tests/test_plugin_manager.py with g++; needs no GPU.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "manager.hpp"

#define PROBE_API extern "C" __attribute__((visibility("default")))

namespace {

struct Probe {
    std::string events;
    Q::plugin::manager* mgr = nullptr;
};

void log_event(void* host, const char* msg) {
    auto* p = static_cast<Probe*>(host);
    p->events += msg;
    p->events += '\n';
}

}  // namespace

PROBE_API void* mgr_new(const char* library, unsigned width, unsigned height, unsigned settle_ms) {
    auto* p = new Probe;
    Q_plugin_context ctx{};
    ctx.viewport_width = width;
    ctx.viewport_height = height;
    ctx.host_data = p;
    ctx.log = log_event;
    Q::plugin::reload_hooks hooks;
    hooks.pre_unload = [p] { p->events += "pre_unload\n"; };
    hooks.post_load = [p] { p->events += "post_load\n"; };
    hooks.on_error = [p](const std::string& m) { p->events += "error: " + m + "\n"; };
    p->mgr = new Q::plugin::manager(library, ctx, hooks, std::chrono::milliseconds(settle_ms));
    return p;
}

PROBE_API void mgr_free(void* h) {
    auto* p = static_cast<Probe*>(h);
    delete p->mgr;
    delete p;
}

// 0 on success, else 1 + the manager's error code.
PROBE_API int mgr_load(void* h) {
    auto r = static_cast<Probe*>(h)->mgr->load();
    return r ? 0 : 1 + int(r.error());
}

PROBE_API int mgr_reload(void* h) {
    auto r = static_cast<Probe*>(h)->mgr->reload();
    return r ? 0 : 1 + int(r.error());
}

PROBE_API int mgr_poll(void* h) { return static_cast<Probe*>(h)->mgr->poll() ? 1 : 0; }

PROBE_API int mgr_is_loaded(void* h) { return static_cast<Probe*>(h)->mgr->is_loaded() ? 1 : 0; }

PROBE_API void mgr_render(void* h, int frames) {
    auto* m = static_cast<Probe*>(h)->mgr;
    for (int i = 0; i < frames; ++i) {
        Q_render_frame f{};
        f.width = m->context().viewport_width;
        f.height = m->context().viewport_height;
        m->update(0.f);
        m->render(&f);
    }
}

// First value of the readback, or -1 when there is none.
PROBE_API float mgr_readback(void* h) {
    auto* pl = static_cast<Probe*>(h)->mgr->plugin();
    if (!pl || !pl->supports_readback()) return -1.f;
    Q_readback_result r = pl->readback();
    const float v = r.data ? r.data[0] : -1.f;
    pl->readback_free(&r);
    return v;
}

// info version major, or -1 when nothing is loaded.
PROBE_API int mgr_version(void* h) {
    auto i = static_cast<Probe*>(h)->mgr->info();
    return i ? int(i->version.major) : -1;
}

PROBE_API void mgr_stats(void* h, unsigned long long* counts, float* last_seconds) {
    const auto& s = static_cast<Probe*>(h)->mgr->stats();
    counts[0] = s.reload_count;
    counts[1] = s.success_count;
    counts[2] = s.failure_count;
    *last_seconds = s.last_reload_time;
}

PROBE_API void mgr_copy_path(void* h, char* out, unsigned len) {
    std::snprintf(out, len, "%s", static_cast<Probe*>(h)->mgr->loaded_copy().c_str());
}

// The event log since the last call (then cleared).
PROBE_API void mgr_events(void* h, char* out, unsigned len) {
    auto* p = static_cast<Probe*>(h);
    std::snprintf(out, len, "%s", p->events.c_str());
    p->events.clear();
}
