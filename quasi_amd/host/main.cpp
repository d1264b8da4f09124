// main.cpp -- headless host for libquasi_hip.so.
//
// The reference host (src/quasi/host/main.cpp:120-330) opens a window, loads
// the Metal plugin through dynamic_library + loader, and in batch mode
// (--render N) renders N frames, reads the AOVs back and writes a
// timestamped EXR.  This host does the same without a window, on a HIP
// context:
//
//   quasi_host [plugin.so] --render N [--out file.exr | --out-dir DIR]
//              [--width W] [--height H] [--device K] [--scene file.qscene]
//              [--spp-per-render S] [--depth D] [--rr-start R] [--seed X]
//              [--integrator reference|nee] [--rng reference|hashed]
//              [--engine megakernel|wavefront|single-launch] [--compression zip|none]
//              [--single-layer] [--watch] [--reload-at K] [--wait-at K --wait-file PATH]
//              [--ppm file.ppm]
//              [--azimuth A] [--elevation E] [--distance D] [--orbit-step S]
//
// Without a plugin path the library next to this executable's directory
// (../libquasi_hip.so) is used.  The camera is the reference's orbit camera at
// rest (main.cpp:38-45, 104-115), or the scene file's; it is marked dirty on
// the first frame only, so the N frames accumulate.  The render parameters and
// the scene go through the plugin's extension exports when given.
//
// The plugin is loaded through the hot-reload manager (manager.hpp, after the
// reference's plugin/manager.hpp): from a private copy, so the library may be
// rebuilt while the host runs.  --watch polls the library between frames and
// reloads it when it changes; --reload-at K reloads it once before frame K;
// --wait-at K --wait-file PATH holds frame K until PATH exists (a fixed point
// for a driver of the host, such as a test that rebuilds the library).
// Around a reload the host waits for the GPU, and afterwards hands the new
// plugin the scene and parameters again and marks the camera dirty, so the
// accumulation restarts from the reload.
//
// The reference's orbit camera (main.cpp:39-112: spherical coordinates around
// a target, elevation clamped to +-1.5 rad, up +y) places the camera when any
// of --azimuth / --elevation / --distance / --orbit-step is given (target and
// fov from the scene file, else (0, 1, 0) and 40 degrees); --orbit-step turns
// it by S radians of azimuth before every frame after the first, standing in
// for a mouse drag, and marks the camera dirty, so the accumulation restarts
// as in the reference (plugin.mm:643-646).  --ppm also writes the display
// image (Q_plugin_readback_display: the tonemap pass of plugin.mm:334-350) as
// a binary PPM.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

#include "dynamic_library.hpp"
#include "exr_writer.hpp"
#include "hip_context.hpp"
#include "loader.hpp"
#include "manager.hpp"
#include "quasi/quasi_hip_ext.h"
#include "scene_file.hpp"

namespace {

void plugin_log(void* /*host*/, const char* msg) { std::printf("[Plugin] %s\n", msg); }
void plugin_request_shutdown(void* /*host*/) {}

struct Options {
    std::filesystem::path plugin, out, out_dir = ".", scene;
    int frames = 16;
    uint32_t width = 720, height = 720;   // the reference window (main.cpp:155)
    int device = 0;
    long spp = -1, depth = -1, rr_start = -1, seed = -1, integrator = -1, rng = -1, engine = -1;
    Q::io::exr_compression compression = Q::io::exr_compression::zip;
    bool single_layer = false;
    bool watch = false;
    int reload_at = -1;
    int wait_at = -1;                      // before this frame, wait until wait_file exists
    std::filesystem::path wait_file;
    std::filesystem::path ppm;
    bool orbit = false;
    float azimuth = 0.f, elevation = 0.f, distance = 3.5f, orbit_step = 0.f;
};

// The reference host's orbit camera (main.cpp:39-112) without the mouse input.
struct orbit_camera {
    float target[3] = {0.0f, 1.0f, 0.0f};
    float distance = 3.5f;
    float azimuth = 0.0f;     // radians
    float elevation = 0.0f;   // radians, clamped to +-1.5
    float fov = 40.0f;

    Q_camera camera() {
        elevation = std::min(std::max(elevation, -1.5f), 1.5f);
        const float ce = std::cos(elevation);
        Q_camera c{};
        c.position[0] = target[0] + distance * std::sin(azimuth) * ce;
        c.position[1] = target[1] + distance * std::sin(elevation);
        c.position[2] = target[2] + distance * std::cos(azimuth) * ce;
        for (int k = 0; k < 3; ++k) c.target[k] = target[k];
        c.up[0] = 0.0f;
        c.up[1] = 1.0f;
        c.up[2] = 0.0f;
        c.fov = fov;
        return c;
    }
};

// Binary PPM (P6) of the RGB channels of an RGBA8 image, top row first.
bool write_ppm(const std::filesystem::path& path, const std::vector<uint8_t>& rgba, uint32_t w, uint32_t h) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "P6\n%u %u\n255\n", w, h);
    std::vector<uint8_t> rgb(3 * size_t(w) * h);
    for (size_t i = 0; i < size_t(w) * h; ++i)
        for (int k = 0; k < 3; ++k) rgb[3 * i + k] = rgba[4 * i + k];
    const bool ok = std::fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
    return std::fclose(f) == 0 && ok;
}

int usage(const char* argv0) {
    std::fprintf(stderr,
                 "usage: %s [plugin.so] --render N [--out file.exr | --out-dir DIR] [--width W] [--height H]\n"
                 "       [--device K] [--scene file.qscene] [--spp-per-render S] [--depth D] [--rr-start R]\n"
                 "       [--seed X] [--integrator reference|nee] [--rng reference|hashed]\n"
                 "       [--engine megakernel|wavefront|single-launch] [--compression zip|none] [--single-layer]\n"
                 "       [--watch] [--reload-at K] [--wait-at K --wait-file PATH] [--ppm file.ppm]\n"
                 "       [--azimuth A] [--elevation E] [--distance D] [--orbit-step S]\n",
                 argv0);
    return EXIT_FAILURE;
}

bool parse(int argc, char** argv, Options& o) {
    for (int i = 1; i < argc; ++i) {
        const std::string_view a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
        const char* v = nullptr;
        if (a == "--single-layer") { o.single_layer = true; continue; }
        if (a == "--watch") { o.watch = true; continue; }
        if (a.size() > 2 && a.substr(0, 2) == "--") {
            v = next();
            if (!v) return false;
        }
        if (a == "--render") o.frames = std::atoi(v);
        else if (a == "--out") o.out = v;
        else if (a == "--out-dir") o.out_dir = v;
        else if (a == "--width") o.width = uint32_t(std::atol(v));
        else if (a == "--height") o.height = uint32_t(std::atol(v));
        else if (a == "--device") o.device = std::atoi(v);
        else if (a == "--scene") o.scene = v;
        else if (a == "--reload-at") o.reload_at = std::atoi(v);
        else if (a == "--wait-at") o.wait_at = std::atoi(v);
        else if (a == "--wait-file") o.wait_file = v;
        else if (a == "--ppm") o.ppm = v;
        else if (a == "--azimuth") { o.azimuth = std::strtof(v, nullptr); o.orbit = true; }
        else if (a == "--elevation") { o.elevation = std::strtof(v, nullptr); o.orbit = true; }
        else if (a == "--distance") { o.distance = std::strtof(v, nullptr); o.orbit = true; }
        else if (a == "--orbit-step") { o.orbit_step = std::strtof(v, nullptr); o.orbit = true; }
        else if (a == "--spp-per-render") o.spp = std::atol(v);
        else if (a == "--depth") o.depth = std::atol(v);
        else if (a == "--rr-start") o.rr_start = std::atol(v);
        else if (a == "--seed") o.seed = std::atol(v);
        else if (a == "--integrator") o.integrator = std::string_view(v) == "nee" ? Q_INTEGRATOR_NEE : Q_INTEGRATOR_REFERENCE;
        else if (a == "--rng") o.rng = std::string_view(v) == "hashed" ? Q_RNG_HASHED : Q_RNG_REFERENCE;
        else if (a == "--engine")
            o.engine = std::string_view(v) == "wavefront"       ? Q_ENGINE_WAVEFRONT
                       : std::string_view(v) == "single-launch" ? Q_ENGINE_PERSISTENT
                                                                : Q_ENGINE_MEGAKERNEL;
        else if (a == "--compression") o.compression = std::string_view(v) == "none" ? Q::io::exr_compression::none : Q::io::exr_compression::zip;
        else if (a[0] != '-') o.plugin = std::string(a);
        else return false;
    }
    return o.frames > 0 && o.width > 0 && o.height > 0;
}

std::filesystem::path default_plugin(const char* argv0) {
    std::error_code ec;
    auto self = std::filesystem::canonical("/proc/self/exe", ec);
    if (ec) self = std::filesystem::absolute(argv0);
    return self.parent_path().parent_path() / "libquasi_hip.so";
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, o)) return usage(argv[0]);
    if (o.plugin.empty()) o.plugin = default_plugin(argv[0]);

    auto ctx_result = Q::gpu::hip::context::create(o.device, o.width, o.height);
    if (!ctx_result) {
        std::fprintf(stderr, "Failed to create HIP context: %s\n", Q::gpu::hip::to_string(ctx_result.error()));
        return EXIT_FAILURE;
    }
    auto& hip = *ctx_result;

    // camera: the reference orbit camera at rest, or the scene's
    Q_camera cam{{0.0f, 1.0f, 3.5f}, {0.0f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, 40.0f};
    Q::io::scene_data scene;
    if (!o.scene.empty()) {
        auto sc = Q::io::read_scene(o.scene);
        if (!sc) {
            std::fprintf(stderr, "Scene %s: %s\n", o.scene.c_str(), Q::io::to_string(sc.error()));
            return EXIT_FAILURE;
        }
        scene = std::move(*sc);
        std::memcpy(cam.position, scene.camera, 3 * sizeof(float));
        std::memcpy(cam.target, scene.camera + 3, 3 * sizeof(float));
        std::memcpy(cam.up, scene.camera + 6, 3 * sizeof(float));
        cam.fov = scene.camera[9];
    }
    orbit_camera orbit;
    if (o.orbit) {
        for (int k = 0; k < 3; ++k) orbit.target[k] = cam.target[k];
        orbit.fov = cam.fov;
        orbit.distance = o.distance;
        orbit.azimuth = o.azimuth;
        orbit.elevation = o.elevation;
        cam = orbit.camera();
    }

    Q_plugin_context ctx{};
    ctx.viewport_width = o.width;
    ctx.viewport_height = o.height;
    ctx.host_data = nullptr;
    ctx.gpu = hip.gpu();
    ctx.log = plugin_log;
    ctx.request_shutdown = plugin_request_shutdown;

    // The scene and parameters the plugin gets after every load.
    Q::plugin::manager* mgr_ptr = nullptr;
    bool camera_dirty = true;
    auto configure = [&]() -> bool {
        Q::plugin::manager& m = *mgr_ptr;
        Q::plugin::loader* plugin = m.plugin();
        const auto& lib = m.library();
        if (!o.scene.empty()) {
            auto set_scene = lib.get_symbol<int32_t (*)(Q_plugin_handle*, const Q_scene_desc*)>("Q_plugin_set_scene");
            if (!set_scene) {
                std::fprintf(stderr, "The plugin has no Q_plugin_set_scene: cannot load %s\n", o.scene.c_str());
                return false;
            }
            Q_scene_desc d{};
            d.struct_size = sizeof d;
            d.num_vertices = uint32_t(scene.vertices.size() / 3);
            d.num_triangles = uint32_t(scene.indices.size() / 3);
            d.vertices = scene.vertices.data();
            d.indices = scene.indices.data();
            d.albedo = scene.albedo.data();
            d.emission = scene.emission.data();
            if ((*set_scene)(plugin->handle(), &d) != Q_EXT_OK) {
                std::fprintf(stderr, "Q_plugin_set_scene failed\n");
                return false;
            }
            std::printf("[Host] Scene %s: %u triangles\n", o.scene.c_str(), d.num_triangles);
        }
        if (o.spp > 0 || o.depth > 0 || o.rr_start >= 0 || o.seed >= 0 || o.integrator >= 0 || o.rng >= 0 ||
            o.engine >= 0) {
            auto get = lib.get_symbol<int32_t (*)(Q_plugin_handle*, Q_render_params*)>("Q_plugin_get_params");
            auto set = lib.get_symbol<int32_t (*)(Q_plugin_handle*, const Q_render_params*)>("Q_plugin_set_params");
            if (!get || !set) {
                std::fprintf(stderr, "The plugin has no render-parameter exports\n");
                return false;
            }
            Q_render_params p{};
            p.struct_size = sizeof p;
            if ((*get)(plugin->handle(), &p) != Q_EXT_OK) {
                std::fprintf(stderr, "Q_plugin_get_params failed\n");
                return false;
            }
            if (o.spp > 0) p.spp_per_render = uint32_t(o.spp);
            if (o.depth > 0) p.max_depth = uint32_t(o.depth);
            if (o.rr_start >= 0) p.rr_start = uint32_t(o.rr_start);
            if (o.seed >= 0) p.seed = uint32_t(o.seed);
            if (o.integrator >= 0) p.integrator = uint32_t(o.integrator);
            if (o.rng >= 0) p.rng_mode = uint32_t(o.rng);
            if (o.engine >= 0) p.engine = uint32_t(o.engine);
            if ((*set)(plugin->handle(), &p) != Q_EXT_OK) {
                std::fprintf(stderr, "Q_plugin_set_params rejected the parameters\n");
                return false;
            }
        }
        camera_dirty = true;   // the accumulation starts again with this plugin
        const Q_plugin_info info = plugin->info();
        std::printf("Loaded plugin: %s v%u.%u.%u\n", info.name, info.version.major, info.version.minor,
                    info.version.patch);
        std::printf("Description: %s\n", info.description);
        return true;
    };
    bool configured = true;
    Q::plugin::reload_hooks hooks;
    hooks.pre_unload = [&] { hip.finish(); };   // nothing of the old plugin may still run
    hooks.post_load = [&] { configured = configure(); };
    hooks.on_error = [](const std::string& msg) { std::fprintf(stderr, "[Host] %s\n", msg.c_str()); };
    Q::plugin::manager mgr(o.plugin, ctx, hooks);
    mgr_ptr = &mgr;
    if (auto r = mgr.load(); !r) {
        std::fprintf(stderr, "Failed to load plugin: %s (%s)\n", Q::plugin::manager::to_string(r.error()),
                     o.plugin.c_str());
        return EXIT_FAILURE;
    }
    if (!configure()) return EXIT_FAILURE;

    std::printf("[Host] Batch mode: rendering %d frames then saving EXR\n", o.frames);
    std::fflush(stdout);   // a watcher of this output learns that the frame loop starts
    auto last = std::chrono::steady_clock::now();
    const auto t0 = last;
    int since_load = 0;   // frames accumulated by the current plugin
    for (int i = 0; i < o.frames; ++i) {
        if (i == o.wait_at && !o.wait_file.empty()) {
            // a fixed point for whoever drives the host (a test changing the
            // watched library): frame i starts once the file exists
            std::printf("[Host] Waiting before frame %d for %s\n", i, o.wait_file.c_str());
            std::fflush(stdout);
            const auto until = std::chrono::steady_clock::now() + std::chrono::seconds(120);
            std::error_code ec;
            while (!std::filesystem::exists(o.wait_file, ec)) {
                if (std::chrono::steady_clock::now() > until) {
                    std::fprintf(stderr, "[Host] %s did not appear within 120 s\n", o.wait_file.c_str());
                    return EXIT_FAILURE;
                }
                std::this_thread::sleep_for(std::chrono::milliseconds(5));
            }
        }
        bool reloaded = false;
        if (i == o.reload_at) {
            (void)mgr.reload();
            reloaded = true;
        } else if (o.watch) {
            reloaded = mgr.poll();
        }
        if (reloaded) {
            const auto& st = mgr.stats();
            std::printf("[Host] Plugin reloaded before frame %d (%llu reloads, %llu ok, %llu failed, %.1f ms)\n", i,
                        (unsigned long long)st.reload_count, (unsigned long long)st.success_count,
                        (unsigned long long)st.failure_count, st.last_reload_time * 1000.f);
            if (!mgr.is_loaded() || !configured) {
                std::fprintf(stderr, "[Host] No plugin after the reload\n");
                return EXIT_FAILURE;
            }
            since_load = 0;
        }
        if (o.orbit && o.orbit_step != 0.f && i > 0) {   // the drag between two frames
            orbit.azimuth += o.orbit_step;
            cam = orbit.camera();
            camera_dirty = true;
            since_load = 0;
        }
        const auto now = std::chrono::steady_clock::now();
        mgr.update(std::chrono::duration<float>(now - last).count());
        last = now;
        auto fr = hip.begin_frame();
        if (!fr) {
            std::fprintf(stderr, "begin_frame failed: %s\n", Q::gpu::hip::to_string(fr.error()));
            return EXIT_FAILURE;
        }
        Q_render_frame& frame = *fr;
        frame.camera = cam;
        frame.camera_dirty = camera_dirty ? 1u : 0u;
        camera_dirty = false;
        mgr.render(&frame);
        hip.end_frame(frame);
        ++since_load;
    }
    hip.finish();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("[Host] %d frames in %.3f s\n", o.frames, secs);

    const auto path = o.out.empty() ? Q::io::make_timestamped_path(o.out_dir) : o.out;
    int rc = EXIT_SUCCESS;
    Q::plugin::loader& plugin = *mgr.plugin();
    if (!o.ppm.empty()) {
        auto display = mgr.library().get_symbol<int32_t (*)(Q_plugin_handle*, uint8_t*, uint64_t)>(
            "Q_plugin_readback_display");
        std::vector<uint8_t> rgba(4 * size_t(o.width) * o.height);
        if (!display || (*display)(plugin.handle(), rgba.data(), rgba.size()) != Q_EXT_OK ||
            !write_ppm(o.ppm, rgba, o.width, o.height)) {
            std::fprintf(stderr, "[Host] Display image not written: %s\n", o.ppm.c_str());
            rc = EXIT_FAILURE;
        } else {
            std::printf("[Host] Saved display image: %s\n", o.ppm.c_str());
        }
    }
    if (plugin.supports_readback_aov() && !o.single_layer) {
        std::printf("[Host] Saving EXR with AOVs (%d samples)...\n", since_load);
        Q_readback_aov_result rb = plugin.readback_aov();
        if (rb.buffers[Q_AOV_BEAUTY].data) {
            auto w = Q::io::write_exr(path, rb, o.compression);
            if (w) std::printf("[Host] Saved: %s\n", path.c_str());
            else { std::fprintf(stderr, "[Host] EXR write failed: %s\n", Q::io::to_string(w.error())); rc = EXIT_FAILURE; }
            plugin.readback_aov_free(&rb);
        } else {
            std::fprintf(stderr, "[Host] AOV readback returned no data\n");
            rc = EXIT_FAILURE;
        }
    } else if (plugin.supports_readback()) {
        std::printf("[Host] Saving EXR (%d samples)...\n", since_load);
        Q_readback_result rb = plugin.readback();
        if (rb.data) {
            auto w = Q::io::write_exr(path, rb, o.compression);
            if (w) std::printf("[Host] Saved: %s\n", path.c_str());
            else { std::fprintf(stderr, "[Host] EXR write failed: %s\n", Q::io::to_string(w.error())); rc = EXIT_FAILURE; }
            plugin.readback_free(&rb);
        } else {
            std::fprintf(stderr, "[Host] Readback returned no data\n");
            rc = EXIT_FAILURE;
        }
    } else {
        std::printf("[Host] Plugin does not support HDR readback\n");
    }
    return rc;
}
