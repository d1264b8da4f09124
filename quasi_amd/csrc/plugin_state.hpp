// plugin_state.hpp -- the plugin's internal state and host helpers, shared by
// plugin.cpp (the reference ABI and single-scene extensions) and
// plugin_batch.cpp (batched scenes, pass graphs, autotuner).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/quasi/plugin_interface.h"
#include "../../include/quasi/quasi_hip_ext.h"
#include "bvh_gpu.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace quasi_hip {
namespace host {

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

enum KernelSlot : uint32_t {
    K_MEGAKERNEL = 0, K_RESOLVE, K_FILL,
    K_RAYGEN, K_EXTEND, K_SHADE, K_SHADOW, K_ACCUMULATE,  // wavefront, in WaveKernel order
    K_BATCH_PASS,                                          // one replay of a captured batch pass
    K_PATHS,                                               // one single-launch pass (Q_ENGINE_PERSISTENT)
    K_COUNT
};
extern const char* const kKernelNames[K_COUNT];
extern const bool kRenderKernel[K_COUNT];

// Wavefront path-state buffers (see WaveState in common.hpp).
struct WaveBuffers {
    DeviceBuffer ray_o[2], ray_d[2], hit, thr, radiance, rng, aov0, aov1, sh_o, sh_d, sh_c, tq[2];
    DeviceBuffer counts, shadow_counts, fetch;
    uint32_t capacity = 0;
};

struct TimedLaunch {
    uint32_t slot;
    hipEvent_t start, stop;
};

// One mesh prepared in the layouts of common.hpp: host arrays to upload
// (host build), or arrays already on the device (GPU build: tri4 .. nodes4
// stay empty and `device` holds them until upload_scene adopts them).
struct SceneHost {
    std::vector<float> tri4, mat4, nodes, nodes4;
    std::vector<LightEntry> lights;
    DeviceScene device;
    uint32_t num_tris = 0;    // input triangles
    uint32_t num_prims = 0;   // BVH-order positions (references of the split triangles)
    uint32_t num_nodes = 0, num_nodes4 = 0, depth = 0, depth4 = 0;
    uint32_t builder = Q_BVH_BUILD_HOST;   // where the BVH was built
    double sah_cost = 0, device_ms = 0;
    double gather_ms = 0, bvh_ms = 0, layout_ms = 0;   // host phases of a host build
    BvhGpuTimes gpu_times;                               // phases of a GPU build
};

// Several scenes traced together (Q_plugin_set_batch).
struct BatchState {
    bool active = false;
    std::vector<SceneDesc> descs;          // host copy with device pointers
    DeviceBuffer tris, mats, nodes, nodes4, lights, table;
    DeviceBuffer qnodes4;                  // quantised BVH4 nodes of the concatenated arrays
    bool quant_ok = false;                 // qnodes4 holds every scene's nodes
    DeviceBuffer accum[Q_AOV_COUNT], ids, resolved, sample_begin;
    uint32_t num_scenes = 0, total_pixels = 0, total_tris = 0;
    uint32_t total_prims = 0;              // BVH-order triangle records of the concatenated arrays
    uint32_t total_nodes = 0, total_nodes4 = 0;   // BVH2 / BVH4 nodes of the concatenated arrays
    uint32_t sample_count = 0;
    uint32_t gstack_entries = 0;
    double build_ms = 0;
    // captured pass
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_key = 0;
};

// Buffers of Q_plugin_adjoint (kept between calls).
struct AdjointState {
    DeviceBuffer indices;        // 3 vertex indices per input triangle
    DeviceBuffer image;          // caller's image: dL/dI or the L2 target
    DeviceBuffer dimg;           // dL/dI formed on the device (L2 target mode)
    DeviceBuffer grads;          // FP32 replicas: albedo then vertices
    DeviceBuffer out;            // FP64 reduced gradients: albedo then vertices
    DeviceBuffer partials;       // FP32 per-block loss sums
    DeviceBuffer rec_hit;        // path records of the replayed pass: hit (t, triangle) per bounce and path
    DeviceBuffer rec_vis;        // ... and shadow-ray visibility (bytes)
    DeviceBuffer counters;       // ray counters of the recorded passes (kept out of the render statistics)
    DeviceBuffer param_map;      // albedo parameter of each triangle (input order)
    DeviceBuffer param_out;      // FP64 gradients of the albedo parameters
    uint64_t indices_generation = 0;
};

struct plugin_state {
    Q_plugin_context* context = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipStream_t last_stream = nullptr;
    hipStream_t capture_stream = nullptr;
    char arch[64] = {0};

    Q_render_params params{};

    // host copy of the scene
    std::vector<float> verts;
    std::vector<uint32_t> indices;
    std::vector<float> albedo, emission;

    // device scene
    DeviceBuffer tris, mats, nodes, nodes4, lights, gstack;
    DeviceBuffer qnodes4;               // the BVH4 as quantised 64-B nodes (bvh_quant.hpp)
    DeviceBuffer quant_flag;            // one word: set when a node is beyond the quantised range
    bool quant_ok = false;              // qnodes4 holds the current scene's nodes
    uint32_t num_tris = 0, num_nodes = 0, bvh_depth = 0, num_lights = 0;
    uint32_t num_prims = 0;   // BVH-order triangle records (references of the split triangles)
    uint32_t num_nodes4 = 0, bvh_depth4 = 0, gstack_entries = 0;
    double bvh_build_ms = 0;
    uint32_t bvh_builder = Q_BVH_BUILD_HOST;
    double bvh_sah_cost = 0, bvh_device_ms = 0;

    // framebuffer
    uint32_t width = 0, height = 0;
    uint32_t sample_count = 0;
    DeviceBuffer accum[Q_AOV_COUNT];
    DeviceBuffer ids;
    DeviceBuffer resolved;
    DeviceBuffer counters;

    // wavefront engine
    WaveBuffers wave;
    TraceConfig trace_cfg{4, 128, kChunkQuant};
    uint32_t trace_blocks = 0;
    uint32_t paths_blocks = 0;          // grid of the single-launch pass kernel for trace_cfg (0: unsupported)
    hipEvent_t pending_start = nullptr;
    int num_cus = 256;

    BatchState batch;
    AdjointState adjoint;
    uint64_t scene_generation = 0;      // incremented by every scene upload
    bool scene_ok = false;              // the device scene is whole (false after a failed upload)
    Q_camera last_camera{};
    bool have_camera = false;
    void* noalloc_scratch = nullptr;    // this device's working area of Q_plugin_render_noalloc

    // stats
    std::vector<hipEvent_t> event_pool;
    std::vector<TimedLaunch> timed;
    uint64_t launches[K_COUNT] = {};
    double kernel_ms[K_COUNT] = {};
};

void log_msg(plugin_state* st, const char* fmt, ...);
bool hip_ok(plugin_state* st, hipError_t e, const char* what);
void free_buffer(DeviceBuffer& b);
bool ensure_buffer(plugin_state* st, DeviceBuffer& b, size_t bytes, const char* what);
// Quantised 64-byte form (bvh_quant.hpp) of n device BVH4 nodes into q;
// in_range is false when some node is beyond the quantised range (the 128-byte
// nodes are then traced).  False on a device or allocation error.
bool quantize_bvh4(plugin_state* st, const void* nodes4, uint32_t n, DeviceBuffer& q, bool& in_range);
hipEvent_t take_event(plugin_state* st);
void harvest_events(plugin_state* st);
bool sync_render(plugin_state* st);

// Brackets one launch with events when kernel timing is on.
struct LaunchTimer {
    plugin_state* st;
    hipStream_t s;
    uint32_t slot;
    hipEvent_t a = nullptr, b = nullptr;
    LaunchTimer(plugin_state* st_, hipStream_t s_, uint32_t slot_);
    ~LaunchTimer();
};

// Prepares one mesh (BVH build, triangle / material / node / light arrays);
// threads = 0 uses up to 16 host threads for a host BVH build and lets
// builder (Q_BVH_BUILD_*) choose the GPU build; threads > 0 always builds on
// the host with that many threads.
int32_t prepare_scene(plugin_state* st, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const float* albedo, const float* emission, unsigned threads, SceneHost& out,
                      uint32_t builder = Q_BVH_BUILD_HOST);

// Overflow entries per thread that the persistent traversal stacks need for
// a BVH2 of `depth` and a BVH4 of `depth4` levels.
uint32_t stack_overflow_entries(uint32_t depth, uint32_t depth4);

// The persistent kernels' overflow stacks for `entries` per thread.
bool ensure_gstack(plugin_state* st, uint32_t entries);

bool ensure_wave(plugin_state* st, uint32_t paths);
// The path state of a pass of `paths` paths (batch: contiguous queue segments).
WaveState wave_state(const WaveBuffers& w, uint32_t paths, bool batch);
void wave_timer_before(void* user, uint32_t kernel, hipStream_t s);
void wave_timer_after(void* user, uint32_t kernel, hipStream_t s);

// Host camera packing (plugin.mm:649-674) and the basis of get_camera_ray.
CameraBasis make_camera(const Q_camera& c, uint32_t w, uint32_t h);

// Mean of npix RGBA sums (divided by n samples) into a malloc'd array.
Q_aov_buffer read_mean_of(plugin_state* st, const void* sums, uint32_t width, uint32_t height, uint32_t n);

// Takes the address of this device's Q_plugin_render_noalloc working area
// (st->device must be current); false when the runtime refuses.
bool noalloc_init(plugin_state* st);

void free_batch(plugin_state* st);
void free_adjoint(plugin_state* st);

// The body of Q_plugin_render (false when a step failed; the failure is logged).
bool render_frame(plugin_state* st, const Q_render_frame* frame);

// Zeroes the single-scene accumulation and its sample count.
bool clear_single(plugin_state* st, hipStream_t s);

}  // namespace host
}  // namespace quasi_hip
