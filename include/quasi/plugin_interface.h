/* plugin_interface.h -- the Quasi plugin C ABI, version 3.
 *
 * Mirrors the reference's src/quasi/plugin/plugin_interface.hpp:20-183: the same
 * structs, the same 10 exported symbol names and the same ABI version (3), so an
 * unmodified Quasi host (loader.hpp:65-137 accepts versions 1..3) can dlopen
 * libquasi_hip.so in place of libquasi_metal.dylib.  Written as C-compatible
 * declarations; the struct layouts are identical to the C++ originals.
 */
#ifndef QUASI_PLUGIN_INTERFACE_H
#define QUASI_PLUGIN_INTERFACE_H

#include <stddef.h>
#include <stdint.h>

#include "gpu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define Q_PLUGIN_ABI_VERSION 3u /* plugin_interface.hpp:183 */

typedef struct Q_plugin_handle Q_plugin_handle; /* plugin_interface.hpp:27 */

/* plugin_interface.hpp:30-34 */
typedef struct Q_plugin_version {
    uint32_t major;
    uint32_t minor;
    uint32_t patch;
} Q_plugin_version;

/* plugin_interface.hpp:37-42 */
typedef struct Q_plugin_info {
    const char* name;
    Q_plugin_version version;
    const char* description;
    const char* author;
} Q_plugin_info;

/* plugin_interface.hpp:48-59 */
typedef struct Q_plugin_context {
    uint32_t viewport_width;
    uint32_t viewport_height;
    void* host_data;
    Q_gpu_context* gpu;
    void (*log)(void* host_data, const char* message);
    void (*request_shutdown)(void* host_data);
} Q_plugin_context;

/* plugin_interface.hpp:65-70: RGBA float, row-major, top-to-bottom */
typedef struct Q_readback_result {
    float* data;
    uint32_t width;
    uint32_t height;
    uint32_t channels;
} Q_readback_result;

/* plugin_interface.hpp:73-79 */
#ifdef __cplusplus
enum Q_aov_type : uint32_t {
#else
typedef uint32_t Q_aov_type;
enum {
#endif
    Q_AOV_BEAUTY = 0,
    Q_AOV_ALBEDO = 1,
    Q_AOV_NORMAL = 2,
    Q_AOV_DEPTH  = 3,
    Q_AOV_COUNT  = 4
};

/* plugin_interface.hpp:82-87 */
typedef struct Q_aov_buffer {
    float* data;
    uint32_t width;
    uint32_t height;
    uint32_t channels;
} Q_aov_buffer;

/* plugin_interface.hpp:90-92 */
typedef struct Q_readback_aov_result {
    Q_aov_buffer buffers[Q_AOV_COUNT];
} Q_readback_aov_result;

/* ---- The 10 exported entry points (plugin_interface.hpp:108-148). ---- */

/* Replaces plugin.mm:543-545.  Returns 3. */
uint32_t Q_plugin_abi_version(void);

/* Replaces plugin.mm:547-554.  Static strings. */
Q_plugin_info Q_plugin_get_info(void);

/* Replaces plugin.mm:556-592.  Returns NULL (and logs through ctx->log) when ctx
 * or ctx->gpu is NULL, when ctx->gpu->backend != Q_GPU_BACKEND_HIP, or when no HIP
 * device can be opened.  Builds the reference Cornell box (cornell_box.hpp:95-134)
 * as the default scene, like plugin.mm:584. */
Q_plugin_handle* Q_plugin_create(Q_plugin_context* ctx);

/* Replaces plugin.mm:594-617.  NULL-safe; frees every device allocation, stream
 * and event the plugin owns. */
void Q_plugin_destroy(Q_plugin_handle* handle);

/* Replaces plugin.mm:619-622.  No-op. */
void Q_plugin_update(Q_plugin_handle* handle, float delta_time);

/* Replaces plugin.mm:624-792.  Adds the configured number of samples per pixel
 * (one by default, as in the reference) to the FP32 accumulation.  Returns
 * silently on a NULL handle or frame.  Honours resize (plugin.mm:638) and
 * camera_dirty (plugin.mm:643).  frame->command_buffer is the hipStream_t to
 * enqueue on (NULL: the plugin's stream); frame->drawable is ignored. */
void Q_plugin_render(Q_plugin_handle* handle, Q_render_frame* frame);

/* Replaces plugin.mm:846-863.  malloc'd RGBA32F mean radiance, top-to-bottom;
 * data == NULL on failure or before the first render. */
Q_readback_result Q_plugin_readback(Q_plugin_handle* handle);

/* Replaces plugin.mm:865-870. */
void Q_plugin_readback_free(Q_readback_result* result);

/* Replaces plugin.mm:872-892.  Beauty, albedo, normal, depth (RGBA32F each). */
Q_readback_aov_result Q_plugin_readback_aov(Q_plugin_handle* handle);

/* Replaces plugin.mm:894-902. */
void Q_plugin_readback_aov_free(Q_readback_aov_result* result);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* QUASI_PLUGIN_INTERFACE_H */
