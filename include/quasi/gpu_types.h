/* gpu_types.h -- backend-agnostic GPU context types of the Quasi plugin ABI.
 *
 * Same layout and names as the reference's src/quasi/gpu/types.hpp:14-70
 * (Q_gpu_backend, Q_gpu_context, Q_camera, Q_render_frame), written so that it
 * compiles as C as well as C++.  One additive change: Q_GPU_BACKEND_HIP = 4
 * (the reference enum ends at WEBGPU = 3, types.hpp:14-19).
 *
 * HIP meaning of the opaque pointers:
 *   Q_gpu_context.device  -> const int*   (HIP device ordinal owned by the host)
 *   Q_gpu_context.queue   -> hipStream_t  (may be NULL: the plugin then makes its own stream)
 *   Q_gpu_context.layer   -> NULL         (headless; no swapchain)
 *   Q_render_frame.drawable       -> NULL (headless) or a host-owned pointer the plugin ignores
 *   Q_render_frame.command_buffer -> hipStream_t for this frame (NULL: the plugin's stream)
 */
#ifndef QUASI_GPU_TYPES_H
#define QUASI_GPU_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef __cplusplus
enum Q_gpu_backend : uint32_t {
#else
typedef uint32_t Q_gpu_backend;
enum {
#endif
    Q_GPU_BACKEND_NONE   = 0,
    Q_GPU_BACKEND_METAL  = 1,
    Q_GPU_BACKEND_VULKAN = 2,
    Q_GPU_BACKEND_WEBGPU = 3,
    Q_GPU_BACKEND_HIP    = 4
};

/* types.hpp:37-42 */
typedef struct Q_gpu_context {
    Q_gpu_backend backend;
    void* device;
    void* queue;
    void* layer;
} Q_gpu_context;

/* types.hpp:45-50 */
typedef struct Q_camera {
    float position[3];
    float target[3];
    float up[3];
    float fov;          /* vertical field of view, degrees */
} Q_camera;

/* types.hpp:63-70 */
typedef struct Q_render_frame {
    void* drawable;
    void* command_buffer;
    uint32_t width;
    uint32_t height;
    Q_camera camera;
    uint32_t camera_dirty;
} Q_render_frame;

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* QUASI_GPU_TYPES_H */
