// hip_context.hpp -- the host's GPU context for the HIP backend.
//
// Counterpart of the reference's Metal context
// (src/quasi/gpu/metal/context.hpp:54-103): create, gpu() for the plugin's
// Q_gpu_context, begin_frame / end_frame around each Q_plugin_render call,
// resize.  Headless: there is no window layer, so frames carry no drawable.
//   Q_gpu_context: backend Q_GPU_BACKEND_HIP, device -> the host-owned int
//     ordinal, queue -> the context's hipStream_t, layer = NULL.
//   Q_render_frame: command_buffer -> the same stream.
// Frames in flight are bounded the way the Metal context bounds them with a
// dispatch semaphore (context.mm:117, :138): begin_frame waits for the event
// recorded by end_frame kFramesInFlight frames earlier.
#pragma once

#include <cstdint>

#include "expected.hpp"
#include "quasi/gpu_types.h"

namespace Q::gpu::hip {

enum class context_error {
    device_not_found,
    stream_failed,
    event_failed,
};

constexpr const char* to_string(context_error e) noexcept {
    switch (e) {
        case context_error::device_not_found: return "HIP device not found";
        case context_error::stream_failed: return "HIP stream creation failed";
        case context_error::event_failed: return "HIP event wait failed";
    }
    return "Unknown error";
}

class context {
public:
    static constexpr int kFramesInFlight = 3;

    template <class T>
    using result = Q::expected<T, context_error>;

    // The GPU with this ordinal; the frame size starts at width x height.
    static result<context> create(int ordinal, uint32_t width, uint32_t height);

    context() = default;
    ~context();
    context(context&& o) noexcept;
    context& operator=(context&& o) noexcept;
    context(const context&) = delete;
    context& operator=(const context&) = delete;

    Q_gpu_context* gpu() noexcept { return &gpu_; }
    const Q_gpu_context* gpu() const noexcept { return &gpu_; }

    result<Q_render_frame> begin_frame();
    void end_frame(Q_render_frame& frame);
    void resize(uint32_t width, uint32_t height);
    // Waits for all frames submitted so far.
    void finish();

    uint32_t width() const noexcept { return width_; }
    uint32_t height() const noexcept { return height_; }
    int ordinal() const noexcept { return *ordinal_; }

private:
    void release();
    void bind();

    int* ordinal_ = nullptr;           // heap-held so gpu_.device stays valid across moves
    void* stream_ = nullptr;           // hipStream_t
    void* events_[kFramesInFlight] = {};
    uint64_t frame_ = 0;
    uint32_t width_ = 0, height_ = 0;
    Q_gpu_context gpu_{};
};

}  // namespace Q::gpu::hip
