// hip_context.cpp -- see hip_context.hpp.
#include "hip_context.hpp"

#include <hip/hip_runtime.h>

#include <utility>

namespace Q::gpu::hip {

context::result<context> context::create(int ordinal, uint32_t width, uint32_t height) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || ordinal < 0 || ordinal >= count)
        return Q::unexpected(context_error::device_not_found);
    if (hipSetDevice(ordinal) != hipSuccess) return Q::unexpected(context_error::device_not_found);
    context c;
    c.ordinal_ = new int(ordinal);
    hipStream_t s = nullptr;
    // default priority (the plugin's kernels and this stream share the queue order)
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return Q::unexpected(context_error::stream_failed);
    c.stream_ = s;
    for (auto& e : c.events_) {
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            return Q::unexpected(context_error::stream_failed);
        e = ev;
    }
    c.width_ = width;
    c.height_ = height;
    c.bind();
    return c;
}

void context::bind() {
    gpu_.backend = Q_GPU_BACKEND_HIP;
    gpu_.device = ordinal_;
    gpu_.queue = stream_;
    gpu_.layer = nullptr;
}

context::~context() { release(); }

context::context(context&& o) noexcept { *this = std::move(o); }

context& context::operator=(context&& o) noexcept {
    if (this != &o) {
        release();
        ordinal_ = std::exchange(o.ordinal_, nullptr);
        stream_ = std::exchange(o.stream_, nullptr);
        for (int i = 0; i < kFramesInFlight; ++i) events_[i] = std::exchange(o.events_[i], nullptr);
        frame_ = o.frame_;
        width_ = o.width_;
        height_ = o.height_;
        o.gpu_ = Q_gpu_context{};
        bind();
    }
    return *this;
}

void context::release() {
    if (stream_) (void)hipStreamSynchronize(static_cast<hipStream_t>(stream_));
    for (auto& e : events_) {
        if (e) (void)hipEventDestroy(static_cast<hipEvent_t>(e));
        e = nullptr;
    }
    if (stream_) (void)hipStreamDestroy(static_cast<hipStream_t>(stream_));
    stream_ = nullptr;
    delete ordinal_;
    ordinal_ = nullptr;
    gpu_ = Q_gpu_context{};
}

context::result<Q_render_frame> context::begin_frame() {
    // the slot's previous frame must have finished before it is reused
    if (frame_ >= uint64_t(kFramesInFlight)) {
        hipEvent_t e = static_cast<hipEvent_t>(events_[frame_ % kFramesInFlight]);
        if (hipEventSynchronize(e) != hipSuccess) return Q::unexpected(context_error::event_failed);
    }
    Q_render_frame f{};
    f.drawable = nullptr;
    f.command_buffer = stream_;
    f.width = width_;
    f.height = height_;
    return f;
}

void context::end_frame(Q_render_frame& frame) {
    (void)frame;
    (void)hipEventRecord(static_cast<hipEvent_t>(events_[frame_ % kFramesInFlight]),
                         static_cast<hipStream_t>(stream_));
    ++frame_;
}

void context::resize(uint32_t width, uint32_t height) {
    width_ = width;
    height_ = height;
}

void context::finish() {
    if (stream_) (void)hipStreamSynchronize(static_cast<hipStream_t>(stream_));
}

}  // namespace Q::gpu::hip
