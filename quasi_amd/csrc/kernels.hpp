// kernels.hpp -- host-callable launchers of the HIP kernels (all asynchronous on
// the given stream; they return the launch status only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.hpp"

namespace quasi_hip {

// Deepest BVH (internal levels) the traversal stacks support.
uint32_t megakernel_max_depth();

// One sample pass of the megakernel engine over the whole framebuffer.
hipError_t launch_megakernel(const PassParams& p, uint32_t bvh_depth, hipStream_t stream);

// Wavefront engine kernels, for timing hooks.
enum WaveKernel : uint32_t {
    WAVE_K_RAYGEN = 0,
    WAVE_K_EXTEND,
    WAVE_K_SHADE,
    WAVE_K_SHADOW,
    WAVE_K_ACCUMULATE,
    WAVE_K_COUNT
};

// Optional callbacks around every wavefront launch (used to record hipEvents).
struct WaveTimers {
    void* user;
    void (*before)(void* user, uint32_t kernel, hipStream_t stream);
    void (*after)(void* user, uint32_t kernel, hipStream_t stream);
};

// Compiled launch shapes of the persistent traversal kernels.
bool trace_config_supported(const TraceConfig& c);
uint32_t trace_config_count();
TraceConfig trace_config_at(uint32_t i);

// Blocks of the persistent extend/shadow kernels that fill the device once
// (never more than kMaxThreadsPerCU threads per CU).
uint32_t persistent_blocks(int device, const TraceConfig& c);

// One sample pass of the wavefront engine over s.paths = W*H*S paths (samples
// p.sample_begin .. p.sample_begin + S - 1): a fixed, host-synchronisation-free
// launch sequence.  p.gstack must hold p.gstack_entries words for each of
// trace_blocks * cfg.block threads, with kShortStack + p.gstack_entries at
// least the deepest traversal stack of the chosen BVH width.
hipError_t launch_wavefront_pass(const PassParams& p, const WaveState& s, const TraceConfig& cfg,
                                 uint32_t trace_blocks, hipStream_t stream, WaveTimers* timers);

// out[i] = sum[i] / n for npix RGBA pixels (the running mean of plugin.mm:326-327).
hipError_t launch_resolve(const float4* sum, float4* out, size_t npix, uint32_t n, hipStream_t stream);
// 8-bit display image (RGBA, 4 bytes per pixel) of the mean of n samples (plugin.mm:334-350).
hipError_t launch_tonemap(const float4* sum, void* out, size_t npix, uint32_t n, hipStream_t stream);

// Path-replay adjoint over the whole framebuffer with `blocks` blocks of
// adjoint_block_size() threads (blocks * adjoint_block_size() must not exceed
// the threads p.pass.gstack was sized for).  Adds into the FP32 replicas,
// which the caller clears first.
uint32_t adjoint_block_size();
uint32_t adjoint_max_depth();
hipError_t launch_adjoint(const AdjointParams& p, uint32_t blocks, hipStream_t stream);

// out[i] = sum over r < replicas of (double) rep[r * n + i].
hipError_t launch_replica_reduce(const float* rep, size_t n, uint32_t replicas, double* out, hipStream_t stream);

// L2 image loss against a target: dimg = 2 (sum / n - target) per pixel (rgb),
// and one FP32 partial sum of squared differences per block in partials
// (l2_image_blocks(npix) of them).
uint32_t l2_image_blocks(size_t npix);
hipError_t launch_l2_image(const float4* sum, uint32_t n, const float4* target, float4* dimg, size_t npix,
                           float* partials, hipStream_t stream);

}  // namespace quasi_hip

#include "bvh_build.hpp"

namespace quasi_hip {

// Triangle count from which Q_BVH_BUILD_AUTO builds on the GPU.
constexpr uint32_t kGpuBuildMinTriangles = 1u << 14;   // the GPU build is as fast as host threads from ~10k

// Fills n 32-bit words with value (used instead of hipMemsetAsync so that the
// clears stay valid inside captured graphs).
hipError_t launch_fill_u32(void* ptr, size_t n, uint32_t value, hipStream_t stream);

}  // namespace quasi_hip
