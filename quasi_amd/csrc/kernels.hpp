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

// The end of a pass: per pixel, the pass's per-path radiance and first-hit
// records summed in sample order into the accumulation (and, for a captured
// pass, the device-side sample index moved on).
hipError_t launch_pass_accumulate(const PassParams& p, const WaveState& s, hipStream_t stream);

// The single-launch form of a pass (Q_ENGINE_PERSISTENT, and batches rendered
// with it): one persistent grid traces every path of the pass from its camera
// ray to its end -- closest hit, shading, the shadow ray, the next segment --
// then launch_pass_accumulate.  Same results as launch_wavefront_pass, bit for
// bit.  cfg.chunk is not used; `blocks` = paths_pass_blocks(device, cfg).
bool paths_config_supported(const TraceConfig& c);
uint32_t paths_pass_blocks(int device, const TraceConfig& c);
hipError_t launch_paths_pass(const PassParams& p, const WaveState& s, const TraceConfig& cfg, uint32_t blocks,
                             hipStream_t stream);

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

// out[param[t]] += per_tri[t] (3 doubles per triangle and parameter) for the
// nt triangles whose param[t] < np; out must be cleared first.
hipError_t launch_param_scatter(const double* per_tri, const uint32_t* param, size_t nt, uint32_t np, double* out,
                               hipStream_t stream);

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

// Writes the quantised 64-byte form (bvh_quant.hpp) of n 128-byte BVH4 nodes
// (nodes4) into q, and sets *failed (a device word the caller cleared) to 1
// when some node's extent is beyond the quantised range.
hipError_t launch_quantize4(const void* nodes4, uint32_t n, void* q, uint32_t* failed, hipStream_t stream);

}  // namespace quasi_hip
