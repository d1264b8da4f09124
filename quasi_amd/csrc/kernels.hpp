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

// out[i] = sum[i] / n for npix RGBA pixels (the running mean of plugin.mm:326-327).
hipError_t launch_resolve(const float4* sum, float4* out, size_t npix, uint32_t n, hipStream_t stream);

// Fills n 32-bit words with value (used instead of hipMemsetAsync so that the
// clears stay valid inside captured graphs).
hipError_t launch_fill_u32(void* ptr, size_t n, uint32_t value, hipStream_t stream);

}  // namespace quasi_hip
