// kernels_util.hip -- small bandwidth kernels: accumulation resolve and fills.
#include "kernels.hpp"

namespace quasi_hip {

namespace {

__global__ __launch_bounds__(256) void k_resolve(const float4* __restrict__ sum, float4* __restrict__ out,
                                                 size_t npix, float n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < npix; i += size_t(gridDim.x) * blockDim.x) {
        float4 s = sum[i];
        out[i] = make_float4(s.x / n, s.y / n, s.z / n, s.w / n);
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t value) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = value;
}

uint32_t grid_for(size_t n) {
    size_t blocks = (n + 255) / 256;
    return uint32_t(blocks < 8192 ? (blocks ? blocks : 1) : 8192);
}

}  // namespace

hipError_t launch_resolve(const float4* sum, float4* out, size_t npix, uint32_t n, hipStream_t stream) {
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_resolve, dim3(grid_for(npix)), dim3(256), 0, stream, sum, out, npix, float(n));
    return hipGetLastError();
}

hipError_t launch_fill_u32(void* ptr, size_t n, uint32_t value, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(256), 0, stream, static_cast<uint32_t*>(ptr), n, value);
    return hipGetLastError();
}

}  // namespace quasi_hip
