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

// The display image of plugin.mm:334-350 (tonemap_frag): mean radiance
// (the sum over n samples divided as in k_resolve), Reinhard c / (c + 1),
// gamma 1/2.2, stored as 8-bit unorm (clamped to [0, 1], times 255, rounded
// to nearest even) with alpha 255, rows top-to-bottom like the accumulation.
__global__ __launch_bounds__(256) void k_tonemap(const float4* __restrict__ sum, uchar4* __restrict__ out,
                                                 size_t npix, float n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < npix; i += size_t(gridDim.x) * blockDim.x) {
        const float4 s = sum[i];
        const float m[3] = {s.x / n, s.y / n, s.z / n};
        unsigned char u[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float c = m[k] / (m[k] + 1.f);
            const float g = powf(c, 1.f / 2.2f);
            u[k] = (unsigned char)rintf(fminf(fmaxf(g, 0.f), 1.f) * 255.f);
        }
        out[i] = make_uchar4(u[0], u[1], u[2], 255);
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

hipError_t launch_tonemap(const float4* sum, void* out, size_t npix, uint32_t n, hipStream_t stream) {
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tonemap, dim3(grid_for(npix)), dim3(256), 0, stream, sum, static_cast<uchar4*>(out), npix,
                       float(n));
    return hipGetLastError();
}

hipError_t launch_fill_u32(void* ptr, size_t n, uint32_t value, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(256), 0, stream, static_cast<uint32_t*>(ptr), n, value);
    return hipGetLastError();
}

}  // namespace quasi_hip
