// kernels_mega.hip -- the megakernel engine: one thread per pixel, every
// sample of the pass traced start to finish in registers, then one
// read-modify-write of the FP32 accumulation.
//
// Replaces the reference's per-frame passes pathtrace_frag (plugin.mm:287-309,
// with path_trace :223-276) and accumulate (:313-330): the four AOVs are
// accumulated in the same kernel instead of four extra dispatches.
#include "device.hpp"
#include "kernels.hpp"

namespace quasi_hip {

namespace {

constexpr int kTileW = 16;

template <int STACK, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_megakernel(PassParams P) {
    __shared__ int s_stack[STACK * BLOCK];
    int* stk = s_stack + threadIdx.x;
    const float4* __restrict__ nodes = static_cast<const float4*>(P.nodes);
    const float4* __restrict__ tris = static_cast<const float4*>(P.tris);
    const float4* __restrict__ mats = static_cast<const float4*>(P.mats);

    // 16 x (BLOCK/16) pixel tiles: a wave covers a 16 x 4 patch, so primary
    // rays of one wave are spatially coherent.
    constexpr int kTileH = BLOCK / kTileW;
    const uint32_t tiles_x = (P.width + kTileW - 1) / kTileW;
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const uint32_t x = tx * kTileW + (threadIdx.x % kTileW);
    const uint32_t y = ty * kTileH + (threadIdx.x / kTileW);
    const bool active = x < P.width && y < P.height;

    uint32_t n_cam = 0, n_ext = 0, n_shadow = 0;
    if (active) {
        const size_t pix = size_t(y) * P.width + x;
        float4 acc_b = {0, 0, 0, 0}, acc_a = {0, 0, 0, 0}, acc_n = {0, 0, 0, 0}, acc_d = {0, 0, 0, 0};
        for (uint32_t s = P.sample_begin; s < P.sample_begin + P.sample_count; ++s) {
            uint32_t rng = path_seed(P.rng_mode, x, y, P.width, P.height, s, P.seed);
            Ray ray = camera_ray(P.cam, P.width, P.height, x, y, rng);
            f3 thr = mk(1.f, 1.f, 1.f);
            f3 L = mk(0.f, 0.f, 0.f);
            FirstHit first{mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f), 0.f, kMissId};
            for (uint32_t b = 0; b < P.max_depth; ++b) {
                TraceHit h{kTFar, -1, 0u};
                if (P.num_nodes) h = trace_closest<BLOCK>(nodes, tris, ray, kTMin, kTFar, stk);
                if (b == 0) ++n_cam; else ++n_ext;
                if (h.tri < 0) break;
                ShadeOut so;
                shade_hit(P, tris, mats, P.lights, P.num_lights, b, ray, h.t, h.tri, thr, L, rng,
                          b == 0 ? &first : nullptr, so);
                if (so.shadow) {
                    ++n_shadow;
                    if (!trace_any<BLOCK>(nodes, tris, so.sray, kTMin, so.stmax, stk)) L = L + so.scontrib;
                }
                if (!so.cont) break;
                ray = so.next;
            }
            // plugin.mm:301-307
            const float nd = fminf(fmaxf((first.t - 1.0f) / 4.0f, 0.0f), 1.0f);
            acc_b.x += L.x; acc_b.y += L.y; acc_b.z += L.z; acc_b.w += 1.0f;
            acc_a.x += first.albedo.x; acc_a.y += first.albedo.y; acc_a.z += first.albedo.z; acc_a.w += 1.0f;
            acc_n.x += first.normal.x; acc_n.y += first.normal.y; acc_n.z += first.normal.z; acc_n.w += 1.0f;
            acc_d.x += nd; acc_d.y += nd; acc_d.z += nd; acc_d.w += 1.0f;
            if (P.record_ids && s == P.id_sample) P.ids[pix] = first.id;
        }
        float4* ab = static_cast<float4*>(P.accum_beauty);
        float4* aa = static_cast<float4*>(P.accum_albedo);
        float4* an = static_cast<float4*>(P.accum_normal);
        float4* ad = static_cast<float4*>(P.accum_depth);
        float4 o;
        o = ab[pix]; o.x += acc_b.x; o.y += acc_b.y; o.z += acc_b.z; o.w += acc_b.w; ab[pix] = o;
        o = aa[pix]; o.x += acc_a.x; o.y += acc_a.y; o.z += acc_a.z; o.w += acc_a.w; aa[pix] = o;
        o = an[pix]; o.x += acc_n.x; o.y += acc_n.y; o.z += acc_n.z; o.w += acc_n.w; an[pix] = o;
        o = ad[pix]; o.x += acc_d.x; o.y += acc_d.y; o.z += acc_d.z; o.w += acc_d.w; ad[pix] = o;
    }
    // one atomic per wave and counter
    n_cam = wave_sum(n_cam);
    n_ext = wave_sum(n_ext);
    n_shadow = wave_sum(n_shadow);
    if ((threadIdx.x & 63u) == 0u) {
        if (n_cam) atomicAdd(&P.counters->rays_camera, (unsigned long long)n_cam);
        if (n_ext) atomicAdd(&P.counters->rays_extension, (unsigned long long)n_ext);
        if (n_shadow) atomicAdd(&P.counters->rays_shadow, (unsigned long long)n_shadow);
    }
}

template <int STACK, int BLOCK>
hipError_t launch_mega_t(const PassParams& p, hipStream_t stream) {
    constexpr int kTileH = BLOCK / kTileW;
    const uint32_t tiles = ((p.width + kTileW - 1) / kTileW) * ((p.height + kTileH - 1) / kTileH);
    hipLaunchKernelGGL((k_megakernel<STACK, BLOCK>), dim3(tiles), dim3(BLOCK), 0, stream, p);
    return hipGetLastError();
}

}  // namespace

uint32_t megakernel_max_depth() { return 64; }

hipError_t launch_megakernel(const PassParams& p, uint32_t bvh_depth, hipStream_t stream) {
    if (p.width == 0 || p.height == 0 || p.sample_count == 0) return hipSuccess;
    // the traversal stack holds at most one entry per internal level
    if (bvh_depth <= 32) return launch_mega_t<32, 128>(p, stream);
    if (bvh_depth <= 64) return launch_mega_t<64, 128>(p, stream);
    return hipErrorInvalidValue;
}

}  // namespace quasi_hip
