// kernels_wave.hip -- the wavefront engine.
//
// One sample pass over N = W*H*S paths runs as
//   raygen -> (extend -> shade -> shadow) x max_depth -> accumulate
// with every queue length kept on the device, so the host enqueues a fixed
// launch sequence without synchronising (and the sequence can be captured in
// a hipGraph).  extend and shadow are persistent kernels: waves fetch 64-ray
// chunks from a per-bounce cursor until the queue is drained.  shade compacts
// the surviving paths and the requested shadow rays with one ballot and one
// atomic per wave.  Radiance lives in a per-path slot and is reduced per pixel
// in sample order by accumulate, so the result does not depend on the order in
// which queues were filled.
#include "device.hpp"
#include "kernels.hpp"

namespace quasi_hip {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Path id p -> (pixel, sample within the pass).
__device__ __forceinline__ void path_pixel(const PassParams& P, uint32_t p, uint32_t& x, uint32_t& y,
                                           uint32_t& s) {
    const uint32_t npix = P.width * P.height;
    s = p / npix;
    const uint32_t pix = p - s * npix;
    y = pix / P.width;
    x = pix - y * P.width;
}

__global__ __launch_bounds__(256) void k_raygen(PassParams P, WaveState S) {
    if (blockIdx.x == 0 && threadIdx.x < 2 * kMaxDepth + 1) {
        if (threadIdx.x <= kMaxDepth) S.counts[threadIdx.x] = threadIdx.x == 0 ? S.paths : 0u;
        if (threadIdx.x < kMaxDepth) S.shadow_counts[threadIdx.x] = 0u;
        if (threadIdx.x < 2 * kMaxDepth) S.fetch[threadIdx.x] = 0u;
    }
    float4* ro = static_cast<float4*>(S.ray_o[0]);
    float4* rd = static_cast<float4*>(S.ray_d[0]);
    float4* thr = static_cast<float4*>(S.thr);
    float4* rad = static_cast<float4*>(S.radiance);
    float4* a0 = static_cast<float4*>(S.aov0);
    float4* a1 = static_cast<float4*>(S.aov1);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < S.paths; p += gridDim.x * blockDim.x) {
        uint32_t x, y, s;
        path_pixel(P, p, x, y, s);
        uint32_t rng = path_seed(P.rng_mode, x, y, P.width, P.height, P.sample_begin + s, P.seed);
        const Ray r = camera_ray(P, x, y, rng);
        ro[p] = make_float4(r.o.x, r.o.y, r.o.z, __uint_as_float(p));
        rd[p] = make_float4(r.d.x, r.d.y, r.d.z, 0.f);
        thr[p] = make_float4(1.f, 1.f, 1.f, 0.f);
        rad[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        a0[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        a1[p] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kMissId));
        S.rng[p] = rng;
    }
}

template <int STACK, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_extend(PassParams P, WaveState S, uint32_t b) {
    __shared__ int s_stack[STACK * BLOCK];
    int* stk = s_stack + threadIdx.x;
    const uint32_t n = S.counts[b];
    if (blockIdx.x == 0 && threadIdx.x == 0 && n)
        atomicAdd(b == 0 ? &P.counters->rays_camera : &P.counters->rays_extension, (unsigned long long)n);
    const float4* __restrict__ nodes = static_cast<const float4*>(P.nodes);
    const float4* __restrict__ tris = static_cast<const float4*>(P.tris);
    const float4* __restrict__ ro = static_cast<const float4*>(S.ray_o[b & 1u]);
    const float4* __restrict__ rd = static_cast<const float4*>(S.ray_d[b & 1u]);
    uint2* hit = static_cast<uint2*>(S.hit);
    uint32_t base;
    while (fetch_chunk(&S.fetch[2 * b], n, base)) {
        const uint32_t i = base + lane_id();
        if (i < n) {
            const float4 o = ro[i], d = rd[i];
            TraceHit h{kTFar, -1, 0u};
            if (P.num_nodes)
                h = trace_closest<BLOCK>(nodes, tris, Ray{xyz(o), xyz(d)}, kTMin, kTFar, stk);
            hit[i] = make_uint2(__float_as_uint(h.t), uint32_t(h.tri));
        }
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_shade(PassParams P, WaveState S, uint32_t b) {
    const uint32_t n = S.counts[b];
    const float4* __restrict__ tris = static_cast<const float4*>(P.tris);
    const float4* __restrict__ mats = static_cast<const float4*>(P.mats);
    const float4* __restrict__ ro = static_cast<const float4*>(S.ray_o[b & 1u]);
    const float4* __restrict__ rd = static_cast<const float4*>(S.ray_d[b & 1u]);
    float4* no = static_cast<float4*>(S.ray_o[(b + 1) & 1u]);
    float4* nd = static_cast<float4*>(S.ray_d[(b + 1) & 1u]);
    const uint2* hit = static_cast<const uint2*>(S.hit);
    float4* thr_a = static_cast<float4*>(S.thr);
    float4* rad_a = static_cast<float4*>(S.radiance);
    float4* a0 = static_cast<float4*>(S.aov0);
    float4* a1 = static_cast<float4*>(S.aov1);
    float4* so_a = static_cast<float4*>(S.sh_o);
    float4* sd_a = static_cast<float4*>(S.sh_d);
    float4* sc_a = static_cast<float4*>(S.sh_c);
    // wave-uniform grid-stride loop so that every lane reaches the ballots
    const uint32_t wave_base = blockIdx.x * BLOCK + (threadIdx.x & ~63u);
    for (uint32_t wb = wave_base; wb < n; wb += gridDim.x * BLOCK) {
        const uint32_t i = wb + lane_id();
        bool cont = false, shadow = false;
        ShadeOut so;
        uint32_t p = 0;
        if (i < n) {
            const float4 o = ro[i], d = rd[i];
            p = __float_as_uint(o.w);
            const uint2 h = hit[i];
            const int tri = int(h.y);
            if (tri >= 0) {
                float4 t4 = thr_a[p], l4 = rad_a[p];
                f3 thr = xyz(t4), L = xyz(l4);
                uint32_t rng = S.rng[p];
                FirstHit first;
                shade_hit(P, tris, mats, b, Ray{xyz(o), xyz(d)}, __uint_as_float(h.x), tri, thr, L, rng,
                          b == 0 ? &first : nullptr, so);
                if (b == 0) {
                    a0[p] = make_float4(first.albedo.x, first.albedo.y, first.albedo.z, first.t);
                    a1[p] = make_float4(first.normal.x, first.normal.y, first.normal.z, __uint_as_float(first.id));
                }
                thr_a[p] = make_float4(thr.x, thr.y, thr.z, 0.f);
                rad_a[p] = make_float4(L.x, L.y, L.z, 0.f);
                S.rng[p] = rng;
                cont = so.cont;
                shadow = so.shadow;
            }
        }
        const uint32_t j = wave_append(cont, &S.counts[b + 1]);
        if (cont) {
            no[j] = make_float4(so.next.o.x, so.next.o.y, so.next.o.z, __uint_as_float(p));
            nd[j] = make_float4(so.next.d.x, so.next.d.y, so.next.d.z, 0.f);
        }
        const uint32_t k = wave_append(shadow, &S.shadow_counts[b]);
        if (shadow) {
            so_a[k] = make_float4(so.sray.o.x, so.sray.o.y, so.sray.o.z, so.stmax);
            sd_a[k] = make_float4(so.sray.d.x, so.sray.d.y, so.sray.d.z, __uint_as_float(p));
            sc_a[k] = make_float4(so.scontrib.x, so.scontrib.y, so.scontrib.z, 0.f);
        }
    }
}

template <int STACK, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_shadow(PassParams P, WaveState S, uint32_t b) {
    __shared__ int s_stack[STACK * BLOCK];
    int* stk = s_stack + threadIdx.x;
    const uint32_t n = S.shadow_counts[b];
    if (blockIdx.x == 0 && threadIdx.x == 0 && n)
        atomicAdd(&P.counters->rays_shadow, (unsigned long long)n);
    const float4* __restrict__ nodes = static_cast<const float4*>(P.nodes);
    const float4* __restrict__ tris = static_cast<const float4*>(P.tris);
    const float4* __restrict__ so_a = static_cast<const float4*>(S.sh_o);
    const float4* __restrict__ sd_a = static_cast<const float4*>(S.sh_d);
    const float4* __restrict__ sc_a = static_cast<const float4*>(S.sh_c);
    float4* rad = static_cast<float4*>(S.radiance);
    uint32_t base;
    while (fetch_chunk(&S.fetch[2 * b + 1], n, base)) {
        const uint32_t k = base + lane_id();
        if (k < n) {
            const float4 o = so_a[k], d = sd_a[k];
            if (!trace_any<BLOCK>(nodes, tris, Ray{xyz(o), xyz(d)}, kTMin, o.w, stk)) {
                const uint32_t p = __float_as_uint(d.w);
                const float4 c = sc_a[k];
                float4 l = rad[p];
                l.x += c.x; l.y += c.y; l.z += c.z;
                rad[p] = l;
            }
        }
    }
}

// Per pixel: the pass's samples in sample order, added to the accumulation.
__global__ __launch_bounds__(256) void k_accumulate(PassParams P, WaveState S) {
    const uint32_t npix = P.width * P.height;
    const uint32_t spp = S.paths / npix;
    const float4* rad = static_cast<const float4*>(S.radiance);
    const float4* a0 = static_cast<const float4*>(S.aov0);
    const float4* a1 = static_cast<const float4*>(S.aov1);
    float4* ab = static_cast<float4*>(P.accum_beauty);
    float4* aa = static_cast<float4*>(P.accum_albedo);
    float4* an = static_cast<float4*>(P.accum_normal);
    float4* ad = static_cast<float4*>(P.accum_depth);
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < npix; pix += gridDim.x * blockDim.x) {
        float4 sb = {0, 0, 0, 0}, sa = {0, 0, 0, 0}, sn = {0, 0, 0, 0}, sd = {0, 0, 0, 0};
        for (uint32_t s = 0; s < spp; ++s) {
            const uint32_t p = s * npix + pix;
            const float4 l = rad[p], x0 = a0[p], x1 = a1[p];
            const float dn = fminf(fmaxf((x0.w - 1.0f) / 4.0f, 0.0f), 1.0f);   // plugin.mm:301
            sb.x += l.x; sb.y += l.y; sb.z += l.z; sb.w += 1.0f;
            sa.x += x0.x; sa.y += x0.y; sa.z += x0.z; sa.w += 1.0f;
            sn.x += x1.x; sn.y += x1.y; sn.z += x1.z; sn.w += 1.0f;
            sd.x += dn; sd.y += dn; sd.z += dn; sd.w += 1.0f;
            if (P.record_ids && P.sample_begin + s == P.id_sample) P.ids[pix] = __float_as_uint(x1.w);
        }
        float4 o;
        o = ab[pix]; o.x += sb.x; o.y += sb.y; o.z += sb.z; o.w += sb.w; ab[pix] = o;
        o = aa[pix]; o.x += sa.x; o.y += sa.y; o.z += sa.z; o.w += sa.w; aa[pix] = o;
        o = an[pix]; o.x += sn.x; o.y += sn.y; o.z += sn.z; o.w += sn.w; an[pix] = o;
        o = ad[pix]; o.x += sd.x; o.y += sd.y; o.z += sd.z; o.w += sd.w; ad[pix] = o;
    }
}

uint32_t grid_for(size_t n, uint32_t block, uint32_t cap) {
    size_t g = (n + block - 1) / block;
    return uint32_t(g < 1 ? 1 : (g > cap ? cap : g));
}

constexpr int kTraceBlock = 256;
constexpr int kShadeBlock = 256;

template <int STACK>
hipError_t launch_trace_pair(const PassParams& p, const WaveState& s, uint32_t b, uint32_t persistent_blocks,
                             uint32_t shade_blocks, bool shadows, hipStream_t stream, WaveTimers* tm) {
    if (tm && tm->before) tm->before(tm->user, WAVE_K_EXTEND, stream);
    hipLaunchKernelGGL((k_extend<STACK, kTraceBlock>), dim3(persistent_blocks), dim3(kTraceBlock), 0, stream, p, s, b);
    if (tm && tm->after) tm->after(tm->user, WAVE_K_EXTEND, stream);
    if (tm && tm->before) tm->before(tm->user, WAVE_K_SHADE, stream);
    hipLaunchKernelGGL((k_shade<kShadeBlock>), dim3(shade_blocks), dim3(kShadeBlock), 0, stream, p, s, b);
    if (tm && tm->after) tm->after(tm->user, WAVE_K_SHADE, stream);
    if (shadows) {
        if (tm && tm->before) tm->before(tm->user, WAVE_K_SHADOW, stream);
        hipLaunchKernelGGL((k_shadow<STACK, kTraceBlock>), dim3(persistent_blocks), dim3(kTraceBlock), 0, stream, p, s, b);
        if (tm && tm->after) tm->after(tm->user, WAVE_K_SHADOW, stream);
    }
    return hipGetLastError();
}

}  // namespace

uint32_t wavefront_persistent_blocks(int device, uint32_t bvh_depth) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    int per_cu = 0;
    hipError_t e = bvh_depth <= 32
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_extend<32, kTraceBlock>, kTraceBlock, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_extend<64, kTraceBlock>, kTraceBlock, 0);
    if (e != hipSuccess || per_cu <= 0) per_cu = 2;
    return uint32_t(cus * per_cu);
}

hipError_t launch_wavefront_pass(const PassParams& p, const WaveState& s, uint32_t bvh_depth,
                                 uint32_t persistent_blocks, hipStream_t stream, WaveTimers* tm) {
    if (s.paths == 0) return hipSuccess;
    if (p.max_depth > kMaxDepth || bvh_depth > 64) return hipErrorInvalidValue;
    const uint32_t gen_blocks = grid_for(s.paths, 256, 16384);
    const uint32_t shade_blocks = grid_for(s.paths, kShadeBlock, 8192);
    if (tm && tm->before) tm->before(tm->user, WAVE_K_RAYGEN, stream);
    hipLaunchKernelGGL(k_raygen, dim3(gen_blocks), dim3(256), 0, stream, p, s);
    if (tm && tm->after) tm->after(tm->user, WAVE_K_RAYGEN, stream);
    const bool shadows = p.integrator != 0u && p.num_lights != 0u;
    for (uint32_t b = 0; b < p.max_depth; ++b) {
        hipError_t e = bvh_depth <= 32
            ? launch_trace_pair<32>(p, s, b, persistent_blocks, shade_blocks, shadows, stream, tm)
            : launch_trace_pair<64>(p, s, b, persistent_blocks, shade_blocks, shadows, stream, tm);
        if (e != hipSuccess) return e;
    }
    if (tm && tm->before) tm->before(tm->user, WAVE_K_ACCUMULATE, stream);
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(size_t(p.width) * p.height, 256, 16384)), dim3(256), 0, stream, p, s);
    if (tm && tm->after) tm->after(tm->user, WAVE_K_ACCUMULATE, stream);
    return hipGetLastError();
}

}  // namespace quasi_hip
