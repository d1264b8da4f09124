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
        const f3 llo = mk(P.cam.llo[0], P.cam.llo[1], P.cam.llo[2]);
        const f3 hor = mk(P.cam.horizontal[0], P.cam.horizontal[1], P.cam.horizontal[2]);
        const f3 ver = mk(P.cam.vertical[0], P.cam.vertical[1], P.cam.vertical[2]);
        const f3 corg = mk(P.cam.origin[0], P.cam.origin[1], P.cam.origin[2]);
        const float u0 = (float(x) + 0.5f) / float(P.width);
        const float v0 = 1.0f - (float(y) + 0.5f) / float(P.height);

        float4 acc_b = {0, 0, 0, 0}, acc_a = {0, 0, 0, 0}, acc_n = {0, 0, 0, 0}, acc_d = {0, 0, 0, 0};
        for (uint32_t s = P.sample_begin; s < P.sample_begin + P.sample_count; ++s) {
            uint32_t rng = path_seed(P.rng_mode, x, y, P.width, P.height, s, P.seed);
            // plugin.mm:131-132 (x jitter drawn first)
            const float jx = random_float(rng);
            const float jy = random_float(rng);
            const float u = u0 + (jx - 0.5f) * 0.001f;
            const float v = v0 + (jy - 0.5f) * 0.001f;
            Ray ray{corg, normalize(llo + hor * u + ver * v)};

            f3 thr = mk(1.f, 1.f, 1.f);
            f3 L = mk(0.f, 0.f, 0.f);
            f3 aov_alb = mk(0.f, 0.f, 0.f), aov_nrm = mk(0.f, 0.f, 0.f);
            float aov_t = 0.f;
            uint32_t prim_id = kMissId;
            for (uint32_t b = 0; b < P.max_depth; ++b) {
                TraceHit h{kTFar, -1, 0u};
                if (P.num_nodes) h = trace_closest<BLOCK>(nodes, tris, ray, kTMin, kTFar, stk);
                if (b == 0) ++n_cam; else ++n_ext;
                if (h.tri < 0) break;
                const float4 t0 = tris[3 * h.tri + 0];
                const float4 t1 = tris[3 * h.tri + 1];
                const float4 t2 = tris[3 * h.tri + 2];
                const f3 n = normalize(cross(xyz(t1), xyz(t2)));
                const f3 nf = dot(n, ray.d) < 0.f ? n : -n;          // plugin.mm:192
                const f3 pt = ray.o + ray.d * h.t;                     // plugin.mm:179
                const float4 m0 = mats[2 * h.tri + 0];
                const float4 m1 = mats[2 * h.tri + 1];
                const f3 alb = xyz(m0), Le = xyz(m1);
                const float emit_str = m0.w;
                (void)t0;
                if (b == 0) {                                          // plugin.mm:242-249
                    aov_alb = emit_str > 0.1f ? Le : alb;
                    aov_nrm = nf * 0.5f + mk(0.5f, 0.5f, 0.5f);
                    aov_t = h.t;
                    prim_id = h.id;
                }
                if (P.integrator == 0u) {
                    L = L + thr * Le;                                  // plugin.mm:252
                    if (emit_str > 0.1f) break;                        // plugin.mm:255-258
                    thr = thr * alb;                                   // plugin.mm:261
                    if (b + 1 >= P.max_depth) break;                   // no further trace
                    if (b >= P.rr_start) {                             // plugin.mm:264-268
                        const float q = fmaxf(0.05f, max3(thr));
                        if (random_float(rng) > q) break;
                        thr = thr * (1.0f / q);
                    }
                    ray.o = pt + nf * kOffset;                         // plugin.mm:271-272
                    ray.d = cosine_sample(nf, rng);
                } else {
                    if (b == 0) L = L + Le;
                    if (emit_str > 0.1f) break;
                    if (b + 1 >= P.max_depth) break;
                    const f3 org = pt + nf * kOffset;
                    if (P.num_lights) {
                        // light choice: first entry with r < cdf (last when none)
                        const float rs = random_float(rng);
                        uint32_t lo = 0, hi = P.num_lights;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (rs < P.lights[mid].cdf) hi = mid; else lo = mid + 1;
                        }
                        if (lo >= P.num_lights) lo = P.num_lights - 1;
                        const LightEntry le = P.lights[lo];
                        const float r1 = random_float(rng);
                        const float r2 = random_float(rng);
                        const float4 l0 = tris[3 * le.tri + 0];
                        const float4 l1 = tris[3 * le.tri + 1];
                        const float4 l2 = tris[3 * le.tri + 2];
                        const float su = sqrtf(r1);
                        const f3 xl = xyz(l0) + xyz(l1) * (su * (1.0f - r2)) + xyz(l2) * (su * r2);
                        // geometry term at the hit point; the offset origin
                        // is used only for the visibility ray
                        const f3 w = xl - pt;
                        const float dist2 = dot(w, w);
                        const f3 wi = w * (1.0f / sqrtf(dist2));
                        const f3 nl = normalize(cross(xyz(l1), xyz(l2)));
                        const float cos_x = dot(nf, wi);
                        const float cos_l = fabsf(dot(nl, wi));
                        if (cos_x > 0.f && cos_l > 0.f) {
                            ++n_shadow;
                            const f3 sv = xl - org;
                            const float sl = sqrtf(dot(sv, sv));
                            const f3 ds = sv * (1.0f / sl);
                            const bool occ = trace_any<BLOCK>(nodes, tris, Ray{org, ds}, kTMin, sl * 0.999f, stk);
                            if (!occ) {
                                const f3 Ll = xyz(mats[2 * le.tri + 1]);
                                const float g = cos_x * cos_l * le.area / (kPi * dist2 * le.psel);
                                L = L + thr * alb * Ll * g;
                            }
                        }
                    }
                    thr = thr * alb;
                    if (b >= P.rr_start) {
                        const float q = fmaxf(0.05f, max3(thr));
                        if (random_float(rng) > q) break;
                        thr = thr * (1.0f / q);
                    }
                    ray.o = org;
                    ray.d = cosine_sample(nf, rng);
                }
            }
            // plugin.mm:301-307
            const float nd = fminf(fmaxf((aov_t - 1.0f) / 4.0f, 0.0f), 1.0f);
            acc_b.x += L.x; acc_b.y += L.y; acc_b.z += L.z; acc_b.w += 1.0f;
            acc_a.x += aov_alb.x; acc_a.y += aov_alb.y; acc_a.z += aov_alb.z; acc_a.w += 1.0f;
            acc_n.x += aov_nrm.x; acc_n.y += aov_nrm.y; acc_n.z += aov_nrm.z; acc_n.w += 1.0f;
            acc_d.x += nd; acc_d.y += nd; acc_d.z += nd; acc_d.w += 1.0f;
            if (P.record_ids && s == P.id_sample) P.ids[pix] = prim_id;
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
