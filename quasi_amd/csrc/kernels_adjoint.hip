// kernels_adjoint.hip -- path-replay adjoint of the two integrators.
//
// For a loss L with image gradient dL/dI, the adjoint computes
//   dL/dtheta = sum_pixels dL/dI . (1/S) sum_s dC_s/dtheta
// for theta = per-triangle albedo and vertex positions, where C_s is the
// radiance of one path with its discrete decisions held fixed (triangles hit,
// facing sides, light choices, shadow visibility, Russian-roulette outcomes
// and probabilities).  Nothing is kept from the forward render.  A wavefront
// pass over the same samples (the forward's own kernels) first records, per
// path and bounce, the hit (distance and triangle) and the shadow ray's
// visibility; then each thread of k_adjoint replays its path from the RNG
// seed with the forward's code (camera_ray, the recorded hits, shade_hit)
// without any BVH traversal, records the path's ray segments in
// LDS (at most max_depth of them), and runs reverse mode over that record in
// the same thread.
//
// Radiance of a path with hit vertices j = 0..K-1:
//   C = sum_j thr_j (E_j + alb_j N_j),  thr_{j+1} = thr_j alb_j / q_j
// with E_j the emission the integrator adds at j and N_j = Ll g_j the
// unoccluded NEE term per unit throughput and albedo.  With the suffix
//   V_j = E_j + alb_j U_j,  U_j = N_j + V_{j+1} / q_j,
// the albedo gradient is dC/dalb_j = thr_j U_j (path replay backpropagation
// for a multiplicative throughput).  Geometry enters through the hit points
// x_j = o_j + t_j d_j (ray-plane distance of the frozen triangle), the facing
// normals, the NEE geometry term g_j and, for the next segment, the offset
// origin and the cosine-sampled direction, whose local frame turns with the
// normal; reverse mode runs over that chain from the last vertex back to the
// camera.
//
// Gradients are accumulated in FP32 replicas (block % replicas) with
// hardware float atomics after a per-triangle wave aggregation; the final
// reduction over replicas is the only FP64 arithmetic.
#include "device.hpp"
#include "kernels.hpp"

namespace quasi_hip {

namespace {

constexpr int kAdjBlock = 128;      // one 16 x 8 pixel tile per block
constexpr int kRecWords = 11;       // o.xyz, d.xyz, thr.xyz, tri | flags << 31, rng
constexpr int kAggRounds = 8;       // per-triangle wave aggregation rounds before direct atomics
constexpr uint32_t kRecNee = 1u;    // the vertex's NEE sample was requested and unoccluded
// triangle indices are below 2^29, so the flag rides in the top bit of the
// triangle word: 11 words per vertex let 7 blocks (14 waves) share a CU's LDS
constexpr uint32_t kRecFlagShift = 31;

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Vector-Jacobian product of cosine_sample's direction with respect to its
// normal argument, for the same draws (r1, r2) and the same frame choice.
__device__ __forceinline__ f3 cosine_sample_vjp(f3 n, float r1, float r2, f3 gbar) {
    const float phi = 2.0f * kPi * r1;
    const float cos_t = sqrtf(1.0f - r2);
    const float sin_t = sqrtf(r2);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float ln = sqrtf(dot(n, n));
    const f3 w = n * (1.0f / ln);
    const f3 a = fabsf(w.x) > 0.9f ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
    const f3 vv = cross(w, a);
    const float lv = sqrtf(dot(vv, vv));
    const f3 v = vv * (1.0f / lv);
    const f3 u = cross(w, v);
    const f3 D = u * (cp * sin_t) + v * (sp * sin_t) + w * cos_t;
    const float lD = sqrtf(dot(D, D));
    const f3 dir = D * (1.0f / lD);
    const f3 Db = (gbar - dir * dot(dir, gbar)) * (1.0f / lD);
    f3 ub = Db * (cp * sin_t);
    f3 vb = Db * (sp * sin_t);
    f3 wb = Db * cos_t;
    wb = wb + cross(v, ub);               // u = w x v
    vb = vb + cross(ub, w);
    const f3 Vb = (vb - v * dot(v, vb)) * (1.0f / lv);
    wb = wb + cross(a, Vb);               // v = normalize(w x a)
    return (wb - w * dot(w, wb)) * (1.0f / ln);
}

// Adds one triangle's gradient: 3 albedo channels and the positions of its
// three vertices.
__device__ __forceinline__ void emit_grads(uint32_t tri, const float (&g)[12], float* ga, float* gv,
                                           const uint32_t* __restrict__ idx) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (g[c] != 0.f) unsafeAtomicAdd(ga + 3 * size_t(tri) + c, g[c]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float gx = g[3 + 3 * k], gy = g[4 + 3 * k], gz = g[5 + 3 * k];
        if (gx == 0.f && gy == 0.f && gz == 0.f) continue;
        const size_t vi = idx[3 * size_t(tri) + k];
        if (gx != 0.f) unsafeAtomicAdd(gv + 3 * vi + 0, gx);
        if (gy != 0.f) unsafeAtomicAdd(gv + 3 * vi + 1, gy);
        if (gz != 0.f) unsafeAtomicAdd(gv + 3 * vi + 2, gz);
    }
}

// Scatter of the active lanes' gradients.  Lanes with the same triangle are
// summed across the wave first (one round per distinct triangle, led by the
// lowest pending lane); lanes still pending after kAggRounds rounds add
// their own values.  Every lane of the wave must call this.
__device__ __forceinline__ void scatter_grads(bool act, uint32_t key, const float (&g)[12], float* ga, float* gv,
                                              const uint32_t* __restrict__ idx) {
    const int lane = int(threadIdx.x & 63u);
    bool pending = act;
    for (int round = 0; round < kAggRounds; ++round) {
        const unsigned long long pm = __ballot(pending);
        if (pm == 0ull) return;
        const int leader = __ffsll((unsigned long long)pm) - 1;
        const uint32_t kl = __shfl(key, leader, 64);
        const bool mine = pending && key == kl;
        float s[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = wave_sum_f(mine ? g[k] : 0.f);
        if (lane == leader) emit_grads(kl, s, ga, gv, idx);
        pending = pending && !mine;
    }
    if (pending) emit_grads(key, g, ga, gv, idx);
}

template <int MAXD>
__global__ __launch_bounds__(kAdjBlock) void k_adjoint(AdjointParams A) {
    __shared__ uint32_t s_rec[MAXD * kRecWords * kAdjBlock];
    const PassParams& P = A.pass;
    const float4* __restrict__ tris = static_cast<const float4*>(P.tris);
    const float4* __restrict__ mats = static_cast<const float4*>(P.mats);
    const uint2* __restrict__ rec_hit = static_cast<const uint2*>(P.rec_hit);
    const uint8_t* __restrict__ rec_vis = P.rec_vis;
    const float4* __restrict__ dimg = static_cast<const float4*>(A.dimg);
    const uint32_t W = P.width, H = P.height;
    // the band of rows [row_begin, row_end) that the recorded pass traced
    const uint32_t row0 = A.row_begin, rows = A.rows ? A.rows : H, row_end = row0 + rows;
    const uint32_t tiles_x = (W + 15u) / 16u;
    const uint32_t tiles = tiles_x * ((rows + 7u) / 8u);
    const uint32_t rep = blockIdx.x % A.replicas;
    float* ga = A.grad_albedo + size_t(rep) * 3 * A.num_tris;
    float* gv = A.grad_vertices + size_t(rep) * 3 * A.num_vertices;
    uint32_t* rec = s_rec + threadIdx.x;
#define REC(j, f) rec[((j) * kRecWords + (f)) * kAdjBlock]

    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t x = (tile % tiles_x) * 16u + (threadIdx.x & 15u);
        const uint32_t y = row0 + (tile / tiles_x) * 8u + (threadIdx.x >> 4);
        f3 delta = mk(0.f, 0.f, 0.f);
        if (x < W && y < row_end) {
            const float4 g = dimg[size_t(y) * W + x];
            delta = mk(g.x, g.y, g.z) * A.scale;
        }
        const bool live = delta.x != 0.f || delta.y != 0.f || delta.z != 0.f;
        for (uint32_t s = 0; s < P.sample_count; ++s) {
            // ---- replay: the forward path, recorded
            int K = 0;
            if (live) {
                uint32_t rng = path_seed(P.rng_mode, x, y, W, H, P.sample_begin + s, P.seed);
                Ray ray = camera_ray(P.cam, W, H, x, y, rng);
                f3 thr = mk(1.f, 1.f, 1.f), L = mk(0.f, 0.f, 0.f);
                // the path's id in the recorded pass (pixel-major, as the wavefront engine numbers paths)
                const size_t pid = (size_t(y - row0) * W + x) * P.sample_count + s;
                for (uint32_t b = 0; b < P.max_depth; ++b) {
                    // the recorded closest hit, distance and triangle as traced
                    const uint2 rh = rec_hit[size_t(b) * A.rec_paths + pid];
                    const TraceHit h{__uint_as_float(rh.x), int(rh.y), 0u};
                    if (h.tri < 0) break;
                    REC(b, 0) = __float_as_uint(ray.o.x);
                    REC(b, 1) = __float_as_uint(ray.o.y);
                    REC(b, 2) = __float_as_uint(ray.o.z);
                    REC(b, 3) = __float_as_uint(ray.d.x);
                    REC(b, 4) = __float_as_uint(ray.d.y);
                    REC(b, 5) = __float_as_uint(ray.d.z);
                    REC(b, 6) = __float_as_uint(thr.x);
                    REC(b, 7) = __float_as_uint(thr.y);
                    REC(b, 8) = __float_as_uint(thr.z);
                    REC(b, 10) = rng;
                    ShadeOut out;
                    shade_hit(P, tris, mats, P.lights, P.num_lights, b, ray, h.t, h.tri, thr, L, rng, nullptr, out);
                    uint32_t fl = 0;
                    if (out.shadow && rec_vis[size_t(b) * A.rec_paths + pid]) fl |= kRecNee;
                    REC(b, 9) = uint32_t(h.tri) | (fl << kRecFlagShift);
                    K = int(b) + 1;
                    if (!out.cont) break;
                    ray = out.next;
                }
            }
            // ---- reverse sweep, last vertex first (wave-uniform loop)
            const int Kmax = wave_max_i(K);
            f3 V = mk(0.f, 0.f, 0.f);     // V_{j+1}
            f3 ob = mk(0.f, 0.f, 0.f);    // adjoint of the next segment's origin
            f3 db = mk(0.f, 0.f, 0.f);    // adjoint of the next segment's direction
            for (int j = Kmax - 1; j >= 0; --j) {
                bool act = false;
                uint32_t key = 0;
                float g[12];
#pragma unroll
                for (int k = 0; k < 12; ++k) g[k] = 0.f;
                if (j < K) {
                    const uint32_t b = uint32_t(j);
                    const f3 o = mk(__uint_as_float(REC(j, 0)), __uint_as_float(REC(j, 1)), __uint_as_float(REC(j, 2)));
                    const f3 d = mk(__uint_as_float(REC(j, 3)), __uint_as_float(REC(j, 4)), __uint_as_float(REC(j, 5)));
                    const f3 thr = mk(__uint_as_float(REC(j, 6)), __uint_as_float(REC(j, 7)), __uint_as_float(REC(j, 8)));
                    const uint32_t w9 = REC(j, 9);
                    const int tri = int(w9 & ((1u << kRecFlagShift) - 1u));
                    uint32_t rng = REC(j, 10);
                    const uint32_t fl = w9 >> kRecFlagShift;
                    const float4 t0 = tris[3 * tri + 0];
                    const float4 t1 = tris[3 * tri + 1];
                    const float4 t2 = tris[3 * tri + 2];
                    const float4 m0 = mats[2 * tri + 0];
                    const float4 m1 = mats[2 * tri + 1];
                    const f3 alb = xyz(m0), Le = xyz(m1);
                    const f3 E = (P.integrator == 0u || b == 0u) ? Le : mk(0.f, 0.f, 0.f);
                    if (m0.w > 0.1f) {
                        // emitter: the path ends here and nothing before it
                        // depends on where it was hit
                        V = E;
                        ob = mk(0.f, 0.f, 0.f);
                        db = mk(0.f, 0.f, 0.f);
                    } else {
                        const f3 P0 = xyz(t0), e1 = xyz(t1), e2 = xyz(t2);
                        const f3 N = cross(e1, e2);
                        const float lN = sqrtf(dot(N, N));
                        const f3 n = N * (1.0f / lN);
                        const float sg = dot(n, d) < 0.f ? 1.0f : -1.0f;
                        const f3 nf = n * sg;
                        const float den = dot(N, d);
                        const float t = dot(N, P0 - o) / den;
                        const f3 xh = o + d * t;
                        // adjoints of the hit point and the facing normal
                        f3 xb = ob;                    // next origin = x + nf * offset
                        f3 nfb = ob * kOffset;
                        f3 Nt = mk(0.f, 0.f, 0.f);
                        const bool more = b + 1u < P.max_depth;
                        if (P.integrator != 0u && more && P.num_lights) {
                            const uint32_t li = select_light(P.lights, P.num_lights, random_float(rng));
                            const float r1 = random_float(rng);
                            const float r2 = random_float(rng);
                            if (fl & kRecNee) {
                                const LightEntry le = P.lights[li];
                                const float4 l0 = tris[3 * le.tri + 0];
                                const float4 l1 = tris[3 * le.tri + 1];
                                const float4 l2 = tris[3 * le.tri + 2];
                                const float su = sqrtf(r1);
                                const f3 xl = xyz(l0) + xyz(l1) * (su * (1.0f - r2)) + xyz(l2) * (su * r2);
                                const f3 nl = normalize(cross(xyz(l1), xyz(l2)));
                                const f3 Ll = xyz(mats[2 * le.tri + 1]);
                                const f3 w = xl - xh;
                                const float w2 = dot(w, w);
                                const float a = dot(nf, w);
                                const float bl = dot(nl, w);
                                const float sig = bl < 0.f ? -1.0f : 1.0f;
                                const float bb = bl * sig;
                                const float c = le.area / (kPi * le.psel);
                                const float i4 = 1.0f / (w2 * w2);
                                const float gg = c * a * bb * i4;      // cos_x cos_l A / (pi d^2 psel)
                                Nt = Ll * gg;
                                const float gb = dot(delta * thr * alb, Ll);
                                nfb = nfb + w * (gb * c * bb * i4);
                                const f3 dgdw = (nf * bb + nl * (sig * a)) * (c * i4) - w * (4.0f * gg / w2);
                                xb = xb - dgdw * gb;
                            }
                        }
                        float qinv = 1.0f;
                        if (more && b >= P.rr_start) {
                            qinv = 1.0f / fmaxf(0.05f, max3(thr * alb));
                            (void)random_float(rng);
                        }
                        if (j + 1 < K) {               // the next segment hit something
                            const float r1 = random_float(rng);
                            const float r2 = random_float(rng);
                            nfb = nfb + cosine_sample_vjp(nf, r1, r2, db);
                        }
                        const f3 U = Nt + V * qinv;
                        const f3 gal = delta * thr * U;
                        V = E + alb * U;
                        // n = N / |N|, facing sign fixed
                        const f3 nb = nfb * sg;
                        f3 Nb = (nb - n * dot(n, nb)) * (1.0f / lN);
                        // x = o + t d,  t = N.(P0 - o) / N.d
                        const float tb = dot(xb, d);
                        f3 obn = xb;
                        f3 dbn = xb * t;
                        const float numb = tb / den;
                        const float denb = -tb * t / den;
                        Nb = Nb + (P0 - o) * numb + d * denb;
                        f3 P0b = N * numb;
                        obn = obn - N * numb;
                        dbn = dbn + N * denb;
                        // N = e1 x e2,  e1 = P1 - P0,  e2 = P2 - P0
                        const f3 e1b = cross(e2, Nb);
                        const f3 e2b = cross(Nb, e1);
                        P0b = P0b - e1b - e2b;
                        ob = obn;
                        db = dbn;
                        act = true;
                        key = __float_as_uint(t0.w);
                        g[0] = gal.x; g[1] = gal.y; g[2] = gal.z;
                        g[3] = P0b.x; g[4] = P0b.y; g[5] = P0b.z;
                        g[6] = e1b.x; g[7] = e1b.y; g[8] = e1b.z;
                        g[9] = e2b.x; g[10] = e2b.y; g[11] = e2b.z;
                    }
                }
                scatter_grads(act, key, g, ga, gv, A.indices);
            }
        }
    }
#undef REC
}

// out[param[t]] += per_tri[t] for every triangle t with a parameter (3
// channels), in FP64: the reduction of per-triangle gradients onto a shared
// parameter map.
__global__ void k_param_scatter(const double* __restrict__ per_tri, const uint32_t* __restrict__ param, size_t nt,
                                uint32_t np, double* __restrict__ out) {
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < nt; t += size_t(gridDim.x) * blockDim.x) {
        const uint32_t k = param[t];
        if (k >= np) continue;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = per_tri[3 * t + c];
            if (v != 0.0) atomicAdd(&out[3 * size_t(k) + c], v);
        }
    }
}

__global__ void k_replica_reduce(const float* __restrict__ rep, size_t n, uint32_t replicas, double* __restrict__ out) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        double s = 0.0;
        for (uint32_t r = 0; r < replicas; ++r) s += double(rep[size_t(r) * n + i]);
        out[i] = s;
    }
}

constexpr int kL2Block = 256;

__global__ __launch_bounds__(kL2Block) void k_l2_image(const float4* __restrict__ sum, float inv_n,
                                                       const float4* __restrict__ target, float4* __restrict__ dimg,
                                                       size_t npix, float* __restrict__ partials) {
    __shared__ float red[kL2Block];
    float acc = 0.f;
    for (size_t i = size_t(blockIdx.x) * kL2Block + threadIdx.x; i < npix; i += size_t(gridDim.x) * kL2Block) {
        const float4 s = sum[i], t = target[i];
        const float dx = s.x * inv_n - t.x, dy = s.y * inv_n - t.y, dz = s.z * inv_n - t.z;
        dimg[i] = make_float4(2.f * dx, 2.f * dy, 2.f * dz, 0.f);
        acc += dx * dx + dy * dy + dz * dz;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kL2Block / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

}  // namespace

uint32_t adjoint_block_size() { return kAdjBlock; }
uint32_t adjoint_max_depth() { return kMaxDepth; }

hipError_t launch_adjoint(const AdjointParams& p, uint32_t blocks, hipStream_t stream) {
    const uint32_t rows = p.rows ? p.rows : p.pass.height;
    if (blocks == 0 || p.pass.max_depth == 0 || p.pass.max_depth > kMaxDepth || p.replicas == 0 ||
        !p.pass.rec_hit || !p.pass.rec_vis || uint64_t(p.row_begin) + rows > p.pass.height ||
        uint64_t(p.pass.sample_count) * p.pass.width * rows > p.rec_paths)
        return hipErrorInvalidValue;
    if (p.pass.max_depth <= 4)
        hipLaunchKernelGGL(k_adjoint<4>, dim3(blocks), dim3(kAdjBlock), 0, stream, p);
    else if (p.pass.max_depth <= 8)
        hipLaunchKernelGGL(k_adjoint<8>, dim3(blocks), dim3(kAdjBlock), 0, stream, p);
    else
        hipLaunchKernelGGL(k_adjoint<16>, dim3(blocks), dim3(kAdjBlock), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_param_scatter(const double* per_tri, const uint32_t* param, size_t nt, uint32_t np, double* out,
                               hipStream_t stream) {
    if (nt == 0 || np == 0) return hipSuccess;
    const size_t want = (nt + 255) / 256;
    const uint32_t blocks = uint32_t(want < 8192 ? want : 8192);
    hipLaunchKernelGGL(k_param_scatter, dim3(blocks), dim3(256), 0, stream, per_tri, param, nt, np, out);
    return hipGetLastError();
}

hipError_t launch_replica_reduce(const float* rep, size_t n, uint32_t replicas, double* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const size_t want = (n + 255) / 256;
    const uint32_t blocks = uint32_t(want < 4096 ? want : 4096);
    hipLaunchKernelGGL(k_replica_reduce, dim3(blocks), dim3(256), 0, stream, rep, n, replicas, out);
    return hipGetLastError();
}

uint32_t l2_image_blocks(size_t npix) {
    const size_t want = (npix + kL2Block - 1) / kL2Block;
    return uint32_t(want < 2048 ? (want ? want : 1) : 2048);
}

hipError_t launch_l2_image(const float4* sum, uint32_t n, const float4* target, float4* dimg, size_t npix,
                           float* partials, hipStream_t stream) {
    if (n == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_l2_image, dim3(l2_image_blocks(npix)), dim3(kL2Block), 0, stream, sum, 1.0f / float(n),
                       target, dimg, npix, partials);
    return hipGetLastError();
}

}  // namespace quasi_hip
