// bvh_gpu.hip -- binned-SAH BVH build on the GPU.
//
// Same algorithm as the host build (bvh_build.cpp): per node, 32 centroid
// bins on each axis, the cheapest of the 31 x 3 candidate planes by the SAH,
// leaves of at most kLeafMax triangles chosen against the best split with the
// node cost kSahNodeCost (common.hpp), a
// median split when every centroid coincides.  The output is a binary tree of
// BinaryNode plus the triangle order, which bvh_layout.hip turns into the
// traversal layouts on the device; the triangle and material arrays are then
// written in BVH order on the device too (build_scene_gpu), so after the mesh
// upload nothing of the scene passes through host memory.
//
// Two phases:
//  - large nodes (more than kLargeNode triangles) are split one tree level at
//    a time with many blocks per node: bounds (block reductions + one atomic
//    per block), binning (sixteen bank-padded LDS copies of the bins per
//    block, folded and merged with global atomics, or stored when the block
//    is the node's only one), split choice (one thread per node), partition
//    (count, then write: one atomic per block and side into the other id
//    buffer);
//  - each remaining subtree is built by one wave, which reserves the
//    subtree's node slots with one atomic: bounds by DPP wave reductions,
//    bins in LDS, the SAH sweep with one bin per lane, stable partition with
//    ballots, the smaller child first so the LDS task stack stays within
//    log2 of the subtree size.  Slots a subtree does not need stay unused
//    and are counted for the layout's reachability check.
// Every index a kernel derives is checked against its buffer; a violation
// sets an error flag and the host falls back to the host build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "bvh_build.hpp"
#include "bvh_gpu.hpp"
#include "bvh_split.hpp"
#include "common.hpp"
#include "device_buffer.hpp"

namespace quasi_hip {

namespace {

constexpr int kBins = kSahBins;             // 32
constexpr uint32_t kLargeNode = 512;        // larger nodes take the level-synchronous path
constexpr uint32_t kChunk = 8192;           // triangles per block in the large-node kernels
constexpr int kBlock = 256;
constexpr int kStack = 48;                  // subtree task stack (smaller child first: <= log2(512) + 1 used)
constexpr uint32_t kIntMax = 0x7FFFFFFF;

// Error flags
constexpr uint32_t kErrRange = 1, kErrNodes = 2, kErrStack = 4;

__device__ __forceinline__ int fkey(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float funkey(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }

__device__ __forceinline__ float bin_scale(float lo, float hi) {
    const float ext = hi - lo;
    const float s = ext > 0.f ? float(kBins) / ext : 0.f;
    return s < INFINITY ? s : 0.f;   // a subnormal extent: no split on this axis
}
__device__ __forceinline__ int bin_of(float c, float lo, float scale) {
    int b = int((c - lo) * scale);
    return min(max(b, 0), kBins - 1);
}

__device__ __forceinline__ float area_of(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    if (!(dx >= 0.f) || !(dy >= 0.f) || !(dz >= 0.f)) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

// One large node of the current level.
struct LargeTask {
    uint32_t node, begin, count, buf;   // buf: which id buffer holds the range
    int box_lo[3], box_hi[3];           // fkey bounds of the triangles
    int cen_lo[3], cen_hi[3];           // fkey bounds of the centroids
    int axis, split;                    // decision (axis -1: median split)
    uint32_t nl;                        // triangles going left
    uint32_t lcur, rcur;                // partition cursors
    uint32_t pad[3];
};

// A node handed to the subtree kernel (or a child of a large node).
struct NodeTask {
    uint32_t node, begin, count, buf;
};

struct Bins {
    uint32_t cnt[3][kBins];
    int lo[3][kBins][3];
    int hi[3][kBins][3];
};

struct WorkItem {
    uint32_t task, start, end;   // prim positions [start, end) of task
};

__global__ void k_prim_bounds(const float* __restrict__ p0, const float* __restrict__ p1,
                              const float* __restrict__ p2, uint32_t n, float4* __restrict__ plo,
                              float4* __restrict__ phi, uint32_t* __restrict__ ids) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            const float a = p0[3 * size_t(t) + k], b = p1[3 * size_t(t) + k], c = p2[3 * size_t(t) + k];
            lo[k] = fminf(a, fminf(b, c));
            hi[k] = fmaxf(a, fmaxf(b, c));
        }
        plo[t] = make_float4(lo[0], lo[1], lo[2], 0.f);
        phi[t] = make_float4(hi[0], hi[1], hi[2], 0.f);
        ids[t] = t;
    }
}

// ---- cross-lane operations on DPP (data-parallel primitives: VALU operands read
// from another lane of the same 16-lane row, or broadcast between rows), which
// keep these reductions and scans off the LDS crossbar.  Every caller runs them
// with all 64 lanes active.
constexpr int kDppQuadXor1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppWaveShr1 = 0x138;    // lane i reads lane i - 1 of the wave
constexpr int kDppRowMirror = 0x140;   // lane i of a row reads lane 15 - i
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowBcast15 = 0x142;  // lane 15 of each row to the next row
constexpr int kDppRowBcast31 = 0x143;  // lane 31 to rows 2 and 3

// The value of lane (per CTRL) in the lanes the masks enable, `identity` where
// the source lane lies outside the row or the row is masked off.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp_i(int identity, int v) {
    return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_f(float identity, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(identity), __float_as_int(v), CTRL, ROW_MASK,
                                                      0xF, false));
}

struct OpMinI { __device__ static int id() { return int(kIntMax); } __device__ static int f(int a, int b) { return min(a, b); } };
struct OpMaxI { __device__ static int id() { return -int(kIntMax) - 1; } __device__ static int f(int a, int b) { return max(a, b); } };

// Reduction over all 64 lanes; the result is in lane 63, returned wave-uniform.
template <class Op>
__device__ __forceinline__ int wave_reduce_i(int v) {
    v = Op::f(v, dpp_i<kDppQuadXor1>(Op::id(), v));
    v = Op::f(v, dpp_i<kDppQuadXor2>(Op::id(), v));
    v = Op::f(v, dpp_i<kDppRowHalfMirror>(Op::id(), v));
    v = Op::f(v, dpp_i<kDppRowMirror>(Op::id(), v));          // every lane of a row: the row's result
    v = Op::f(v, dpp_i<kDppRowBcast15, 0xA>(Op::id(), v));    // rows 1, 3: rows 0-1, 2-3
    v = Op::f(v, dpp_i<kDppRowBcast31, 0xC>(Op::id(), v));    // rows 2, 3: all rows
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min_i(int v) { return wave_reduce_i<OpMinI>(v); }
__device__ __forceinline__ int wave_max_i(int v) { return wave_reduce_i<OpMaxI>(v); }
struct OpAddI { __device__ static int id() { return 0; } __device__ static int f(int a, int b) { return a + b; } };
__device__ __forceinline__ int wave_reduce_add(int v) { return wave_reduce_i<OpAddI>(v); }

// Inclusive scans within each 32-lane half of the wave (lanes 0-31, 32-63).
__device__ __forceinline__ uint32_t half_scan_add(uint32_t v) {
    int x = int(v);
    x += dpp_i<kDppRowShr1>(0, x);
    x += dpp_i<kDppRowShr2>(0, x);
    x += dpp_i<kDppRowShr4>(0, x);
    x += dpp_i<kDppRowShr8>(0, x);
    x += dpp_i<kDppRowBcast15, 0xA>(0, x);
    return uint32_t(x);
}
__device__ __forceinline__ float half_scan_min(float x) {
    x = fminf(x, dpp_f<kDppRowShr1>(INFINITY, x));
    x = fminf(x, dpp_f<kDppRowShr2>(INFINITY, x));
    x = fminf(x, dpp_f<kDppRowShr4>(INFINITY, x));
    x = fminf(x, dpp_f<kDppRowShr8>(INFINITY, x));
    return fminf(x, dpp_f<kDppRowBcast15, 0xA>(INFINITY, x));
}
__device__ __forceinline__ float half_scan_max(float x) {
    x = fmaxf(x, dpp_f<kDppRowShr1>(-INFINITY, x));
    x = fmaxf(x, dpp_f<kDppRowShr2>(-INFINITY, x));
    x = fmaxf(x, dpp_f<kDppRowShr4>(-INFINITY, x));
    x = fmaxf(x, dpp_f<kDppRowShr8>(-INFINITY, x));
    return fmaxf(x, dpp_f<kDppRowBcast15, 0xA>(-INFINITY, x));
}

// Minimum over lanes 0-31, returned wave-uniform.
__device__ __forceinline__ int half_min_i(int v) {
    v = min(v, dpp_i<kDppQuadXor1>(int(kIntMax), v));
    v = min(v, dpp_i<kDppQuadXor2>(int(kIntMax), v));
    v = min(v, dpp_i<kDppRowHalfMirror>(int(kIntMax), v));
    v = min(v, dpp_i<kDppRowMirror>(int(kIntMax), v));
    v = min(v, dpp_i<kDppRowBcast15, 0xA>(int(kIntMax), v));
    return __builtin_amdgcn_readlane(v, 31);
}

// ---------------------------------------------------------------- large nodes
__global__ __launch_bounds__(kBlock) void k_large_bounds(LargeTask* tasks, const WorkItem* items, uint32_t nitems,
                                                         const float4* __restrict__ plo,
                                                         const float4* __restrict__ phi,
                                                         const uint32_t* __restrict__ ids_a,
                                                         const uint32_t* __restrict__ ids_b, uint32_t n,
                                                         uint32_t* err) {
    __shared__ int red[12][kBlock / 64];
    if (blockIdx.x >= nitems) return;
    const WorkItem it = items[blockIdx.x];
    LargeTask& t = tasks[it.task];
    const uint32_t* ids = t.buf ? ids_b : ids_a;
    int v[12];
    for (int k = 0; k < 3; ++k) {
        v[k] = kIntMax; v[3 + k] = -int(kIntMax) - 1; v[6 + k] = kIntMax; v[9 + k] = -int(kIntMax) - 1;
    }
    if (it.end > n || it.start > it.end) {
        if (threadIdx.x == 0) atomicOr(err, kErrRange);
        return;
    }
    for (uint32_t i = it.start + threadIdx.x; i < it.end; i += kBlock) {
        const uint32_t p = ids[i];
        if (p >= n) { atomicOr(err, kErrRange); continue; }
        const float4 l = plo[p], h = phi[p];
        const float lo[3] = {l.x, l.y, l.z}, hi[3] = {h.x, h.y, h.z};
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * (lo[k] + hi[k]);
            v[k] = min(v[k], fkey(lo[k]));
            v[3 + k] = max(v[3 + k], fkey(hi[k]));
            v[6 + k] = min(v[6 + k], fkey(c));
            v[9 + k] = max(v[9 + k], fkey(c));
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int k = 0; k < 12; ++k) {
        const int r = (k < 3 || (k >= 6 && k < 9)) ? wave_min_i(v[k]) : wave_max_i(v[k]);
        if (lane == 0) red[k][wave] = r;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        const bool is_min = k < 3 || (k >= 6 && k < 9);
        int r = red[k][0];
        for (int w = 1; w < kBlock / 64; ++w) r = is_min ? min(r, red[k][w]) : max(r, red[k][w]);
        int* dst = k < 3 ? &t.box_lo[k] : k < 6 ? &t.box_hi[k - 3] : k < 9 ? &t.cen_lo[k - 6] : &t.cen_hi[k - 9];
        if (is_min) atomicMin(dst, r); else atomicMax(dst, r);
    }
}

// Binning of one chunk of a large node: every lane adds its triangles to one
// of kBinCopies copies of the bins in LDS (lane & 15; the copies are padded so
// a word's copies fall in different banks), then the block folds the copies
// and merges them into the node's bins.  Neighbouring triangles of the order
// are mostly neighbours in space; with one copy, the 64 lanes of a wave adding
// to the same bin serialised the LDS atomics.
constexpr int kBinCopies = 16;
constexpr int kBinWords = 3 * kBins * 7;            // cnt, then lo, then hi (the Bins layout)
constexpr int kBinStride = kBinWords + 1;
__global__ __launch_bounds__(kBlock) void k_large_bin(const LargeTask* tasks, const WorkItem* items, uint32_t nitems,
                                                      Bins* bins, const float4* __restrict__ plo,
                                                      const float4* __restrict__ phi,
                                                      const uint32_t* __restrict__ ids_a,
                                                      const uint32_t* __restrict__ ids_b, uint32_t n) {
    static_assert(sizeof(Bins) == kBinWords * 4, "Bins is cnt[3][32], lo[3][32][3], hi[3][32][3]");
    constexpr int kCnt = 3 * kBins, kBox = 3 * kBins * 3;
    __shared__ int lb[kBinCopies * kBinStride];
    if (blockIdx.x >= nitems) return;
    const WorkItem it = items[blockIdx.x];
    const LargeTask& t = tasks[it.task];
    const uint32_t* ids = t.buf ? ids_b : ids_a;
    for (int i = threadIdx.x; i < kBinCopies * kBinStride; i += kBlock) {
        const int w = i % kBinStride;
        lb[i] = w < kCnt ? 0 : w < kCnt + kBox ? int(kIntMax) : -int(kIntMax) - 1;
    }
    __syncthreads();
    float clo[3], sc[3];
    for (int k = 0; k < 3; ++k) {
        clo[k] = funkey(t.cen_lo[k]);
        sc[k] = bin_scale(clo[k], funkey(t.cen_hi[k]));
    }
    int* mine = lb + (threadIdx.x & (kBinCopies - 1)) * kBinStride;
    if (it.end <= n && it.start <= it.end) {
        for (uint32_t i = it.start + threadIdx.x; i < it.end; i += kBlock) {
            const uint32_t p = ids[i];
            if (p >= n) continue;
            const float4 l = plo[p], h = phi[p];
            const float lo[3] = {l.x, l.y, l.z}, hi[3] = {h.x, h.y, h.z};
            for (int k = 0; k < 3; ++k) {
                if (sc[k] == 0.f) continue;
                const int b = bin_of(0.5f * (lo[k] + hi[k]), clo[k], sc[k]);
                atomicAdd(&mine[k * kBins + b], 1);
                for (int a = 0; a < 3; ++a) {
                    atomicMin(&mine[kCnt + (k * kBins + b) * 3 + a], fkey(lo[a]));
                    atomicMax(&mine[kCnt + kBox + (k * kBins + b) * 3 + a], fkey(hi[a]));
                }
            }
        }
    }
    __syncthreads();
    Bins& g = bins[it.task];
    int* gw = reinterpret_cast<int*>(&g);
    const bool sole = t.count <= kChunk;   // the node's only block: store instead of merging
    for (int w = threadIdx.x; w < kBinWords; w += kBlock) {
        int v = lb[w];
        for (int c = 1; c < kBinCopies; ++c) {
            const int x = lb[c * kBinStride + w];
            v = w < kCnt ? v + x : w < kCnt + kBox ? min(v, x) : max(v, x);
        }
        if (sole) gw[w] = v;
        else if (w < kCnt) { if (v) atomicAdd(&gw[w], v); }
        else if (w < kCnt + kBox) { if (v != int(kIntMax)) atomicMin(&gw[w], v); }
        else if (v != -int(kIntMax) - 1) atomicMax(&gw[w], v);
    }
}

// One thread per large node: the SAH split, the node's box and its two children.
__global__ void k_large_split(LargeTask* tasks, uint32_t ntasks, const Bins* bins, BinaryNode* nodes,
                              uint32_t node_cap, uint32_t* node_counter, NodeTask* children, uint32_t* err) {
    const uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= ntasks) return;
    LargeTask& t = tasks[ti];
    float blo[3], bhi[3], clo[3], chi[3], sc[3];
    for (int k = 0; k < 3; ++k) {
        blo[k] = funkey(t.box_lo[k]);
        bhi[k] = funkey(t.box_hi[k]);
        clo[k] = funkey(t.cen_lo[k]);
        chi[k] = funkey(t.cen_hi[k]);
        sc[k] = bin_scale(clo[k], chi[k]);
    }
    const Bins& g = bins[ti];
    const float parent = fmaxf(area_of(blo[0], blo[1], blo[2], bhi[0], bhi[1], bhi[2]), 1e-30f);
    int best_axis = -1, best_split = -1;
    float best_cost = INFINITY;
    uint32_t best_nl = 0;
    for (int k = 0; k < 3; ++k) {
        if (sc[k] == 0.f) continue;
        float rarea[kBins];
        uint32_t rcnt[kBins];
        float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t run = 0;
        for (int b = kBins - 1; b >= 1; --b) {
            if (g.cnt[k][b]) {
                for (int a = 0; a < 3; ++a) {
                    rl[a] = fminf(rl[a], funkey(g.lo[k][b][a]));
                    rh[a] = fmaxf(rh[a], funkey(g.hi[k][b][a]));
                }
            }
            run += g.cnt[k][b];
            rarea[b] = area_of(rl[0], rl[1], rl[2], rh[0], rh[1], rh[2]);
            rcnt[b] = run;
        }
        float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t lrun = 0;
        for (int s = 1; s < kBins; ++s) {
            if (g.cnt[k][s - 1]) {
                for (int a = 0; a < 3; ++a) {
                    ll[a] = fminf(ll[a], funkey(g.lo[k][s - 1][a]));
                    lh[a] = fmaxf(lh[a], funkey(g.hi[k][s - 1][a]));
                }
            }
            lrun += g.cnt[k][s - 1];
            if (lrun == 0 || rcnt[s] == 0) continue;
            const float cost =
                1.f + (area_of(ll[0], ll[1], ll[2], lh[0], lh[1], lh[2]) * float(lrun) + rarea[s] * float(rcnt[s])) /
                          parent;
            if (cost < best_cost) { best_cost = cost; best_axis = k; best_split = s; best_nl = lrun; }
        }
    }
    const uint32_t nl = best_axis >= 0 ? best_nl : t.count / 2;
    t.axis = best_axis;
    t.split = best_split;
    t.nl = nl;
    t.lcur = 0;
    t.rcur = 0;
    const uint32_t l = atomicAdd(node_counter, 2u);
    if (t.node >= node_cap || l + 1 >= node_cap) {
        atomicOr(err, kErrNodes);
        children[2 * ti] = NodeTask{0, 0, 0, 0};
        children[2 * ti + 1] = NodeTask{0, 0, 0, 0};
        return;
    }
    BinaryNode& nd = nodes[t.node];
    for (int k = 0; k < 3; ++k) { nd.lo[k] = blo[k]; nd.hi[k] = bhi[k]; }
    nd.left = l;
    nd.right = l + 1;
    nd.first = 0;
    nd.count = 0;
    children[2 * ti] = NodeTask{l, t.begin, nl, t.buf ^ 1u};
    children[2 * ti + 1] = NodeTask{l + 1, t.begin + nl, t.count - nl, t.buf ^ 1u};
}

__global__ __launch_bounds__(kBlock) void k_large_partition(LargeTask* tasks, const WorkItem* items, uint32_t nitems,
                                                            const float4* __restrict__ plo,
                                                            const float4* __restrict__ phi, uint32_t* ids_a,
                                                            uint32_t* ids_b, uint32_t n, uint32_t* err) {
    static_assert(kChunk <= 32 * kBlock, "one side bit per element in a 32-bit mask");
    __shared__ uint32_t wl[kBlock / 64], wr[kBlock / 64];
    __shared__ uint32_t base_l, base_r;
    if (blockIdx.x >= nitems) return;
    const WorkItem it = items[blockIdx.x];
    LargeTask& t = tasks[it.task];
    const uint32_t* in = t.buf ? ids_b : ids_a;
    uint32_t* out = t.buf ? ids_a : ids_b;
    if (it.end > n || it.start > it.end || it.end - it.start > kChunk || t.begin + t.count > n) {
        if (threadIdx.x == 0) atomicOr(err, kErrRange);
        return;
    }
    const int axis = t.axis;
    float clo = 0.f, sc = 0.f;
    if (axis >= 0) {
        clo = funkey(t.cen_lo[axis]);
        sc = bin_scale(clo, funkey(t.cen_hi[axis]));
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // pass 1: the side of each of this thread's elements (bit j: element
    // it.start + threadIdx.x + j * kBlock goes left) and the chunk's counts,
    // then one pair of atomics for the whole chunk
    uint32_t side = 0, nleft = 0, nright = 0;
    for (uint32_t j = 0, i = it.start + threadIdx.x; i < it.end; ++j, i += kBlock) {
        bool left = false;
        if (axis < 0) {
            left = (i - t.begin) < t.nl;
        } else {
            const uint32_t p = in[i];
            if (p < n) {
                const float4 l = plo[p], h = phi[p];
                const float c = axis == 0 ? 0.5f * (l.x + h.x) : axis == 1 ? 0.5f * (l.y + h.y) : 0.5f * (l.z + h.z);
                left = bin_of(c, clo, sc) < t.split;
            }
        }
        side |= uint32_t(left) << j;
        nleft += left;
        nright += !left;
    }
    const uint32_t cl = uint32_t(wave_reduce_add(int(nleft))), cr = uint32_t(wave_reduce_add(int(nright)));
    if (lane == 0) { wl[wave] = cl; wr[wave] = cr; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sl = 0, sr = 0;
        for (int w = 0; w < kBlock / 64; ++w) { sl += wl[w]; sr += wr[w]; }
        base_l = sl ? atomicAdd(&t.lcur, sl) : 0u;
        base_r = sr ? atomicAdd(&t.rcur, sr) : 0u;
    }
    __syncthreads();
    // pass 2: stable placement inside the chunk's ranges, in element order
    uint32_t off_l = base_l, off_r = base_r;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t r0 = it.start, j = 0; r0 < it.end; r0 += kBlock, ++j) {
        const uint32_t i = r0 + threadIdx.x;
        const bool valid = i < it.end;
        const bool left = valid && ((side >> j) & 1u);
        const unsigned long long bl = __ballot(left);
        const unsigned long long br = __ballot(valid && !left);
        __syncthreads();   // the previous round's wl / wr are read
        if (lane == 0) { wl[wave] = __popcll(bl); wr[wave] = __popcll(br); }
        __syncthreads();
        uint32_t ol = off_l, orr = off_r;
        for (int w = 0; w < wave; ++w) { ol += wl[w]; orr += wr[w]; }
        if (valid) {
            const uint32_t p = in[i];
            const uint32_t dst = left ? t.begin + ol + __popcll(bl & below)
                                      : t.begin + t.nl + orr + __popcll(br & below);
            if (dst < t.begin + t.count) out[dst] = p;
            else atomicOr(err, kErrRange);
        }
        for (int w = 0; w < kBlock / 64; ++w) { off_l += wl[w]; off_r += wr[w]; }
    }
}

// Per-wave LDS for the sweep: the unions of bins s..31, indexed by s.
struct SweepScratch {
    uint32_t cnt[kBins];
    float lo[3][kBins], hi[3][kBins];
};

// SAH sweep over the 32 LDS bins of axis k, bin b on lane b (each half-wave
// holds the whole axis): unions and counts of bins 0..b by DPP prefix scans,
// of bins b..31 by the same scans over the bins in reverse order (handed to
// their lanes through LDS), then the first cheapest split plane s (bins < s go
// left) by a reduction.  Returns split -1 when no plane has triangles on both
// sides.  The costs are the serial sweep's: 1 + (A_left n_left + A_right
// n_right) / parent.  Called by all 64 lanes.
__device__ __forceinline__ void sweep_axis(const Bins& lb, SweepScratch& sx, int k, float parent, float& out_cost,
                                           int& out_split, uint32_t& out_nl) {
    const int lane = threadIdx.x & 63;
    const int b = lane & 31, rb = kBins - 1 - b;
    const uint32_t cl = lb.cnt[k][b], cr = lb.cnt[k][rb];
    float pl[3], ph[3], rl[3], rh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        pl[a] = half_scan_min(cl ? funkey(lb.lo[k][b][a]) : INFINITY);
        ph[a] = half_scan_max(cl ? funkey(lb.hi[k][b][a]) : -INFINITY);
        rl[a] = half_scan_min(cr ? funkey(lb.lo[k][rb][a]) : INFINITY);
        rh[a] = half_scan_max(cr ? funkey(lb.hi[k][rb][a]) : -INFINITY);
    }
    const uint32_t pc = half_scan_add(cl), rcnt = half_scan_add(cr);
    // lane b holds the union of bins rb..31: store it under rb
    if (lane < 32) {
        sx.cnt[rb] = rcnt;
#pragma unroll
        for (int a = 0; a < 3; ++a) { sx.lo[a][rb] = rl[a]; sx.hi[a][rb] = rh[a]; }
    }
    __syncthreads();
    const uint32_t scnt = sx.cnt[b];
    float sl[3], sh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) { sl[a] = sx.lo[a][b]; sh[a] = sx.hi[a][b]; }
    __syncthreads();
    // plane s = b: left = bins 0..b-1, the prefix one lane down
    const uint32_t lrun = uint32_t(dpp_i<kDppWaveShr1>(0, int(pc)));
    float ll[3], lh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) { ll[a] = dpp_f<kDppWaveShr1>(INFINITY, pl[a]); lh[a] = dpp_f<kDppWaveShr1>(-INFINITY, ph[a]); }
    float c = INFINITY;
    if (b >= 1 && lrun != 0 && scnt != 0)
        c = 1.f + (area_of(ll[0], ll[1], ll[2], lh[0], lh[1], lh[2]) * float(lrun) +
                   area_of(sl[0], sl[1], sl[2], sh[0], sh[1], sh[2]) * float(scnt)) / parent;
    // costs are >= 1 or +inf, so their bit patterns order like the values
    const int best = half_min_i(__float_as_int(c));
    const int split = half_min_i(c < INFINITY && __float_as_int(c) == best ? b : 64);
    out_cost = __int_as_float(best);
    out_split = split < 32 ? split : -1;
    out_nl = split < 32 ? uint32_t(__builtin_amdgcn_readlane(int(lrun), split)) : 0u;
}

// ---------------------------------------------------------------- subtrees (one wave each)
struct SubTask {
    uint32_t node, begin, count, buf;
};

__global__ __launch_bounds__(64) void k_subtree(const NodeTask* tasks, uint32_t ntasks,
                                                const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                uint32_t* ids_a, uint32_t* ids_b, uint32_t* order, uint32_t n,
                                                BinaryNode* nodes, uint32_t node_cap, uint32_t* node_counter,
                                                uint32_t* unused, uint32_t* err) {
    __shared__ Bins lb;
    __shared__ SweepScratch sx;
    __shared__ SubTask stack[kStack];
    __shared__ struct { float4 lo[64], hi[64]; uint32_t id[64]; } cache;   // one slot per lane
    const int lane = threadIdx.x;
    for (uint32_t ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
        const NodeTask root = tasks[ti];
        if (root.count == 0) continue;
        // The child slots of the whole subtree, taken with one atomic: a subtree
        // of c triangles has at most 2c - 1 nodes, its root allocated already.
        // (One atomic per split on the one counter serialised the kernel.)
        const uint32_t reserve = 2u * root.count - 2u;
        uint32_t slot = 0;
        if (lane == 0 && reserve) slot = atomicAdd(node_counter, reserve);
        slot = __shfl(slot, 0, 64);
        const uint32_t slot_end = slot + reserve;
        if (slot_end > node_cap || slot_end < slot) {
            if (lane == 0) atomicOr(err, kErrNodes);
            continue;
        }
        int sp = 0;
        SubTask cur{root.node, root.begin, root.count, root.buf};
        bool have = true;
        while (have) {
            // guard every range before touching memory (uniform across the wave)
            if (cur.begin + cur.count > n || cur.begin + cur.count < cur.begin || cur.node >= node_cap) {
                if (lane == 0) atomicOr(err, kErrRange);
                break;
            }
            const uint32_t* in = cur.buf ? ids_b : ids_a;
            uint32_t* out = cur.buf ? ids_a : ids_b;
            // A node of at most 64 triangles keeps triangle `lane` in the lane's own LDS
            // slot, loaded once for the bounds, the binning and the partition (one
            // dependent id -> box load chain from memory per node instead of three).
            const bool cached = cur.count <= 64u;
            if (cached && uint32_t(lane) < cur.count) {
                const uint32_t id = in[cur.begin + lane];
                cache.id[lane] = id;
                if (id < n) {
                    cache.lo[lane] = plo[id];
                    cache.hi[lane] = phi[id];
                }
            }
            // triangle at position i (i = cur.begin + lane for a cached node)
            auto tri = [&](uint32_t i, uint32_t& p, float4& l, float4& h) -> bool {
                if (cached) {
                    p = cache.id[lane];
                    if (p >= n) return false;
                    l = cache.lo[lane];
                    h = cache.hi[lane];
                    return true;
                }
                p = in[i];
                if (p >= n) return false;
                l = plo[p];
                h = phi[p];
                return true;
            };
            // bounds
            int v[12];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[k] = kIntMax; v[3 + k] = -int(kIntMax) - 1; v[6 + k] = kIntMax; v[9 + k] = -int(kIntMax) - 1;
            }
            for (uint32_t i = cur.begin + lane; i < cur.begin + cur.count; i += 64) {
                uint32_t p;
                float4 l, h;
                if (!tri(i, p, l, h)) continue;
                const float lo[3] = {l.x, l.y, l.z}, hi[3] = {h.x, h.y, h.z};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float c = 0.5f * (lo[k] + hi[k]);
                    v[k] = min(v[k], fkey(lo[k]));
                    v[3 + k] = max(v[3 + k], fkey(hi[k]));
                    v[6 + k] = min(v[6 + k], fkey(c));
                    v[9 + k] = max(v[9 + k], fkey(c));
                }
            }
#pragma unroll
            for (int k = 0; k < 12; ++k)
                v[k] = (k < 3 || (k >= 6 && k < 9)) ? wave_min_i(v[k]) : wave_max_i(v[k]);
            float blo[3], bhi[3], clo[3], sc[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                blo[k] = funkey(v[k]);
                bhi[k] = funkey(v[3 + k]);
                clo[k] = funkey(v[6 + k]);
                sc[k] = bin_scale(clo[k], funkey(v[9 + k]));
            }
            const uint32_t cnt = cur.count;
            int best_axis = -1, best_split = -1;
            float best_cost = INFINITY;
            uint32_t best_nl = 0;
            if (cnt == 2u && cached && cache.id[0] < n && cache.id[1] < n) {
                // Two triangles: every plane of an axis whose two bins differ splits
                // them the same way, so the binned sweep's answer is the first such
                // axis, at plane min(b0, b1) + 1, with the cost 1 + (A0 + A1) / A of
                // the two triangles' own boxes (the bins hold exactly those boxes).
                const float4 l0 = cache.lo[0], h0 = cache.hi[0], l1 = cache.lo[1], h1 = cache.hi[1];
                const float a0 = area_of(l0.x, l0.y, l0.z, h0.x, h0.y, h0.z);
                const float a1 = area_of(l1.x, l1.y, l1.z, h1.x, h1.y, h1.z);
                const float parent = fmaxf(area_of(blo[0], blo[1], blo[2], bhi[0], bhi[1], bhi[2]), 1e-30f);
                const float lo0[3] = {l0.x, l0.y, l0.z}, hi0[3] = {h0.x, h0.y, h0.z};
                const float lo1[3] = {l1.x, l1.y, l1.z}, hi1[3] = {h1.x, h1.y, h1.z};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (sc[k] == 0.f || best_axis >= 0) continue;
                    const int b0 = bin_of(0.5f * (lo0[k] + hi0[k]), clo[k], sc[k]);
                    const int b1 = bin_of(0.5f * (lo1[k] + hi1[k]), clo[k], sc[k]);
                    if (b0 == b1) continue;
                    // the sweep's expression, left side first (the sum does not depend on the order)
                    const float al = b0 < b1 ? a0 : a1, ar = b0 < b1 ? a1 : a0;
                    best_cost = 1.f + (al * float(1u) + ar * float(1u)) / parent;
                    best_axis = k;
                    best_split = min(b0, b1) + 1;
                    best_nl = 1u;
                }
            } else if (cnt > 1) {
                uint32_t* lw = &lb.cnt[0][0];
                for (int i = lane; i < 3 * kBins; i += 64) lw[i] = 0u;
                int* llo = &lb.lo[0][0][0];
                int* lhi = &lb.hi[0][0][0];
                for (int i = lane; i < 3 * kBins * 3; i += 64) { llo[i] = kIntMax; lhi[i] = -int(kIntMax) - 1; }
                __syncthreads();
                for (uint32_t i = cur.begin + lane; i < cur.begin + cnt; i += 64) {
                    uint32_t p;
                    float4 l, h;
                    if (!tri(i, p, l, h)) continue;
                    const float lo[3] = {l.x, l.y, l.z}, hi[3] = {h.x, h.y, h.z};
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        if (sc[k] == 0.f) continue;
                        const int b = bin_of(0.5f * (lo[k] + hi[k]), clo[k], sc[k]);
                        atomicAdd(&lb.cnt[k][b], 1u);
                        for (int a = 0; a < 3; ++a) {
                            atomicMin(&lb.lo[k][b][a], fkey(lo[a]));
                            atomicMax(&lb.hi[k][b][a], fkey(hi[a]));
                        }
                    }
                }
                __syncthreads();
                // the sweep: one axis at a time, bins on lanes (both half-waves hold the same axis)
                const float parent = fmaxf(area_of(blo[0], blo[1], blo[2], bhi[0], bhi[1], bhi[2]), 1e-30f);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (sc[k] == 0.f) continue;
                    float c;
                    int split;
                    uint32_t nl;
                    sweep_axis(lb, sx, k, parent, c, split, nl);
                    if (split >= 0 && c < best_cost) { best_cost = c; best_axis = k; best_split = split; best_nl = nl; }
                }
                __syncthreads();
            }
            const bool leaf = cnt <= 1u || (cnt <= uint32_t(kLeafMax) &&
                                             (best_axis < 0 || float(cnt) <= best_cost - (1.f - kSahNodeCost)));
            BinaryNode& nd = nodes[cur.node];
            if (leaf) {
                if (lane == 0) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) { nd.lo[k] = blo[k]; nd.hi[k] = bhi[k]; }
                    nd.left = 0;
                    nd.right = 0;
                    nd.first = cur.begin;
                    nd.count = cnt;
                }
                for (uint32_t i = cur.begin + lane; i < cur.begin + cnt; i += 64) order[i] = cached ? cache.id[lane] : in[i];
            } else {
                const uint32_t nl = best_axis >= 0 ? best_nl : cnt / 2;
                // stable partition into the other buffer
                uint32_t lc = 0, rc = 0;
                // (selects rather than indexing keeps clo / sc in registers)
                const float pclo = best_axis == 0 ? clo[0] : best_axis == 1 ? clo[1] : best_axis == 2 ? clo[2] : 0.f;
                const float psc = best_axis == 0 ? sc[0] : best_axis == 1 ? sc[1] : best_axis == 2 ? sc[2] : 0.f;
                const unsigned long long below = (1ull << lane) - 1ull;
                for (uint32_t r0 = cur.begin; r0 < cur.begin + cnt; r0 += 64) {
                    const uint32_t i = r0 + lane;
                    const bool valid = i < cur.begin + cnt;
                    uint32_t p = 0;
                    bool left = false;
                    if (valid) {
                        float4 l, h;
                        const bool ok = tri(i, p, l, h);
                        if (!ok) p = cached ? cache.id[lane] : in[i];   // an out-of-range id is carried along unchanged
                        if (best_axis < 0) {
                            left = (i - cur.begin) < nl;
                        } else if (ok) {
                            const float c = best_axis == 0 ? 0.5f * (l.x + h.x)
                                          : best_axis == 1 ? 0.5f * (l.y + h.y) : 0.5f * (l.z + h.z);
                            left = bin_of(c, pclo, psc) < best_split;
                        }
                    }
                    const unsigned long long bl = __ballot(valid && left);
                    const unsigned long long br = __ballot(valid && !left);
                    if (valid) {
                        const uint32_t dst = left ? cur.begin + lc + __popcll(bl & below)
                                                  : cur.begin + nl + rc + __popcll(br & below);
                        if (dst < cur.begin + cnt) out[dst] = p;
                        else atomicOr(err, kErrRange);
                    }
                    lc += __popcll(bl);
                    rc += __popcll(br);
                }
                const uint32_t l = slot;
                slot += 2u;
                if (slot > slot_end) {
                    if (lane == 0) atomicOr(err, kErrNodes);
                    break;
                }
                if (lane == 0) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) { nd.lo[k] = blo[k]; nd.hi[k] = bhi[k]; }
                    nd.left = l;
                    nd.right = l + 1;
                    nd.first = 0;
                    nd.count = 0;
                }
                // children: left (l, first nl triangles) and right (l + 1, the rest); continue
                // with the smaller, keep the larger (field-wise selects keep both in registers)
                const bool left_small = nl <= cnt - nl;
                const uint32_t buf = cur.buf ^ 1u;
                const SubTask big{left_small ? l + 1 : l, left_small ? cur.begin + nl : cur.begin,
                                  left_small ? cnt - nl : nl, buf};
                const SubTask small{left_small ? l : l + 1, left_small ? cur.begin : cur.begin + nl,
                                    left_small ? nl : cnt - nl, buf};
                if (sp >= kStack) {
                    if (lane == 0) atomicOr(err, kErrStack);
                    break;
                }
                if (lane == 0) stack[sp] = big;
                ++sp;
                cur = small;
                __syncthreads();
                continue;
            }
            if (sp == 0) {
                have = false;
            } else {
                --sp;
                __syncthreads();
                cur = stack[sp];
                __syncthreads();
            }
        }
        // slots the subtree did not need (leaves of several triangles) stay unused
        if (lane == 0 && slot < slot_end) atomicAdd(unused, slot_end - slot);
    }
}

__global__ void k_fill_u32_bvh(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

__global__ void k_init_bins(Bins* bins, uint32_t ntasks) {
    for (uint32_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        Bins& g = bins[t];
        uint32_t* c = &g.cnt[0][0];
        int* lo = &g.lo[0][0][0];
        int* hi = &g.hi[0][0][0];
        for (int i = threadIdx.x; i < 3 * kBins; i += blockDim.x) c[i] = 0u;
        for (int i = threadIdx.x; i < 3 * kBins * 3; i += blockDim.x) { lo[i] = kIntMax; hi[i] = -int(kIntMax) - 1; }
    }
}

#define BVH_CHECK(x)                     \
    do {                                 \
        if ((x) != hipSuccess) return false; \
    } while (0)

}  // namespace

namespace {

// The binary tree of one build, on the device.
struct DeviceTree {
    DevBuf<BinaryNode> nodes;
    DevBuf<uint32_t> prims;   // tree triangle position -> input triangle
    uint32_t used = 0;        // node slots taken, root 0
    uint32_t unused = 0;      // of which reserved by a subtree but not needed (never reached)
};

// Builds the tree of triangles d0/d1/d2 (device arrays, 3 floats per
// triangle) on stream s; device_ms receives the kernels' time.
bool build_tree(const float* d0, const float* d1, const float* d2, uint32_t n, hipStream_t s, DeviceTree& tree,
                double* device_ms, const char** why) {
    if (n < 8 || n > (1u << 29)) { *why = "size"; return false; }
    const uint32_t cap = 2 * n;
    DevBuf<float4> plo, phi;
    DevBuf<uint32_t> ids_a, ids_b, counter, err, unused;
    DevBuf<BinaryNode>& nodes = tree.nodes;
    DevBuf<uint32_t>& order = tree.prims;
    if (!plo.alloc(n) || !phi.alloc(n) || !ids_a.alloc(n) || !ids_b.alloc(n) || !order.alloc(n) ||
        !counter.alloc(1) || !err.alloc(1) || !unused.alloc(1) || !nodes.alloc(cap)) {
        *why = "allocation";
        return false;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BVH_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return false; }
    struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
    BVH_CHECK(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_prim_bounds, dim3(std::min<uint32_t>((n + 255) / 256, 65536)), dim3(256), 0, s, d0, d1, d2,
                       n, plo.p, phi.p, ids_a.p);
    hipLaunchKernelGGL(k_fill_u32_bvh, dim3(1), dim3(64), 0, s, counter.p, size_t(1), 1u);
    hipLaunchKernelGGL(k_fill_u32_bvh, dim3(1), dim3(64), 0, s, err.p, size_t(1), 0u);
    hipLaunchKernelGGL(k_fill_u32_bvh, dim3(1), dim3(64), 0, s, unused.p, size_t(1), 0u);
    BVH_CHECK(hipGetLastError());
    // large-node levels
    std::vector<NodeTask> medium;
    std::vector<NodeTask> level{NodeTask{0, 0, n, 0}};
    DevBuf<LargeTask> dtasks;
    DevBuf<WorkItem> ditems;
    DevBuf<Bins> dbins;
    DevBuf<NodeTask> dchildren;
    for (int levels = 0; !level.empty(); ++levels) {
        if (levels >= 64) { *why = "level count"; return false; }
        std::vector<NodeTask> large;
        for (const NodeTask& t : level) {
            if (t.count > kLargeNode) large.push_back(t);
            else if (t.count > 0) medium.push_back(t);
        }
        if (large.empty()) break;
        const uint32_t nt = uint32_t(large.size());
        std::vector<LargeTask> ht(nt);
        std::vector<WorkItem> hi;
        for (uint32_t i = 0; i < nt; ++i) {
            LargeTask& t = ht[i];
            std::memset(&t, 0, sizeof t);
            t.node = large[i].node;
            t.begin = large[i].begin;
            t.count = large[i].count;
            t.buf = large[i].buf;
            for (int k = 0; k < 3; ++k) {
                t.box_lo[k] = int(kIntMax); t.box_hi[k] = -int(kIntMax) - 1;
                t.cen_lo[k] = int(kIntMax); t.cen_hi[k] = -int(kIntMax) - 1;
            }
            for (uint32_t c = t.begin; c < t.begin + t.count; c += kChunk)
                hi.push_back(WorkItem{i, c, std::min(t.begin + t.count, c + kChunk)});
        }
        const uint32_t ni = uint32_t(hi.size());
        if (dtasks.n < nt && !dtasks.alloc(nt)) { *why = "allocation"; return false; }
        if (ditems.n < ni && !ditems.alloc(ni)) { *why = "allocation"; return false; }
        if (dbins.n < nt && !dbins.alloc(nt)) { *why = "allocation"; return false; }
        if (dchildren.n < 2 * nt && !dchildren.alloc(2 * nt)) { *why = "allocation"; return false; }
        BVH_CHECK(hipMemcpyAsync(dtasks.p, ht.data(), nt * sizeof(LargeTask), hipMemcpyHostToDevice, s));
        BVH_CHECK(hipMemcpyAsync(ditems.p, hi.data(), ni * sizeof(WorkItem), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_init_bins, dim3(std::min<uint32_t>(nt, 4096)), dim3(256), 0, s, dbins.p, nt);
        hipLaunchKernelGGL(k_large_bounds, dim3(ni), dim3(kBlock), 0, s, dtasks.p, ditems.p, ni, plo.p, phi.p,
                           ids_a.p, ids_b.p, n, err.p);
        hipLaunchKernelGGL(k_large_bin, dim3(ni), dim3(kBlock), 0, s, dtasks.p, ditems.p, ni, dbins.p, plo.p, phi.p,
                           ids_a.p, ids_b.p, n);
        hipLaunchKernelGGL(k_large_split, dim3((nt + 63) / 64), dim3(64), 0, s, dtasks.p, nt, dbins.p, nodes.p, cap,
                           counter.p, dchildren.p, err.p);
        hipLaunchKernelGGL(k_large_partition, dim3(ni), dim3(kBlock), 0, s, dtasks.p, ditems.p, ni, plo.p, phi.p,
                           ids_a.p, ids_b.p, n, err.p);
        BVH_CHECK(hipGetLastError());
        level.resize(2 * nt);
        BVH_CHECK(hipMemcpyAsync(level.data(), dchildren.p, 2 * nt * sizeof(NodeTask), hipMemcpyDeviceToHost, s));
        uint32_t herr = 0;
        BVH_CHECK(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
        BVH_CHECK(hipStreamSynchronize(s));
        if (herr) { *why = "large-node pass"; return false; }
        for (const NodeTask& c : level)
            if (c.begin + c.count > n || c.node >= cap) { *why = "large-node children"; return false; }
    }
    // subtrees
    if (!medium.empty()) {
        DevBuf<NodeTask> dmed;
        if (!dmed.alloc(medium.size())) { *why = "allocation"; return false; }
        BVH_CHECK(hipMemcpyAsync(dmed.p, medium.data(), medium.size() * sizeof(NodeTask), hipMemcpyHostToDevice, s));
        const uint32_t grid = uint32_t(std::min<size_t>(medium.size(), 65536));
        hipLaunchKernelGGL(k_subtree, dim3(grid), dim3(64), 0, s, dmed.p, uint32_t(medium.size()), plo.p, phi.p,
                           ids_a.p, ids_b.p, order.p, n, nodes.p, cap, counter.p, unused.p, err.p);
        BVH_CHECK(hipGetLastError());
        BVH_CHECK(hipStreamSynchronize(s));
    }
    BVH_CHECK(hipEventRecord(e1, s));
    uint32_t used = 0, herr = 0, spare = 0;
    BVH_CHECK(hipMemcpyAsync(&used, counter.p, 4, hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipMemcpyAsync(&spare, unused.p, 4, hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (device_ms) *device_ms = ms;
    if (herr || used == 0 || used > cap || spare >= used) { *why = "subtree pass"; return false; }
    tree.used = used;
    tree.unused = spare;
    return true;
}

// Corner positions of each triangle (3 floats per corner) from the indexed
// mesh; an index out of range sets err and reads vertex 0.
__global__ void k_gather_corners(const float* __restrict__ verts, uint32_t nv, const uint32_t* __restrict__ idx,
                                 uint32_t nt, float* __restrict__ d0, float* __restrict__ d1, float* __restrict__ d2,
                                 uint32_t* __restrict__ err) {
    bool bad = false;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        float* dst[3] = {d0, d1, d2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            uint32_t v = idx[3 * size_t(t) + c];
            if (v >= nv) { bad = true; v = 0; }
#pragma unroll
            for (int k = 0; k < 3; ++k) dst[c][3 * size_t(t) + k] = verts[3 * size_t(v) + k];
        }
    }
    if (bad) atomicOr(err, 1u);
}

// Triangle and material arrays in BVH order (layouts of common.hpp): position
// i holds triangle ref_tri[r] of reference r = order[i]; inv[r] = i.  A
// reference that occurs twice or out of range sets err.
__global__ void k_scene_arrays(const uint32_t* __restrict__ order, uint32_t n, const uint32_t* __restrict__ ref_tri,
                               uint32_t nt, const float* __restrict__ d0, const float* __restrict__ d1,
                               const float* __restrict__ d2, const float* __restrict__ albedo,
                               const float* __restrict__ emission, float4* __restrict__ tris,
                               float4* __restrict__ mats, uint32_t* __restrict__ inv, uint32_t* __restrict__ err) {
    bool bad = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = order[i];
        if (r >= n || atomicExch(&inv[r], i) != 0xFFFFFFFFu) { bad = true; continue; }
        const uint32_t t = ref_tri[r];
        if (t >= nt) { bad = true; continue; }
        const size_t o = 3 * size_t(t);
        const float ax = d0[o], ay = d0[o + 1], az = d0[o + 2];
        tris[3 * size_t(i)] = make_float4(ax, ay, az, __uint_as_float(t));
        tris[3 * size_t(i) + 1] = make_float4(d1[o] - ax, d1[o + 1] - ay, d1[o + 2] - az, 0.f);
        tris[3 * size_t(i) + 2] = make_float4(d2[o] - ax, d2[o + 1] - ay, d2[o + 2] - az, 0.f);
        const float e0 = emission[o], e1 = emission[o + 1], e2 = emission[o + 2];
        const float m12 = e1 < e2 ? e2 : e1;          // std::max(e0, std::max(e1, e2)) of the host layout
        const float m = e0 < m12 ? m12 : e0;
        mats[2 * size_t(i)] = make_float4(albedo[o], albedo[o + 1], albedo[o + 2], m);
        mats[2 * size_t(i) + 1] = make_float4(e0, e1, e2, 0.f);
    }
    if (bad) atomicOr(err, 2u);
}

// fkey bounds of the finite corner coordinates, per axis, into b[0..2] (min)
// and b[3..5] (max), which hold kIntMax and -kIntMax - 1 beforehand: the scene
// box of the split threshold.  Blocks of 256 threads over a grid of at most
// kBoundsBlocks, one atomic per block and bound.
constexpr uint32_t kBoundsBlocks = 512;
__global__ __launch_bounds__(256) void k_corner_bounds(const float* __restrict__ d0, const float* __restrict__ d1,
                                                       const float* __restrict__ d2, uint32_t nt,
                                                       int* __restrict__ b) {
    __shared__ int s_b[6];
    if (threadIdx.x < 6u) s_b[threadIdx.x] = threadIdx.x < 3u ? int(kIntMax) : -int(kIntMax) - 1;
    __syncthreads();
    int lo[3] = {int(kIntMax), int(kIntMax), int(kIntMax)};
    int hi[3] = {-int(kIntMax) - 1, -int(kIntMax) - 1, -int(kIntMax) - 1};
    const float* corners[3] = {d0, d1, d2};
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) {
                const float v = corners[c][3 * size_t(t) + k];
                if (!isfinite(v)) continue;
                lo[k] = min(lo[k], fkey(v));
                hi[k] = max(hi[k], fkey(v));
            }
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off >= 1; off >>= 1) {
            lo[k] = min(lo[k], __shfl_xor(lo[k], off));
            hi[k] = max(hi[k], __shfl_xor(hi[k], off));
        }
    }
    if ((threadIdx.x & 63u) == 0u)
        for (int k = 0; k < 3; ++k) {
            atomicMin(&s_b[k], lo[k]);
            atomicMax(&s_b[3 + k], hi[k]);
        }
    __syncthreads();
    if (threadIdx.x < 3u) atomicMin(&b[threadIdx.x], s_b[threadIdx.x]);
    else if (threadIdx.x < 6u) atomicMax(&b[threadIdx.x], s_b[threadIdx.x]);
}

constexpr int kSplitBlock = 256;

__device__ __forceinline__ bool emits(const float* e) {
    return (__float_as_uint(e[0]) | __float_as_uint(e[1]) | __float_as_uint(e[2])) != 0u;
}

// References per triangle (bvh_split.hpp), one triangle per thread; the sum
// of each block's into bsum[block].
__global__ __launch_bounds__(kSplitBlock) void k_split_count(const float* __restrict__ d0,
                                                             const float* __restrict__ d1,
                                                             const float* __restrict__ d2,
                                                             const float* __restrict__ emission, uint32_t nt, float T,
                                                             uint32_t* __restrict__ bsum) {
    __shared__ uint32_t s_sum;
    if (threadIdx.x == 0) s_sum = 0u;
    __syncthreads();
    const uint32_t t = blockIdx.x * kSplitBlock + threadIdx.x;
    uint32_t c = 0;
    if (t < nt) {
        PieceBox pc[kMaxPieces];
        const size_t o = 3 * size_t(t);
        c = split_triangle(d0 + o, d1 + o, d2 + o, T, !emits(emission + o), pc);
    }
    atomicAdd(&s_sum, c);
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = s_sum;
}

// The references themselves: triangle t's at [tri_off[t], tri_off[t] +
// count), block offsets from the host's scan of k_split_count's sums, in
// triangle order; boxes as lo / hi corners, ref_tri = t.
__global__ __launch_bounds__(kSplitBlock) void k_split_write(const float* __restrict__ d0,
                                                             const float* __restrict__ d1,
                                                             const float* __restrict__ d2,
                                                             const float* __restrict__ emission, uint32_t nt, float T,
                                                             const uint32_t* __restrict__ boff,
                                                             float* __restrict__ rlo, float* __restrict__ rhi,
                                                             uint32_t* __restrict__ ref_tri,
                                                             uint32_t* __restrict__ tri_off) {
    __shared__ uint32_t s_scan[kSplitBlock];
    const uint32_t t = blockIdx.x * kSplitBlock + threadIdx.x;
    PieceBox pc[kMaxPieces];
    uint32_t c = 0;
    if (t < nt) {
        const size_t o = 3 * size_t(t);
        c = split_triangle(d0 + o, d1 + o, d2 + o, T, !emits(emission + o), pc);
    }
    s_scan[threadIdx.x] = c;
    __syncthreads();
    for (uint32_t off = 1; off < uint32_t(kSplitBlock); off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0u;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    if (t >= nt) return;
    const uint32_t base = boff[blockIdx.x] + s_scan[threadIdx.x] - c;
    tri_off[t] = base;
    for (uint32_t j = 0; j < c; ++j) {
        for (int k = 0; k < 3; ++k) {
            rlo[3 * size_t(base + j) + k] = pc[j].lo[k];
            rhi[3 * size_t(base + j) + k] = pc[j].hi[k];
        }
        ref_tri[base + j] = t;
    }
}

// Emission of the listed triangles (ids checked on the host) into a zeroed array.
__global__ void k_scatter_emission(const uint32_t* __restrict__ ids, const float* __restrict__ vals, uint32_t m,
                                   float* __restrict__ emission) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        const size_t o = 3 * size_t(ids[j]);
        emission[o] = vals[3 * size_t(j)];
        emission[o + 1] = vals[3 * size_t(j) + 1];
        emission[o + 2] = vals[3 * size_t(j) + 2];
    }
}

// BVH-order position of each queried triangle: that of its first reference
// (an emitter has only one).
__global__ void k_gather_positions(const uint32_t* __restrict__ inv, const uint32_t* __restrict__ tri_off,
                                   const uint32_t* __restrict__ query, uint32_t nq, uint32_t* __restrict__ pos) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x)
        pos[q] = inv[tri_off[query[q]]];
}

uint32_t grid256(size_t items) { return uint32_t(std::max<size_t>(1, std::min<size_t>((items + 255) / 256, 16384))); }

}  // namespace

void DeviceScene::release() {
    if (tris) (void)hipFree(tris);
    if (mats) (void)hipFree(mats);
    if (nodes) (void)hipFree(nodes);
    if (nodes4) (void)hipFree(nodes4);
    tris = mats = nullptr;
    nodes = nodes4 = nullptr;
}

bool build_scene_gpu(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* albedo,
                     const float* emission, const uint32_t* emitters, uint32_t n_emit, const uint32_t* query,
                     uint32_t nquery, uint32_t* query_pos, hipStream_t s, DeviceScene& out, BvhGpuTimes* times,
                     const char** why) {
    *why = "";
    out.release();
    if (nt < 8 || nt > (1u << 29) || nv == 0) { *why = "size"; return false; }
    for (uint32_t q = 0; q < nquery; ++q)
        if (query[q] >= nt) { *why = "query"; return false; }
    BvhGpuTimes tm;
    using clock = std::chrono::steady_clock;
    clock::time_point tp = clock::now();
    auto lap = [&tp](double& into) {
        const clock::time_point now = clock::now();
        into = std::chrono::duration<double, std::milli>(now - tp).count();
        tp = now;
    };
    const size_t n3 = 3 * size_t(nt);
    DevBuf<float> dverts, dalb, demi, dp[3];
    DevBuf<uint32_t> didx, err;
    if (!dverts.alloc(3 * size_t(nv)) || !didx.alloc(n3) || !dalb.alloc(n3) || !demi.alloc(n3) || !dp[0].alloc(n3) ||
        !dp[1].alloc(n3) || !dp[2].alloc(n3) || !err.alloc(1)) {
        *why = "allocation";
        return false;
    }
    lap(tm.alloc_ms);
    BVH_CHECK(hipMemcpyAsync(dverts.p, verts, 12 * size_t(nv), hipMemcpyHostToDevice, s));
    BVH_CHECK(hipMemcpyAsync(didx.p, idx, 4 * n3, hipMemcpyHostToDevice, s));
    BVH_CHECK(hipMemcpyAsync(dalb.p, albedo, 4 * n3, hipMemcpyHostToDevice, s));
    // emission: the whole array, or zeros and the few nonzero triangles' values
    const bool sparse = emitters && n_emit <= nt / 8;
    std::vector<float> evals;
    DevBuf<uint32_t> deids;
    DevBuf<float> devals;
    if (!sparse) {
        BVH_CHECK(hipMemcpyAsync(demi.p, emission, 4 * n3, hipMemcpyHostToDevice, s));
    } else {
        evals.resize(3 * size_t(n_emit));
        for (uint32_t j = 0; j < n_emit; ++j) {
            if (emitters[j] >= nt) { *why = "emitter id"; return false; }
            std::memcpy(&evals[3 * size_t(j)], &emission[3 * size_t(emitters[j])], 12);
        }
        if (!deids.alloc(n_emit) || !devals.alloc(3 * size_t(n_emit))) { *why = "allocation"; return false; }
        hipLaunchKernelGGL(k_fill_u32_bvh, dim3(grid256(n3)), dim3(256), 0, s, reinterpret_cast<uint32_t*>(demi.p),
                           n3, 0u);
        if (n_emit) {
            BVH_CHECK(hipMemcpyAsync(deids.p, emitters, 4 * size_t(n_emit), hipMemcpyHostToDevice, s));
            BVH_CHECK(hipMemcpyAsync(devals.p, evals.data(), 12 * size_t(n_emit), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_scatter_emission, dim3(grid256(n_emit)), dim3(256), 0, s, deids.p, devals.p, n_emit,
                               demi.p);
        }
    }
    hipLaunchKernelGGL(k_fill_u32_bvh, dim3(1), dim3(64), 0, s, err.p, size_t(1), 0u);
    hipLaunchKernelGGL(k_gather_corners, dim3(grid256(nt)), dim3(256), 0, s, dverts.p, nv, didx.p, nt, dp[0].p,
                       dp[1].p, dp[2].p, err.p);
    BVH_CHECK(hipGetLastError());
    uint32_t herr = 0;
    BVH_CHECK(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipStreamSynchronize(s));
    if (herr) { *why = "vertex index out of range"; return false; }
    dverts.release();
    didx.release();
    lap(tm.upload_ms);

    // references (bvh_split.hpp): the threshold from the scene box, then the
    // pieces of each triangle in triangle order
    DevBuf<int> dbounds;
    if (!dbounds.alloc(6)) { *why = "allocation"; return false; }
    hipLaunchKernelGGL(k_fill_u32_bvh, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(dbounds.p), size_t(3),
                       kIntMax);
    hipLaunchKernelGGL(k_fill_u32_bvh, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(dbounds.p) + 3, size_t(3),
                       kIntMax + 1u);
    hipLaunchKernelGGL(k_corner_bounds, dim3(std::min(grid256(nt), kBoundsBlocks)), dim3(256), 0, s, dp[0].p, dp[1].p,
                       dp[2].p, nt, dbounds.p);
    BVH_CHECK(hipGetLastError());
    int hb[6];
    BVH_CHECK(hipMemcpyAsync(hb, dbounds.p, sizeof hb, hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipStreamSynchronize(s));
    float slo[3], shi[3];
    bool finite_box = true;
    for (int k = 0; k < 3; ++k) {
        finite_box = finite_box && hb[k] <= hb[3 + k];
        const int a = hb[k] >= 0 ? hb[k] : hb[k] ^ int(kIntMax), b = hb[3 + k] >= 0 ? hb[3 + k] : hb[3 + k] ^ int(kIntMax);
        std::memcpy(&slo[k], &a, 4);
        std::memcpy(&shi[k], &b, 4);
    }
    float T = finite_box ? split_threshold(slo, shi, nt) : 3.0e38f;
    const uint32_t nblk = (nt + kSplitBlock - 1) / kSplitBlock;
    DevBuf<uint32_t> bsum, boff, tri_off, ref_tri;
    DevBuf<float> rlo, rhi;
    if (!bsum.alloc(nblk) || !boff.alloc(nblk) || !tri_off.alloc(nt)) { *why = "allocation"; return false; }
    std::vector<uint32_t> hsum(nblk);
    uint64_t total = 0;
    // references beyond the builder's limit: build the triangles unsplit
    for (int attempt = 0; attempt < 2; ++attempt) {
        hipLaunchKernelGGL(k_split_count, dim3(nblk), dim3(kSplitBlock), 0, s, dp[0].p, dp[1].p, dp[2].p, demi.p, nt,
                           T, bsum.p);
        BVH_CHECK(hipGetLastError());
        BVH_CHECK(hipMemcpyAsync(hsum.data(), bsum.p, 4 * size_t(nblk), hipMemcpyDeviceToHost, s));
        BVH_CHECK(hipStreamSynchronize(s));
        total = 0;
        for (uint32_t b = 0; b < nblk; ++b) {
            const uint32_t c = hsum[b];
            hsum[b] = uint32_t(total);
            total += c;
        }
        if (total <= (1u << 29)) break;
        T = 3.0e38f;
    }
    if (total < nt || total > (1u << 29)) { *why = "reference count"; return false; }
    const uint32_t nr = uint32_t(total);
    if (!rlo.alloc(3 * size_t(nr)) || !rhi.alloc(3 * size_t(nr)) || !ref_tri.alloc(nr)) {
        *why = "allocation";
        return false;
    }
    BVH_CHECK(hipMemcpyAsync(boff.p, hsum.data(), 4 * size_t(nblk), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_split_write, dim3(nblk), dim3(kSplitBlock), 0, s, dp[0].p, dp[1].p, dp[2].p, demi.p, nt, T,
                       boff.p, rlo.p, rhi.p, ref_tri.p, tri_off.p);
    BVH_CHECK(hipGetLastError());
    bsum.release();
    lap(tm.split_ms);

    // the tree over the reference boxes (each given as the triangle lo, hi, lo)
    DeviceTree tree;
    if (!build_tree(rlo.p, rhi.p, rlo.p, nr, s, tree, &tm.device_ms, why)) {
        if (!**why) *why = "HIP call";
        return false;
    }
    rlo.release();
    rhi.release();
    lap(tm.build_ms);

    DeviceLayout lay;
    struct LayGuard { DeviceLayout& l; ~LayGuard() { l.release(); } } lg{lay};
    if (!layout_tree_gpu(tree.nodes.p, tree.used, tree.unused, tree.prims.p, nr, s, lay, why)) return false;
    tree.nodes.release();
    tree.prims.release();
    DevBuf<float4> tris, mats;
    DevBuf<uint32_t> inv, dq, dpos;
    if (!tris.alloc(3 * size_t(nr)) || !mats.alloc(2 * size_t(nr)) || !inv.alloc(nr) ||
        (nquery && (!dq.alloc(nquery) || !dpos.alloc(nquery)))) {
        *why = "allocation";
        return false;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BVH_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return false; }
    struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
    BVH_CHECK(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_fill_u32_bvh, dim3(grid256(nr)), dim3(256), 0, s, inv.p, size_t(nr), 0xFFFFFFFFu);
    hipLaunchKernelGGL(k_scene_arrays, dim3(grid256(nr)), dim3(256), 0, s, lay.order, nr, ref_tri.p, nt, dp[0].p,
                       dp[1].p, dp[2].p, dalb.p, demi.p, tris.p, mats.p, inv.p, err.p);
    if (nquery) {
        BVH_CHECK(hipMemcpyAsync(dq.p, query, 4 * size_t(nquery), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gather_positions, dim3(grid256(nquery)), dim3(256), 0, s, inv.p, tri_off.p, dq.p, nquery,
                           dpos.p);
        BVH_CHECK(hipMemcpyAsync(query_pos, dpos.p, 4 * size_t(nquery), hipMemcpyDeviceToHost, s));
    }
    BVH_CHECK(hipGetLastError());
    BVH_CHECK(hipEventRecord(e1, s));
    BVH_CHECK(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipStreamSynchronize(s));
    if (herr) { *why = "triangle order"; return false; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    tm.layout_device_ms = lay.device_ms + ms;
    out.tris = tris.detach();
    out.mats = mats.detach();
    out.nodes = lay.nodes;
    out.nodes4 = lay.nodes4;
    lay.nodes = lay.nodes4 = nullptr;
    out.num_prims = nr;
    out.num_nodes = lay.num_nodes;
    out.num_leaves = lay.num_leaves;
    out.depth = lay.depth;
    out.num_nodes4 = lay.num_nodes4;
    out.depth4 = lay.depth4;
    out.sah_cost = lay.sah_cost;
    lap(tm.layout_ms);
    if (times) *times = tm;
    return true;
}

bool build_tree_gpu_host(const float* p0, const float* p1, const float* p2, uint32_t ntri, hipStream_t s,
                         std::vector<BinaryNode>& tree_out, std::vector<uint32_t>& prims_out, uint32_t* unused,
                         const char** why) {
    *why = "";
    DevBuf<float> dp[3];
    const float* src[3] = {p0, p1, p2};
    for (int k = 0; k < 3; ++k) {
        if (!dp[k].alloc(3 * size_t(ntri))) { *why = "allocation"; return false; }
        BVH_CHECK(hipMemcpyAsync(dp[k].p, src[k], 12 * size_t(ntri), hipMemcpyHostToDevice, s));
    }
    DeviceTree tree;
    if (!build_tree(dp[0].p, dp[1].p, dp[2].p, ntri, s, tree, nullptr, why)) {
        if (!**why) *why = "HIP call";
        return false;
    }
    tree_out.resize(tree.used);
    prims_out.resize(ntri);
    if (unused) *unused = tree.unused;
    BVH_CHECK(hipMemcpyAsync(tree_out.data(), tree.nodes.p, tree.used * sizeof(BinaryNode), hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipMemcpyAsync(prims_out.data(), tree.prims.p, 4 * size_t(ntri), hipMemcpyDeviceToHost, s));
    BVH_CHECK(hipStreamSynchronize(s));
    return true;
}

}  // namespace quasi_hip
