// bvh_layout.hip -- the pair-node and 4-wide traversal layouts of a binary
// build tree, computed on the GPU.
//
// The output is the one convert_binary_tree (bvh_build.cpp) writes for the
// same tree, byte for byte: both layouts in depth-first preorder, left child
// before right, a node's leaf children appended to the triangle order when
// the node is visited, and each 4-wide node made from a binary node by
// opening, while fewer than four children remain, the internal child of
// largest surface area (first such child on ties).  Preorder numbers follow
// from subtree sizes: pair node x's left internal child is x + 1 and its
// right internal child x + 1 + (internal nodes under the left child); the
// same holds for triangle positions and for 4-wide nodes.  So the tree is
// walked level by level:
//  1. top-down, breadth first: each level's internal nodes into a list, with
//     every index checked (children in range and above their parent, every
//     node reached once, leaves of 1..kLeafMax triangles inside prims that
//     cover each prims entry once, prims a permutation);
//  2. bottom-up: subtree sizes (internal nodes, triangles, 4-wide nodes),
//     heights, SAH cost and each node's 4-wide children;
//  3. top-down: preorder numbers and triangle positions; each pair node and
//     its leaves' triangle order are written as soon as its numbers are known;
//  4. one pass over the internal nodes writes the 4-wide nodes.
// The host reads back the level sizes once, between steps 1 and 2.
//
// k_quantize4 derives the 64-byte quantised BVH4 nodes (bvh_quant.hpp) from
// the 128-byte ones, whichever builder made them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "bvh_gpu.hpp"
#include "bvh_quant.hpp"
#include "common.hpp"
#include "device_buffer.hpp"
#include "kernels.hpp"

namespace quasi_hip {

namespace {

constexpr int kMaxLevels = 64;            // binary levels; deeper trees are rejected
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kLayBlock = 256;
// ctl words
constexpr int kCtlErr = 0, kCtlLeaves = 1, kCtlCount = 2, kCtlBegin = kCtlCount + kMaxLevels + 2;
constexpr int kCtlWords = kCtlBegin + kMaxLevels + 2;
// error bits
constexpr uint32_t kLayRoot = 1, kLayChild = 2, kLayTwice = 4, kLayLeaf = 8, kLayWrite = 16;

// Surface area in double, rounded like the host converter's (no contraction
// into fused multiply-adds).
__device__ __forceinline__ double area_d(const BinaryNode& b) {
#pragma clang fp contract(off)
    const double dx = double(b.hi[0]) - double(b.lo[0]);
    const double dy = double(b.hi[1]) - double(b.lo[1]);
    const double dz = double(b.hi[2]) - double(b.lo[2]);
    if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

__global__ void k_lay_prep(const BinaryNode* __restrict__ tree, uint32_t used, uint32_t n, double* __restrict__ area,
                           uint32_t* __restrict__ visited, uint32_t* __restrict__ idx4, uint32_t* __restrict__ cover,
                           uint32_t* __restrict__ seen) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < used; i += gridDim.x * blockDim.x) {
        area[i] = area_d(tree[i]);
        visited[i] = 0u;
        idx4[i] = kNone;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        cover[i] = 0u;
        seen[i] = 0u;
    }
}

// prims must be a permutation of 0 .. n-1.
__global__ void k_lay_perm(const uint32_t* __restrict__ prims, uint32_t n, uint32_t* __restrict__ seen,
                           uint32_t* __restrict__ err) {
    bool bad = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t v = prims[i];
        if (v >= n || atomicAdd(&seen[v], 1u) != 0u) bad = true;
    }
    if (bad) atomicOr(err, kLayLeaf);
}

__global__ void k_lay_root(const BinaryNode* __restrict__ tree, uint32_t* __restrict__ ctl,
                           uint32_t* __restrict__ bfs, uint32_t* __restrict__ visited, uint2* __restrict__ place,
                           uint32_t* __restrict__ idx4) {
    for (int i = threadIdx.x; i < kCtlWords; i += blockDim.x) ctl[i] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl[kCtlCount + 0] = 1u;   // level 0: the root
        if (tree[0].count != 0) ctl[kCtlErr] = kLayRoot;
        bfs[0] = 0u;
        visited[0] = 1u;
        place[0] = make_uint2(0u, 0u);
        idx4[0] = 0u;
    }
}

// Step 1, level L: internal children of the level's nodes into level L + 1.
__global__ void k_lay_bfs(const BinaryNode* __restrict__ tree, uint32_t used, uint32_t n, int L,
                          uint32_t* __restrict__ ctl, uint32_t* __restrict__ bfs, uint32_t* __restrict__ visited,
                          uint32_t* __restrict__ cover) {
    const uint32_t b = ctl[kCtlBegin + L], c = ctl[kCtlCount + L], e = b + c;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[kCtlBegin + L + 1] = e;
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t bad = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < c; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool act = i < c && b + i < used;
        const uint32_t x = act ? bfs[b + i] : 0u;
        for (int side = 0; side < 2; ++side) {
            bool inner = false, leaf = false;
            uint32_t ch = 0;
            if (act) {
                ch = side ? tree[x].right : tree[x].left;
                if (ch <= x || ch >= used) {
                    bad |= kLayChild;
                } else if (atomicAdd(&visited[ch], 1u) != 0u) {
                    bad |= kLayTwice;
                } else {
                    const uint32_t cnt = tree[ch].count, first = tree[ch].first;
                    if (cnt == 0) {
                        inner = true;
                    } else if (cnt <= uint32_t(kLeafMax) && first <= n && cnt <= n - first) {
                        leaf = true;
                        for (uint32_t k = 0; k < cnt; ++k)
                            if (atomicAdd(&cover[first + k], 1u) != 0u) bad |= kLayTwice;   // ranges overlap
                    } else {
                        bad |= kLayLeaf;
                    }
                }
            }
            const unsigned long long mi = __ballot(inner);
            const unsigned long long ml = __ballot(leaf);
            uint32_t slot0 = 0;
            if (mi && lane == uint32_t(__ffsll(mi) - 1)) slot0 = atomicAdd(&ctl[kCtlCount + L + 1], uint32_t(__popcll(mi)));
            if (mi) slot0 = __shfl(slot0, __ffsll(mi) - 1, 64);
            if (ml && lane == uint32_t(__ffsll(ml) - 1)) atomicAdd(&ctl[kCtlLeaves], uint32_t(__popcll(ml)));
            if (inner) {
                const uint32_t slot = e + slot0 + uint32_t(__popcll(mi & below));
                if (slot < used) bfs[slot] = ch;
                else bad |= kLayWrite;
            }
        }
    }
    if (bad) atomicOr(&ctl[kCtlErr], bad);
}

// The 4-wide children of binary node x (kNone where fewer than four).
__device__ __forceinline__ uint4 kids4_of(const BinaryNode* __restrict__ tree, const double* __restrict__ area,
                                          uint32_t x) {
    uint32_t k[4] = {tree[x].left, tree[x].right, kNone, kNone};
    uint32_t nk = 2;
#pragma unroll
    for (int round = 0; round < 2; ++round) {
        int pick = -1;
        double best = -1.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (uint32_t(j) < nk && tree[k[j]].count == 0 && area[k[j]] > best) { best = area[k[j]]; pick = j; }
        if (pick < 0) break;
        const uint32_t l = tree[k[pick]].left, r = tree[k[pick]].right;
#pragma unroll
        for (int j = 3; j >= 1; --j)
            if (j > pick + 1 && uint32_t(j) <= nk) k[j] = k[j - 1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j == pick) k[j] = l;
            if (j == pick + 1) k[j] = r;
        }
        ++nk;
    }
    return make_uint4(k[0], k[1], k[2], k[3]);
}

// Step 2, level L: sizes, heights, cost and 4-wide children of its nodes.
// sums = (internal nodes, triangles, 4-wide nodes, height | height4 << 16).
__global__ void k_lay_up(const BinaryNode* __restrict__ tree, const double* __restrict__ area, double root_area,
                         const uint32_t* __restrict__ bfs, uint32_t begin, uint32_t count, uint4* __restrict__ sums,
                         double* __restrict__ cost, uint4* __restrict__ kids4) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t x = bfs[begin + i];
        uint32_t internal = 1, prims = 0, h2 = 0;
        double c = area[x] / root_area;
        const uint32_t ch[2] = {tree[x].left, tree[x].right};
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const uint32_t y = ch[side];
            const uint32_t cnt = tree[y].count;
            if (cnt) {
                prims += cnt;
                c += area[y] / root_area * double(cnt);
            } else {
                const uint4 s = sums[y];
                internal += s.x;
                prims += s.y;
                h2 = max(h2, s.w & 0xFFFFu);
                c += cost[y];
            }
        }
        const uint4 k = kids4_of(tree, area, x);
        const uint32_t kk[4] = {k.x, k.y, k.z, k.w};
        uint32_t n4 = 1, h4 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (kk[j] != kNone && tree[kk[j]].count == 0) {
                const uint4 s = sums[kk[j]];
                n4 += s.z;
                h4 = max(h4, s.w >> 16);
            }
        sums[x] = make_uint4(internal, prims, n4, (h2 + 1u) | ((h4 + 1u) << 16));
        cost[x] = c;
        kids4[x] = k;
    }
}

__device__ __forceinline__ int32_t leaf_ref_of(uint32_t pos, uint32_t cnt) {
    return ~int32_t((pos << 2) | (cnt - 1u));
}

// Step 3, level L: preorder numbers of the children, pair node and leaf
// triangle order of each node.  place = (pair-node index, first position).
__global__ void k_lay_down(const BinaryNode* __restrict__ tree, const uint32_t* __restrict__ prims, uint32_t n,
                           const uint32_t* __restrict__ bfs, uint32_t begin, uint32_t count,
                           const uint4* __restrict__ sums, const uint4* __restrict__ kids4,
                           uint2* __restrict__ place, uint32_t* __restrict__ idx4, int32_t* __restrict__ leaf_ref,
                           uint32_t num_nodes, uint32_t num_nodes4, float4* __restrict__ nodes,
                           uint32_t* __restrict__ order, uint32_t* __restrict__ err) {
    uint32_t bad = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t x = bfs[begin + i];
        const uint2 p = place[x];
        const BinaryNode nd = tree[x];
        const BinaryNode L = tree[nd.left], R = tree[nd.right];
        const uint32_t pos_l = p.y;
        const uint32_t pos_r = p.y + L.count;            // L.count is 0 for an internal child
        const uint32_t inner0 = pos_r + R.count;          // first position below the internal children
        int32_t ref[2];
        const uint32_t ch[2] = {nd.left, nd.right};
        const BinaryNode* cn[2] = {&L, &R};
        const uint32_t pos[2] = {pos_l, pos_r};
        uint32_t next_idx = p.x + 1u, next_pos = inner0;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const BinaryNode& c = *cn[side];
            if (c.count) {
                if (pos[side] > n || c.count > n - pos[side]) { bad |= kLayWrite; ref[side] = 0; continue; }
                ref[side] = leaf_ref_of(pos[side], c.count);
                leaf_ref[ch[side]] = ref[side];
                for (uint32_t k = 0; k < c.count; ++k) order[pos[side] + k] = prims[c.first + k];
            } else {
                if (next_idx >= num_nodes) { bad |= kLayWrite; ref[side] = 0; continue; }
                ref[side] = int32_t(next_idx);
                place[ch[side]] = make_uint2(next_idx, next_pos);
                const uint4 s = sums[ch[side]];
                next_idx += s.x;
                next_pos += s.y;
            }
        }
        if (p.x < num_nodes) {
            float4* o = nodes + 4 * size_t(p.x);
            o[0] = make_float4(L.lo[0], L.hi[0], L.lo[1], L.hi[1]);
            o[1] = make_float4(R.lo[0], R.hi[0], R.lo[1], R.hi[1]);
            o[2] = make_float4(L.lo[2], L.hi[2], R.lo[2], R.hi[2]);
            o[3] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), 0.f, 0.f);
        } else {
            bad |= kLayWrite;
        }
        const uint32_t q = idx4[x];
        if (q != kNone) {
            const uint4 k = kids4[x];
            const uint32_t kk[4] = {k.x, k.y, k.z, k.w};
            uint32_t next4 = q + 1u;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (kk[j] != kNone && tree[kk[j]].count == 0) {
                    if (next4 >= num_nodes4) { bad |= kLayWrite; break; }
                    idx4[kk[j]] = next4;
                    next4 += sums[kk[j]].z;
                }
        }
    }
    if (bad) atomicOr(err, bad);
}

// Step 4: the 4-wide nodes, one per binary node that heads one.
__global__ void k_lay_write4(const BinaryNode* __restrict__ tree, const uint32_t* __restrict__ bfs, uint32_t count,
                             const uint4* __restrict__ kids4, const uint32_t* __restrict__ idx4,
                             const int32_t* __restrict__ leaf_ref, uint32_t num_nodes4, float4* __restrict__ nodes4,
                             uint32_t* __restrict__ err) {
    uint32_t bad = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t x = bfs[i];
        const uint32_t q = idx4[x];
        if (q == kNone) continue;
        if (q >= num_nodes4) { bad |= kLayWrite; continue; }
        const uint4 k = kids4[x];
        const uint32_t kk[4] = {k.x, k.y, k.z, k.w};
        float lo[3][4], hi[3][4];
        int32_t ref[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (kk[j] == kNone) {
#pragma unroll
                for (int a = 0; a < 3; ++a) { lo[a][j] = kEmptyLo; hi[a][j] = kEmptyHi; }
                ref[j] = kEmptyChild;
            } else {
                const BinaryNode& c = tree[kk[j]];
#pragma unroll
                for (int a = 0; a < 3; ++a) { lo[a][j] = c.lo[a]; hi[a][j] = c.hi[a]; }
                if (c.count) {
                    ref[j] = leaf_ref[kk[j]];
                } else {
                    const uint32_t ci = idx4[kk[j]];
                    if (ci >= num_nodes4) bad |= kLayWrite;
                    ref[j] = int32_t(ci);
                }
            }
        }
        float4* o = nodes4 + 8 * size_t(q);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
            o[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
        }
        o[6] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]),
                           __int_as_float(ref[3]));
        o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (bad) atomicOr(err, bad);
}

uint32_t grid_for(size_t items) {
    return uint32_t(std::max<size_t>(1, std::min<size_t>((items + kLayBlock - 1) / kLayBlock, 8192)));
}

#define LAY_CHECK(x)                                              \
    do {                                                          \
        if ((x) != hipSuccess) { *why = "HIP call"; return false; } \
    } while (0)

// One thread per BVH4 node: its quantised form (bvh_quant.hpp).
__global__ void k_quantize4(const float4* __restrict__ n4, uint32_t n, float4* __restrict__ q, uint32_t* failed) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (!quantize_node4(n4 + size_t(i) * 8, q + size_t(i) * 4)) *failed = 1u;
}

}  // namespace

hipError_t launch_quantize4(const void* nodes4, uint32_t n, void* q, uint32_t* failed, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_quantize4, dim3(blocks), dim3(256), 0, stream, static_cast<const float4*>(nodes4), n,
                       static_cast<float4*>(q), failed);
    return hipGetLastError();
}

void DeviceLayout::release() {
    if (nodes) (void)hipFree(nodes);
    if (nodes4) (void)hipFree(nodes4);
    if (order) (void)hipFree(order);
    nodes = nodes4 = nullptr;
    order = nullptr;
}

bool layout_tree_gpu(const BinaryNode* tree, uint32_t used, uint32_t unused, const uint32_t* prims, uint32_t n,
                     hipStream_t s, DeviceLayout& out, const char** why) {
    out.release();
    *why = "";
    if (used < 3 || n == 0 || n > (1u << 29)) { *why = "layout size"; return false; }
    DevBuf<double> area, cost;
    DevBuf<uint32_t> visited, idx4, bfs, ctl, cover, seen;
    DevBuf<uint2> place;
    DevBuf<uint4> sums, kids4;
    DevBuf<int32_t> leaf_ref;
    if (!area.alloc(used) || !cost.alloc(used) || !visited.alloc(used) || !idx4.alloc(used) || !bfs.alloc(used) ||
        !ctl.alloc(kCtlWords) || !place.alloc(used) || !sums.alloc(used) || !kids4.alloc(used) ||
        !leaf_ref.alloc(used) || !cover.alloc(n) || !seen.alloc(n)) {
        *why = "layout allocation";
        return false;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    LAY_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); *why = "HIP call"; return false; }
    struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
    LAY_CHECK(hipEventRecord(e0, s));

    // 1. breadth-first levels
    hipLaunchKernelGGL(k_lay_prep, dim3(grid_for(std::max(used, n))), dim3(kLayBlock), 0, s, tree, used, n, area.p,
                       visited.p, idx4.p, cover.p, seen.p);
    hipLaunchKernelGGL(k_lay_root, dim3(1), dim3(kLayBlock), 0, s, tree, ctl.p, bfs.p, visited.p, place.p, idx4.p);
    hipLaunchKernelGGL(k_lay_perm, dim3(grid_for(n)), dim3(kLayBlock), 0, s, prims, n, seen.p, ctl.p + kCtlErr);
    for (int L = 0; L < kMaxLevels; ++L)
        hipLaunchKernelGGL(k_lay_bfs, dim3(grid_for(used / 2 + 1)), dim3(kLayBlock), 0, s, tree, used, n, L, ctl.p,
                           bfs.p, visited.p, cover.p);
    LAY_CHECK(hipGetLastError());
    std::vector<uint32_t> hctl(kCtlWords);
    LAY_CHECK(hipMemcpyAsync(hctl.data(), ctl.p, kCtlWords * 4, hipMemcpyDeviceToHost, s));
    LAY_CHECK(hipStreamSynchronize(s));
    if (hctl[kCtlErr]) { *why = "layout: malformed tree"; return false; }
    int levels = 0;
    uint64_t internal = 0;
    while (levels < kMaxLevels && hctl[kCtlCount + levels]) internal += hctl[kCtlCount + levels++];
    if (levels == kMaxLevels && hctl[kCtlCount + kMaxLevels]) { *why = "layout: tree deeper than 64 levels"; return false; }
    const uint32_t leaves = hctl[kCtlLeaves];
    if (internal + leaves + unused != used) { *why = "layout: unreachable nodes"; return false; }
    std::vector<uint32_t> begin(levels);
    for (int L = 0; L < levels; ++L) begin[L] = hctl[kCtlBegin + L];

    // 2. bottom-up sizes
    BinaryNode root;
    LAY_CHECK(hipMemcpyAsync(&root, tree, sizeof root, hipMemcpyDeviceToHost, s));
    LAY_CHECK(hipStreamSynchronize(s));
    double root_area;
    {
        const double dx = double(root.hi[0]) - root.lo[0], dy = double(root.hi[1]) - root.lo[1],
                     dz = double(root.hi[2]) - root.lo[2];
        root_area = (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) ? 0.0 : 2.0 * (dx * dy + dy * dz + dz * dx);
    }
    root_area = std::max(root_area, 1e-30);
    for (int L = levels - 1; L >= 0; --L) {
        const uint32_t c = hctl[kCtlCount + L];
        hipLaunchKernelGGL(k_lay_up, dim3(grid_for(c)), dim3(kLayBlock), 0, s, tree, area.p, root_area, bfs.p,
                           begin[L], c, sums.p, cost.p, kids4.p);
    }
    LAY_CHECK(hipGetLastError());
    uint4 rs;
    double rcost = 0;
    LAY_CHECK(hipMemcpyAsync(&rs, sums.p, sizeof rs, hipMemcpyDeviceToHost, s));
    LAY_CHECK(hipMemcpyAsync(&rcost, cost.p, sizeof rcost, hipMemcpyDeviceToHost, s));
    LAY_CHECK(hipStreamSynchronize(s));
    if (rs.x != internal || rs.y != n || (rs.w & 0xFFFFu) != uint32_t(levels) || rs.z == 0 || rs.z > rs.x) {
        *why = "layout: subtree sizes";
        return false;
    }
    const uint32_t num_nodes = rs.x, num_nodes4 = rs.z;

    // 3. top-down numbering, pair nodes and triangle order
    DevBuf<float4> dnodes, dnodes4;
    DevBuf<uint32_t> dorder;
    if (!dnodes.alloc(4 * size_t(num_nodes)) || !dnodes4.alloc(8 * size_t(num_nodes4)) || !dorder.alloc(n)) {
        *why = "layout allocation";
        return false;
    }
    for (int L = 0; L < levels; ++L) {
        const uint32_t c = hctl[kCtlCount + L];
        hipLaunchKernelGGL(k_lay_down, dim3(grid_for(c)), dim3(kLayBlock), 0, s, tree, prims, n, bfs.p, begin[L], c,
                           sums.p, kids4.p, place.p, idx4.p, leaf_ref.p, num_nodes, num_nodes4, dnodes.p, dorder.p,
                           ctl.p + kCtlErr);
    }
    // 4. 4-wide nodes
    hipLaunchKernelGGL(k_lay_write4, dim3(grid_for(num_nodes)), dim3(kLayBlock), 0, s, tree, bfs.p, num_nodes,
                       kids4.p, idx4.p, leaf_ref.p, num_nodes4, dnodes4.p, ctl.p + kCtlErr);
    LAY_CHECK(hipGetLastError());
    LAY_CHECK(hipEventRecord(e1, s));
    uint32_t herr = 0;
    LAY_CHECK(hipMemcpyAsync(&herr, ctl.p + kCtlErr, 4, hipMemcpyDeviceToHost, s));
    LAY_CHECK(hipStreamSynchronize(s));
    if (herr) { *why = "layout: index out of range"; return false; }
    (void)hipEventElapsedTime(&out.device_ms, e0, e1);
    out.nodes = reinterpret_cast<float*>(dnodes.detach());
    out.nodes4 = reinterpret_cast<float*>(dnodes4.detach());
    out.order = dorder.detach();
    out.num_nodes = num_nodes;
    out.num_leaves = leaves;
    out.depth = uint32_t(levels);
    out.num_nodes4 = num_nodes4;
    out.depth4 = rs.w >> 16;
    out.sah_cost = rcost;
    return true;
}

}  // namespace quasi_hip
