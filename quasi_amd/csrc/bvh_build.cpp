// bvh_build.cpp -- parallel binned-SAH build (see bvh_build.hpp).
//
// Work is a set of independent (node, prim range) tasks on a shared queue;
// each task bins its range on all three axes in one pass, picks the cheapest
// of the 31 x 3 candidate planes, partitions in place and pushes both halves.
// Node slots come from an atomic counter, so the temporary tree depends on
// scheduling; the final pair-node layout is produced by a depth-first walk
// from the root that orders children left before right, which makes the
// output independent of the thread count.
#include "bvh_build.hpp"

#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

namespace quasi_hip {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    double area() const {
        double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
        if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct TmpNode {
    Box box;
    uint32_t left = 0, right = 0;   // children (internal)
    uint32_t first = 0, count = 0;  // leaf range in prims (count > 0)
};

struct Builder {
    const float* p0;
    const float* p1;
    const float* p2;
    uint32_t ntri;
    std::vector<float> tlo, thi, cen;   // 3 per triangle
    std::vector<uint32_t> prims;
    std::vector<TmpNode> nodes;
    std::atomic<uint32_t> next_node{0};

    struct Task { uint32_t node, begin, end; };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Task> queue;
    uint32_t pending = 0;  // queued + running tasks

    void prepare(unsigned threads) {
        tlo.resize(3 * size_t(ntri));
        thi.resize(3 * size_t(ntri));
        cen.resize(3 * size_t(ntri));
        prims.resize(ntri);
        auto work = [&](size_t b, size_t e) {
            for (size_t t = b; t < e; ++t) {
                for (int k = 0; k < 3; ++k) {
                    float a = p0[3 * t + k], c = p1[3 * t + k], d = p2[3 * t + k];
                    float lo = std::min(a, std::min(c, d)), hi = std::max(a, std::max(c, d));
                    tlo[3 * t + k] = lo;
                    thi[3 * t + k] = hi;
                    cen[3 * t + k] = 0.5f * (lo + hi);
                }
                prims[t] = uint32_t(t);
            }
        };
        run_parallel(ntri, threads, work);
        nodes.resize(std::max<size_t>(1, 2 * size_t(ntri)));
    }

    template <class F>
    static void run_parallel(size_t n, unsigned threads, F&& f) {
        if (threads <= 1 || n < 65536) { f(0, n); return; }
        std::vector<std::thread> pool;
        size_t chunk = (n + threads - 1) / threads;
        for (unsigned i = 0; i < threads; ++i) {
            size_t b = i * chunk, e = std::min(n, b + chunk);
            if (b >= e) break;
            pool.emplace_back([&f, b, e] { f(b, e); });
        }
        for (auto& t : pool) t.join();
    }

    void process(const Task& task, std::vector<Task>& out) {
        TmpNode& nd = nodes[task.node];
        const uint32_t n = task.end - task.begin;
        Box box, cbox;
        for (uint32_t i = task.begin; i < task.end; ++i) {
            uint32_t p = prims[i];
            for (int k = 0; k < 3; ++k) {
                box.lo[k] = std::min(box.lo[k], tlo[3 * size_t(p) + k]);
                box.hi[k] = std::max(box.hi[k], thi[3 * size_t(p) + k]);
                cbox.lo[k] = std::min(cbox.lo[k], cen[3 * size_t(p) + k]);
                cbox.hi[k] = std::max(cbox.hi[k], cen[3 * size_t(p) + k]);
            }
        }
        nd.box = box;
        constexpr int B = kSahBins;
        int best_axis = -1, best_split = -1;
        double best_cost = INFINITY;
        float scale[3] = {0, 0, 0};
        if (n > 1) {
            Box bins[3][B];
            uint32_t cnt[3][B] = {};
            for (int k = 0; k < 3; ++k) {
                float ext = cbox.hi[k] - cbox.lo[k];
                scale[k] = ext > 0.f ? float(B) / ext : 0.f;
            }
            for (uint32_t i = task.begin; i < task.end; ++i) {
                uint32_t p = prims[i];
                Box tb;
                for (int k = 0; k < 3; ++k) { tb.lo[k] = tlo[3 * size_t(p) + k]; tb.hi[k] = thi[3 * size_t(p) + k]; }
                for (int k = 0; k < 3; ++k) {
                    if (scale[k] == 0.f) continue;
                    int b = int((cen[3 * size_t(p) + k] - cbox.lo[k]) * scale[k]);
                    b = std::min(std::max(b, 0), B - 1);
                    cnt[k][b]++;
                    bins[k][b].grow(tb);
                }
            }
            const double inv_parent = 1.0 / std::max(box.area(), 1e-30);
            for (int k = 0; k < 3; ++k) {
                if (scale[k] == 0.f) continue;
                double rarea[B];
                uint32_t rcnt[B];
                Box acc;
                uint32_t run = 0;
                for (int b = B - 1; b >= 1; --b) {
                    acc.grow(bins[k][b]);
                    run += cnt[k][b];
                    rarea[b] = acc.area();
                    rcnt[b] = run;
                }
                Box lacc;
                uint32_t lrun = 0;
                for (int s = 1; s < B; ++s) {
                    lacc.grow(bins[k][s - 1]);
                    lrun += cnt[k][s - 1];
                    if (lrun == 0 || rcnt[s] == 0) continue;
                    double cost = 1.0 + (lacc.area() * lrun + rarea[s] * rcnt[s]) * inv_parent;
                    if (cost < best_cost) { best_cost = cost; best_axis = k; best_split = s; }
                }
            }
        }
        bool leaf = (n <= 1) || (n <= uint32_t(kLeafMax) && (best_axis < 0 || double(n) <= best_cost));
        if (leaf) {
            nd.first = task.begin;
            nd.count = n;
            return;
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = task.begin + n / 2;
        } else {
            const int k = best_axis;
            const float lo = cbox.lo[k], sc = scale[k];
            auto* m = std::partition(prims.data() + task.begin, prims.data() + task.end, [&](uint32_t p) {
                int b = int((cen[3 * size_t(p) + k] - lo) * sc);
                b = std::min(std::max(b, 0), B - 1);
                return b < best_split;
            });
            mid = uint32_t(m - prims.data());
            if (mid == task.begin || mid == task.end) mid = task.begin + n / 2;
        }
        uint32_t l = next_node.fetch_add(2);
        nd.left = l;
        nd.right = l + 1;
        nd.count = 0;
        out.push_back({l, task.begin, mid});
        out.push_back({l + 1, mid, task.end});
    }

    void worker() {
        std::vector<Task> local_out;
        std::vector<Task> local_stack;
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !queue.empty() || pending == 0; });
                if (queue.empty()) return;
                t = queue.front();
                queue.pop_front();
            }
            // Small subtrees are finished locally without touching the queue.
            local_stack.clear();
            local_stack.push_back(t);
            uint32_t done = 0;
            while (!local_stack.empty()) {
                Task cur = local_stack.back();
                local_stack.pop_back();
                local_out.clear();
                process(cur, local_out);
                done++;
                for (const Task& c : local_out) {
                    if (c.end - c.begin > 32768) {
                        std::lock_guard<std::mutex> lk(mu);
                        queue.push_back(c);
                        pending++;
                        cv.notify_one();
                    } else {
                        local_stack.push_back(c);
                    }
                }
            }
            (void)done;
            {
                std::lock_guard<std::mutex> lk(mu);
                pending--;
                if (pending == 0) cv.notify_all();
            }
        }
    }

    void build(unsigned threads) {
        next_node = 1;
        queue.push_back({0, 0, ntri});
        pending = 1;
        unsigned nth = std::max(1u, threads);
        std::vector<std::thread> pool;
        for (unsigned i = 1; i < nth; ++i) pool.emplace_back([this] { worker(); });
        worker();
        for (auto& t : pool) t.join();
    }
};

}  // namespace

BvhBuildResult build_bvh(const float* p0, const float* p1, const float* p2, uint32_t ntri,
                         unsigned threads) {
    if (ntri == 0) return BvhBuildResult{};
    if (threads == 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    Builder b;
    b.p0 = p0; b.p1 = p1; b.p2 = p2; b.ntri = ntri;
    b.prepare(threads);
    b.build(threads);
    const uint32_t used = b.next_node.load();
    std::vector<BinaryNode> tree(used);
    for (uint32_t i = 0; i < used; ++i) {
        const TmpNode& t = b.nodes[i];
        BinaryNode& o = tree[i];
        for (int k = 0; k < 3; ++k) { o.lo[k] = t.box.lo[k]; o.hi[k] = t.box.hi[k]; }
        o.left = t.left;
        o.right = t.right;
        o.first = t.first;
        o.count = t.count;
    }
    return convert_binary_tree(tree.data(), used, b.prims.data(), ntri);
}

bool validate_binary_tree(const BinaryNode* nodes, uint32_t num_nodes, const uint32_t* prims, uint32_t ntri) {
    if (num_nodes == 0 || ntri == 0) return false;
    std::vector<uint8_t> seen_node(num_nodes, 0), seen_pos(ntri, 0), seen_prim(ntri, 0);
    std::vector<uint32_t> st{0};
    uint32_t visited = 0;
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        if (i >= num_nodes || seen_node[i]) return false;
        seen_node[i] = 1;
        ++visited;
        const BinaryNode& n = nodes[i];
        if (n.count) {
            if (n.count > uint32_t(kLeafMax) || n.first >= ntri || n.count > ntri - n.first) return false;
            for (uint32_t k = n.first; k < n.first + n.count; ++k) {
                if (seen_pos[k]) return false;
                seen_pos[k] = 1;
            }
        } else {
            st.push_back(n.right);
            st.push_back(n.left);
        }
    }
    if (visited != num_nodes) return false;
    for (uint32_t k = 0; k < ntri; ++k) {
        if (!seen_pos[k] || prims[k] >= ntri || seen_prim[prims[k]]) return false;
        seen_prim[prims[k]] = 1;
    }
    return true;
}

namespace {

Box box_of(const BinaryNode& n) {
    Box b;
    for (int k = 0; k < 3; ++k) { b.lo[k] = n.lo[k]; b.hi[k] = n.hi[k]; }
    return b;
}

}  // namespace

BvhBuildResult convert_binary_tree(const BinaryNode* tnodes, uint32_t num_nodes, const uint32_t* prims,
                                   uint32_t ntri) {
    BvhBuildResult res;
    if (ntri == 0 || num_nodes == 0) return res;
    struct Tmp {
        Box box;
        double area;     // surface area of box
        uint32_t left, right, first, count;
    };
    std::vector<Tmp> nodes(num_nodes);
    {
        const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        auto fill = [&](size_t b0, size_t b1) {
            for (size_t i = b0; i < b1; ++i) {
                const Box bx = box_of(tnodes[i]);
                nodes[i] = Tmp{bx, bx.area(), tnodes[i].left, tnodes[i].right, tnodes[i].first, tnodes[i].count};
            }
        };
        if (num_nodes < (1u << 16) || nth == 1) {
            fill(0, num_nodes);
        } else {
            std::vector<std::thread> pool;
            const size_t chunk = (size_t(num_nodes) + nth - 1) / nth;
            for (unsigned t = 0; t < nth; ++t) {
                const size_t b0 = t * chunk, b1 = std::min<size_t>(num_nodes, b0 + chunk);
                if (b0 < b1) pool.emplace_back(fill, b0, b1);
            }
            for (auto& th : pool) th.join();
        }
    }
    struct {
        const std::vector<Tmp>& nodes;
        const uint32_t* prims;
        uint32_t next_count;
    } b{nodes, prims, num_nodes};

    // Depth-first conversion to pair nodes.
    const Tmp& root = b.nodes[0];
    res.order.reserve(ntri);
    std::vector<float>& out = res.nodes;
    // leaf reference of every temporary leaf node, shared by both layouts
    std::vector<int32_t> leaf_ref(b.next_count, 0);
    auto encode_leaf = [&](uint32_t tmp) -> int32_t {
        const Tmp& lf = b.nodes[tmp];
        uint32_t first = uint32_t(res.order.size());
        for (uint32_t i = 0; i < lf.count; ++i) res.order.push_back(b.prims[lf.first + i]);
        res.num_leaves++;
        leaf_ref[tmp] = ~int32_t((first << 2) | (lf.count - 1));
        return leaf_ref[tmp];
    };
    auto write_node = [&](uint32_t idx, const Box& L, const Box& R, int32_t cl, int32_t cr) {
        float* n = &out[16 * size_t(idx)];
        n[0] = L.lo[0]; n[1] = L.hi[0]; n[2] = L.lo[1]; n[3] = L.hi[1];
        n[4] = R.lo[0]; n[5] = R.hi[0]; n[6] = R.lo[1]; n[7] = R.hi[1];
        n[8] = L.lo[2]; n[9] = L.hi[2]; n[10] = R.lo[2]; n[11] = R.hi[2];
        std::memcpy(&n[12], &cl, 4);
        std::memcpy(&n[13], &cr, 4);
        n[14] = 0.f; n[15] = 0.f;
    };
    if (root.count > 0) {
        // The whole scene fits one leaf: split it into two leaves under a
        // single pair node (both children are the same one-triangle leaf when
        // there is only one triangle).  Both children get the root box.
        out.assign(16, 0.f);
        for (uint32_t i = 0; i < root.count; ++i) res.order.push_back(b.prims[root.first + i]);
        const uint32_t n = root.count, half = n > 1 ? n / 2 : 1;
        const int32_t cl = ~int32_t((0u << 2) | (half - 1));
        const int32_t cr = n > 1 ? ~int32_t((half << 2) | (n - half - 1)) : cl;
        write_node(0, root.box, root.box, cl, cr);
        res.num_nodes = 1;
        res.num_leaves = n > 1 ? 2 : 1;
        res.depth = 1;
        res.sah_cost = 1.0 + n;
        res.nodes4.assign(32, 0.f);
        const int32_t refs[4] = {cl, n > 1 ? cr : kEmptyChild, kEmptyChild, kEmptyChild};
        for (int k = 0; k < 4; ++k) {
            const bool used = refs[k] != kEmptyChild;
            for (int a = 0; a < 3; ++a) {
                res.nodes4[8 * a + k] = used ? root.box.lo[a] : kEmptyLo;
                res.nodes4[8 * a + 4 + k] = used ? root.box.hi[a] : kEmptyHi;
            }
            std::memcpy(&res.nodes4[24 + k], &refs[k], 4);
        }
        res.num_nodes4 = 1;
        res.depth4 = 1;
        return res;
    }
    // Count internal nodes for allocation.
    uint32_t internal = 0;
    {
        std::vector<uint32_t> st{0};
        while (!st.empty()) {
            uint32_t i = st.back(); st.pop_back();
            if (b.nodes[i].count) continue;
            internal++;
            st.push_back(b.nodes[i].right);
            st.push_back(b.nodes[i].left);
        }
    }
    out.assign(16 * size_t(internal), 0.f);
    // Preorder walk, left subtree first: a node's output index is assigned when
    // it is popped, and the parent's child slot is patched at that moment.
    struct Walk { uint32_t tmp; int64_t parent; uint32_t side; uint32_t depth; };
    uint32_t next_out = 0;
    double root_area = std::max(root.area, 1e-30);
    double cost = 0;
    std::vector<Walk> st{{0, -1, 0, 1}};
    while (!st.empty()) {
        Walk it = st.back(); st.pop_back();
        const uint32_t idx = next_out++;
        if (it.parent >= 0) {
            int32_t c = int32_t(idx);
            std::memcpy(&out[16 * size_t(it.parent) + 12 + it.side], &c, 4);
        }
        const Tmp& nd = b.nodes[it.tmp];
        res.depth = std::max(res.depth, it.depth);
        cost += nd.area / root_area;
        const Tmp& L = b.nodes[nd.left];
        const Tmp& R = b.nodes[nd.right];
        int32_t cl = 0, cr = 0;
        if (L.count) { cl = encode_leaf(nd.left); cost += L.area / root_area * L.count; }
        if (R.count) { cr = encode_leaf(nd.right); cost += R.area / root_area * R.count; }
        write_node(idx, L.box, R.box, cl, cr);
        if (!R.count) st.push_back({nd.right, int64_t(idx), 1, it.depth + 1});
        if (!L.count) st.push_back({nd.left, int64_t(idx), 0, it.depth + 1});
    }
    res.num_nodes = internal;
    res.sah_cost = cost;

    // BVH4: collapse the binary tree.  Each 4-wide node takes its binary
    // node's two children and repeatedly opens the internal child with the
    // largest box until it has four children (or only leaves remain).  Leaves
    // keep the references assigned above, so both layouts share `order`.
    struct Walk4 { uint32_t tmp; int64_t parent; uint32_t slot; uint32_t depth; };
    std::vector<Walk4> st4{{0, -1, 0, 1}};
    std::vector<float>& o4 = res.nodes4;
    o4.reserve(size_t(internal) * 32 / 2);
    uint32_t next4 = 0;
    while (!st4.empty()) {
        Walk4 it = st4.back(); st4.pop_back();
        const uint32_t idx = next4++;
        o4.resize(32 * size_t(next4), 0.f);
        if (it.parent >= 0) {
            int32_t c = int32_t(idx);
            std::memcpy(&o4[32 * size_t(it.parent) + 24 + it.slot], &c, 4);
        }
        res.depth4 = std::max(res.depth4, it.depth);
        uint32_t kids[4] = {b.nodes[it.tmp].left, b.nodes[it.tmp].right, 0, 0};
        uint32_t nk = 2;
        while (nk < 4) {
            int pick = -1;
            double best_area = -1.0;
            for (uint32_t k = 0; k < nk; ++k) {
                const Tmp& c = b.nodes[kids[k]];
                if (c.count == 0 && c.area > best_area) { best_area = c.area; pick = int(k); }
            }
            if (pick < 0) break;
            const Tmp& c = b.nodes[kids[pick]];
            for (uint32_t k = nk; k > uint32_t(pick) + 1; --k) kids[k] = kids[k - 1];
            kids[pick + 1] = c.right;
            kids[pick] = c.left;
            ++nk;
        }
        float* n = &o4[32 * size_t(idx)];
        for (uint32_t k = 0; k < 4; ++k) {
            int32_t ref = kEmptyChild;
            if (k < nk) {
                const Tmp& c = b.nodes[kids[k]];
                for (int a = 0; a < 3; ++a) { n[8 * a + k] = c.box.lo[a]; n[8 * a + 4 + k] = c.box.hi[a]; }
                ref = c.count ? leaf_ref[kids[k]] : 0;
            } else {
                for (int a = 0; a < 3; ++a) { n[8 * a + k] = kEmptyLo; n[8 * a + 4 + k] = kEmptyHi; }
            }
            std::memcpy(&n[24 + k], &ref, 4);
        }
        for (int k = int(nk) - 1; k >= 0; --k)
            if (b.nodes[kids[k]].count == 0) st4.push_back({kids[k], int64_t(idx), uint32_t(k), it.depth + 1});
    }
    res.num_nodes4 = next4;
    return res;
}

}  // namespace quasi_hip
