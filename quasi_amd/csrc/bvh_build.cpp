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
        bool leaf = (n <= 1) || (n <= uint32_t(kLeafMax) &&
                                (best_axis < 0 || double(n) <= best_cost - (1.0 - double(kSahNodeCost))));
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
                         unsigned threads, uint32_t cut_depth) {
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
    return convert_binary_tree(tree.data(), used, b.prims.data(), ntri, threads, cut_depth);
}

namespace {

// Runs f(b, e) over [0, n) split into `threads` contiguous ranges.
template <class F>
void for_ranges(size_t n, unsigned threads, F&& f) {
    if (threads <= 1 || n < (1u << 16)) { f(size_t(0), n); return; }
    std::vector<std::thread> pool;
    const size_t chunk = (n + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t) {
        const size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) pool.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& th : pool) th.join();
}

// Runs f(k) for k in [0, n) on `threads` threads taking items from a shared counter.
template <class F>
void for_items(size_t n, unsigned threads, F&& f) {
    if (threads <= 1 || n <= 1) { for (size_t k = 0; k < n; ++k) f(k); return; }
    std::atomic<size_t> next{0};
    auto work = [&] { for (size_t k; (k = next.fetch_add(1)) < n;) f(k); };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

unsigned default_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

}  // namespace

bool validate_binary_tree(const BinaryNode* nodes, uint32_t num_nodes, const uint32_t* prims, uint32_t ntri) {
    if (num_nodes == 0 || ntri == 0) return false;
    // Every child has a larger index than its parent and every node but the
    // root has exactly one parent: following parents then strictly decreases
    // the index and ends at the root, so the structure is one tree rooted at
    // node 0 that holds all num_nodes nodes.  All checks are per node, so they
    // run in parallel.
    const unsigned threads = default_threads();
    std::vector<uint8_t> parents(num_nodes, 0), seen_pos(ntri, 0), seen_prim(ntri, 0);
    std::atomic<bool> ok{true};
    auto bump = [](uint8_t* c) { return __atomic_fetch_add(c, uint8_t(1), __ATOMIC_RELAXED) == 0; };
    for_ranges(num_nodes, threads, [&](size_t b, size_t e) {
        bool good = true;
        for (size_t i = b; i < e && good; ++i) {
            const BinaryNode& n = nodes[i];
            if (n.count) {
                if (n.count > uint32_t(kLeafMax) || n.first >= ntri || n.count > ntri - n.first) { good = false; break; }
                for (uint32_t k = n.first; k < n.first + n.count; ++k) good = good && bump(&seen_pos[k]);
            } else {
                if (n.left <= i || n.right <= i || n.left >= num_nodes || n.right >= num_nodes || n.left == n.right) {
                    good = false;
                    break;
                }
                good = bump(&parents[n.left]) && bump(&parents[n.right]);
            }
        }
        if (!good) ok = false;
    });
    if (!ok || parents[0] != 0) return false;
    for_ranges(ntri, threads, [&](size_t b, size_t e) {
        bool good = true;
        for (size_t k = b; k < e && good; ++k) good = prims[k] < ntri && bump(&seen_prim[prims[k]]);
        if (!good) ok = false;
    });
    if (!ok) return false;
    for_ranges(num_nodes, threads, [&](size_t b, size_t e) {
        for (size_t i = std::max<size_t>(b, 1); i < e; ++i)
            if (parents[i] != 1) { ok = false; return; }
    });
    for_ranges(ntri, threads, [&](size_t b, size_t e) {
        for (size_t k = b; k < e; ++k)
            if (!seen_pos[k]) { ok = false; return; }
    });
    return ok;
}

namespace {

Box box_of(const BinaryNode& n) {
    Box b;
    for (int k = 0; k < 3; ++k) { b.lo[k] = n.lo[k]; b.hi[k] = n.hi[k]; }
    return b;
}

// Conversion of a binary build tree to the two traversal layouts.  Both are
// written in depth-first preorder (children left before right; a node's leaf
// children are appended to the triangle order when the node is visited).  The
// walks split at a cut depth: the nodes above it are walked by one thread,
// which gives every subtree below the cut its first output index (from the
// subtree sizes counted beforehand), and the subtrees are then written by all
// threads.  The result is the same as one sequential walk.
struct Converter {
    struct Tmp {
        Box box;
        double area;     // surface area of box
        uint32_t left, right, first, count;
    };
    std::vector<Tmp> nodes;
    const uint32_t* prims;
    uint32_t ntri;
    unsigned threads;
    double root_area;
    BvhBuildResult& res;
    std::vector<int32_t> leaf_ref;     // leaf reference of every leaf node, shared by both layouts

    Converter(BvhBuildResult& r) : res(r) {}

    void write_pair(uint32_t idx, const Box& L, const Box& R, int32_t cl, int32_t cr) {
        float* n = &res.nodes[16 * size_t(idx)];
        n[0] = L.lo[0]; n[1] = L.hi[0]; n[2] = L.lo[1]; n[3] = L.hi[1];
        n[4] = R.lo[0]; n[5] = R.hi[0]; n[6] = R.lo[1]; n[7] = R.hi[1];
        n[8] = L.lo[2]; n[9] = L.hi[2]; n[10] = R.lo[2]; n[11] = R.hi[2];
        std::memcpy(&n[12], &cl, 4);
        std::memcpy(&n[13], &cr, 4);
        n[14] = 0.f; n[15] = 0.f;
    }
    void patch(std::vector<float>& out, size_t stride, int64_t parent, uint32_t slot, size_t offset, uint32_t idx) {
        if (parent < 0) return;
        const int32_t c = int32_t(idx);
        std::memcpy(&out[stride * size_t(parent) + offset + slot], &c, 4);
    }

    // ---- pair nodes
    struct Count2 { uint32_t internal = 0, prims = 0, depth = 0; double cost = 0; };
    // Internal nodes, leaf triangles, depth (root = 1) and SAH cost of a subtree, in preorder.
    Count2 count2(uint32_t root) const {
        Count2 c;
        std::vector<std::pair<uint32_t, uint32_t>> st{{root, 1}};
        while (!st.empty()) {
            const auto [i, d] = st.back();
            st.pop_back();
            const Tmp& nd = nodes[i];
            ++c.internal;
            c.depth = std::max(c.depth, d);
            c.cost += nd.area / root_area;
            const Tmp& L = nodes[nd.left];
            const Tmp& R = nodes[nd.right];
            if (L.count) { c.prims += L.count; c.cost += L.area / root_area * L.count; }
            if (R.count) { c.prims += R.count; c.cost += R.area / root_area * R.count; }
            if (!R.count) st.push_back({nd.right, d + 1});
            if (!L.count) st.push_back({nd.left, d + 1});
        }
        return c;
    }
    // Writes the subtree of internal node `root` from output index out0 and
    // triangle-order position pos0.  Returns the number of leaves written.
    uint32_t write2(uint32_t root, uint32_t out0, uint32_t pos0) {
        struct Walk { uint32_t tmp; int64_t parent; uint32_t side; };
        uint32_t next_out = out0, pos = pos0, leaves = 0;
        std::vector<Walk> st{{root, -1, 0}};
        while (!st.empty()) {
            const Walk it = st.back();
            st.pop_back();
            const uint32_t idx = next_out++;
            patch(res.nodes, 16, it.parent, it.side, 12, idx);
            const Tmp& nd = nodes[it.tmp];
            int32_t cref[2] = {0, 0};
            for (int side = 0; side < 2; ++side) {
                const uint32_t ch = side ? nd.right : nd.left;
                const Tmp& c = nodes[ch];
                if (!c.count) continue;
                for (uint32_t k = 0; k < c.count; ++k) res.order[pos + k] = prims[c.first + k];
                leaf_ref[ch] = ~int32_t((pos << 2) | (c.count - 1));
                cref[side] = leaf_ref[ch];
                pos += c.count;
                ++leaves;
            }
            write_pair(idx, nodes[nd.left].box, nodes[nd.right].box, cref[0], cref[1]);
            if (!nodes[nd.right].count) st.push_back({nd.right, int64_t(idx), 1});
            if (!nodes[nd.left].count) st.push_back({nd.left, int64_t(idx), 0});
        }
        return leaves;
    }

    // ---- 4-wide nodes
    // The binary node's 4-wide children: its two children, then repeatedly the
    // internal child of largest area replaced by its two children.
    uint32_t kids4(uint32_t x, uint32_t (&kids)[4]) const {
        kids[0] = nodes[x].left;
        kids[1] = nodes[x].right;
        uint32_t nk = 2;
        while (nk < 4) {
            int pick = -1;
            double best_area = -1.0;
            for (uint32_t k = 0; k < nk; ++k) {
                const Tmp& c = nodes[kids[k]];
                if (c.count == 0 && c.area > best_area) { best_area = c.area; pick = int(k); }
            }
            if (pick < 0) break;
            const Tmp& c = nodes[kids[pick]];
            for (uint32_t k = nk; k > uint32_t(pick) + 1; --k) kids[k] = kids[k - 1];
            kids[pick + 1] = c.right;
            kids[pick] = c.left;
            ++nk;
        }
        return nk;
    }
    void write4(uint32_t idx, const uint32_t (&kids)[4], uint32_t nk) {
        float* n = &res.nodes4[32 * size_t(idx)];
        for (uint32_t k = 0; k < 4; ++k) {
            int32_t ref = kEmptyChild;
            if (k < nk) {
                const Tmp& c = nodes[kids[k]];
                for (int a = 0; a < 3; ++a) { n[8 * a + k] = c.box.lo[a]; n[8 * a + 4 + k] = c.box.hi[a]; }
                ref = c.count ? leaf_ref[kids[k]] : 0;
            } else {
                for (int a = 0; a < 3; ++a) { n[8 * a + k] = kEmptyLo; n[8 * a + 4 + k] = kEmptyHi; }
            }
            std::memcpy(&n[24 + k], &ref, 4);
        }
    }
    // 4-wide nodes and depth (root = 1) of the collapse rooted at binary node x.
    std::pair<uint32_t, uint32_t> count4(uint32_t x) const {
        uint32_t count = 0, depth = 0;
        std::vector<std::pair<uint32_t, uint32_t>> st{{x, 1}};
        while (!st.empty()) {
            const auto [i, d] = st.back();
            st.pop_back();
            ++count;
            depth = std::max(depth, d);
            uint32_t kids[4];
            const uint32_t nk = kids4(i, kids);
            for (int k = int(nk) - 1; k >= 0; --k)
                if (nodes[kids[k]].count == 0) st.push_back({kids[k], d + 1});
        }
        return {count, depth};
    }
    void write4_tree(uint32_t x, uint32_t out0) {
        struct Walk4 { uint32_t tmp; int64_t parent; uint32_t slot; };
        uint32_t next4 = out0;
        std::vector<Walk4> st{{x, -1, 0}};
        while (!st.empty()) {
            const Walk4 it = st.back();
            st.pop_back();
            const uint32_t idx = next4++;
            patch(res.nodes4, 32, it.parent, it.slot, 24, idx);
            uint32_t kids[4];
            const uint32_t nk = kids4(it.tmp, kids);
            write4(idx, kids, nk);
            for (int k = int(nk) - 1; k >= 0; --k)
                if (nodes[kids[k]].count == 0) st.push_back({kids[k], int64_t(idx), uint32_t(k)});
        }
    }

    void run(uint32_t cut) {
        // ---- pair nodes: the walk above the cut
        struct Walk { uint32_t tmp; int64_t parent; uint32_t side; uint32_t depth; };
        struct Sub { uint32_t tmp, depth; int64_t parent; uint32_t side; Count2 c; uint32_t out0 = 0, pos0 = 0; };
        std::vector<Sub> subs;
        {
            std::vector<Walk> st{{0, -1, 0, 1}};
            while (!st.empty()) {
                const Walk it = st.back();
                st.pop_back();
                if (it.depth > cut) { subs.push_back({it.tmp, it.depth, it.parent, it.side, {}}); continue; }
                const Tmp& nd = nodes[it.tmp];
                if (!nodes[nd.right].count) st.push_back({nd.right, 0, 1, it.depth + 1});
                if (!nodes[nd.left].count) st.push_back({nd.left, 0, 0, it.depth + 1});
            }
        }
        for_items(subs.size(), threads, [&](size_t k) { subs[k].c = count2(subs[k].tmp); });
        // number the nodes above the cut and give every subtree its first
        // output index and triangle-order position, in preorder
        uint32_t next_out = 0, pos = 0, sub_k = 0;
        double cost = 0;
        uint32_t total_internal = 0;
        {
            std::vector<Walk> st{{0, -1, 0, 1}};
            uint32_t cap_nodes = 0;
            while (!st.empty()) {
                const Walk it = st.back();
                st.pop_back();
                if (it.depth > cut) continue;
                ++cap_nodes;
                const Tmp& nd = nodes[it.tmp];
                if (!nodes[nd.right].count) st.push_back({nd.right, 0, 1, it.depth + 1});
                if (!nodes[nd.left].count) st.push_back({nd.left, 0, 0, it.depth + 1});
            }
            total_internal = cap_nodes;
            for (const Sub& sb : subs) total_internal += sb.c.internal;
        }
        res.nodes.assign(16 * size_t(total_internal), 0.f);
        res.order.assign(ntri, 0u);
        {
            std::vector<Walk> st{{0, -1, 0, 1}};
            while (!st.empty()) {
                const Walk it = st.back();
                st.pop_back();
                if (it.depth > cut) {
                    Sub& sb = subs[sub_k++];
                    sb.out0 = next_out;
                    sb.pos0 = pos;
                    patch(res.nodes, 16, it.parent, it.side, 12, next_out);
                    next_out += sb.c.internal;
                    pos += sb.c.prims;
                    cost += sb.c.cost;
                    res.depth = std::max(res.depth, it.depth - 1 + sb.c.depth);
                    continue;
                }
                const uint32_t idx = next_out++;
                patch(res.nodes, 16, it.parent, it.side, 12, idx);
                const Tmp& nd = nodes[it.tmp];
                res.depth = std::max(res.depth, it.depth);
                cost += nd.area / root_area;
                int32_t cref[2] = {0, 0};
                for (int side = 0; side < 2; ++side) {
                    const uint32_t ch = side ? nd.right : nd.left;
                    const Tmp& c = nodes[ch];
                    if (!c.count) continue;
                    for (uint32_t k = 0; k < c.count; ++k) res.order[pos + k] = prims[c.first + k];
                    leaf_ref[ch] = ~int32_t((pos << 2) | (c.count - 1));
                    cref[side] = leaf_ref[ch];
                    pos += c.count;
                    cost += c.area / root_area * c.count;
                    res.num_leaves++;
                }
                write_pair(idx, nodes[nd.left].box, nodes[nd.right].box, cref[0], cref[1]);
                if (!nodes[nd.right].count) st.push_back({nd.right, int64_t(idx), 1, it.depth + 1});
                if (!nodes[nd.left].count) st.push_back({nd.left, int64_t(idx), 0, it.depth + 1});
            }
        }
        std::atomic<uint32_t> sub_leaves{0};
        for_items(subs.size(), threads, [&](size_t k) { sub_leaves += write2(subs[k].tmp, subs[k].out0, subs[k].pos0); });
        res.num_leaves += sub_leaves.load();
        res.num_nodes = total_internal;
        res.sah_cost = cost;

        // ---- 4-wide nodes: the same split, at 4-wide depth cut4 (a 4-wide
        // level spans about two binary levels)
        const uint32_t cut4 = cut == ~0u ? ~0u : (cut + 1) / 2;
        struct Walk4 { uint32_t tmp; int64_t parent; uint32_t slot; uint32_t depth; };
        struct Sub4 { uint32_t tmp, depth; uint32_t count = 0, sdepth = 0, out0 = 0; };
        std::vector<Sub4> subs4;
        {
            std::vector<Walk4> st{{0, -1, 0, 1}};
            while (!st.empty()) {
                const Walk4 it = st.back();
                st.pop_back();
                if (it.depth > cut4) { subs4.push_back({it.tmp, it.depth}); continue; }
                uint32_t kids[4];
                const uint32_t nk = kids4(it.tmp, kids);
                for (int k = int(nk) - 1; k >= 0; --k)
                    if (nodes[kids[k]].count == 0) st.push_back({kids[k], 0, 0, it.depth + 1});
            }
        }
        for_items(subs4.size(), threads, [&](size_t k) {
            const auto [c, d] = count4(subs4[k].tmp);
            subs4[k].count = c;
            subs4[k].sdepth = d;
        });
        uint32_t total4 = 0;
        {
            std::vector<Walk4> st{{0, -1, 0, 1}};
            while (!st.empty()) {
                const Walk4 it = st.back();
                st.pop_back();
                if (it.depth > cut4) continue;
                ++total4;
                uint32_t kids[4];
                const uint32_t nk = kids4(it.tmp, kids);
                for (int k = int(nk) - 1; k >= 0; --k)
                    if (nodes[kids[k]].count == 0) st.push_back({kids[k], 0, 0, it.depth + 1});
            }
            for (const Sub4& sb : subs4) total4 += sb.count;
        }
        res.nodes4.assign(32 * size_t(total4), 0.f);
        {
            uint32_t next4 = 0, k4 = 0;
            std::vector<Walk4> st{{0, -1, 0, 1}};
            while (!st.empty()) {
                const Walk4 it = st.back();
                st.pop_back();
                if (it.depth > cut4) {
                    Sub4& sb = subs4[k4++];
                    sb.out0 = next4;
                    patch(res.nodes4, 32, it.parent, it.slot, 24, next4);
                    next4 += sb.count;
                    res.depth4 = std::max(res.depth4, it.depth - 1 + sb.sdepth);
                    continue;
                }
                const uint32_t idx = next4++;
                patch(res.nodes4, 32, it.parent, it.slot, 24, idx);
                res.depth4 = std::max(res.depth4, it.depth);
                uint32_t kids[4];
                const uint32_t nk = kids4(it.tmp, kids);
                write4(idx, kids, nk);
                for (int k = int(nk) - 1; k >= 0; --k)
                    if (nodes[kids[k]].count == 0) st.push_back({kids[k], int64_t(idx), uint32_t(k), it.depth + 1});
            }
        }
        for_items(subs4.size(), threads, [&](size_t k) { write4_tree(subs4[k].tmp, subs4[k].out0); });
        res.num_nodes4 = total4;
    }
};

}  // namespace

BvhBuildResult convert_binary_tree(const BinaryNode* tnodes, uint32_t num_nodes, const uint32_t* prims,
                                   uint32_t ntri, unsigned threads, uint32_t cut_depth) {
    BvhBuildResult res;
    if (ntri == 0 || num_nodes == 0) return res;
    if (threads == 0) threads = default_threads();
    Converter cv(res);
    cv.nodes.resize(num_nodes);
    for_ranges(num_nodes, threads, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const Box bx = box_of(tnodes[i]);
            cv.nodes[i] = Converter::Tmp{bx, bx.area(), tnodes[i].left, tnodes[i].right, tnodes[i].first,
                                         tnodes[i].count};
        }
    });
    cv.prims = prims;
    cv.ntri = ntri;
    cv.threads = threads;
    const Converter::Tmp& root = cv.nodes[0];
    if (root.count > 0) {
        // The whole scene fits one leaf: split it into two leaves under a
        // single pair node (both children are the same one-triangle leaf when
        // there is only one triangle).  Both children get the root box.
        res.nodes.assign(16, 0.f);
        for (uint32_t i = 0; i < root.count; ++i) res.order.push_back(prims[root.first + i]);
        const uint32_t n = root.count, half = n > 1 ? n / 2 : 1;
        const int32_t cl = ~int32_t((0u << 2) | (half - 1));
        const int32_t cr = n > 1 ? ~int32_t((half << 2) | (n - half - 1)) : cl;
        cv.write_pair(0, root.box, root.box, cl, cr);
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
    cv.root_area = std::max(root.area, 1e-30);
    cv.leaf_ref.assign(num_nodes, 0);
    // about 4096 subtrees below the cut when the tree is large enough to pay
    // for the threads; otherwise one walk
    if (cut_depth == 0) cut_depth = (threads > 1 && num_nodes >= (1u << 18)) ? 12u : ~0u;
    cv.run(cut_depth);
    return res;
}

}  // namespace quasi_hip
