// st_bvh.cpp — host SAH BVH builder producing the 64-byte two-box nodes and
// 48-byte world-space triangles of st_layout.h.
//
// Replaces the Vulkan driver's acceleration-structure build that the reference
// requests through vkCmdBuildAccelerationStructuresKHR
// (src/App/Scene.cpp:17-55, 543-754).  All mesh instances are flattened into one
// world-space triangle set (each triangle keeps its 16|16 instance|primitive
// id), so traversal needs no instance transforms.  Binned SAH (32 bins per axis),
// leaves of at most 4 triangles, subtrees built in parallel; when the remaining
// depth budget gets tight the builder switches to object-median splits, so the
// depth never exceeds 64 (the traversal stack bound).
#include "st_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <future>
#include <limits>
#include <stdexcept>

namespace st {
namespace {

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::numeric_limits<float>::infinity();
            hi[k] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        float d[3];
        for (int k = 0; k < 3; k++) d[k] = std::max(0.0f, hi[k] - lo[k]);
        return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
};

struct Ref {
    Box b;
    float c[3];
    uint32_t tri;
};

struct Builder {
    const float* verts;  // 9 floats per triangle
    std::vector<Ref> refs;
    std::vector<NodeOut> nodes;
    std::atomic<uint32_t> next_node{1};
    std::atomic<int> max_depth{0};

    struct Child {
        Box box;
        int32_t index;
        int32_t count;
    };

    Box range_box(uint32_t b, uint32_t e) const {
        Box r;
        r.reset();
        for (uint32_t i = b; i < e; i++) r.grow(refs[i].b);
        return r;
    }

    // Partition [b, e) and return the split position.
    uint32_t split(uint32_t b, uint32_t e, int depth) {
        const uint32_t n = e - b;
        Box cb;
        cb.reset();
        for (uint32_t i = b; i < e; i++) cb.grow(refs[i].c);
        int axis = 0;
        float ext[3];
        for (int k = 0; k < 3; k++) ext[k] = cb.hi[k] - cb.lo[k];
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        auto median = [&](int ax) {
            uint32_t mid = b + n / 2;
            std::nth_element(refs.begin() + b, refs.begin() + mid, refs.begin() + e,
                             [ax](const Ref& x, const Ref& y) { return x.c[ax] < y.c[ax]; });
            return mid;
        };
        // keep room for a balanced finish: median splits need ceil(log2(n / 4)) more levels
        int need = 0;
        while ((n >> need) > (uint32_t)kBvhMaxLeaf) need++;
        if (ext[axis] <= 0.0f || depth + need + 2 >= kBvhMaxDepth - 4) return median(axis);
        constexpr int B = 32;
        float best = std::numeric_limits<float>::infinity();
        int best_axis = -1, best_bin = -1;
        for (int ax = 0; ax < 3; ax++) {
            if (ext[ax] <= 0.0f) continue;
            Box bb[B];
            uint32_t cnt[B] = {};
            for (auto& x : bb) x.reset();
            const float scale = B / ext[ax];
            for (uint32_t i = b; i < e; i++) {
                int k = std::min(B - 1, (int)((refs[i].c[ax] - cb.lo[ax]) * scale));
                cnt[k]++;
                bb[k].grow(refs[i].b);
            }
            float ra[B];
            uint32_t rc[B];
            Box acc;
            acc.reset();
            uint32_t c = 0;
            for (int k = B - 1; k > 0; k--) {
                acc.grow(bb[k]);
                c += cnt[k];
                ra[k] = acc.area();
                rc[k] = c;
            }
            acc.reset();
            c = 0;
            for (int k = 0; k < B - 1; k++) {
                acc.grow(bb[k]);
                c += cnt[k];
                if (c == 0 || rc[k + 1] == 0) continue;
                float cost = acc.area() * c + ra[k + 1] * rc[k + 1];
                if (cost < best) {
                    best = cost;
                    best_axis = ax;
                    best_bin = k;
                }
            }
        }
        if (best_axis < 0) return median(axis);
        const float scale = B / ext[best_axis];
        const float lo = cb.lo[best_axis];
        auto it = std::partition(refs.begin() + b, refs.begin() + e, [&](const Ref& r) {
            return std::min(B - 1, (int)((r.c[best_axis] - lo) * scale)) <= best_bin;
        });
        uint32_t m = (uint32_t)(it - refs.begin());
        if (m == b || m == e) return median(axis);
        return m;
    }

    Child make_child(uint32_t b, uint32_t e, int depth) {
        Child c;
        c.box = range_box(b, e);
        if (e - b <= (uint32_t)kBvhMaxLeaf) {
            c.index = (int32_t)b;
            c.count = (int32_t)(e - b);
            int d = depth;
            int cur = max_depth.load();
            while (d > cur && !max_depth.compare_exchange_weak(cur, d)) {}
            return c;
        }
        c.index = (int32_t)next_node.fetch_add(1);
        c.count = 0;
        build_node((uint32_t)c.index, b, e, depth);
        return c;
    }

    void build_node(uint32_t ni, uint32_t b, uint32_t e, int depth) {
        uint32_t m = split(b, e, depth);
        Child l, r;
        if (e - b > 65536) {
            auto fut = std::async(std::launch::async, [&] { return make_child(b, m, depth + 1); });
            r = make_child(m, e, depth + 1);
            l = fut.get();
        } else {
            l = make_child(b, m, depth + 1);
            r = make_child(m, e, depth + 1);
        }
        write_node(ni, l, r);
    }

    static float down(float x) { return std::nextafter(x, -std::numeric_limits<float>::infinity()); }
    static float up(float x) { return std::nextafter(x, std::numeric_limits<float>::infinity()); }

    void write_node(uint32_t ni, const Child& l, const Child& r) {
        NodeOut& o = nodes[ni];
        const Box& a = l.box;
        const Box& c = r.box;
        float f[12] = {down(a.lo[0]), up(a.hi[0]), down(a.lo[1]), up(a.hi[1]), down(a.lo[2]), up(a.hi[2]),
                       down(c.lo[0]), up(c.hi[0]), down(c.lo[1]), up(c.hi[1]), down(c.lo[2]), up(c.hi[2])};
        std::memcpy(o.f, f, sizeof f);
        o.i[0] = l.index;
        o.i[1] = r.index;
        o.i[2] = l.count;
        o.i[3] = r.count;
    }
};

// ---------------------------------------------------------------- 4-wide collapse
struct Kid {
    float lo[3], hi[3];
    int32_t index, count;  // count > 0 leaf (triangle start = index), 0 inner (binary node), < 0 empty
};

static float area(const Kid& k) {
    float d[3];
    for (int a = 0; a < 3; a++) d[a] = std::max(0.0f, k.hi[a] - k.lo[a]);
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
}

static void binary_kids(const std::vector<NodeOut>& nodes, uint32_t ni, std::vector<Kid>& out) {
    const NodeOut& n = nodes[ni];
    for (int c = 0; c < 2; c++) {
        Kid k;
        for (int a = 0; a < 3; a++) {
            k.lo[a] = n.f[6 * c + 2 * a];
            k.hi[a] = n.f[6 * c + 2 * a + 1];
        }
        k.index = n.i[c];
        k.count = n.i[2 + c];
        if (k.count >= 0) out.push_back(k);
    }
}

// decoded coordinate o + q * 2^(e-127), as the traversal computes it
static float decode(float o, uint32_t q, int e) {
    uint32_t bits = (uint32_t)e << 23;
    float s;
    std::memcpy(&s, &bits, 4);
    const float qs = (float)q * s;  // exact: q < 256, s a power of two
    return o + qs;
}

struct Collapser {
    const std::vector<NodeOut>& bn;
    const std::vector<TriOut>& bt;
    std::vector<Node4Out> nodes;
    std::vector<TriOut> tris;
    int depth = 0;

    void make(std::vector<Kid> kids, uint32_t slot, int d) {
        depth = std::max(depth, d);
        while (kids.size() < 4) {
            int best = -1;
            for (int i = 0; i < (int)kids.size(); i++)
                if (kids[i].count == 0 && (best < 0 || area(kids[i]) > area(kids[best]))) best = i;
            if (best < 0) break;
            std::vector<Kid> sub;
            binary_kids(bn, (uint32_t)kids[best].index, sub);
            kids.erase(kids.begin() + best);
            kids.insert(kids.begin() + best, sub.begin(), sub.end());
        }
        Node4Out n;
        std::memset(&n, 0, sizeof n);
        float L[3], U[3];
        for (int a = 0; a < 3; a++) {
            L[a] = std::numeric_limits<float>::infinity();
            U[a] = -std::numeric_limits<float>::infinity();
            for (const Kid& k : kids) {
                L[a] = std::min(L[a], k.lo[a]);
                U[a] = std::max(U[a], k.hi[a]);
            }
        }
        int e[3];
        for (int a = 0; a < 3; a++) {
            n.o[a] = L[a];
            const double ext = (double)U[a] - (double)L[a];
            int k = ext > 0 ? (int)std::ceil(std::log2(ext / 255.0)) : -100;
            k = std::max(-126, std::min(127, k));
            while (k < 127 && decode(L[a], 255, k + 127) < U[a]) k++;
            e[a] = k + 127;
            if (decode(L[a], 255, e[a]) < U[a]) throw std::runtime_error("BVH quantisation range");
        }
        uint32_t types = 0;
        uint32_t n_inner = 0;
        for (const Kid& k : kids) n_inner += k.count == 0;
        const uint32_t base_inner = (uint32_t)nodes.size();
        nodes.resize(nodes.size() + n_inner);
        const uint32_t base_tri = (uint32_t)tris.size();
        for (int c = 0; c < 4; c++) {
            uint32_t qlo[3] = {255, 255, 255}, qhi[3] = {0, 0, 0}, type = 0;
            if (c < (int)kids.size()) {
                const Kid& k = kids[c];
                for (int a = 0; a < 3; a++) {
                    uint32_t bits = (uint32_t)e[a] << 23;
                    float sc;
                    std::memcpy(&sc, &bits, 4);
                    double ql = std::floor(((double)k.lo[a] - L[a]) / sc), qh = std::ceil(((double)k.hi[a] - L[a]) / sc);
                    qlo[a] = (uint32_t)std::max(0.0, std::min(255.0, ql));
                    qhi[a] = (uint32_t)std::max(0.0, std::min(255.0, qh));
                    while (qlo[a] > 0 && decode(L[a], qlo[a], e[a]) > k.lo[a]) qlo[a]--;
                    while (qhi[a] < 255 && decode(L[a], qhi[a], e[a]) < k.hi[a]) qhi[a]++;
                    if (decode(L[a], qlo[a], e[a]) > k.lo[a] || decode(L[a], qhi[a], e[a]) < k.hi[a])
                        throw std::runtime_error("BVH quantisation not conservative");
                }
                if (k.count == 0) {
                    type = 7;
                } else {
                    type = (uint32_t)k.count;
                    for (int t = 0; t < k.count; t++) tris.push_back(bt[(size_t)k.index + t]);
                }
            }
            for (int a = 0; a < 3; a++) {
                n.q[4 * a + c] = (uint8_t)qlo[a];
                n.q[12 + 4 * a + c] = (uint8_t)qhi[a];
            }
            types |= type << (3 * c);
        }
        n.x0 = (uint32_t)e[0] | ((uint32_t)e[1] << 8) | ((uint32_t)e[2] << 16) | ((types & 0xffu) << 24);
        if (base_inner >= (1u << 28)) throw std::runtime_error("too many BVH nodes");
        n.inner = base_inner | ((types >> 8) << 28);
        n.tri = base_tri;
        nodes[slot] = n;
        uint32_t r = 0;
        for (const Kid& k : kids) {
            if (k.count != 0) continue;
            std::vector<Kid> sub;
            binary_kids(bn, (uint32_t)k.index, sub);
            queue.push_back({std::move(sub), base_inner + r, d + 1});
            r++;
        }
    }

    // Breadth-first: node i's children get the next free slots, so the top
    // levels of the tree occupy the lowest node indices.
    struct Job {
        std::vector<Kid> kids;
        uint32_t slot;
        int depth;
    };
    std::deque<Job> queue;
    void run(std::vector<Kid> root) {
        nodes.resize(1);
        queue.push_back({std::move(root), 0, 1});
        while (!queue.empty()) {
            Job j = std::move(queue.front());
            queue.pop_front();
            make(std::move(j.kids), j.slot, j.depth);
        }
    }
};

}  // namespace

BvhOut build_bvh(const float* verts, const uint32_t* ids, uint32_t n) {
    BvhOut out;
    if (n == 0) {
        out.depth = 0;
        return out;
    }
    Builder bd;
    bd.verts = verts;
    bd.refs.resize(n);
    for (uint32_t t = 0; t < n; t++) {
        Ref& r = bd.refs[t];
        r.b.reset();
        for (int j = 0; j < 3; j++) r.b.grow(verts + 9 * (size_t)t + 3 * j);
        for (int k = 0; k < 3; k++) r.c[k] = 0.5f * (r.b.lo[k] + r.b.hi[k]);
        r.tri = t;
    }
    bd.nodes.resize(std::max<uint32_t>(1, n));
    if (n <= (uint32_t)kBvhMaxLeaf) {
        Builder::Child l, r;
        l.box = bd.range_box(0, n);
        l.index = 0;
        l.count = (int32_t)n;
        for (int k = 0; k < 3; k++) r.box.lo[k] = r.box.hi[k] = 0.0f;
        r.index = 0;
        r.count = -1;  // empty child: never entered
        bd.write_node(0, l, r);
        bd.max_depth = 1;
    } else {
        bd.build_node(0, 0, n, 0);
    }
    uint32_t nn = bd.next_node.load();
    out.nodes.assign(bd.nodes.begin(), bd.nodes.begin() + nn);
    out.tris.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t t = bd.refs[i].tri;
        TriOut& o = out.tris[i];
        const float* v = verts + 9 * (size_t)t;
        o.f[0] = v[0]; o.f[1] = v[1]; o.f[2] = v[2];
        std::memcpy(&o.f[3], &ids[t], 4);
        o.f[4] = v[3]; o.f[5] = v[4]; o.f[6] = v[5]; o.f[7] = 0;
        o.f[8] = v[6]; o.f[9] = v[7]; o.f[10] = v[8]; o.f[11] = 0;
    }
    out.depth = bd.max_depth.load();
    if (out.depth > kBvhMaxDepth) throw std::runtime_error("BVH deeper than the traversal stack");
    Collapser col{out.nodes, out.tris, {}, {}, 0, {}};
    std::vector<Kid> root;
    binary_kids(out.nodes, 0, root);
    col.run(std::move(root));
    out.nodes4 = std::move(col.nodes);
    out.tris4 = std::move(col.tris);
    out.depth4 = col.depth;
    if (out.tris4.size() != out.tris.size()) throw std::runtime_error("4-wide collapse lost triangles");
    return out;
}

}  // namespace st
