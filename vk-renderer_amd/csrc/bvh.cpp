// bvh.cpp - binned-SAH BVH builder (host, C++17, threads for big subtrees).
#include "bvh.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace vkr {
namespace {

struct Prim {
    float lo[3], hi[3], c[3];
    uint32_t id;
    uint32_t sphere;
};

struct Sub {
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> tri_ids;
    std::vector<uint32_t> sph_ids;
    uint32_t leaves = 0;
    uint32_t depth = 0;
};

constexpr float kEmpty = 1e30f;
constexpr int kBins = 16;
constexpr size_t kParallelMin = 1u << 15;

struct Box {
    float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                   std::numeric_limits<float>::infinity()};
    float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                   -std::numeric_limits<float>::infinity()};
    void grow(const float* l, const float* h)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    void grow_pt(const float* p) { grow(p, p); }
    float area() const
    {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

inline uint32_t make_leaf(uint32_t first, uint32_t count, bool sphere)
{
    return kRefLeaf | (sphere ? kRefSphere : 0u) | ((count - 1u) << kRefCountShift) | (first & kRefFirstMask);
}

// Centre and half-extent of [lo, hi] in float, with the extent rounded up so
// that [c - e, c + e] contains [lo, hi] exactly.
inline void centre_extent(float lo, float hi, float& c, float& e)
{
    const double cd = 0.5 * ((double)lo + (double)hi);
    c = (float)cd;
    const double ed = 0.5 * ((double)hi - (double)lo) + std::fabs(cd - (double)c);
    e = (float)ed;
    if ((double)e < ed) e = std::nextafter(e, std::numeric_limits<float>::infinity());
}

inline void set_child(BvhNode& n, int slot, const Box& b, uint32_t ref)
{
    float c[3], e[3];
    for (int k = 0; k < 3; ++k) centre_extent(b.lo[k], b.hi[k], c[k], e[k]);
    if (slot == 0) {
        n.a[0] = c[0]; n.a[1] = e[0]; n.a[2] = c[1]; n.a[3] = e[1];
        n.b[0] = c[2]; n.b[1] = e[2];
    } else {
        n.b[2] = c[0]; n.b[3] = e[0];
        n.c[0] = c[1]; n.c[1] = e[1]; n.c[2] = c[2]; n.c[3] = e[2];
    }
    n.d[slot] = ref;
}

// A child box every ray enters (centre 0, infinite extent).
inline void set_child_always(BvhNode& n, int slot, uint32_t ref)
{
    const float inf = std::numeric_limits<float>::infinity();
    if (slot == 0) {
        n.a[0] = 0.0f; n.a[1] = inf; n.a[2] = 0.0f; n.a[3] = inf;
        n.b[0] = 0.0f; n.b[1] = inf;
    } else {
        n.b[2] = 0.0f; n.b[3] = inf;
        n.c[0] = 0.0f; n.c[1] = inf; n.c[2] = 0.0f; n.c[3] = inf;
    }
    n.d[slot] = ref;
}

// A child box no ray enters (the point (1e30, 1e30, 1e30)).
inline void set_child_never(BvhNode& n, int slot, uint32_t ref)
{
    if (slot == 0) {
        n.a[0] = kEmpty; n.a[1] = 0.0f; n.a[2] = kEmpty; n.a[3] = 0.0f;
        n.b[0] = kEmpty; n.b[1] = 0.0f;
    } else {
        n.b[2] = kEmpty; n.b[3] = 0.0f;
        n.c[0] = kEmpty; n.c[1] = 0.0f; n.c[2] = kEmpty; n.c[3] = 0.0f;
    }
    n.d[slot] = ref;
}

struct Builder {
    uint32_t max_leaf;
    uint32_t max_depth;
    float ctrav = 1.0f;
    std::atomic<int> threads_left;

    static bool pure(const Prim* p, size_t n)
    {
        for (size_t i = 1; i < n; ++i)
            if (p[i].sphere != p[0].sphere) return false;
        return true;
    }

    // Appends a leaf to s and returns its ref.
    uint32_t leaf(Sub& s, const Prim* p, size_t n, uint32_t dep)
    {
        const bool sph = p[0].sphere != 0;
        std::vector<uint32_t>& ids = sph ? s.sph_ids : s.tri_ids;
        const uint32_t first = (uint32_t)ids.size();
        for (size_t i = 0; i < n; ++i) ids.push_back(p[i].id);
        s.leaves++;
        s.depth = std::max(s.depth, dep);
        return make_leaf(first, (uint32_t)n, sph);
    }

    // Splits p[0..n) and returns the size of the left part (0 < m < n).
    size_t split(Prim* p, size_t n, uint32_t dep, bool* make_leaf_out)
    {
        *make_leaf_out = false;
        Box cb;
        for (size_t i = 0; i < n; ++i) cb.grow_pt(p[i].c);
        const bool is_pure = pure(p, n);

        // Mixed small ranges: separate the kinds first.
        if (n <= max_leaf && !is_pure) {
            Prim* mid = std::partition(p, p + n, [](const Prim& q) { return q.sphere == 0; });
            return (size_t)(mid - p);
        }

        int axis = 0;
        float ext[3];
        for (int k = 0; k < 3; ++k) ext[k] = cb.hi[k] - cb.lo[k];
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;

        // Depth cap: median splits below a range of k primitives reach
        // leaves of max_leaf after levels(k) more levels (one more for a range
        // that mixes kinds).  SAH splits are taken only when both parts can
        // still finish within the cap that way; a range with no such split,
        // or no room left, is split at the median.
        auto levels = [this](size_t k) -> uint32_t {
            uint32_t l = 0;
            while (((size_t)max_leaf << l) < k) ++l;
            return l;
        };
        const uint32_t mixed = is_pure ? 0u : 1u;
        const bool forced_median = dep + levels(n) + mixed >= max_depth;

        if (!forced_median && ext[axis] > 0.0f) {
            float best_cost = std::numeric_limits<float>::infinity();
            int best_axis = -1, best_bin = -1;
            Box pb;
            for (size_t i = 0; i < n; ++i) pb.grow(p[i].lo, p[i].hi);
            const float parea = std::max(pb.area(), 1e-30f);
            for (int ax = 0; ax < 3; ++ax) {
                if (!(ext[ax] > 0.0f)) continue;
                Box bb[kBins];
                uint32_t bc[kBins] = {0};
                const float scale = (float)kBins / ext[ax];
                for (size_t i = 0; i < n; ++i) {
                    int b = (int)((p[i].c[ax] - cb.lo[ax]) * scale);
                    b = std::min(std::max(b, 0), kBins - 1);
                    bb[b].grow(p[i].lo, p[i].hi);
                    bc[b]++;
                }
                float ra[kBins];
                uint32_t rn[kBins];
                Box acc;
                uint32_t cnt = 0;
                for (int b = kBins - 1; b > 0; --b) {
                    if (bc[b]) acc.grow(bb[b].lo, bb[b].hi);
                    cnt += bc[b];
                    ra[b] = acc.area();
                    rn[b] = cnt;
                }
                Box lacc;
                uint32_t lcnt = 0;
                for (int b = 0; b < kBins - 1; ++b) {
                    if (bc[b]) lacc.grow(bb[b].lo, bb[b].hi);
                    lcnt += bc[b];
                    if (lcnt == 0 || rn[b + 1] == 0) continue;
                    if (dep + 1u + levels(lcnt) > max_depth || dep + 1u + levels(rn[b + 1]) > max_depth) continue;
                    const float cost = ctrav + (lacc.area() * (float)lcnt + ra[b + 1] * (float)rn[b + 1]) / parea;
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = ax;
                        best_bin = b;
                    }
                }
            }
            if (n <= max_leaf && is_pure && (float)n <= best_cost) {
                *make_leaf_out = true;
                return 0;
            }
            if (best_axis >= 0) {
                const float scale = (float)kBins / ext[best_axis];
                const float lo = cb.lo[best_axis];
                Prim* mid = std::partition(p, p + n, [&](const Prim& q) {
                    int b = (int)((q.c[best_axis] - lo) * scale);
                    b = std::min(std::max(b, 0), kBins - 1);
                    return b <= best_bin;
                });
                const size_t m = (size_t)(mid - p);
                if (m > 0 && m < n) return m;
            }
        } else if (n <= max_leaf && is_pure) {
            *make_leaf_out = true;
            return 0;
        }
        // Median split on the widest centroid axis (also the fallback for
        // coincident centroids).
        const size_t m = n / 2;
        std::nth_element(p, p + m, p + n, [axis](const Prim& x, const Prim& y) {
            if (x.c[axis] != y.c[axis]) return x.c[axis] < y.c[axis];
            if (x.sphere != y.sphere) return x.sphere < y.sphere;
            return x.id < y.id;
        });
        return m;
    }

    // Builds the range into s; returns the ref (leaf or node index in s) and
    // the range's bounds.
    uint32_t build(Sub& s, Prim* p, size_t n, uint32_t dep, Box* out_box)
    {
        Box box;
        for (size_t i = 0; i < n; ++i) box.grow(p[i].lo, p[i].hi);
        *out_box = box;
        bool as_leaf = false;
        const size_t m = split(p, n, dep, &as_leaf);
        if (as_leaf) return leaf(s, p, n, dep);

        const uint32_t idx = (uint32_t)s.nodes.size();
        s.nodes.push_back(BvhNode{});
        Box lb, rb;
        uint32_t lref, rref;
        if (n >= kParallelMin && threads_left.fetch_sub(1) > 0) {
            Sub rs;
            auto fut = std::async(std::launch::async, [&]() { return build(rs, p + m, n - m, dep + 1, &rb); });
            lref = build(s, p, m, dep + 1, &lb);
            const uint32_t rlocal = fut.get();
            rref = append(s, rs, rlocal);
        } else {
            lref = build(s, p, m, dep + 1, &lb);
            rref = build(s, p + m, n - m, dep + 1, &rb);
        }
        set_child(s.nodes[idx], 0, lb, lref);
        set_child(s.nodes[idx], 1, rb, rref);
        return idx;
    }

    static uint32_t fix(uint32_t ref, uint32_t node_off, uint32_t tri_off, uint32_t sph_off)
    {
        if (!(ref & kRefLeaf)) return ref + node_off;
        const uint32_t first = (ref & kRefFirstMask) + ((ref & kRefSphere) ? sph_off : tri_off);
        return (ref & ~kRefFirstMask) | first;
    }

    // Appends subtree r to s, relocating its refs; returns the relocated ref.
    static uint32_t append(Sub& s, const Sub& r, uint32_t rref)
    {
        const uint32_t node_off = (uint32_t)s.nodes.size();
        const uint32_t tri_off = (uint32_t)s.tri_ids.size();
        const uint32_t sph_off = (uint32_t)s.sph_ids.size();
        for (BvhNode nd : r.nodes) {
            nd.d[0] = fix(nd.d[0], node_off, tri_off, sph_off);
            nd.d[1] = fix(nd.d[1], node_off, tri_off, sph_off);
            s.nodes.push_back(nd);
        }
        s.tri_ids.insert(s.tri_ids.end(), r.tri_ids.begin(), r.tri_ids.end());
        s.sph_ids.insert(s.sph_ids.end(), r.sph_ids.begin(), r.sph_ids.end());
        s.leaves += r.leaves;
        s.depth = std::max(s.depth, r.depth);
        return fix(rref, node_off, tri_off, sph_off);
    }
};

inline bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// Conservative padding: the float slab test and the float primitive tests
// must never disagree about a primitive that lies on a box face.
inline void pad(float* lo, float* hi)
{
    float m = 0.0f;
    for (int k = 0; k < 3; ++k) m = std::max(m, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
    for (int k = 0; k < 3; ++k) {
        const float e = 1e-5f * (m + (hi[k] - lo[k])) + 1e-6f;
        lo[k] -= e;
        hi[k] += e;
    }
}

}  // namespace

BvhNode join_root(const float lo[3], const float hi[3], uint32_t tri_ref, uint32_t sph_ref)
{
    BvhNode n{};
    if (lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]) {
        Box b;
        b.grow(lo, hi);
        set_child(n, 0, b, tri_ref);
    } else {
        set_child_never(n, 0, tri_ref);
    }
    set_child_always(n, 1, sph_ref);
    return n;
}

void offset_inner_refs(std::vector<BvhNode>& nodes, uint32_t off)
{
    for (BvhNode& n : nodes)
        for (int s = 0; s < 2; ++s)
            if (!(n.d[s] & kRefLeaf)) n.d[s] += off;
}

BuildOutput build_bvh(const BuildInput& in)
{
    BuildOutput out;
    std::vector<Prim> prims;
    prims.reserve((size_t)in.n_tris + in.n_spheres);
    for (uint32_t i = 0; i < in.n_tris; ++i) {
        const float* t = in.tri_v + (size_t)i * 12;
        const float* v[3] = {t, t + 4, t + 8};
        if (!finite3(v[0]) || !finite3(v[1]) || !finite3(v[2])) continue;  // never hit (NaN tests fail)
        Prim p;
        for (int k = 0; k < 3; ++k) {
            p.lo[k] = std::min(v[0][k], std::min(v[1][k], v[2][k]));
            p.hi[k] = std::max(v[0][k], std::max(v[1][k], v[2][k]));
        }
        pad(p.lo, p.hi);
        for (int k = 0; k < 3; ++k) p.c[k] = 0.5f * (p.lo[k] + p.hi[k]);
        p.id = i;
        p.sphere = 0;
        prims.push_back(p);
    }
    for (uint32_t i = 0; i < in.n_spheres; ++i) {
        const float* s = in.sph + (size_t)i * 16;
        const float c[3] = {s[12], s[13], s[14]};
        const float r = std::fabs(s[15]);
        if (!finite3(c) || !std::isfinite(r)) continue;
        Prim p;
        for (int k = 0; k < 3; ++k) {
            p.lo[k] = c[k] - r;
            p.hi[k] = c[k] + r;
        }
        pad(p.lo, p.hi);
        for (int k = 0; k < 3; ++k) p.c[k] = c[k];
        p.id = i;
        p.sphere = 1;
        prims.push_back(p);
    }
    if (prims.empty()) return out;

    Builder b;
    b.max_leaf = std::max(1u, std::min(in.max_leaf ? in.max_leaf : kDefaultLeaf, kMaxLeaf));
    b.max_depth = std::max(8u, std::min(in.max_depth ? in.max_depth : kDefaultMaxDepth, 63u));
    // A node visit (four 128-bit loads, two slab tests, a stack step) costs
    // about three triangle tests on gfx950 (measured on the 100k soup).
    b.ctrav = in.traversal_cost > 0.0f ? in.traversal_cost : 3.0f;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    b.threads_left = (int)std::min(hw, 16u) - 1;

    // Triangles first, spheres after; each kind gets its own subtree.
    const auto mid = std::stable_partition(prims.begin(), prims.end(), [](const Prim& q) { return q.sphere == 0; });
    const size_t n_tri = (size_t)(mid - prims.begin()), n_sph = prims.size() - n_tri;
    if (n_sph) {
        double lo[3], hi[3];
        for (int k = 0; k < 3; ++k) { lo[k] = 1e300; hi[k] = -1e300; }
        for (size_t i = n_tri; i < prims.size(); ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], (double)prims[i].c[k]);
                hi[k] = std::max(hi[k], (double)prims[i].c[k]);
            }
        double r2 = 0.0;
        for (int k = 0; k < 3; ++k) out.sph_c[k] = (float)(0.5 * (lo[k] + hi[k]));
        for (size_t i = n_tri; i < prims.size(); ++i) {
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) d2 += ((double)prims[i].c[k] - out.sph_c[k]) * ((double)prims[i].c[k] - out.sph_c[k]);
            r2 = std::max(r2, d2);
        }
        out.sph_r = (float)(std::sqrt(r2) * (1.0 + 1e-6) + 1e-6);
    }

    Sub s;
    Box rootbox;
    uint32_t ref;
    if (n_tri && n_sph) {
        s.nodes.push_back(BvhNode{});   // root, filled below
        Box tb, sb;
        const uint32_t tref = b.build(s, prims.data(), n_tri, 1, &tb);
        Sub ss;
        const uint32_t sref_local = b.build(ss, prims.data() + n_tri, n_sph, 1, &sb);
        out.sph_node0 = (uint32_t)s.nodes.size();
        const uint32_t sref = Builder::append(s, ss, sref_local);
        set_child(s.nodes[0], 0, tb, tref);
        set_child_always(s.nodes[0], 1, sref);
        s.depth = std::max(s.depth, 1u);
        ref = 0;
    } else {
        ref = b.build(s, prims.data(), prims.size(), 0, &rootbox);
        if (n_sph) out.sph_node0 = 0;
    }
    if (ref & kRefLeaf) {
        // Tiny scene: wrap the single leaf in a root node (its leaf sits at depth 1).
        BvhNode root{};
        set_child(root, 0, rootbox, ref);
        // The empty slot's box can never be hit (a point at 1e30); it still
        // names the real leaf so no ref ever points outside the records.
        set_child_never(root, 1, ref);
        s.nodes.insert(s.nodes.begin(), root);
        s.depth = std::max(s.depth, 1u);
    }
    if (out.sph_node0 != 0xFFFFFFFFu) {
        const float one = 1.0f;
        uint32_t m;
        std::memcpy(&m, &one, sizeof m);
        for (size_t i = out.sph_node0; i < s.nodes.size(); ++i) s.nodes[i].d[2] = m;
    }
    out.nodes = std::move(s.nodes);
    out.leaves = s.leaves;
    out.depth = s.depth;
    out.empty = false;

    out.tris.resize(s.tri_ids.size());
    for (size_t j = 0; j < s.tri_ids.size(); ++j) {
        const uint32_t i = s.tri_ids[j];
        const float* t = in.tri_v + (size_t)i * 12;
        TriRecord& r = out.tris[j];
        for (int k = 0; k < 3; ++k) {
            r.v0[k] = t[k];
            r.e1[k] = t[4 + k] - t[k];
            r.e2[k] = t[8 + k] - t[k];
        }
        r.key = i;
        r.mat_index = in.tri_mat ? in.tri_mat[(size_t)i * 12] : 0u;
        r.pad = 0;
    }
    out.sph.resize(s.sph_ids.size() * 4);
    out.sph_key = s.sph_ids;
    for (size_t j = 0; j < s.sph_ids.size(); ++j) {
        const float* sp = in.sph + (size_t)s.sph_ids[j] * 16;
        out.sph[j * 4 + 0] = sp[12];
        out.sph[j * 4 + 1] = sp[13];
        out.sph[j * 4 + 2] = sp[14];
        out.sph[j * 4 + 3] = sp[15];
    }
    return out;
}

}  // namespace vkr
