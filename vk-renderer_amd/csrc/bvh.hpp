// bvh.hpp - host-side BVH builder for the gfx950 traversal kernels.
//
// The reference has no acceleration structure: trace_ray scans every
// primitive linearly (Assets/Tracer.comp:374-431).  This BVH replaces that
// scan; the kernels recover the scan's exact order-dependent acceptance rule
// from the hits it reports (see trace_kernels.hip).
//
// Layout (all records 16-byte aligned for 128-bit loads):
//   Node (64 B): the bounds of BOTH children, as centre and half-extent per
//   axis (rounded so the float box contains the exact one), plus their refs,
//   so one node fetch tests two boxes:
//     a = (c0.cx, c0.ex, c0.cy, c0.ey)
//     b = (c0.cz, c0.ez, c1.cx, c1.ex)
//     c = (c1.cy, c1.ey, c1.cz, c1.ez)
//     d = (ref0, ref1, m, 0)
//   m is the float 1.0 in the nodes of the sphere subtree and 0 elsewhere;
//   the traversal multiplies its per-ray widening by it.
//   With a centre/extent box the slab test is t = (c - o) / d -+ e / |d|, and
//   a per-ray widening w of the box is one more term: (e + w) / |d|.
//   Triangles and spheres live in separate subtrees: when both exist, node 0
//   is a root whose child 0 is the triangle subtree and child 1 the sphere
//   subtree (an always-entered box), and the nodes from index sph_node0 on
//   belong to the sphere subtree.  The traversal widens the boxes of the
//   sphere subtree per ray (the reference's sphere test is imprecise far from
//   the sphere, see traverse.hpp).
//   ref: bit31 = leaf; for a leaf bit30 = sphere leaf, bits 27..29 = count-1,
//   bits 0..26 = first record; for an inner node the node index.
//   Triangle record (48 B): (v0, key) (e1, mat_index) (e2, 0) where
//   e1 = fl(v1 - v0), e2 = fl(v2 - v0) are exactly the edges the reference
//   computes (Tracer.comp:342-343) and key is the triangle's scan index.
//   Sphere record (16 B + 4 B key): (centre, r) and the sphere's scan index.
#pragma once

#include <cstdint>
#include <vector>

namespace vkr {

constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kRefSphere = 0x40000000u;
constexpr uint32_t kRefCountShift = 27;
constexpr uint32_t kRefFirstMask = (1u << 27) - 1u;
constexpr uint32_t kMaxLeaf = 8;
// Default leaf size cap: small leaves keep the lanes of a wave's leaf pass
// busy.  Scenes with several emitters are traced by the pooled kernel,
// mostly shadow rays: there single-primitive leaves run fastest (1080p /
// 100k-triangle soup: 113 ms against 122 ms with 2, 135 ms with 3).  With
// one emitter, 2 is best on the 2M-triangle heightfield (9.9 ms against
// 10.3 ms with 1).
constexpr uint32_t kDefaultLeaf = 2;
inline uint32_t default_leaf(uint32_t n_emitters) { return n_emitters >= 2u ? 1u : kDefaultLeaf; }
// Default depth cap: trees of at most 23 levels run the path kernel with a
// 24-entry stack (one entry holds the end marker), which keeps 5 waves per
// SIMD resident instead of 4.  The SAH split search respects the cap.
constexpr uint32_t kDefaultMaxDepth = 23;

struct alignas(16) BvhNode {
    float a[4], b[4], c[4];
    uint32_t d[4];
};
static_assert(sizeof(BvhNode) == 64, "node is 64 bytes");

struct alignas(16) TriRecord {
    float v0[3]; uint32_t key;
    float e1[3]; uint32_t mat_index;
    float e2[3]; uint32_t pad;
};
static_assert(sizeof(TriRecord) == 48, "triangle record is 48 bytes");

struct BuildInput {
    const float* tri_v;       // n_tris * 12 floats: v0(3) pad v1(3) pad v2(3) pad (std430)
    const uint32_t* tri_mat;  // optional, stride 12 u32 (the v0 padding word)
    uint32_t n_tris = 0;
    const float* sph;         // n_spheres * 16 floats (std430 Sphere: P at 12, r at 15)
    uint32_t n_spheres = 0;
    uint32_t max_leaf = kDefaultLeaf;
    uint32_t max_depth = kDefaultMaxDepth;
    float traversal_cost = 0.0f;  // SAH cost of a node visit relative to one primitive test (0 = 3)
};

struct BuildOutput {
    std::vector<BvhNode> nodes;
    std::vector<TriRecord> tris;
    std::vector<float> sph;        // 4 floats per sphere record (centre, r)
    std::vector<uint32_t> sph_key; // scan index per sphere record
    uint32_t leaves = 0;
    uint32_t depth = 0;            // deepest leaf (root children are depth 1)
    uint32_t sph_node0 = 0xFFFFFFFFu;  // first node of the sphere subtree (none: all ones)
    float sph_c[3] = {0.0f, 0.0f, 0.0f};  // sphere centres lie within sph_r of sph_c
    float sph_r = 0.0f;
    bool empty = true;
};

// The node above a device-built triangle subtree and a sphere subtree:
// child 0 the triangle box [lo, hi] (an empty box: never entered) with ref
// tri_ref, child 1 always entered with ref sph_ref.
BvhNode join_root(const float lo[3], const float hi[3], uint32_t tri_ref, uint32_t sph_ref);

// Adds off to every inner-node ref of nodes (leaf refs are unchanged).
void offset_inner_refs(std::vector<BvhNode>& nodes, uint32_t off);

// Builds a binned-SAH BVH.  Leaves hold one primitive kind only and at most
// max_leaf primitives; no leaf is deeper than max_depth (splits fall back to
// balanced median splits near the cap), so a traversal stack of max_depth + 1
// entries (one holds the end marker) never overflows.
BuildOutput build_bvh(const BuildInput& in);

}  // namespace vkr
