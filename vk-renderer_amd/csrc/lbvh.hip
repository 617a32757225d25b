// lbvh.hip - device build of the triangle subtree (linear BVH over Morton codes).
//
// The reference uploads its triangles once and scans them (Source/
// GraphicsDevice.cpp:796-822, Assets/Tracer.comp:378-396); the host SAH
// builder (bvh.cpp) replaces the scan with a tree, and this builder makes the
// same tree format on the GPU for scenes whose upload time matters (a 2M
// triangle scene builds in milliseconds instead of ~300 ms on the host).
//
// Method: Karras, "Maximizing Parallelism in the Construction of BVHs,
// Octrees, and k-d Trees" (HPG 2012): 63-bit Morton codes of the triangle box
// centres, a radix sort, and one thread per internal node to find its key
// range, split and children.  Boxes are then filled level by level, deepest
// level first, one launch per level, so every child box a node reads was
// written by an earlier launch.  (The usual bottom-up climb with per-node
// atomics needs device-scope fences at every step, which on gfx950 write
// back and invalidate the L2 of each XCD: 17 ms instead of 0.3 ms on 2M
// triangles.)  A subtree of at most max_leaf triangles (a contiguous range of
// records) becomes one leaf: its parent refers to the records directly and
// its own nodes stay unused ("dead") in the array.
//
// Output, in the layout of bvh.hpp: internal node i of the Karras tree is node
// node_off + i (node 0 of the tree is its root); leaf j is triangle record j,
// records in Morton order.  Triangles with a non-finite vertex sort last and
// get boxes no ray enters (the reference's test never accepts them either).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

#include "kernels.hpp"

namespace vkr {
namespace {

constexpr int kBuildBlock = 256;

// Order-preserving map of floats to unsigned ints (for atomic min / max).
__device__ __forceinline__ uint32_t ord(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Spreads the low 21 bits of x to every third bit (bit k -> bit 3k).
__device__ __forceinline__ uint64_t spread3(uint64_t x)
{
    x &= 0x1FFFFFull;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ bool finite3(const float* v)
{
    return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}

// Centre and half-extent with the extent rounded up so that [c - e, c + e]
// contains [lo, hi] (the host builder's rule, bvh.cpp).
__device__ __forceinline__ void centre_extent(float lo, float hi, float& c, float& e)
{
    const double cd = 0.5 * ((double)lo + (double)hi);
    c = (float)cd;
    const double ed = 0.5 * ((double)hi - (double)lo) + fabs(cd - (double)c);
    e = (float)ed;
    if ((double)e < ed) e = nextafterf(e, __builtin_huge_valf());
}

// Writes the box of child `slot` into a node (refs are left alone).  An empty
// box (lo > hi: only non-finite triangles below) becomes the point
// (1e30, 1e30, 1e30) with zero extent, which no ray enters.
__device__ __forceinline__ void put_child_box(BvhNode* n, int slot, const float* lo, const float* hi)
{
    float c[3], e[3];
    const bool empty = !(lo[0] <= hi[0]) || !(lo[1] <= hi[1]) || !(lo[2] <= hi[2]);
    for (int k = 0; k < 3; ++k) {
        if (empty) {
            c[k] = 1e30f;
            e[k] = 0.0f;
        } else {
            centre_extent(lo[k], hi[k], c[k], e[k]);
        }
    }
    if (slot == 0) {
        n->a[0] = c[0]; n->a[1] = e[0]; n->a[2] = c[1]; n->a[3] = e[1];
        n->b[0] = c[2]; n->b[1] = e[2];
    } else {
        n->b[2] = c[0]; n->b[3] = e[0];
        n->c[0] = c[1]; n->c[1] = e[1]; n->c[2] = c[2]; n->c[3] = e[2];
    }
}

// Block-wide min (k < 3) / max (k >= 3) of six ordered values; thread 0
// returns them.
__device__ __forceinline__ void block_minmax6(uint32_t* v)
{
    __shared__ uint32_t red[kBuildBlock / 64][6];
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 6; ++k) {
            const uint32_t o = (uint32_t)__shfl_xor((int)v[k], off);
            v[k] = k < 3 ? min(v[k], o) : max(v[k], o);
        }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u)
        for (int k = 0; k < 6; ++k) red[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t j = 1; j < blockDim.x / 64u; ++j)
            for (int k = 0; k < 6; ++k) v[k] = k < 3 ? min(v[k], red[j][k]) : max(v[k], red[j][k]);
}

// Per triangle: its padded box (the host builder's padding); per block: the
// ordered min / max of the box centres of its finite triangles (part[6 b ..]).
__global__ __launch_bounds__(kBuildBlock) void k_prim_boxes(const vkr_triangle* __restrict__ tris, uint32_t n,
                                                              float4* __restrict__ blo, float4* __restrict__ bhi,
                                                              uint32_t* __restrict__ part)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    uint32_t* mn = v;
    uint32_t* mx = v + 3;
    if (i < n) {
        const vkr_triangle t = tris[i];
        const bool ok = finite3(t.v0) && finite3(t.v1) && finite3(t.v2);
        float lo[3], hi[3];
        if (ok) {
            float m = 0.0f;
            for (int k = 0; k < 3; ++k) {
                lo[k] = fminf(t.v0[k], fminf(t.v1[k], t.v2[k]));
                hi[k] = fmaxf(t.v0[k], fmaxf(t.v1[k], t.v2[k]));
                m = fmaxf(m, fmaxf(fabsf(lo[k]), fabsf(hi[k])));
            }
            for (int k = 0; k < 3; ++k) {
                const float e = 1e-5f * (m + (hi[k] - lo[k])) + 1e-6f;
                lo[k] -= e;
                hi[k] += e;
                const uint32_t oc = ord(0.5f * (lo[k] + hi[k]));
                mn[k] = oc;
                mx[k] = oc;
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                lo[k] = __builtin_huge_valf();
                hi[k] = -__builtin_huge_valf();
            }
        }
        blo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        bhi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    block_minmax6(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[blockIdx.x * 6u + k] = v[k];
}

// One block: reduces nb six-value partials into cb[0..5] (centre bounds).
__global__ __launch_bounds__(kBuildBlock) void k_reduce_bounds(const uint32_t* __restrict__ part, uint32_t nb,
                                                                 uint32_t* __restrict__ cb)
{
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (uint32_t j = threadIdx.x; j < nb; j += blockDim.x)
        for (int k = 0; k < 6; ++k) v[k] = k < 3 ? min(v[k], part[j * 6u + k]) : max(v[k], part[j * 6u + k]);
    block_minmax6(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) cb[k] = v[k];
}

// Morton code of each box centre over the centre bounds, one scale for all
// three axes (the largest extent), so a flat scene is not split along its
// thin axis as often as along the others; non-finite triangles get the
// largest key and sort last.
__global__ __launch_bounds__(kBuildBlock) void k_morton(uint32_t n, const float4* __restrict__ blo,
                                                          const float4* __restrict__ bhi, const uint32_t* cb,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 l = blo[i], h = bhi[i];
    vals[i] = i;
    if (!(l.x <= h.x)) {
        keys[i] = ~0ull;
        return;
    }
    const float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    const float lo[3] = {unord(cb[0]), unord(cb[1]), unord(cb[2])};
    const float ext = fmaxf(fmaxf(unord(cb[3]) - lo[0], unord(cb[4]) - lo[1]), unord(cb[5]) - lo[2]);
    const float s = ext > 0.0f ? 2097151.0f / ext : 0.0f;
    uint64_t key = 0;
    for (int k = 0; k < 3; ++k) {
        const float q = fminf(fmaxf((c[k] - lo[k]) * s, 0.0f), 2097151.0f);
        key |= spread3((uint64_t)q) << (2 - k);
    }
    keys[i] = key;
}

// Length of the common prefix of sorted keys i and j (ties broken by index);
// -1 outside [0, n).
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll(a ^ b);
    return 64 + __clz((uint32_t)i ^ (uint32_t)j);
}

// Internal node i: its key range, split and children (written to the node's
// refs; a child range of at most max_leaf records becomes a leaf ref);
// par_in[c] = parent of internal node c; alive[i] = 0 when node i lies in a
// collapsed range (then nothing refers to it).
__global__ __launch_bounds__(kBuildBlock) void k_topology(int n, const uint64_t* __restrict__ k, uint32_t node_off,
                                                            uint32_t max_leaf, BvhNode* __restrict__ nodes,
                                                            uint32_t* __restrict__ par_in, uint32_t* __restrict__ alive)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n - 1) return;
    const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(k, n, i, i - d);
    int lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(k, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int g = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    // Children: [lo, g] and [g + 1, hi]; a range of one record is always a
    // leaf (and not a Karras internal node).
    const uint32_t nl = (uint32_t)(g - lo + 1), nr = (uint32_t)(hi - g);
    const uint32_t r0 = nl <= max_leaf ? (kRefLeaf | ((nl - 1u) << kRefCountShift) | (uint32_t)lo)
                                       : node_off + (uint32_t)g;
    const uint32_t r1 = nr <= max_leaf ? (kRefLeaf | ((nr - 1u) << kRefCountShift) | (uint32_t)(g + 1))
                                       : node_off + (uint32_t)(g + 1);
    if (nl > 1u) par_in[g] = (uint32_t)i;
    if (nr > 1u) par_in[g + 1] = (uint32_t)i;
    alive[i] = (i == 0 || (uint32_t)(hi - lo + 1) > max_leaf) ? 1u : 0u;
    BvhNode* nd = nodes + node_off + i;
    nd->d[0] = r0;
    nd->d[1] = r1;
    nd->d[2] = 0u;   // m = 0: triangle subtree, no per-ray widening
    nd->d[3] = 0u;
}

constexpr uint32_t kDepthLimit = 255;   // deeper trees are reported as 255 (and rejected)
constexpr uint32_t kDead = 0xFFFFFFFFu;  // depth of a node in a collapsed range (no level)

// Depth of live internal node i below the root (the root, node 0, is at 0),
// by following parents (a live node's parent is live).  Per block, to
// part[3 b ..]: the deepest live node, the live nodes and their leaf refs.
__global__ __launch_bounds__(kBuildBlock) void k_depth(uint32_t ni, uint32_t node_off,
                                                         const BvhNode* __restrict__ nodes,
                                                         const uint32_t* __restrict__ par_in,
                                                         const uint32_t* __restrict__ alive,
                                                         uint32_t* __restrict__ depth, uint32_t* __restrict__ part)
{
    __shared__ uint32_t red[kBuildBlock / 64][3];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t v[3] = {0u, 0u, 0u};
    if (i < ni) {
        if (alive[i]) {
            uint32_t p = i, d = 0;
            while (p != 0u && d < kDepthLimit) {
                p = par_in[p];
                ++d;
            }
            depth[i] = d;
            const BvhNode& nd = nodes[node_off + i];
            v[0] = d;
            v[1] = 1u;
            v[2] = ((nd.d[0] & kRefLeaf) ? 1u : 0u) + ((nd.d[1] & kRefLeaf) ? 1u : 0u);
        } else {
            depth[i] = kDead;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        v[0] = max(v[0], (uint32_t)__shfl_xor((int)v[0], off));
        v[1] += (uint32_t)__shfl_xor((int)v[1], off);
        v[2] += (uint32_t)__shfl_xor((int)v[2], off);
    }
    if ((threadIdx.x & 63u) == 0u)
        for (int k = 0; k < 3; ++k) red[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t j = 1; j < blockDim.x / 64u; ++j) {
            v[0] = max(v[0], red[j][0]);
            v[1] += red[j][1];
            v[2] += red[j][2];
        }
        for (int k = 0; k < 3; ++k) part[blockIdx.x * 3u + k] = v[k];
    }
}

// One block: reduces nb (max, sum, sum) partials into out[0..2].
__global__ __launch_bounds__(kBuildBlock) void k_reduce3(const uint32_t* __restrict__ part, uint32_t nb,
                                                           uint32_t* __restrict__ out)
{
    __shared__ uint32_t red[kBuildBlock / 64][3];
    uint32_t v[3] = {0u, 0u, 0u};
    for (uint32_t j = threadIdx.x; j < nb; j += blockDim.x) {
        v[0] = max(v[0], part[j * 3u]);
        v[1] += part[j * 3u + 1u];
        v[2] += part[j * 3u + 2u];
    }
    for (int off = 32; off > 0; off >>= 1) {
        v[0] = max(v[0], (uint32_t)__shfl_xor((int)v[0], off));
        v[1] += (uint32_t)__shfl_xor((int)v[1], off);
        v[2] += (uint32_t)__shfl_xor((int)v[2], off);
    }
    if ((threadIdx.x & 63u) == 0u)
        for (int k = 0; k < 3; ++k) red[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t j = 1; j < blockDim.x / 64u; ++j) {
            v[0] = max(v[0], red[j][0]);
            v[1] += red[j][1];
            v[2] += red[j][2];
        }
        for (int k = 0; k < 3; ++k) out[k] = v[k];
    }
}

// The internal nodes at depth `level`: both child boxes into the node, and
// the node's own box (their union) to nlo / nhi for its parent, which the
// next launch (one level up) reads.
__global__ __launch_bounds__(kBuildBlock) void k_level(uint32_t ni, uint32_t level, const uint32_t* __restrict__ depth,
                                                         uint32_t node_off, BvhNode* __restrict__ nodes,
                                                         const uint32_t* __restrict__ order,
                                                         const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                         float4* __restrict__ nlo, float4* __restrict__ nhi)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni || depth[i] != level) return;
    BvhNode* nd = nodes + node_off + i;
    float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
    float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
    for (int slot = 0; slot < 2; ++slot) {
        const uint32_t ref = nd->d[slot];
        float4 l, h;
        if (ref & kRefLeaf) {
            const uint32_t first = ref & kRefFirstMask, cnt = ((ref >> kRefCountShift) & 7u) + 1u;
            l = blo[order[first]];
            h = bhi[order[first]];
            for (uint32_t c = 1; c < cnt; ++c) {
                const float4 l2 = blo[order[first + c]], h2 = bhi[order[first + c]];
                l = make_float4(fminf(l.x, l2.x), fminf(l.y, l2.y), fminf(l.z, l2.z), 0.0f);
                h = make_float4(fmaxf(h.x, h2.x), fmaxf(h.y, h2.y), fmaxf(h.z, h2.z), 0.0f);
            }
        } else {
            l = nlo[ref - node_off];
            h = nhi[ref - node_off];
        }
        const float cl[3] = {l.x, l.y, l.z}, ch[3] = {h.x, h.y, h.z};
        put_child_box(nd, slot, cl, ch);
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], cl[k]);
            hi[k] = fmaxf(hi[k], ch[k]);
        }
    }
    nlo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    nhi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

// Triangle records in Morton order: (v0, key) (e1, mat_index) (e2, 0) with
// the edges the reference computes (Tracer.comp:342-343).
__global__ __launch_bounds__(kBuildBlock) void k_records(uint32_t n, const vkr_triangle* __restrict__ tris,
                                                           const uint32_t* __restrict__ order,
                                                           TriRecord* __restrict__ recs)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = order[j];
    const vkr_triangle t = tris[i];
    TriRecord r;
    for (int k = 0; k < 3; ++k) {
        r.v0[k] = t.v0[k];
        r.e1[k] = t.v1[k] - t.v0[k];
        r.e2[k] = t.v2[k] - t.v0[k];
    }
    r.key = i;
    r.mat_index = t.mat_index;
    r.pad = 0u;
    recs[j] = r;
}

inline size_t up256(size_t b) { return (b + 255u) & ~(size_t)255u; }

}  // namespace

namespace {

// Scratch layout of a build over n triangles: 256-byte aligned pieces of one
// allocation.
struct ScratchPlan {
    size_t b_f4n, b_f4i, b_key, b_u32n, b_u32i, b_part, b_part2, b_sort, total;
};

hipError_t plan_scratch(uint32_t n, ScratchPlan& sp)
{
    const uint32_t ni = n - 1u;
    const uint32_t gn = (n + kBuildBlock - 1u) / kBuildBlock;
    const uint32_t gi = (ni + kBuildBlock - 1u) / kBuildBlock;
    size_t sort_bytes = 0;
    const hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint64_t*)nullptr,
                                                            (uint64_t*)nullptr, (const uint32_t*)nullptr,
                                                            (uint32_t*)nullptr, (int)n, 0, 64, (hipStream_t)0);
    if (e != hipSuccess) return e;
    sp.b_f4n = up256((size_t)n * sizeof(float4));
    sp.b_f4i = up256((size_t)ni * sizeof(float4));
    sp.b_key = up256((size_t)n * sizeof(uint64_t));
    sp.b_u32n = up256((size_t)n * sizeof(uint32_t));
    sp.b_u32i = up256((size_t)ni * sizeof(uint32_t));
    sp.b_part = up256((size_t)gn * 6u * sizeof(uint32_t));
    sp.b_part2 = up256((size_t)gi * 3u * sizeof(uint32_t));
    sp.b_sort = up256(sort_bytes);
    sp.total = 2 * sp.b_f4n + 2 * sp.b_f4i + 2 * sp.b_key + 2 * sp.b_u32n + 3 * sp.b_u32i + sp.b_part +
               sp.b_part2 + 256 + sp.b_sort;
    return hipSuccess;
}

}  // namespace

size_t lbvh_scratch_bytes(uint32_t n)
{
    ScratchPlan sp{};
    if (n < 2u || plan_scratch(n, sp) != hipSuccess) return 0;
    return sp.total;
}

hipError_t lbvh_build(const vkr_triangle* d_tris, uint32_t n, uint32_t node_off, uint32_t max_leaf,
                      uint32_t height_cap, void* scratch, size_t scratch_bytes, BvhNode* d_nodes, TriRecord* d_recs,
                      LbvhRoot* root, hipStream_t s)
{
    if (n < 2u || max_leaf < 1u || max_leaf > kMaxLeaf) return hipErrorInvalidValue;
    const uint32_t ni = n - 1u;
    const uint32_t gn = (n + kBuildBlock - 1u) / kBuildBlock;
    const uint32_t gi = (ni + kBuildBlock - 1u) / kBuildBlock;
    ScratchPlan sp{};
    hipError_t e = plan_scratch(n, sp);
    if (e != hipSuccess) return e;
    if (!scratch || scratch_bytes < sp.total) return hipErrorInvalidValue;
    size_t sort_bytes = sp.b_sort;
    const size_t b_f4n = sp.b_f4n;
    const size_t b_f4i = sp.b_f4i, b_key = sp.b_key, b_u32n = sp.b_u32n, b_u32i = sp.b_u32i;
    const size_t b_part = sp.b_part, b_part2 = sp.b_part2;
    char* q = static_cast<char*>(scratch);
    auto take = [&](size_t b) {
        char* r = q;
        q += b;
        return r;
    };
    float4* blo = (float4*)take(b_f4n);
    float4* bhi = (float4*)take(b_f4n);
    float4* nlo = (float4*)take(b_f4i);
    float4* nhi = (float4*)take(b_f4i);
    uint64_t* keys = (uint64_t*)take(b_key);
    uint64_t* keys_s = (uint64_t*)take(b_key);
    uint32_t* vals = (uint32_t*)take(b_u32n);
    uint32_t* order = (uint32_t*)take(b_u32n);
    uint32_t* par_in = (uint32_t*)take(b_u32i);
    uint32_t* depth = (uint32_t*)take(b_u32i);
    uint32_t* alive = (uint32_t*)take(b_u32i);
    uint32_t* part = (uint32_t*)take(b_part);
    uint32_t* part2 = (uint32_t*)take(b_part2);
    uint32_t* small = (uint32_t*)take(256);   // [0..5] centre bounds, [6..8] deepest, live nodes, leaves
    void* sort_tmp = take(sp.b_sort);

    const dim3 blk(kBuildBlock);
    hipLaunchKernelGGL(k_prim_boxes, dim3(gn), blk, 0, s, d_tris, n, blo, bhi, part);
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), blk, 0, s, part, gn, small);
    hipLaunchKernelGGL(k_morton, dim3(gn), blk, 0, s, n, blo, bhi, small, keys, vals);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, keys, keys_s, vals, order, (int)n, 0, 64, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_topology, dim3(gi), blk, 0, s, (int)n, keys_s, node_off, max_leaf, d_nodes, par_in, alive);
    hipLaunchKernelGGL(k_depth, dim3(gi), blk, 0, s, ni, node_off, d_nodes, par_in, alive, depth, part2);
    hipLaunchKernelGGL(k_reduce3, dim3(1), blk, 0, s, part2, gi, small + 6);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t st[3] = {0u, 0u, 0u};
    if ((e = hipMemcpyAsync(st, small + 6, sizeof st, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    const uint32_t maxd = st[0];
    root->height = maxd + 1u;   // the deepest live node's children are leaves
    root->nodes = st[1];
    root->leaves = st[2];
    if (root->height > height_cap) return hipSuccess;   // rejected: no boxes, no records
    for (uint32_t level = maxd + 1u; level-- > 0u;)
        hipLaunchKernelGGL(k_level, dim3(gi), blk, 0, s, ni, level, depth, node_off, d_nodes, order, blo, bhi, nlo,
                           nhi);
    hipLaunchKernelGGL(k_records, dim3(gn), blk, 0, s, n, d_tris, order, d_recs);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    float4 rb[2];
    if ((e = hipMemcpyAsync(&rb[0], nlo, sizeof(float4), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&rb[1], nhi, sizeof(float4), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    root->lo[0] = rb[0].x; root->lo[1] = rb[0].y; root->lo[2] = rb[0].z;
    root->hi[0] = rb[1].x; root->hi[1] = rb[1].y; root->hi[2] = rb[1].z;
    return hipSuccess;
}

}  // namespace vkr
