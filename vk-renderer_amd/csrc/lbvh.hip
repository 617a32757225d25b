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
// centres, a radix sort, one thread per internal node to find its key range
// and split, then a bottom-up pass in which the second thread to reach a node
// writes its two child boxes and continues to the parent (the first one stops:
// no thread ever waits on another).  Leaves hold one triangle.
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

__device__ __forceinline__ void st(float* p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ float ld(float* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void stu(uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ uint32_t ldu(uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// cb[0..2]: ordered min of the box centres, cb[3..5]: ordered max.
__global__ void k_init(uint32_t* cb, uint32_t* flags, uint32_t n_flags)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3) cb[i] = 0xFFFFFFFFu;
    else if (i < 6) cb[i] = 0u;
    if (i < n_flags) flags[i] = 0u;
}

// Per triangle: its padded box (the host builder's padding) and the centre
// bounds of all finite triangles (one atomic per wave).
__global__ __launch_bounds__(kBuildBlock) void k_prim_boxes(const vkr_triangle* __restrict__ tris, uint32_t n,
                                                              float4* __restrict__ blo, float4* __restrict__ bhi,
                                                              uint32_t* cb)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
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
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 3; ++k) {
            mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], off));
            mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], off));
        }
    if ((threadIdx.x & 63u) == 0u)
        for (int k = 0; k < 3; ++k) {
            if (mn[k] != 0xFFFFFFFFu) atomicMin(&cb[k], mn[k]);
            if (mx[k] != 0u) atomicMax(&cb[3 + k], mx[k]);
        }
}

// Morton code of each box centre over the centre bounds; non-finite
// triangles get the largest key and sort last.
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
    uint64_t key = 0;
    for (int k = 0; k < 3; ++k) {
        const float lo = unord(cb[k]), ext = unord(cb[3 + k]) - lo;
        const float s = ext > 0.0f ? 2097151.0f / ext : 0.0f;
        const float q = fminf(fmaxf((c[k] - lo) * s, 0.0f), 2097151.0f);
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

// Internal node i: its key range, split and children.  Parents are stored as
// (node << 1) | slot.
__global__ __launch_bounds__(kBuildBlock) void k_topology(int n, const uint64_t* __restrict__ k, uint32_t node_off,
                                                            BvhNode* __restrict__ nodes, uint32_t* __restrict__ par_in,
                                                            uint32_t* __restrict__ par_leaf)
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
    const uint32_t me = (uint32_t)i << 1;
    uint32_t r0, r1;
    if (lo == g) {
        r0 = kRefLeaf | (uint32_t)g;
        par_leaf[g] = me | 0u;
    } else {
        r0 = node_off + (uint32_t)g;
        par_in[g] = me | 0u;
    }
    if (hi == g + 1) {
        r1 = kRefLeaf | (uint32_t)(g + 1);
        par_leaf[g + 1] = me | 1u;
    } else {
        r1 = node_off + (uint32_t)(g + 1);
        par_in[g + 1] = me | 1u;
    }
    BvhNode* nd = nodes + node_off + i;
    nd->d[0] = r0;
    nd->d[1] = r1;
    nd->d[2] = 0u;   // m = 0: triangle subtree, no per-ray widening
    nd->d[3] = 0u;
}

// Bottom-up boxes: one thread per leaf climbs while it is the second to reach
// a node.  slot_box holds 2 x 6 floats per node, slot_h 2 heights per node.
__global__ __launch_bounds__(kBuildBlock) void k_boxes(uint32_t n, const uint32_t* __restrict__ order,
                                                         const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                         const uint32_t* __restrict__ par_in,
                                                         const uint32_t* __restrict__ par_leaf, float* slot_box,
                                                         uint32_t* slot_h, uint32_t* flags, uint32_t node_off,
                                                         BvhNode* __restrict__ nodes, float* root_out)
{
    const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= n) return;
    const uint32_t tri = order[pos];
    const float4 l4 = blo[tri], h4 = bhi[tri];
    float lo[3] = {l4.x, l4.y, l4.z}, hi[3] = {h4.x, h4.y, h4.z};
    uint32_t h = 0;
    uint32_t p = par_leaf[pos];
    for (;;) {
        const uint32_t node = p >> 1, slot = p & 1u;
        float* mine = slot_box + (size_t)node * 12u + slot * 6u;
        for (int k = 0; k < 3; ++k) {
            st(mine + k, lo[k]);
            st(mine + 3 + k, hi[k]);
        }
        stu(slot_h + node * 2u + slot, h);
        __threadfence();
        if (atomicAdd(&flags[node], 1u) == 0u) return;   // the sibling finishes this node
        __threadfence();
        float* other = slot_box + (size_t)node * 12u + (slot ^ 1u) * 6u;
        float olo[3], ohi[3];
        for (int k = 0; k < 3; ++k) {
            olo[k] = ld(other + k);
            ohi[k] = ld(other + 3 + k);
        }
        const uint32_t oh = ldu(slot_h + node * 2u + (slot ^ 1u));
        BvhNode* nd = nodes + node_off + node;
        put_child_box(nd, (int)slot, lo, hi);
        put_child_box(nd, (int)(slot ^ 1u), olo, ohi);
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], olo[k]);
            hi[k] = fmaxf(hi[k], ohi[k]);
        }
        h = 1u + (h > oh ? h : oh);
        if (node == 0u) {
            for (int k = 0; k < 3; ++k) {
                root_out[k] = lo[k];
                root_out[3 + k] = hi[k];
            }
            root_out[6] = __uint_as_float(h);
            return;
        }
        p = par_in[node];
    }
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

struct Scratch {
    void* p = nullptr;
    ~Scratch()
    {
        if (p) (void)hipFree(p);
    }
};

inline size_t up256(size_t b) { return (b + 255u) & ~(size_t)255u; }

}  // namespace

hipError_t lbvh_build(const vkr_triangle* d_tris, uint32_t n, uint32_t node_off, BvhNode* d_nodes,
                      TriRecord* d_recs, LbvhRoot* root, hipStream_t s)
{
    if (n < 2u) return hipErrorInvalidValue;
    const uint32_t ni = n - 1u;
    size_t sort_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint64_t*)nullptr,
                                                      (uint64_t*)nullptr, (const uint32_t*)nullptr,
                                                      (uint32_t*)nullptr, (int)n, 0, 64, s);
    if (e != hipSuccess) return e;
    // One scratch allocation, carved into 256-byte aligned pieces.
    const size_t b_box = up256((size_t)n * sizeof(float4));
    const size_t b_key = up256((size_t)n * sizeof(uint64_t));
    const size_t b_u32n = up256((size_t)n * sizeof(uint32_t));
    const size_t b_u32i = up256((size_t)ni * sizeof(uint32_t));
    const size_t b_slot = up256((size_t)ni * 12u * sizeof(float));
    const size_t b_h = up256((size_t)ni * 2u * sizeof(uint32_t));
    const size_t b_cb = 256, b_root = 256;
    const size_t total = 2 * b_box + 2 * b_key + 3 * b_u32n + 2 * b_u32i + b_slot + b_h + b_cb + b_root +
                         up256(sort_bytes);
    Scratch sc;
    if ((e = hipMalloc(&sc.p, total)) != hipSuccess) return e;
    char* q = static_cast<char*>(sc.p);
    auto take = [&](size_t b) {
        char* r = q;
        q += b;
        return r;
    };
    float4* blo = (float4*)take(b_box);
    float4* bhi = (float4*)take(b_box);
    uint64_t* keys = (uint64_t*)take(b_key);
    uint64_t* keys_s = (uint64_t*)take(b_key);
    uint32_t* vals = (uint32_t*)take(b_u32n);
    uint32_t* order = (uint32_t*)take(b_u32n);
    uint32_t* par_leaf = (uint32_t*)take(b_u32n);
    uint32_t* par_in = (uint32_t*)take(b_u32i);
    uint32_t* flags = (uint32_t*)take(b_u32i);
    float* slot_box = (float*)take(b_slot);
    uint32_t* slot_h = (uint32_t*)take(b_h);
    uint32_t* cb = (uint32_t*)take(b_cb);
    float* root_out = (float*)take(b_root);
    void* sort_tmp = take(up256(sort_bytes));

    const uint32_t gn = (n + kBuildBlock - 1u) / kBuildBlock;
    const uint32_t gi = (ni + kBuildBlock - 1u) / kBuildBlock;
    const uint32_t g0 = ((ni > 6u ? ni : 6u) + kBuildBlock - 1u) / kBuildBlock;
    hipLaunchKernelGGL(k_init, dim3(g0), dim3(kBuildBlock), 0, s, cb, flags, ni);
    hipLaunchKernelGGL(k_prim_boxes, dim3(gn), dim3(kBuildBlock), 0, s, d_tris, n, blo, bhi, cb);
    hipLaunchKernelGGL(k_morton, dim3(gn), dim3(kBuildBlock), 0, s, n, blo, bhi, cb, keys, vals);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, keys, keys_s, vals, order, (int)n, 0, 64, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_topology, dim3(gi), dim3(kBuildBlock), 0, s, (int)n, keys_s, node_off, d_nodes, par_in,
                       par_leaf);
    hipLaunchKernelGGL(k_boxes, dim3(gn), dim3(kBuildBlock), 0, s, n, order, blo, bhi, par_in, par_leaf, slot_box,
                       slot_h, flags, node_off, d_nodes, root_out);
    hipLaunchKernelGGL(k_records, dim3(gn), dim3(kBuildBlock), 0, s, n, d_tris, order, d_recs);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    float host_root[7];
    if ((e = hipMemcpyAsync(host_root, root_out, sizeof host_root, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    for (int k = 0; k < 3; ++k) {
        root->lo[k] = host_root[k];
        root->hi[k] = host_root[3 + k];
    }
    std::memcpy(&root->height, &host_root[6], sizeof(uint32_t));
    return hipSuccess;
}

}  // namespace vkr
