// traverse.hpp - BVH traversal shared by the path and adjoint kernels.
//
// trace_ray of Assets/Tracer.comp:374-431 scans every triangle, then every
// sphere, then every plane, accepting `t > EPSILON && t < best + EPSILON` for
// triangles and spheres and `t < best - EPSILON` for planes.  The BVH visits
// primitives in another order, so the scan's result is rebuilt exactly: the
// smallest hit t* is always accepted by the scan, hence only hits within a
// window above t* can still change the winner; they are kept in a 4-entry
// list and replayed in scan order after traversal.  A window or list
// overflow is counted (documented near-tie).  Planes stay a linear list.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "vkr_device.hpp"

namespace vkr {

constexpr uint32_t kDone = 0xFFFFFFFFu;


// A wave leaves the node loop for the leaf pass once at least kLeafExit8/8
// of its lanes that are still traversing wait at a leaf (or, in the pooled
// loop, for a new ray); the others keep their place in the tree and go on in
// the next round.  Every ray still visits the same nodes and primitives in
// the same order (results do not depend on it); only the grouping of lanes
// into wave iterations changes.  4/8 measured best on incoherent 1080p /
// 100k-triangle frames with leaves of at most 2 primitives (2/8 and 6/8 are
// 12% and 4% slower with the pooled loop).
constexpr uint32_t kLeafExit8 = 4;
constexpr uint32_t kPhTrace = 0;
constexpr uint32_t kPhNee = 1;
constexpr int kCand = 4;

struct WorkCnt {
    uint32_t nodes, tris, spheres, planes, overflow;
    uint32_t wave_nodes, wave_leaves;   // wave-level loop iterations (counting variant)
};

// Counts one wave-level iteration (on the first active lane).
__device__ __forceinline__ void count_wave(uint32_t& c)
{
    const unsigned long long act = __ballot(1);
    if (__lane_id() == __ffsll((long long)act) - 1) ++c;
}

struct TravOut {
    uint32_t prim;  // (kind << 30) | index, kMiss when nothing was accepted
    float t;        // accepted t (t_max when missed)
    bool hit;       // closest: something accepted; any: occluded
};

__device__ __forceinline__ void cswap(float& ta, uint32_t& ka, float& tb, uint32_t& kb)
{
    const bool s = kb < ka;
    const float t0 = s ? tb : ta, t1 = s ? ta : tb;
    const uint32_t k0 = s ? kb : ka, k1 = s ? ka : kb;
    ta = t0; tb = t1; ka = k0; kb = k1;
}

// One ray in flight: trace_ray of Assets/Tracer.comp:374-431 over the BVH.
//   any == false: closest hit with the reference's acceptance rule.
//   any == true : only "was anything accepted" (shadow rays, :473).
struct Ray {
    f3 o, d;
    float tmax, tmax_eps, cb, tmin;     // cb: the box test's far bound (cull distance plus slack)
    float ix, iy, iz, ox, oy, oz, pad;   // |ix| etc. come free as source modifiers
    float ct0, ct1, ct2, ct3;
    uint32_t ck0, ck1, ck2, ck3;
    bool any, ovf, occluded, nan_ray;
    uint32_t ref;   // >= 0 (as int): inner node; < -1: leaf; kDone: traversal over
    int sp;         // stack entries; entry 0 is always kDone, so a pop needs no bounds check

    // Boxes entered beyond `cull` hold no hit the ray can still accept; the
    // box test compares against it with a relative and absolute slack, kept
    // here so that node steps do not recompute it (it changes only at leaves).
    __device__ __forceinline__ void set_cull(float cull) { cb = cull + fabsf(cull) * 1e-5f + 1e-5f; }

    __device__ __forceinline__ void start(const DevScene& sc, f3 o_, f3 d_, float tmax_, float eps, bool any_)
    {
        const float inf = __builtin_huge_valf();
        o = o_;
        d = d_;
        tmax = tmax_;
        any = any_;
        tmax_eps = tmax + eps;
        set_cull(any ? tmax_eps : tmax_eps + 4.0f * eps);
        tmin = inf;
        ct0 = ct1 = ct2 = ct3 = inf;
        ck0 = ck1 = ck2 = ck3 = 0;
        ovf = false;
        occluded = false;
        nan_ray = (d.x != d.x) || (d.y != d.y) || (d.z != d.z) || (o.x != o.x) || (o.y != o.y) || (o.z != o.z);
        ix = 1.0f / (fabsf(d.x) > 1e-20f ? d.x : copysignf(1e-20f, d.x));
        iy = 1.0f / (fabsf(d.y) > 1e-20f ? d.y : copysignf(1e-20f, d.y));
        iz = 1.0f / (fabsf(d.z) > 1e-20f ? d.z : copysignf(1e-20f, d.z));
        ox = o.x * ix;
        oy = o.y * iy;
        oz = o.z * iz;
        // The sphere test (sphere_t, Assets/Tracer.comp:314-330) rounds its
        // discriminant with an error up to ~60 ulp of |o - c|^2, so from far
        // away it reports hits up to ~1e-3 |o - c| outside the sphere.  The
        // sphere subtree's boxes are widened by 2e-3 (|o - C| + R), which
        // bounds that distance for every centre c within R of C.
        const f3 oc = sub(o, f3{sc.sph_c[0], sc.sph_c[1], sc.sph_c[2]});
        pad = fminf(2e-3f * (sqrtf(dot(oc, oc)) + sc.sph_r), 3e38f);
        ref = (sc.has_bvh && !nan_ray) ? 0u : kDone;
        sp = 1;   // above the kDone entry the traversal loops keep at the bottom
    }

    // Visits inner node `ref`: near-first descent with a branch-free step.
    template <int BLOCK, bool COUNT>
    __device__ __forceinline__ void node(const BvhNode* __restrict__ nodes, uint32_t* stk, WorkCnt& wc)
    {
        // 32-bit byte offsets (node count < 2^26, checked at upload) let the
        // loads use a scalar base address.
        const float4* np = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(nodes) + (ref << 6));
        const float4 A = np[0], B = np[1], C = np[2];
        const uint4 D = reinterpret_cast<const uint4*>(np)[3];
        if (COUNT) {
            wc.nodes++;
            count_wave(wc.wave_nodes);
        }
        // Slab test on centre/half-extent boxes: t = (c - o) i -+ (e + w) |i|,
        // with the widening w only in the sphere subtree (m = 1 there).
        const float mp = __uint_as_float(D.z) * pad;
        const float ax = fabsf(ix), ay = fabsf(iy), az = fabsf(iz);
        const float qx = mp * ax, qy = mp * ay, qz = mp * az;
        const float a0 = __builtin_fmaf(A.x, ix, -ox), a1 = __builtin_fmaf(A.y, ax, qx);
        const float a2 = __builtin_fmaf(A.z, iy, -oy), a3 = __builtin_fmaf(A.w, ay, qy);
        const float a4 = __builtin_fmaf(B.x, iz, -oz), a5 = __builtin_fmaf(B.y, az, qz);
        const float b0 = __builtin_fmaf(B.z, ix, -ox), b1 = __builtin_fmaf(B.w, ax, qx);
        const float b2 = __builtin_fmaf(C.x, iy, -oy), b3 = __builtin_fmaf(C.y, ay, qy);
        const float b4 = __builtin_fmaf(C.z, iz, -oz), b5 = __builtin_fmaf(C.w, az, qz);
        const float n0 = fmaxf(fmaxf(a0 - a1, a2 - a3), fmaxf(a4 - a5, 0.0f));
        const float f0 = fminf(fminf(a0 + a1, a2 + a3), fminf(a4 + a5, cb));
        const float n1 = fmaxf(fmaxf(b0 - b1, b2 - b3), fmaxf(b4 - b5, 0.0f));
        const float f1 = fminf(fminf(b0 + b1, b2 + b3), fminf(b4 + b5, cb));
        const bool h0 = n0 <= f0, h1 = n1 <= f1;
        // The far child is always written to the free slot above the top
        // (sp < STACK at any inner node: the stack holds the kDone entry and
        // at most one entry per level above, and the tree is at most
        // STACK - 1 levels deep), and only counted as pushed when both
        // children are hit.  With neither hit the top is popped; the bottom
        // kDone entry ends the traversal.
        const bool first0 = n0 <= n1;
        const uint32_t nearr = (h0 && (!h1 || first0)) ? D.x : D.y;
        stk[sp * BLOCK] = first0 ? D.y : D.x;
        const uint32_t top = stk[(sp - 1) * BLOCK];
        const bool none = !h0 && !h1;
        ref = none ? top : nearr;
        sp += (h0 && h1 ? 1 : 0) - (none ? 1 : 0);
    }

    // Tests the primitives of leaf `ref`, then pops the stack.
    template <int BLOCK, bool COUNT>
    __device__ __forceinline__ void leaf(const DevScene& sc, float eps, uint32_t* stk, WorkCnt& wc)
    {
        const float inf = __builtin_huge_valf();
        const float win = 4.0f * eps;
        if (COUNT) count_wave(wc.wave_leaves);
        const uint32_t count = ((ref >> kRefCountShift) & 7u) + 1u;
        const uint32_t first = ref & kRefFirstMask;
        for (uint32_t i = 0; i < count; ++i) {
            float t;
            uint32_t key;
            if (ref & kRefSphere) {
                const float4 s4 = sc.sph[first + i];
                if (COUNT) wc.spheres++;
                t = sphere_t(o, d, f3{s4.x, s4.y, s4.z}, s4.w);
                key = 0x80000000u | sc.sph_key[first + i];
            } else {
                const float4* tp = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sc.tris) +
                                                                    (first + i) * (uint32_t)sizeof(TriRecord));
                const float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
                if (COUNT) wc.tris++;
                t = tri_t(o, d, f3{p0.x, p0.y, p0.z}, f3{p1.x, p1.y, p1.z}, f3{p2.x, p2.y, p2.z}, eps);
                key = __float_as_uint(p0.w);
            }
            if (!(t > eps)) continue;
            if (any) {
                if (t < tmax_eps) occluded = true;
            } else if (t < fminf(tmin, tmax_eps) + win) {
                if (t < tmin) {
                    tmin = t;
                    const float lim = tmin + win;
                    ct0 = ct0 < lim ? ct0 : inf;
                    ct1 = ct1 < lim ? ct1 : inf;
                    ct2 = ct2 < lim ? ct2 : inf;
                    ct3 = ct3 < lim ? ct3 : inf;
                }
                if (ct0 == inf) { ct0 = t; ck0 = key; }
                else if (ct1 == inf) { ct1 = t; ck1 = key; }
                else if (ct2 == inf) { ct2 = t; ck2 = key; }
                else if (ct3 == inf) { ct3 = t; ck3 = key; }
                else {
                    // List full: keep the four smallest t (documented near-tie).
                    ovf = true;
                    float mx = ct0; int mi = 0;
                    if (ct1 > mx) { mx = ct1; mi = 1; }
                    if (ct2 > mx) { mx = ct2; mi = 2; }
                    if (ct3 > mx) { mx = ct3; mi = 3; }
                    if (t < mx) {
                        if (mi == 0) { ct0 = t; ck0 = key; }
                        else if (mi == 1) { ct1 = t; ck1 = key; }
                        else if (mi == 2) { ct2 = t; ck2 = key; }
                        else { ct3 = t; ck3 = key; }
                    }
                }
                set_cull(fminf(tmin, tmax_eps) + win);
            }
        }
        if (occluded) {
            ref = kDone;
        } else {
            ref = stk[(sp - 1) * BLOCK];   // the bottom entry is kDone
            --sp;
        }
    }

    // After traversal: the planes (a linear list) and, for closest-hit rays,
    // the replay of the candidates in scan order.
    template <bool COUNT>
    __device__ __forceinline__ TravOut finish(const DevScene& sc, float eps, WorkCnt& wc)
    {
        TravOut r;
        const uint32_t np = sc.n_planes;
        const vkr_plane* __restrict__ planes = sc.planes;
        if (any) {
            if (!occluded && !nan_ray) {
                const float lim = tmax - eps;
                for (uint32_t i = 0; i < np; ++i) {
                    if (COUNT) wc.planes++;
                    const float t = plane_t(o, d, ld3(planes[i].N), planes[i].len);
                    if ((t > eps) && (t < lim)) { occluded = true; break; }
                }
            }
            r.hit = occluded;
            r.prim = kMiss;
            r.t = tmax;
            return r;
        }
        cswap(ct0, ck0, ct1, ck1);
        cswap(ct2, ck2, ct3, ck3);
        cswap(ct0, ck0, ct2, ck2);
        cswap(ct1, ck1, ct3, ck3);
        cswap(ct1, ck1, ct2, ck2);
        float best = tmax;
        uint32_t win_key = kMiss;
        const float top = tmin + 4.0f * eps - eps;
#define VKR_REPLAY(T, K)                          \
    if ((T) < best + eps) {                       \
        best = (T);                               \
        win_key = (K);                            \
        if (best >= top) ovf = true;              \
    }
        VKR_REPLAY(ct0, ck0)
        VKR_REPLAY(ct1, ck1)
        VKR_REPLAY(ct2, ck2)
        VKR_REPLAY(ct3, ck3)
#undef VKR_REPLAY
        uint32_t prim = kMiss;
        if (win_key != kMiss) prim = (win_key & 0x80000000u) ? ((1u << 30) | (win_key & 0x7FFFFFFFu)) : win_key;
        if (!nan_ray) {
            for (uint32_t i = 0; i < np; ++i) {
                if (COUNT) wc.planes++;
                const float t = plane_t(o, d, ld3(planes[i].N), planes[i].len);
                if ((t > eps) && (t < best - eps)) {
                    best = t;
                    prim = (2u << 30) | i;
                }
            }
        }
        if (ovf) wc.overflow++;
        r.hit = prim != kMiss;
        r.prim = prim;
        r.t = best;
        return r;
    }
};

// The node loop's exit test: a wave leaves it when no lane is at an inner
// node, or when at least kLeafExit8/8 of the lanes that still traverse wait
// (at a leaf, or for a new ray).  Both loops take two node steps per test
// (the test's ballots cost about as much as a third of a step): 3-5% faster
// on 1080p frames of the 100k-triangle soup and the 2M-triangle heightfield.
__device__ __forceinline__ bool leave_nodes(uint32_t nn, uint32_t nw)
{
    return nn == 0 || (nw > 0 && nw * 8u >= (nw + nn) * kLeafExit8);
}

// One ray per lane (lanes with active == false only take part in the
// wave's loop).
template <int BLOCK, bool COUNT>
__device__ __forceinline__ TravOut traverse(const DevScene& sc, f3 o, f3 d, float tmax, float eps, bool any,
                                            bool active, uint32_t* stk, WorkCnt& wc)
{
    Ray r;
    r.start(sc, o, d, tmax, eps, any);
    stk[0] = kDone;
    if (!active) r.ref = kDone;
    if (r.ref != kDone) {   // (only traversing lanes in the loop: 11 fewer VGPRs in k_path)
        while (true) {
            while (true) {
                const uint32_t nn = (uint32_t)__popcll(__ballot((int)r.ref >= 0));
                const uint32_t nl = (uint32_t)__popcll(__ballot((int)r.ref < -1));
                if (leave_nodes(nn, nl)) break;
                if ((int)r.ref >= 0) r.node<BLOCK, COUNT>(sc.nodes, stk, wc);
                if ((int)r.ref >= 0) r.node<BLOCK, COUNT>(sc.nodes, stk, wc);
            }
            if (__ballot(r.ref != kDone) == 0) break;
            if ((int)r.ref < -1) r.leaf<BLOCK, COUNT>(sc, eps, stk, wc);
        }
    }
    // Lanes without a ray skip the plane list (their result is not read).
    if (!active) return TravOut{kMiss, tmax, false};
    return r.finish<COUNT>(sc, eps, wc);
}

constexpr uint32_t kNone = 0xFFFFu;

// Numbers a wave's work items slot by slot: `bits` holds the slots this lane
// has an item in.  Items cpre[k] .. cpre[k+1]-1 are slot k of the lanes
// whose bit is set in colm[k], in lane order; slots past the last occupied
// one start at the total.  Returns the wave's total (wave-uniform).  The
// caller orders these writes before the lookups (pool_item).
__device__ __forceinline__ uint32_t pool_columns(uint32_t bits, uint32_t* cpre, unsigned long long* colm)
{
    uint32_t tot = 0, my_pre = 0;
    unsigned long long my_col = 0;
    uint32_t k = 0;
#pragma unroll 1
    for (; k < kRaySlots; ++k) {
        if (__ballot((bits >> k) != 0u) == 0) break;   // no lane holds slot k or above
        const unsigned long long c = __ballot((bits >> k) & 1u);
        if (__lane_id() == k) {
            my_col = c;
            my_pre = tot;
        }
        tot += (uint32_t)__popcll(c);
    }
    if (__lane_id() >= k) my_pre = tot;
    if (__lane_id() < kRaySlots) {
        cpre[__lane_id()] = my_pre;
        colm[__lane_id()] = my_col;
    }
    return tot;
}

// Item e (< the total of pool_columns) is slot `slot` of lane `lane`.
__device__ __forceinline__ void pool_item(const uint32_t* cpre, const unsigned long long* colm, uint32_t e,
                                          uint32_t& slot, uint32_t& lane)
{
    // Item e is in the last slot whose first item is <= e (empty slots share
    // their successor's start) ...
    slot = 0;
    for (uint32_t step = kRaySlots / 2; step > 0; step >>= 1)
        if (cpre[slot + step] <= e) slot += step;
    // ... held by the lane of the (e - cpre[slot])-th set bit of colm[slot].
    unsigned long long m = colm[slot];
    uint32_t rem = e - cpre[slot], lo = 0;
    for (uint32_t wdt = 32; wdt > 0; wdt >>= 1) {
        const uint32_t c = (uint32_t)__popcll(m & ((1ull << wdt) - 1ull));
        if (rem >= c) {
            rem -= c;
            m >>= wdt;
            lo += wdt;
        }
    }
    lane = lo;
}

// Traces every ray the lanes of a wave queued for this step.  Lanes take rays
// 0..total of the wave until none is left, so a step costs the wave's total
// work plus the last rays' tail instead of its longest lane.  Rays are
// numbered slot by slot: rays cpre[k] .. cpre[k+1]-1 are slot k of the lanes
// whose bit is set in colm[k], in lane order.  A shadow ray sets bit `slot`
// of occ[owner] when occluded; the closest-hit ray of lane `owner` writes
// (prim, t) to res[owner].  `wrec` is the wave's block of ray records
// (kWHead, kWSlotD: the lanes of a wave reading one slot of neighbouring
// owners read neighbouring addresses).
template <int BLOCK, bool COUNT, bool BATCH>
__device__ __forceinline__ void traverse_pool(const JobDesc* jobs, uint32_t total, const uint32_t* cpre,
                                              const unsigned long long* colm, uint32_t* occ, uint2* res,
                                              const char* __restrict__ wrec, uint32_t* stk, WorkCnt& wc)
{
    // The job (scene, epsilon) of the ray in flight: a batched wave can hold
    // pixels of several jobs, and a lane may trace another lane's ray.
    const JobDesc* rj = jobs;
    Ray r;
    r.ref = kDone;
    stk[0] = kDone;
    uint32_t cur = kNone;
    uint32_t next = 0;   // wave-uniform pool cursor
    while (true) {
        // Post finished rays and hand out new ones until every lane is busy
        // or the pool is empty.
        while (true) {
            if (cur != kNone && r.ref == kDone) {
                const TravOut t = r.finish<COUNT>(rj->sc, rj->eps, wc);
                const uint32_t owner = cur >> 5;
                if (r.any) {
                    if (t.hit) atomicOr(occ + owner, 1u << (cur & 31u));
                } else {
                    res[owner] = make_uint2(t.prim, __float_as_uint(t.t));
                }
                cur = kNone;
            }
            const unsigned long long want = __ballot(cur == kNone);
            if (want != 0 && next < total) {
                const uint32_t rank = (uint32_t)__popcll(want & ((1ull << __lane_id()) - 1ull));
                if (cur == kNone && next + rank < total) {
                    uint32_t slot, lo;
                    pool_item(cpre, colm, next + rank, slot, lo);
                    cur = (lo << 5) | slot;
                    const float4 h = *reinterpret_cast<const float4*>(wrec + (kWHead + lo * 16u));
                    const float4 s = *reinterpret_cast<const float4*>(wrec + (kWSlotD + (slot * 64u + lo) * 16u));
                    const uint32_t hw = __float_as_uint(h.w);   // (owner's job << 1) | shadow ray
                    if constexpr (BATCH) rj = jobs + (hw >> 1);
                    r.start(rj->sc, f3{h.x, h.y, h.z}, f3{s.x, s.y, s.z}, s.w, rj->eps, (hw & 1u) != 0u);
                }
                next += (uint32_t)__popcll(want);
            }
            if (__ballot(cur != kNone && r.ref == kDone) == 0) break;
        }
        if (__ballot(cur != kNone) == 0) break;
        const bool more = next < total;
        while (true) {
            const uint32_t nn = (uint32_t)__popcll(__ballot((int)r.ref >= 0));
            const uint32_t nw = (uint32_t)__popcll(__ballot((int)r.ref < -1 || (more && r.ref == kDone)));
            if (leave_nodes(nn, nw)) break;
            if ((int)r.ref >= 0) r.node<BLOCK, COUNT>(rj->sc.nodes, stk, wc);
            if ((int)r.ref >= 0) r.node<BLOCK, COUNT>(rj->sc.nodes, stk, wc);
        }
        if ((int)r.ref < -1) r.leaf<BLOCK, COUNT>(rj->sc, rj->eps, stk, wc);
    }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void flush_counter(unsigned long long* dst, uint32_t v)
{
    const uint32_t s = wave_sum(v);
    if (__lane_id() == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

// XCD-aware remap: consecutive workgroup ids land on different XCDs, so give
// each XCD a contiguous range of tiles (its L2 then sees neighbouring rays).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t g)
{
    const uint32_t q = g / 8u, r = g % 8u;
    const uint32_t x = b % 8u, i = b / 8u;
    return (x < r) ? x * (q + 1u) + i : r * (q + 1u) + (x - r) * q + i;
}

}  // namespace vkr
