// trace_kernels.hip - gfx950 kernels for the vk-renderer compute path tracer.
//
// The reference (Assets/Tracer.comp) runs one invocation per pixel that
// loops SAMPLES paths of up to DEPTH bounces and scans every primitive
// linearly per ray.  Here:
//   * each lane runs one pixel as a small state machine; in every loop
//     iteration a lane queues either its path segment or the shadow rays of
//     up to kRaySlots emitters (with their contributions, evaluated first),
//     and the wave traces all queued rays from a shared pool: a lane whose
//     ray ends takes the next one, so the wave waits for its total work, not
//     for its longest lane;
//   * persistent waves refill idle lanes from a global pixel pool with one
//     wave-wide ballot + one atomic per refill (8x8 tiles in pool order);
//   * the linear scan is replaced by a BVH (two child boxes per 64-byte node,
//     128-bit loads) with a short per-lane stack in LDS;
//   * the reference's order-dependent acceptance rule
//     (`t > EPSILON && t < best + EPSILON` for triangles then spheres,
//     `t < best - EPSILON` for planes, Tracer.comp:378-428) is reproduced
//     exactly: the smallest hit t* is always accepted by the scan, so only
//     hits within a window above t* can still change the winner; they are
//     kept in a 4-entry list and replayed in scan order after traversal.
//     A window or list overflow is counted (documented near-tie).
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.hpp"
#include "traverse.hpp"
#include "vkr_device.hpp"

namespace vkr {

// A diffuse hit as next-event estimation sees it.
constexpr uint32_t kNeeZero = 1u;     // every emitter contributes exactly +-0 (mirror, |N|^2 < 0.9999)
constexpr uint32_t kNeeFin = 2u;      // the BRDF factor is bounded for any sample
constexpr uint32_t kNeeMirror = 4u;   // metalness 1, roughness 0: zero unless (N.H)^2 rounds to 1
struct NeeHit {
    f3 P, N, V, alb;
    float rough, metal;
    uint32_t flags;                   // kNee*
};

// The next-event sample of emitter `ei` at hit h (Assets/Tracer.comp:464-502)
// with the random stream at `salt` before its two draws (the caller advances
// its own copy).  Returns false when the contribution is exactly +-0: then
// adding it cannot change e whatever the shadow ray finds, so the ray is not
// traced.  Else L is the shadow ray's direction, tl its length and c the
// contribution added when it is unoccluded.
__device__ __forceinline__ bool nee_sample(const JobDesc& j, const NeeHit& h, uint32_t ei, float salt, float uvx,
                                           float uvy, f3& L, float& tl, f3& c)
{
    const vkr_sphere* S = j.sc.spheres + j.sc.emitters[ei];
    const f3 P = h.P, N = h.N, V = h.V;
    // Emitters certified to contribute exactly +-0 before sampling need no
    // draws.  The tests carry margins far above the rounding of the fast
    // reciprocal / square-root estimates they use; an emitter they cannot
    // certify is evaluated in full.
    auto certified = [&]() __attribute__((always_inline)) -> bool {
        const float nn = dot(N, N);
        const f3 l0 = sub(ld3(S->P), P);
        const float sr = S->r;
        const float ll2 = dot(l0, l0);
        // Attenuation < 2^41: the hit point is at least 2^-9 radii from the centre.
        if (!(ll2 >= (sr * sr) * 0x1p-18f)) return false;
        // A sample L has cosa >= cos_a_max and sina <= sig, sig^2 = s2 + 1e-6
        // >= the kernel's 1 - cos_a_max^2, so it lies within t = 1.5708 sig of
        // w = l0 / |l0|.
        const float il = __builtin_amdgcn_rsqf(ll2);
        const float sw = sr * il;
        const float sig2 = fminf(sw * sw + 1e-6f, 1.0f);
        const float sig = __builtin_amdgcn_sqrtf(sig2) * (1.0f + 1e-6f);
        const float t = 1.5708f * sig;
        // Every certificate needs the Fresnel term finite: with V at least
        // 0.01 rad outside the cone, dot(H, V) <= 1 - 1e-5, so
        // pow(1 - dot(H, V), 5) never sees a negative base (NaN, :276).
        const f3 w = muls(l0, il);
        const float cph = dot(V, w);
        const float tv = t + 0.01f;
        if (!(tv < 1.5f && cph < 1.0f - 0.5f * tv * tv - 1e-6f)) return false;
        if (h.flags & kNeeZero) return true;
        if (h.flags & kNeeFin) {
            // (b) cone below the horizon: with c = cos(N, l0) < 0 and
            // s = sin(N, l0) (from |N x l0|, no cancellation),
            // cos(N, L) <= sig s + cl c < -1e-4.
            const float inv = __builtin_amdgcn_rsqf(nn * ll2);
            const float cc = dot(N, l0) * inv;
            const f3 cr = cross(N, l0);
            const float ss = __builtin_amdgcn_sqrtf(dot(cr, cr)) * inv;
            const float cl = __builtin_amdgcn_sqrtf(1.0f - sig2) * (1.0f - 1e-6f);
            if (cc < 0.0f && sig * ss + cl * cc < -1e-4f) return true;
        }
        if (h.flags & kNeeMirror) {
            // (c) mirror hits: H = normalize(V + L) stays within
            // dl >= t / (2 cos((phi0 + t) / 2)) of H0 = normalize(V + w)
            // (|dH| <= |dL| / |V + L|; phi0 = angle(V, w) < 107.5 deg, t < 1,
            // so the path never nears V + L = 0).  If |N| cos(N, H) then stays
            // at least 2e-4 away from 1, NdotH^2 cannot round to 1 and the
            // sample is +-0.
            const f3 vw = cross(V, w);
            const float sph = __builtin_amdgcn_sqrtf(dot(vw, vw));
            const float csum = (cph >= 0.0f ? cph * (1.0f - 0.5f * t * t) : cph) - sph * t;
            const float half = __builtin_amdgcn_sqrtf(fmaxf(0.5f * (1.0f + csum), 0.0f));
            if (cph > -0.3f && half >= 0.05f && t < 1.0f) {
                const float dl = t * __builtin_amdgcn_rcpf(2.0f * half) * 1.001f + 1e-5f;
                const f3 hv = add(V, w);
                const float hh = dot(hv, hv);
                const float ihn = __builtin_amdgcn_rsqf(hh * nn);
                const float c0 = dot(N, hv) * ihn;
                const f3 nh = cross(N, hv);
                const float s0 = __builtin_amdgcn_sqrtf(dot(nh, nh)) * ihn;
                const float nrm = __builtin_amdgcn_sqrtf(nn);
                const float cd = 1.0f - 0.5f * dl * dl;
                const float hi = c0 >= cd - 1e-4f ? 1.0f : (c0 >= 0.0f ? c0 : c0 * cd) + s0 * dl;
                const float lo = (c0 >= 0.0f ? c0 * cd : c0) - s0 * dl;
                if (dl < 1.0f && hh > 1e-6f && (nrm * hi < 1.0f - 2e-4f || nrm * lo > 1.0f + 2e-4f)) return true;
            }
        }
        if (h.flags & kNeeFin) {
            // (a) attenuation exactly 0 (:478-479): with the kernel's own x,
            // emis / x^2 <= kAttSub in every channel, tested with a 2^-16 margin.
            const float tl0 = length(l0) - sr;
            const float x = tl0 / sr + 1.0f;
            const float lim = (kAttSub * (x * x)) * (1.0f - 0x1p-16f);
            const f3 em = ld3(S->mat.emissive);
            if (em.x <= lim && em.y <= lim && em.z <= lim) return true;
        }
        return false;
    };
    if (h.flags && certified()) return false;

    const f3 sP = ld3(S->P);
    const float sr = S->r;
    tl = length(sub(sP, P)) - sr;
    const f3 l0 = sub(sP, P);
    const float cos_a_max = sqrtf(1.0f - gclamp(sr * sr / dot(l0, l0), 0.0f, 1.0f));
    const float cosa = gmix(cos_a_max, 1.0f, rng_next(salt, uvx, uvy));
    const float phi = kTwoPi * rng_next(salt, uvx, uvy);
    L = jitter(l0, phi, sqrtf(1.0f - cosa * cosa), cosa);

    const float x = tl / sr + 1.0f;
    f3 att = divs(muls(ld3(S->mat.emissive), 1.0f), x * x);
    att = divs(subs(att, kAttSub), kAttDiv);
    att = maxs(att, 0.0f);
    const f3 F0 = mix3(f3{0.04f, 0.04f, 0.04f}, h.alb, h.metal);
    const f3 H = normalize(add(V, L));
    const float NDF = distribution_ggx(N, H, h.rough);
    const float G = geometry_smith(N, V, L, h.rough);
    const f3 F = fresnel_schlick(gmax(dot(H, V), 0.0f), F0);
    f3 kD = ssub(1.0f, F);
    kD = muls(kD, 1.0f - h.metal);
    const f3 numerator = muls(F, NDF * G);
    const float denominator = 4.0f * gmax(dot(N, V), 0.0f) * gmax(dot(N, L), 0.0f);
    const f3 specular = divs(numerator, gmax(denominator, 0.001f));
    const float NdotL = gmax(dot(N, L), 0.0f);
    c = muls(mul(add(divs(mul(kD, h.alb), kPi), specular), att), NdotL);
    return !(c.x == 0.0f && c.y == 0.0f && c.z == 0.0f);
}

// The path program of Assets/Tracer.comp:433-593 as a per-lane state machine.
// Each loop iteration: (1) refill idle lanes, (2) queue this lane's rays
// (a path segment, or the shadow rays of the next emitters), (3) trace the
// wave's rays together, (4) shade, then run the shared tail blocks (finish
// diffuse -> Russian roulette -> end of path -> end of pixel) at most once
// each.
template <int BLOCK, int STACK, bool COUNT, bool BATCH, bool POOL>
__device__ __forceinline__ void path_program(const PathLaunch& A)
{
    __shared__ uint32_t s_stack[STACK * BLOCK];
    __shared__ uint32_t s_occ[BLOCK];                // per lane: occluded shadow rays (bit per slot)
    __shared__ uint2 s_res[BLOCK];                   // per lane: closest hit (prim, t)
    __shared__ uint32_t s_msk[BLOCK];                // per lane: slots holding a ray to trace
    __shared__ uint32_t s_cpre[BLOCK / 64 * kRaySlots];            // per wave and slot: items in lower slots
    __shared__ unsigned long long s_colm[BLOCK / 64 * kRaySlots];  // per wave and slot: lanes holding one
    uint32_t* stk = s_stack + threadIdx.x;
    const uint32_t wbase = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    const uint32_t wslot = (wbase >> 6) * kRaySlots;
    char* const wrec = A.lanes + ((size_t)blockIdx.x * (BLOCK / 64) + (wbase >> 6)) * kWaveBytes;
    // This lane's element of a one-per-lane field f is at wrec + (f + lo16):
    // a wave-uniform base and a 32-bit offset, one address form for all fields.
    const uint32_t lo16 = __lane_id() * 16u;

    // Lane state.
    bool has = false;
    bool exhausted = !A.persistent;
    uint32_t job = 0, px = 0, py = 0;
    float uvx = 0.0f, uvy = 0.0f, salt = 0.0f;
    uint32_t smp = 0, depth = 0, phase = kPhTrace;
    f3 ro{}, rd{}, acc{}, mask{}, accum{};
    f3 P{}, N{}, dj{}, e{};
    const vkr_material* hm = nullptr;   // material of the current diffuse hit
    uint32_t ei = 0;
    uint32_t rays_px = 0, drec = 0, drep = 0;
    uint32_t c_rays = 0, c_shadow = 0, c_traced = 0, wave_steps = 0;
    bool new_path = false;
    uint32_t nee_flags = 0;             // kNee* of the current diffuse hit
    WorkCnt wc{};

    // Claims pool index g for this lane (no-op for indices outside the image).
    auto claim = [&](uint32_t g) __attribute__((always_inline)) {
        uint32_t local = g;
        if constexpr (BATCH) {
            uint32_t lo = 0, hi = A.n_jobs;
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (A.job_start[mid] <= g) lo = mid; else hi = mid;
            }
            job = lo;
            local = g - A.job_start[lo];
        }
        const JobDesc& j = BATCH ? A.jobs[job] : A.job;
        const uint32_t tile = local >> 6, within = local & 63u;
        if (tile >= j.n_tiles) return;
        const uint32_t x = (tile % j.tiles_x) * 8u + (within & 7u);
        const uint32_t y = (tile / j.tiles_x) * 8u + (within >> 3);
        if (x >= j.W || y >= j.H) return;
        px = x;
        py = y;
        has = true;
        uvx = (float)x / (float)j.W;                 // :563
        uvy = (float)y / (float)j.H;
        salt = uvy / uvx + j.seed;                   // :565-566
        smp = 0;
        // The sum of finished samples: in the lane's record with the ray
        // pool (fewer registers live across the step), else in registers.
        if constexpr (POOL) *reinterpret_cast<float4*>(wrec + (kWPix + lo16)) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        else accum = f3{0.0f, 0.0f, 0.0f};
        rays_px = 0;
        new_path = true;
    };

    if (!A.persistent) {
        const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
        const uint32_t g = A.base + b * BLOCK + threadIdx.x;
        if (g < A.total) claim(g);
    }

    while (true) {
        if (A.persistent && !exhausted) {
            const unsigned long long need = __ballot(!has);
            const unsigned long long busy = __ballot(has);
            const uint32_t nneed = (uint32_t)__popcll(need);
            if (nneed > 0 && (nneed >= A.chunk || busy == 0)) {
                const int leader = __ffsll((long long)need) - 1;
                uint32_t base = 0;
                if (__lane_id() == leader) base = atomicAdd(A.pool, nneed);
                base = __shfl(base, leader, 64);
                if (base + nneed >= A.total) exhausted = true;
                if (!has) {
                    const uint32_t rank = (uint32_t)__popcll(need & ((1ull << __lane_id()) - 1ull));
                    const uint32_t g = base + rank;
                    if (g < A.total) claim(g);
                }
            }
        }
        const unsigned long long live = __ballot(has);
        if (live == 0) {
            if (exhausted) break;
            continue;
        }
        if (COUNT && __lane_id() == 0) ++wave_steps;

        const JobDesc& j = BATCH ? A.jobs[job] : A.job;
        if (has && new_path) {
            // :561-574 primary ray; every sample restarts from it (:578-581).
            new_path = false;
            const float tx = 2.0f * uvx - 1.0f;
            const float ty = 2.0f * uvy - 1.0f;
            f3 dir = add(add(ld3(j.cam_dir), muls(ld3(j.cam_right), tx)), muls(ld3(j.cam_up), ty));
            dir = mul(dir, f3{j.aspect, 1.0f, j.aspect});
            ro = ld3(j.cam_pos);
            rd = normalize(dir);
            acc = f3{0.0f, 0.0f, 0.0f};
            mask = f3{1.0f, 1.0f, 1.0f};
            depth = 0;
            phase = kPhTrace;
            drec = 0;
            drep = 0;
            if ((j.flags & VKR_TRACE_REPLAY_DECISIONS) && j.decisions)
                drep = j.decisions[((size_t)py * j.W + px) * j.spp + smp];
            // The adjoint's first pass, kept on request: where this sample's
            // random stream starts (its clamp bits follow, per bounce).
            if (j.adj_samples) j.adj_samples[((size_t)py * j.W + px) * j.spp + smp] = make_uint2(__float_as_uint(salt), 0u);
        }

        // This step's rays: with POOL up to kRaySlots in the lane's record,
        // else at most one, kept in registers (o0, d0, t0max; c0 = its
        // next-event contribution).
        uint32_t nray = 0;
        uint32_t nitem = 0;  // POOL: emitters this lane published for the wave to evaluate
        f3 o0{}, d0{}, c0{};
        float t0max = 0.0f;
        if (has) {
            if (phase == kPhTrace) {
                if (j.adj_samples) {
                    // Channels that clamp(acc) saturates at this bounce, 10 bits per channel.
                    const uint32_t bit = 1u << depth;
                    const uint32_t zb = ((acc.x > 1.0f || acc.x < 0.0f) ? bit : 0u) |
                                        ((acc.y > 1.0f || acc.y < 0.0f) ? bit << 10 : 0u) |
                                        ((acc.z > 1.0f || acc.z < 0.0f) ? bit << 20 : 0u);
                    if (zb) j.adj_samples[((size_t)py * j.W + px) * j.spp + smp].y |= zb;
                }
                acc = clamps(acc, 0.0f, 1.0f);                      // :441
                const float dp1 = (float)(depth + 1u);
                const float tmax = 3000.0f / (dp1 * dp1);            // :444
                if constexpr (POOL) {
                    *reinterpret_cast<float4*>(wrec + (kWHead + lo16)) =
                        make_float4(ro.x, ro.y, ro.z, __uint_as_float(BATCH ? job << 1 : 0u));
                    *reinterpret_cast<float4*>(wrec + (kWSlotD + lo16)) = make_float4(rd.x, rd.y, rd.z, tmax);
                } else {
                    o0 = ro;
                    d0 = rd;
                    t0max = tmax;
                }
                nray = 1;
                ++rays_px;
                ++c_rays;
                ++c_traced;
            } else {
                // Next-event samples (:464-469) and their unoccluded
                // contributions (:475-502), evaluated before the shadow rays
                // (nee_sample).  Rays that must be traced are added to e in
                // emitter order once their outcomes are known: the same
                // additions in the same order as the reference's loop.
                NeeHit h;
                h.P = P;
                h.N = N;
                h.V = normalize(sub(ld3(j.cam_pos), P));
                h.alb = ld3(hm->albedo);
                h.rough = hm->roughness;
                h.metal = hm->metalness;
                h.flags = nee_flags;
                if constexpr (POOL) {
                    // Publish the hit and its next window of up to kRaySlots
                    // emitters; the wave evaluates all (lane, emitter) pairs
                    // together below.  Each sample draws twice (:464-465), so
                    // the stream position before each one is known now.
                    const uint32_t w = min(kRaySlots, j.sc.n_emitters - ei);
                    *reinterpret_cast<float4*>(wrec + (kWHit + lo16)) = make_float4(N.x, N.y, N.z, uvx);
                    *reinterpret_cast<float4*>(wrec + (kWHit + 1024u + lo16)) = make_float4(h.alb.x, h.alb.y, h.alb.z, h.rough);
                    *reinterpret_cast<float4*>(wrec + (kWHit + 2048u + lo16)) =
                        make_float4(h.metal, uvy, __uint_as_float(ei), __uint_as_float(nee_flags));
                    float* sk = reinterpret_cast<float*>(wrec + (kWSalt + __lane_id() * 4u));
                    float sv = salt;
                    for (uint32_t k = 0; k < w; ++k) {
                        sk[k * 64u] = sv;
                        sv = sv + 1.0f;
                        sv = sv + 1.0f;
                    }
                    salt = sv;
                    rays_px += w;
                    c_rays += w;
                    c_shadow += w;
                    ei += w;
                    nitem = w;
                    *reinterpret_cast<float4*>(wrec + (kWHead + lo16)) =
                        make_float4(P.x, P.y, P.z, __uint_as_float(BATCH ? (job << 1) | 1u : 1u));
                } else {
                    while (ei < j.sc.n_emitters) {
                        const float s0 = salt;
                        salt = salt + 1.0f;      // the sample's two draws (:464-465)
                        salt = salt + 1.0f;
                        ++rays_px;
                        ++c_rays;
                        ++c_shadow;
                        if (nee_sample(j, h, ei++, s0, uvx, uvy, d0, t0max, c0)) {
                            ++c_traced;
                            o0 = P;
                            nray = 1;
                            break;
                        }
                    }
                }
            }
        }

        if constexpr (POOL) {
            // Path state that the wave's evaluation and tracing do not need
            // waits in the record, so it holds no registers meanwhile.
            float4* park = reinterpret_cast<float4*>(wrec + (kWPark + lo16));
            park[0] = make_float4(ro.x, ro.y, ro.z, dj.x);
            park[64] = make_float4(rd.x, rd.y, rd.z, dj.y);
            park[128] = make_float4(P.x, P.y, P.z, dj.z);
            park[192] = make_float4(N.x, N.y, N.z, 0.0f);
            // The wave evaluates every published (lane, emitter) pair, slot
            // by slot (neighbouring lanes take one emitter of neighbouring
            // owners, whose records are neighbours).  A nonzero contribution
            // goes to slot k of the owner (emitter ei0 + k) and sets bit k
            // of its mask.
            const uint32_t tot = pool_columns(nitem >= 32u ? ~0u : (1u << nitem) - 1u, s_cpre + wslot, s_colm + wslot);
            s_msk[threadIdx.x] = 0u;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t base = 0; base < tot; base += 64u) {
                const uint32_t e = base + __lane_id();
                if (e < tot) {
                    uint32_t k, lo;
                    pool_item(s_cpre + wslot, s_colm + wslot, e, k, lo);
                    const uint32_t oo = lo * 16u;
                    const float4 hd = *reinterpret_cast<const float4*>(wrec + (kWHead + oo));
                    const float4 h0 = *reinterpret_cast<const float4*>(wrec + (kWHit + oo));
                    const float4 h1 = *reinterpret_cast<const float4*>(wrec + (kWHit + 1024u + oo));
                    const float4 h2 = *reinterpret_cast<const float4*>(wrec + (kWHit + 2048u + oo));
                    const float sk = *reinterpret_cast<const float*>(wrec + (kWSalt + (k * 64u + lo) * 4u));
                    const JobDesc& oj = BATCH ? A.jobs[__float_as_uint(hd.w) >> 1] : A.job;
                    NeeHit h;
                    h.P = f3{hd.x, hd.y, hd.z};
                    h.N = f3{h0.x, h0.y, h0.z};
                    h.V = normalize(sub(ld3(oj.cam_pos), h.P));
                    h.alb = f3{h1.x, h1.y, h1.z};
                    h.rough = h1.w;
                    h.metal = h2.x;
                    h.flags = __float_as_uint(h2.w);
                    f3 L, c;
                    float tl;
                    if (nee_sample(oj, h, __float_as_uint(h2.z) + k, sk, h0.w, h2.y, L, tl, c)) {
                        const uint32_t so = (k * 64u + lo) * 16u;
                        *reinterpret_cast<float4*>(wrec + (kWSlotD + so)) = make_float4(L.x, L.y, L.z, tl);
                        *reinterpret_cast<float4*>(wrec + (kWSlotC + so)) = make_float4(c.x, c.y, c.z, 0.0f);
                        atomicOr(s_msk + wbase + lo, 1u << k);
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (has && phase != kPhTrace) {
                nray = (uint32_t)__popc(s_msk[threadIdx.x]);
                c_traced += nray;
            } else if (nray) {
                s_msk[threadIdx.x] = 1u;   // the path segment is slot 0
            }
        }

        TravOut tr{kMiss, 0.0f, false};
        uint32_t occ = 0;    // NEE lanes: bit k set when the shadow ray of slot k is occluded
        if constexpr (!POOL) {
            // One ray per lane, from registers.
            const TravOut t = traverse<BLOCK, COUNT>(j.sc, o0, d0, t0max, j.eps, phase != kPhTrace, nray != 0u, stk, wc);
            if (phase == kPhTrace) tr = t;
            else occ = t.hit ? 1u : 0u;
        } else {
            // The wave traces all queued rays together from its pool, slot
            // by slot: column k holds slot k of every lane that queued one,
            // lanes in order.  Neighbouring lanes then take rays of one slot
            // (one emitter, or the path segments) from neighbouring pixels,
            // which visit much the same nodes.
            const float4* park = reinterpret_cast<const float4*>(wrec + (kWPark + lo16));
            const uint32_t tot = pool_columns(s_msk[threadIdx.x], s_cpre + wslot, s_colm + wslot);
            s_occ[threadIdx.x] = 0u;
            // The pool, the records and the result words are shared by the
            // wave's lanes: order the writes above before the reads below.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            traverse_pool<BLOCK, COUNT, BATCH>(BATCH ? A.jobs : &A.job, tot, s_cpre + wslot, s_colm + wslot,
                                               s_occ + wbase, s_res + wbase, wrec, stk, wc);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const float4 k0 = park[0], k1 = park[64], k2 = park[128], k3 = park[192];
            ro = f3{k0.x, k0.y, k0.z};
            rd = f3{k1.x, k1.y, k1.z};
            P = f3{k2.x, k2.y, k2.z};
            N = f3{k3.x, k3.y, k3.z};
            dj = f3{k0.w, k1.w, k2.w};
            if (phase == kPhTrace) {
                const uint2 r = s_res[threadIdx.x];
                tr.prim = r.x;
                tr.t = __uint_as_float(r.y);
                tr.hit = r.x != kMiss;
            } else {
                occ = s_occ[threadIdx.x];
            }
        }
        if (!has) continue;

        bool fin_diffuse = false, do_rr = false, path_end = false;
        if (phase == kPhTrace) {
            if (smp == 0 && depth == 0) {                            // the primary hit, written at once
                const size_t pi = (size_t)py * j.W + px;
                if (j.prim_id) j.prim_id[pi] = tr.prim;
                if (j.hit_t) j.hit_t[pi] = tr.t;
            }
            if (!tr.hit) {
                path_end = true;                                     // :445
            } else {
                P = add(ro, muls(rd, tr.t));
                const uint32_t kind = tr.prim >> 30, idx = tr.prim & 0x3FFFFFFFu;
                const vkr_material* m;
                if (kind == 0) {
                    const vkr_triangle* T = j.sc.tris_orig + idx;
                    const f3 v0 = ld3(T->v0), v1 = ld3(T->v1), v2 = ld3(T->v2);
                    const f3 u = sub(v1, v0), v = sub(v2, v0);
                    N = f3{(u.y * v.z) - (u.z * v.y), (u.z * v.x) - (u.x * v.z), (u.x * v.y) - (u.y * v.x)};  // :392
                    const uint32_t mi = T->mat_index;
                    m = ((j.sc.flags & VKR_SCENE_TRI_MATERIAL_INDEX) && mi < j.sc.n_materials) ? j.sc.materials + mi
                                                                                               : j.sc.tri_material;
                } else if (kind == 1) {
                    const vkr_sphere* S = j.sc.spheres + idx;
                    N = divs(sub(P, ld3(S->P)), S->r);               // :408
                    m = &S->mat;
                } else {
                    const vkr_plane* Q = j.sc.planes + idx;
                    N = ld3(Q->N);                                   // :424
                    m = &Q->mat;
                }
                hm = m;
                const uint32_t type = m->type;
                if (type == VKR_MAT_DIFFUSE) {                       // :451-513
                    const float r2 = rng_next(salt, uvx, uvy);
                    const float phi = kTwoPi * rng_next(salt, uvx, uvy);
                    dj = muls(jitter(N, phi, sqrtf(r2), sqrtf(1.0f - r2)), 1.0f - m->metalness);
                    e = f3{0.0f, 0.0f, 0.0f};
                    ei = 0;
                    // Metalness 1 and roughness 0 (the reference's `mirror`):
                    // kD = (1 - F) * 0 and NDF = 0 / (pi * denom^2), where
                    // denom = 1 - NdotH^2 >= 8e-5 once |N|^2 < 0.9999 (|H| = 1),
                    // so with finite albedo, F, G and attenuation an emitter's
                    // contribution is exactly +-0 whatever its sample (F is
                    // finite when V stays outside the emitter's cone, which is
                    // checked per emitter).
                    //
                    // More generally, with |albedo|, |metalness| <= 16,
                    // k = (r+1)^2/8 in [2^-10, 1], |N|^2 <= 2^10 and the GGX
                    // denominator 1 - NdotH^2 (1 - a2) >= 2^-10 for every
                    // |H| = 1 (or a2 >= 1), the factor (kD alb/pi + specular)
                    // is below 2^72 for any sample: F <= 2^10, G <= 2^30,
                    // NDF <= 2^22.  An emitter whose attenuation is exactly 0,
                    // or whose whole sampling cone lies below the horizon
                    // (NdotL = +0), then contributes exactly +-0 as well.
                    {
                        const f3 alb = ld3(m->albedo);
                        const float nn = dot(N, N);
                        const float rough = m->roughness, metal = m->metalness;
                        const bool nee_zero = j.sc.emitters_bounded && metal == 1.0f && rough == 0.0f &&
                                   fabsf(alb.x) < 1e30f && fabsf(alb.y) < 1e30f && fabsf(alb.z) < 1e30f &&
                                   nn < 0.9999f;
                        const float rp = rough + 1.0f;
                        const float k = (rp * rp) / 8.0f;
                        const float a = rough * rough;
                        const float a2 = a * a;
                        const bool nee_fin = j.sc.emitters_bounded && fabsf(alb.x) <= 16.0f && fabsf(alb.y) <= 16.0f &&
                                  fabsf(alb.z) <= 16.0f && fabsf(metal) <= 16.0f && k >= 0x1p-10f && k <= 1.0f &&
                                  nn <= 1024.0f &&
                                  (a2 >= 1.0f || 1.0f - (nn * (1.0f + 0x1p-16f)) * (1.0f - a2) >= 0x1p-10f);
                        // Other mirror hits: kD = 0 and NDF = 0 / (pi (1 - NdotH^2)^2)
                        // is 0 unless NdotH^2 rounds to exactly 1 (then NaN).
                        const bool nee_mirror = !nee_zero && j.sc.emitters_bounded && metal == 1.0f &&
                                                rough == 0.0f && fabsf(alb.x) <= 16.0f && fabsf(alb.y) <= 16.0f &&
                                                fabsf(alb.z) <= 16.0f && nn <= 1024.0f;
                        nee_flags = (nee_zero ? kNeeZero : 0u) | (nee_fin ? kNeeFin : 0u) | (nee_mirror ? kNeeMirror : 0u);
                    }
                    if (j.sc.n_emitters > 0) phase = kPhNee;
                    else fin_diffuse = true;
                } else if (type == VKR_MAT_DIELECTRIC) {             // :514-542
                    const f3 emis = ld3(m->emissive);
                    const float rough = m->roughness;
                    acc = add(acc, mul(emis, mask));
                    mask = mul(mask, ld3(m->albedo));
                    const float cosine = -dot(rd, N) / length(rd);
                    const f3 reflected = reflect(rd, N);
                    const f3 refracted = refract(rd, N, rough);
                    const float prob = eq3(refracted, f3{0.0f, 0.0f, 0.0f}) ? 1.0f : schlick(cosine, rough);
                    const float u = rng_next(salt, uvx, uvy);
                    bool refl = u < prob;
                    if (depth < 8u) {
                        if (j.flags & VKR_TRACE_REPLAY_DECISIONS) refl = (drep >> (2u * depth + 1u)) & 1u;
                        if (refl) drec |= 1u << (2u * depth + 1u);
                    }
                    rd = normalize(refl ? reflected : refracted);
                    ro = P;
                    do_rr = true;
                } else {
                    do_rr = true;                                    // no case in the switch
                }
            }
        } else {
            if constexpr (POOL) {
                uint32_t bits = s_msk[threadIdx.x] & ~occ;           // unoccluded (:473), in emitter order
                while (bits) {
                    const uint32_t k = (uint32_t)__builtin_ctz(bits);
                    bits &= bits - 1u;
                    const float4 c = *reinterpret_cast<const float4*>(wrec + (kWSlotC + k * 1024u + lo16));
                    e = add(e, f3{c.x, c.y, c.z});                   // :502
                }
            } else if (nray && !(occ & 1u)) {
                e = add(e, c0);                                      // :473, :502
            }
            if (ei >= j.sc.n_emitters) fin_diffuse = true;
        }

        if (fin_diffuse) {                                           // :507-511
            const f3 emis = ld3(hm->emissive);
            const bool all_pos = gmax(gsign(emis.x), 0.0f) == 1.0f && gmax(gsign(emis.y), 0.0f) == 1.0f &&
                                 gmax(gsign(emis.z), 0.0f) == 1.0f;
            const f3 em = all_pos ? normalize(emis) : f3{0.0f, 0.0f, 0.0f};
            acc = add(acc, mul(mask, add(em, e)));
            mask = mul(mask, ld3(hm->albedo));
            rd = normalize(add(reflect(rd, N), dj));
            ro = P;
            phase = kPhTrace;
            do_rr = true;
        }
        if (do_rr) {                                                 // :545-549
            const float p = max3(mask);
            const float u = rng_next(salt, uvx, uvy);
            bool term = u > p;
            if (depth < 8u) {
                if (j.flags & VKR_TRACE_REPLAY_DECISIONS) term = !((drep >> (2u * depth)) & 1u);
                if (!term) drec |= 1u << (2u * depth);
            }
            if (term) {
                path_end = true;
            } else {
                mask = muls(mask, 1.0f / p);
                ++depth;
                if (depth >= j.max_depth) path_end = true;
                phase = kPhTrace;
            }
        }
        if (path_end) {                                              // :578-581
            if ((j.flags & VKR_TRACE_RECORD_DECISIONS) && j.decisions)
                j.decisions[((size_t)py * j.W + px) * j.spp + smp] = (uint16_t)drec;
            if constexpr (POOL) {
                float4* ap = reinterpret_cast<float4*>(wrec + (kWPix + lo16));
                const float4 a = *ap;
                accum = add(f3{a.x, a.y, a.z}, acc);
                *ap = make_float4(accum.x, accum.y, accum.z, 0.0f);
            } else {
                accum = add(accum, acc);
            }
            ++smp;
            if (smp < j.spp) {
                new_path = true;
            } else {                                                 // :585-592
                const size_t pi = (size_t)py * j.W + px;
                f3 c = divs(accum, (float)j.spp);
                if (j.radiance) reinterpret_cast<float4*>(j.radiance)[pi] = make_float4(c.x, c.y, c.z, 1.0f);
                c = divv(c, adds(c, 1.0f));
                c = f3{pow_f(c.x, kInvGamma), pow_f(c.y, kInvGamma), pow_f(c.z, kInvGamma)};
                const float dither = rng_next(salt, uvx, uvy) / 64.0f;
                c = adds(c, dither);
                if (j.rgba8)
                    reinterpret_cast<uint32_t*>(j.rgba8)[pi] =
                        unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (255u << 24);
                if (j.ray_count) j.ray_count[pi] = rays_px;
                has = false;
            }
        }
    }

    flush_counter(&A.counters->rays, c_rays);
    flush_counter(&A.counters->shadow, c_shadow);
    flush_counter(&A.counters->traced, c_traced);
    flush_counter(&A.counters->overflow, wc.overflow);
    if (COUNT) {
        flush_counter(&A.counters->nodes, wc.nodes);
        flush_counter(&A.counters->tris, wc.tris);
        flush_counter(&A.counters->spheres, wc.spheres);
        flush_counter(&A.counters->planes, wc.planes);
        flush_counter(&A.counters->wave_nodes, wc.wave_nodes);
        flush_counter(&A.counters->wave_leaves, wc.wave_leaves);
        flush_counter(&A.counters->wave_steps, wave_steps);
    }
}

// The timed kernel and the work-counting variant (same traversal order, more
// registers).  The traversal waits on dependent node fetches, so resident
// waves set the speed: with a 24-entry stack a 128-lane block needs 15.1 KB
// of LDS, ten fit a CU, and the kernel is held to 96 VGPRs for 5 waves per
// SIMD (its spills are all in per-step code, outside the pool's traversal
// loops); with deeper stacks LDS allows 4 waves (128 VGPRs).
template <int BLOCK, int STACK, bool BATCH, bool POOL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(STACK <= 24 ? 5 : 4))) void k_path(const PathLaunch A)
{
    path_program<BLOCK, STACK, false, BATCH, POOL>(A);
}

template <int BLOCK, int STACK, bool BATCH, bool POOL>
__global__ __launch_bounds__(BLOCK) void k_path_count(const PathLaunch A)
{
    path_program<BLOCK, STACK, true, BATCH, POOL>(A);
}

// ---------------------------------------------------------------------------
// Whitted kernel, Assets/Raytracer.comp (its own constant scene, :98-127).
// ---------------------------------------------------------------------------
namespace whitted {
constexpr float kEps = 0.01f;  // :89
struct Sph { uint32_t mat; float o[3]; float r; };
struct Pln { uint32_t mat; float n[3]; float d; };
__constant__ Sph kSph[2] = {{5, {-14.0f, 12.0f, 32.0f}, 5.0f}, {4, {32.0f, 24.0f, 25.0f}, 12.0f}};
__constant__ Pln kPln[5] = {{0, {0.0f, 1.0f, 0.0f}, 0.0f},
                            {0, {0.0f, -1.0f, 0.0f}, 128.0f},
                            {2, {0.0f, 0.0f, -1.0f}, 64.0f},
                            {1, {1.0f, 0.0f, 0.0f}, 64.0f},
                            {3, {-1.0f, 0.0f, 0.0f}, 64.0f}};
__constant__ float kTri[9] = {10.0f, 10.0f, 0.0f, 0.0f, 20.0f, 0.0f, -10.0f, 10.0f, 0.0f};
__constant__ uint32_t kRefl[6] = {1, 0, 1, 0, 0, 0};
__constant__ float kDiff[18] = {1, 1, 1, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 1, 0, 1, 0, 1};

struct Hit { uint32_t mat; float dist; f3 P, N; uint32_t prim; };

__device__ __forceinline__ float wplane_t(f3 o, f3 d, f3 n, float dist)  // :180-192
{
    const float dd = dot(d, n);
    if (dd == 0.0f) return 0.0f;
    return gmax(-(dist + dot(o, n)) / dd, 0.0f);
}

__device__ bool trace(f3 o, f3 d, Hit& h)                                  // :224-278
{
    bool found = false;
    const f3 v0 = ld3(kTri), v1 = ld3(kTri + 3), v2 = ld3(kTri + 6);
    {
        const float t = tri_t(o, d, v0, sub(v1, v0), sub(v2, v0), kEps);
        if (t > kEps && t < h.dist) {
            h.dist = t;
            h.mat = 1;
            h.P = add(o, muls(d, t));
            const f3 u = sub(v1, v0), v = sub(v2, v0);
            h.N = f3{(u.y * v.z) - (u.z * v.y), (u.z * v.x) - (u.x * v.z), (u.x * v.y) - (u.y * v.x)};
            h.prim = 0;
            found = true;
        }
    }
    for (int i = 0; i < 2; ++i) {
        const f3 c = ld3(kSph[i].o);
        const float t = sphere_t(o, d, c, kSph[i].r);
        if (t > kEps && t < h.dist) {
            h.dist = t;
            h.mat = kSph[i].mat;
            h.P = add(o, muls(d, t));
            h.N = divs(sub(h.P, c), kSph[i].r);
            h.prim = (1u << 30) | (uint32_t)i;
            found = true;
        }
    }
    for (int i = 0; i < 5; ++i) {
        const float t = wplane_t(o, d, ld3(kPln[i].n), kPln[i].d);
        if (t > kEps && t < h.dist) {
            h.dist = t;
            h.mat = kPln[i].mat;
            h.P = add(o, muls(d, t));
            h.N = ld3(kPln[i].n);
            h.prim = (2u << 30) | (uint32_t)i;
            found = true;
        }
    }
    return found;
}


struct C4 { float r, g, b, a; };

__device__ C4 render(const JobDesc& j, f3& ro, f3& rd, uint32_t& bounce, uint32_t& rays, bool& cap,
                     uint32_t& prim, float& pt)                             // :280-355
{
    Hit h;
    h.dist = 1000.0f;                                                       // :88,290
    h.prim = kMiss;
    h.mat = 0;
    ++rays;
    const bool hit = trace(ro, rd, h);
    if (cap) { cap = false; prim = hit ? h.prim : kMiss; pt = h.dist; }
    if (!hit) return C4{0.0f, 0.0f, 0.0f, 1.0f};
    const f3 lp = ld3(j.light_pos);
    const f3 light_vec = normalize(sub(lp, h.P));
    const float dist_to_light = length(sub(lp, h.P));
    const f3 mdiff = ld3(kDiff + 3 * h.mat);
    const float li = 540.0f / (12.56637072f * dist_to_light);               // :308
    const f3 diffuse = muls(muls(mdiff, li), gmax(dot(h.N, light_vec), 0.0f));
    const f3 half_vec = normalize(add(light_vec, normalize(ld3(j.cam_pos))));
    const float spec = li * pow16(gclamp(dot(h.N, half_vec), 0.0f, 1.0f));
    f3 color = adds(diffuse, spec);
    Hit sh;
    sh.dist = dist_to_light;
    sh.prim = kMiss;
    sh.mat = 0;
    ++rays;
    if (trace(h.P, light_vec, sh)) {                                        // :322-340
        color = muls(color, 0.5f);
        bounce = 3;
    }
    if (kRefl[h.mat]) {                                                     // :343-351
        rd = sub(rd, muls(h.N, 2.0f * dot(h.N, rd)));
        ro = h.P;
    } else {
        bounce = 3;
    }
    return C4{color.x, color.y, color.z, 1.0f};
}
}  // namespace whitted

__global__ __launch_bounds__(256) void k_whitted(const JobDesc j, DevCounters* counters)
{
    const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
    const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
    uint32_t rays = 0;
    if (x < j.W && y < j.H) {
        const float uvx = (float)x / (float)j.W, uvy = (float)y / (float)j.H;   // :362
        const float tx = 2.0f * uvx - 1.0f, ty = 2.0f * uvy - 1.0f;             // :366
        f3 dir = add(add(ld3(j.cam_dir), muls(ld3(j.cam_right), tx)), muls(ld3(j.cam_up), ty));
        dir = mul(dir, f3{j.aspect, 1.0f, j.aspect});                           // :370
        f3 ro = ld3(j.cam_pos), rd = normalize(dir);
        uint32_t bounce = 0, prim = kMiss;
        bool cap = true;
        float pt = 0.0f;
        whitted::C4 fc = whitted::render(j, ro, rd, bounce, rays, cap, prim, pt);
        float s = 0.4f;                                                         // :382
        while (++bounce <= 2u) {                                                // :384-394
            const whitted::C4 rc = whitted::render(j, ro, rd, bounce, rays, cap, prim, pt);
            const float a = 1.0f - s;
            const float ma = 1.0f - a;
            const whitted::C4 m = {rc.r * ma + fc.r * a, rc.g * ma + fc.g * a, rc.b * ma + fc.b * a,
                                   rc.a * ma + fc.a * a};
            fc = whitted::C4{fc.r * (1.0f - s) + m.r * s, fc.g * (1.0f - s) + m.g * s,
                             fc.b * (1.0f - s) + m.b * s, fc.a * (1.0f - s) + m.a * s};
            s = s * 0.5f;
        }
        const size_t pi = (size_t)y * j.W + x;
        if (j.radiance) reinterpret_cast<float4*>(j.radiance)[pi] = make_float4(fc.r, fc.g, fc.b, fc.a);
        if (j.rgba8)
            reinterpret_cast<uint32_t*>(j.rgba8)[pi] =
                unorm8(fc.r) | (unorm8(fc.g) << 8) | (unorm8(fc.b) << 16) | (unorm8(fc.a) << 24);
        if (j.prim_id) j.prim_id[pi] = prim;
        if (j.hit_t) j.hit_t[pi] = pt;
        if (j.ray_count) j.ray_count[pi] = rays;
    }
    flush_counter(&counters->rays, rays);
    flush_counter(&counters->traced, rays);
}

// ---------------------------------------------------------------------------
// Present pass, Assets/Fullscreen.frag:14-31 (+ Fullscreen.vert:6-11).
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 sample_bilinear(const uint8_t* img, uint32_t w, uint32_t h, float u, float v)
{
    // VK_FILTER_LINEAR, CLAMP_TO_BORDER with opaque black (GraphicsDevice.cpp:770-790).
    const float x = u * (float)w - 0.5f;
    const float y = v * (float)h - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int x0 = (int)fx0, y0 = (int)fy0;
    f3 acc{0.0f, 0.0f, 0.0f};
    for (int dy = 0; dy < 2; ++dy) {
        for (int dx = 0; dx < 2; ++dx) {
            const int xi = x0 + dx, yi = y0 + dy;
            const float wgt = (dx ? ax : 1.0f - ax) * (dy ? ay : 1.0f - ay);
            if (xi < 0 || yi < 0 || xi >= (int)w || yi >= (int)h) continue;
            const uint32_t t = reinterpret_cast<const uint32_t*>(img)[(size_t)yi * w + (size_t)xi];
            acc.x = acc.x + ((float)(t & 255u) / 255.0f) * wgt;
            acc.y = acc.y + ((float)((t >> 8) & 255u) / 255.0f) * wgt;
            acc.z = acc.z + ((float)((t >> 16) & 255u) / 255.0f) * wgt;
        }
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_filter(const uint8_t* cur, const uint8_t* prev, uint32_t sw, uint32_t sh,
                                                uint8_t* out, uint32_t dw, uint32_t dh)
{
    const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
    const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
    if (x >= dw || y >= dh) return;
    const float inv_size = 1.0f / 2048.0f;                                  // :12
    const float u = ((float)x + 0.5f) / (float)dw;
    const float v = 1.0f - ((float)y + 0.5f) / (float)dh;                   // :16
    const f3 c0 = sample_bilinear(cur, sw, sh, u, v);
    const f3 c1 = sample_bilinear(prev, sw, sh, u, v);
    const f3 n = sample_bilinear(cur, sw, sh, u + 0.0f * inv_size, v + 1.0f * inv_size);
    const f3 s = sample_bilinear(cur, sw, sh, u + 0.0f * inv_size, v + -1.0f * inv_size);
    const f3 e = sample_bilinear(cur, sw, sh, u + 1.0f * inv_size, v + 0.0f * inv_size);
    const f3 w = sample_bilinear(cur, sw, sh, u + -1.0f * inv_size, v + 0.0f * inv_size);
    const f3 dl = sub(c0, c1);
    const float vT = dot(dl, dl);                                           // :26
    f3 fin;
    if (vT > 0.0005f) fin = divs(add(add(add(n, s), e), w), 4.0f);         // :28
    else fin = c0;
    reinterpret_cast<uint32_t*>(out)[(size_t)y * dw + x] =
        unorm8(fin.x) | (unorm8(fin.y) << 8) | (unorm8(fin.z) << 16) | (255u << 24);
}

// ---------------------------------------------------------------------------
// Building-block evaluation for parity tests (op codes in vkr.h).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_debug(uint32_t op, const float* in, float* out, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float* a = in + (size_t)i * 16;
    float* r = out + (size_t)i * 8;
    switch (op) {
    case VKR_DEBUG_RAND: {  // in: cx cy salt0 -> out: 8 successive rand()
        float salt = a[2];
        for (int k = 0; k < 8; ++k) r[k] = rng_next(salt, a[0], a[1]);
        break;
    }
    case VKR_DEBUG_TRI:     // in: o(3) d(3) v0(3) v1(3) eps ... (v2 at 12..14 via a[12]) -> t
        r[0] = tri_t(f3{a[0], a[1], a[2]}, f3{a[3], a[4], a[5]}, f3{a[6], a[7], a[8]},
                     sub(f3{a[9], a[10], a[11]}, f3{a[6], a[7], a[8]}),
                     sub(f3{a[12], a[13], a[14]}, f3{a[6], a[7], a[8]}), a[15]);
        break;
    case VKR_DEBUG_SPHERE:  // in: o(3) d(3) c(3) r -> t
        r[0] = sphere_t(f3{a[0], a[1], a[2]}, f3{a[3], a[4], a[5]}, f3{a[6], a[7], a[8]}, a[9]);
        break;
    case VKR_DEBUG_PLANE:   // in: o(3) d(3) N(3) len -> t
        r[0] = plane_t(f3{a[0], a[1], a[2]}, f3{a[3], a[4], a[5]}, f3{a[6], a[7], a[8]}, a[9]);
        break;
    case VKR_DEBUG_BRDF: {  // in: N(3) H(3) V(3) L(3) rough cos F0 -> NDF, G, F(3), schlick(cos, rough)
        const f3 Nn{a[0], a[1], a[2]}, H{a[3], a[4], a[5]}, V{a[6], a[7], a[8]}, L{a[9], a[10], a[11]};
        r[0] = distribution_ggx(Nn, H, a[12]);
        r[1] = geometry_smith(Nn, V, L, a[12]);
        const f3 F = fresnel_schlick(a[13], f3{a[14], a[14], a[14]});
        r[2] = F.x; r[3] = F.y; r[4] = F.z;
        r[5] = schlick(a[13], a[12]);
        break;
    }
    case VKR_DEBUG_JITTER: {  // in: d(3) phi sina cosa -> out(3)
        const f3 v = jitter(f3{a[0], a[1], a[2]}, a[3], a[4], a[5]);
        r[0] = v.x; r[1] = v.y; r[2] = v.z;
        break;
    }
    case VKR_DEBUG_COS:     // in: x -> cos, sin, pow(x, 1/2.2)
        r[0] = cos_f(a[0]);
        r[1] = sin_f(a[0]);
        r[2] = pow_f(a[0], kInvGamma);
        break;
    default:
        break;
    }
}

__global__ void k_zero(DevCounters* c, uint32_t* pool)
{
    if (threadIdx.x == 0) {
        c->rays = c->shadow = c->traced = c->nodes = c->tris = c->spheres = c->planes = c->overflow = 0ull;
        c->wave_nodes = c->wave_leaves = c->wave_steps = 0ull;
        if (pool) {
            pool[0] = 0u;
            pool[1] = 0u;   // the adjoint's second pass
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
namespace {
template <int BLOCK, int STACK, bool COUNT, bool BATCH, bool POOL>
const void* fn()
{
    if constexpr (COUNT) return reinterpret_cast<const void*>(&k_path_count<BLOCK, STACK, BATCH, POOL>);
    else return reinterpret_cast<const void*>(&k_path<BLOCK, STACK, BATCH, POOL>);
}

template <int BLOCK, int STACK, bool COUNT, bool BATCH>
const void* fn_pool(bool pool) { return pool ? fn<BLOCK, STACK, COUNT, BATCH, true>() : fn<BLOCK, STACK, COUNT, BATCH, false>(); }

template <int BLOCK, int STACK>
const void* fn_flags(const PathVariant& v)
{
    if (v.count) return v.batch ? fn_pool<BLOCK, STACK, true, true>(v.pool) : fn_pool<BLOCK, STACK, true, false>(v.pool);
    return v.batch ? fn_pool<BLOCK, STACK, false, true>(v.pool) : fn_pool<BLOCK, STACK, false, false>(v.pool);
}

const void* kernel_ptr(const PathVariant& v)
{
    if (v.block == 64) return v.stack <= 24 ? fn_flags<64, 24>(v) : v.stack <= 32 ? fn_flags<64, 32>(v) : fn_flags<64, 64>(v);
    if (v.block == 128) return v.stack <= 24 ? fn_flags<128, 24>(v) : v.stack <= 32 ? fn_flags<128, 32>(v) : fn_flags<128, 64>(v);
    return v.stack <= 24 ? fn_flags<256, 24>(v) : v.stack <= 32 ? fn_flags<256, 32>(v) : fn_flags<256, 64>(v);
}
}  // namespace

uint32_t stack_variant(uint32_t needed)
{
    if (needed <= 24) return 24;   // 5 waves per SIMD
    if (needed <= 32) return 32;
    if (needed <= 64) return 64;
    return 0;
}

hipError_t launch_path(const PathLaunch& a, const PathVariant& v, uint32_t grid, hipStream_t s)
{
    void* args[] = {const_cast<PathLaunch*>(&a)};
    return hipLaunchKernel(kernel_ptr(v), dim3(grid), dim3(v.block), args, 0, s);
}

uint32_t persistent_grid(const PathVariant& v, int device)
{
    // Resident workgroups per CU x CUs, cached per (device, variant).
    struct Entry { int device; PathVariant v; uint32_t grid; };
    static thread_local Entry cache[64];
    static thread_local int n = 0;
    for (int i = 0; i < n; ++i) {
        const Entry& c = cache[i];
        if (c.device == device && c.v.block == v.block && c.v.stack == v.stack && c.v.count == v.count &&
            c.v.batch == v.batch && c.v.pool == v.pool)
            return c.grid;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_ptr(v), (int)v.block, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
    const uint32_t grid = (uint32_t)(per_cu * cus);
    if (n < 64) cache[n++] = Entry{device, v, grid};
    return grid;
}

hipError_t launch_whitted(const JobDesc& j, DevCounters* c, hipStream_t s)
{
    hipLaunchKernelGGL(k_whitted, dim3((j.W + 15) / 16, (j.H + 15) / 16), dim3(256), 0, s, j, c);
    return hipGetLastError();
}

hipError_t launch_filter(const uint8_t* cur, const uint8_t* prev, uint32_t sw, uint32_t sh, uint8_t* out,
                         uint32_t dw, uint32_t dh, hipStream_t s)
{
    hipLaunchKernelGGL(k_filter, dim3((dw + 15) / 16, (dh + 15) / 16), dim3(256), 0, s, cur, prev, sw, sh, out, dw,
                       dh);
    return hipGetLastError();
}

hipError_t launch_debug(uint32_t op, const float* in, float* out, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_debug, dim3((n + 255) / 256), dim3(256), 0, s, op, in, out, n);
    return hipGetLastError();
}

__global__ void k_set_frame(JobDesc* job, vkr_frame_data fd)
{
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; ++k) {
        job->cam_pos[k] = fd.camera.pos[k];
        job->cam_dir[k] = fd.camera.dir[k];
        job->cam_right[k] = fd.camera.right[k];
        job->cam_up[k] = fd.camera.up[k];
        job->light_pos[k] = fd.light_pos[k];
    }
    job->aspect = fd.aspect_ratio;
    job->seed = fd.seed;
}

hipError_t launch_set_frame(JobDesc* job, const vkr_frame_data& fd, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(64), 0, s, job, fd);
    return hipGetLastError();
}

// Scratch for callers that cannot allocate (vkr_trace_frame_static): one
// array per device, held until the process ends.  Host code never passes
// `scratch` itself to a kernel: its name is a host-side stand-in, and the
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) char scratch[kStaticScratchBytes];

hipError_t static_scratch(int device, char** out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static char* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        addr[device] = static_cast<char*>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

hipError_t launch_zero_counters(DevCounters* c, uint32_t* pool, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, c, pool);
    return hipGetLastError();
}

}  // namespace vkr
