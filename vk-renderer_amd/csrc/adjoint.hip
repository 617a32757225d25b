// adjoint.hip - path-replay adjoint of the path kernel (build extension).
//
// Parameters (vkr_scene_set_sphere_params): per sphere s albedo a_s (3),
// emission e_s (3) and roughness r_s, plus one light scale L; the traced
// emissive is E_s = L * e_s.  Given dLoss/d(mean radiance) per pixel, the
// kernel returns dLoss/dparams with every branch decision held fixed (the
// emitter set, Russian roulette, the dielectric choice), i.e. the exact
// derivative of the per-sample estimate of Assets/Tracer.comp:433-593.
//
// Path replay: each sample is traced twice from its RNG state; of the path
// only 8 bytes per sample are stored between the two passes.
//   pass 1 records which channels of clamp(acc) (Tracer.comp:441) saturate,
//          3 bits per bounce; a saturated channel blocks every earlier term.
//          A pixel's samples share one random stream, so this pass runs them
//          in order, one lane per pixel, and keeps for each sample the stream
//          position at its start and its saturation bits;
//   pass 2 re-traces the same path (same random numbers, same shadow rays)
//          from that position, one lane per (pixel, sample), and accumulates,
//          with W_k = dLoss/dL masked by the later clamps:
//   * direct terms of q_k (normalize(emissive), the next-event term, the
//     dielectric emissive) in a_s, r_s, e, L at each vertex k;
//   * the throughput chain.  Because Russian roulette divides the mask by
//     its largest channel (:545-549), mask_k = P_k / max3(P_k) with P_k the
//     product of the albedos hit before k, so
//       d mask_k,c / d a_s,c' = mask_k,c * n_s (delta_cc' - delta_c'c*) / a_s,c'
//     (c* = argmax3 P_k, n_s = occurrences of s before k), a closed form that
//     needs only the sphere ids of earlier vertices.
// Gradients are reduced per workgroup in LDS and added once per block.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "traverse.hpp"
#include "vkr_device.hpp"

namespace vkr {

constexpr uint32_t kMaxAdjDepth = 10;

__device__ __forceinline__ uint32_t argmax3(f3 v)
{
    return (v.x >= v.y && v.x >= v.z) ? 0u : (v.y >= v.z ? 1u : 2u);
}
__device__ __forceinline__ float comp(f3 v, uint32_t c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

// MODE 1: pass 1 over every sample of a pixel, recording A.samples;
// MODE 2: pass 2 of one sample per work item (items: per 8x8 tile, sample by
// sample, so a wave replays one sample of 64 neighbouring pixels).
template <int BLOCK, int STACK, bool LDSGRAD, int MODE>
__global__ __launch_bounds__(BLOCK) void k_adjoint(const AdjointLaunch A)
{
    extern __shared__ uint32_t s_dyn[];
    uint32_t* stk = s_dyn + threadIdx.x;
    uint32_t* chain = s_dyn + STACK * BLOCK + threadIdx.x;          // kMaxAdjDepth entries, stride BLOCK
    float* sgrad = reinterpret_cast<float*>(s_dyn + (STACK + kMaxAdjDepth) * BLOCK);
    const JobDesc& j = A.job;
    const uint32_t ns = j.sc.n_spheres;
    const uint32_t np = 7u * ns + 1u;
    const float* th = A.params;
    const float lscale = th[7u * ns];
    float* grad = LDSGRAD ? sgrad : A.d_params;
    if (LDSGRAD) {
        for (uint32_t i = threadIdx.x; i < np; i += BLOCK) sgrad[i] = 0.0f;
        __syncthreads();
    }

    bool has = false, exhausted = false;
    float uvx = 0.0f, uvy = 0.0f, salt = 0.0f, salt0 = 0.0f;
    uint32_t smp = 0, depth = 0, phase = kPhTrace, pass = MODE, zb = 0, nchain = 0, cstar = 0, hs = kMiss;
    uint32_t pix = 0;
    uint32_t* const cursor = A.pool + (MODE == 2 ? 1 : 0);
    f3 ro{}, rd{}, acc{}, mask{}, P{}, N{}, dj{}, e{}, cnee{}, gp{};
    f3 gA{}, gE{};
    float gR = 0.0f, gL = 0.0f;
    uint32_t emi = 0;
    const vkr_material* hm = nullptr;
    uint32_t ei = 0;
    bool new_path = false;
    uint32_t c_rays = 0;
    WorkCnt wc{};

    auto claim = [&](uint32_t g) __attribute__((always_inline)) {
        uint32_t tile = g >> 6, s = 0;
        if (MODE == 2) {
            tile = g / (64u * j.spp);
            s = (g >> 6) % j.spp;
        }
        const uint32_t within = g & 63u;
        if (tile >= j.n_tiles) return;
        const uint32_t x = (tile % j.tiles_x) * 8u + (within & 7u);
        const uint32_t y = (tile / j.tiles_x) * 8u + (within >> 3);
        if (x >= j.W || y >= j.H) return;
        has = true;
        pix = y * j.W + x;
        uvx = (float)x / (float)j.W;
        uvy = (float)y / (float)j.H;
        smp = s;
        pass = MODE;
        if (MODE == 2) {
            const uint2 rec = A.samples[(size_t)pix * j.spp + s];
            salt0 = __uint_as_float(rec.x);
            zb = rec.y;
        } else {
            salt = uvy / uvx + j.seed;
        }
        const float4 g4 = reinterpret_cast<const float4*>(A.d_radiance)[pix];
        const float inv = 1.0f / (float)j.spp;      // d(mean)/d(sample) (:585)
        gp = f3{g4.x * inv, g4.y * inv, g4.z * inv};
        new_path = true;
    };
    // W_k: dLoss/dL masked by the clamps that follow vertex k.
    auto weight = [&](uint32_t k) __attribute__((always_inline)) {
        const uint32_t above = (0x3FFu >> (k + 1u)) << (k + 1u);
        return f3{((zb >> 0) & above) ? 0.0f : gp.x, ((zb >> 10) & above) ? 0.0f : gp.y,
                  ((zb >> 20) & above) ? 0.0f : gp.z};
    };
    auto gadd = [&](uint32_t i, float v) __attribute__((always_inline)) {
        if (v != 0.0f) atomicAdd(grad + i, v);
    };
    // Throughput chain at vertex k with term T = mask_k (.) q_k.
    auto chain_grad = [&](f3 W, f3 T) __attribute__((always_inline)) {
        const float tot = W.x * T.x + W.y * T.y + W.z * T.z;
        for (uint32_t i = 0; i < nchain; ++i) {
            const uint32_t s = chain[i * BLOCK];
            const float* a = th + 7u * s;
            const float t0 = W.x * T.x - (cstar == 0 ? tot : 0.0f);
            const float t1 = W.y * T.y - (cstar == 1 ? tot : 0.0f);
            const float t2 = W.z * T.z - (cstar == 2 ? tot : 0.0f);
            if (a[0] != 0.0f) gadd(7u * s + 0u, t0 / a[0]);
            if (a[1] != 0.0f) gadd(7u * s + 1u, t1 / a[1]);
            if (a[2] != 0.0f) gadd(7u * s + 2u, t2 / a[2]);
        }
    };

    while (true) {
        if (!exhausted) {
            const unsigned long long need = __ballot(!has);
            const unsigned long long busy = __ballot(has);
            const uint32_t nneed = (uint32_t)__popcll(need);
            if (nneed > 0 && (nneed >= 32u || busy == 0)) {
                const int leader = __ffsll((long long)need) - 1;
                uint32_t base = 0;
                if (__lane_id() == leader) base = atomicAdd(cursor, nneed);
                base = __shfl(base, leader, 64);
                if (base + nneed >= A.total) exhausted = true;
                if (!has) {
                    const uint32_t rank = (uint32_t)__popcll(need & ((1ull << __lane_id()) - 1ull));
                    const uint32_t g = base + rank;
                    if (g < A.total) claim(g);
                }
            }
        }
        if (__ballot(has) == 0) {
            if (exhausted) break;
            continue;
        }

        if (has && new_path) {
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
            nchain = 0;
            cstar = 0;
            if (pass == 1) {
                salt0 = salt;
                zb = 0;
            } else {
                salt = salt0;
            }
        }

        f3 o{}, d{};
        float tmax = 0.0f;
        bool any = false, need = false;
        if (has) {
            if (phase == kPhTrace) {
                if (pass == 1) {
                    const uint32_t bit = 1u << depth;
                    if (acc.x > 1.0f || acc.x < 0.0f) zb |= bit;
                    if (acc.y > 1.0f || acc.y < 0.0f) zb |= bit << 10;
                    if (acc.z > 1.0f || acc.z < 0.0f) zb |= bit << 20;
                }
                acc = clamps(acc, 0.0f, 1.0f);
                const float dp1 = (float)(depth + 1u);
                tmax = 3000.0f / (dp1 * dp1);
                o = ro;
                d = rd;
                need = true;
            } else {
                emi = j.sc.emitters[ei];
                const vkr_sphere* S = j.sc.spheres + emi;
                const f3 sP = ld3(S->P);
                const float sr = S->r;
                const float tl = length(sub(sP, P)) - sr;
                const f3 l0 = sub(sP, P);
                const float cos_a_max = sqrtf(1.0f - gclamp(sr * sr / dot(l0, l0), 0.0f, 1.0f));
                const float cosa = gmix(cos_a_max, 1.0f, rng_next(salt, uvx, uvy));
                const float phi = kTwoPi * rng_next(salt, uvx, uvy);
                const f3 L = jitter(l0, phi, sqrtf(1.0f - cosa * cosa), cosa);
                const float x = tl / sr + 1.0f;
                const float x2 = x * x;
                const f3 Ee = ld3(S->mat.emissive);
                f3 att = divs(muls(Ee, 1.0f), x2);
                att = divs(subs(att, kAttSub), kAttDiv);
                att = maxs(att, 0.0f);
                const f3 alb = ld3(hm->albedo);
                const float rough = hm->roughness, metal = hm->metalness;
                const f3 F0 = mix3(f3{0.04f, 0.04f, 0.04f}, alb, metal);
                const f3 V = normalize(sub(ld3(j.cam_pos), P));
                const f3 H = normalize(add(V, L));
                const float NDF = distribution_ggx(N, H, rough);
                const float G = geometry_smith(N, V, L, rough);
                const float cHV = gmax(dot(H, V), 0.0f);
                const float p5 = pow5(1.0f - cHV);
                const f3 F = fresnel_schlick(cHV, F0);
                f3 kD = ssub(1.0f, F);
                kD = muls(kD, 1.0f - metal);
                const float NdotV = gmax(dot(N, V), 0.0f);
                const float NdotL = gmax(dot(N, L), 0.0f);
                const float den = gmax(4.0f * NdotV * NdotL, 0.001f);
                const f3 specular = divs(muls(F, NDF * G), den);
                const f3 br = add(divs(mul(kD, alb), kPi), specular);
                cnee = muls(mul(br, att), NdotL);
                need = !(cnee.x == 0.0f && cnee.y == 0.0f && cnee.z == 0.0f);
                if (pass == 2 && need) {
                    // Pending gradient of this next-event term (added if unoccluded).
                    const f3 u = mul(weight(depth), mask);
                    // albedo of the shaded sphere: kD*a/pi and F0 = mix(0.04, a, metal)
                    const float dF = metal * (1.0f - p5);
                    const float dkd = -(1.0f - metal) * dF;
                    const float sp = NDF * G / den;
                    gA = f3{u.x * (((kD.x + alb.x * dkd) / kPi + sp * dF) * att.x * NdotL),
                            u.y * (((kD.y + alb.y * dkd) / kPi + sp * dF) * att.y * NdotL),
                            u.z * (((kD.z + alb.z * dkd) / kPi + sp * dF) * att.z * NdotL)};
                    // roughness: NDF (a2 = r^4) and Smith-Schlick (k = (r+1)^2/8)
                    const float a2 = rough * rough * rough * rough;
                    const float NdotH = gmax(dot(N, H), 0.0f);
                    const float nh2 = NdotH * NdotH;
                    const float dd = nh2 * (a2 - 1.0f) + 1.0f;
                    const float dNdf = (dd - 2.0f * a2 * nh2) / (kPi * dd * dd * dd) * 4.0f * rough * rough * rough;
                    const float k = (rough + 1.0f) * (rough + 1.0f) / 8.0f;
                    const float qv = NdotV * (1.0f - k) + k, ql = NdotL * (1.0f - k) + k;
                    const float gv = NdotV / qv, gl = NdotL / ql;
                    const float dgv = -NdotV * (1.0f - NdotV) / (qv * qv), dgl = -NdotL * (1.0f - NdotL) / (ql * ql);
                    const float dG = (dgv * gl + gv * dgl) * (rough + 1.0f) / 4.0f;
                    const float dsp = (dNdf * G + NDF * dG) / den;
                    gR = (u.x * F.x * att.x + u.y * F.y * att.y + u.z * F.z * att.z) * dsp * NdotL;
                    // emitter emission and light scale through att = (L e / x^2 - 0.001) / 0.999
                    const float* eo = th + 7u * emi + 3u;
                    const float s = 1.0f / (x2 * kAttDiv);
                    const f3 w = muls(mul(u, br), NdotL);
                    gE = f3{att.x > 0.0f ? w.x * lscale * s : 0.0f, att.y > 0.0f ? w.y * lscale * s : 0.0f,
                            att.z > 0.0f ? w.z * lscale * s : 0.0f};
                    gL = (att.x > 0.0f ? w.x * eo[0] * s : 0.0f) + (att.y > 0.0f ? w.y * eo[1] * s : 0.0f) +
                         (att.z > 0.0f ? w.z * eo[2] * s : 0.0f);
                }
                o = P;
                d = L;
                tmax = tl;
                any = true;
            }
            ++c_rays;
        }

        TravOut tr{kMiss, 0.0f, false};
        if (has && need) tr = traverse<BLOCK, false>(j.sc, o, d, tmax, j.eps, any, true, stk, wc);
        if (!has) continue;

        bool fin_diffuse = false, do_rr = false, path_end = false;
        if (phase == kPhTrace) {
            if (!tr.hit) {
                path_end = true;
            } else {
                P = add(ro, muls(rd, tr.t));
                const uint32_t kind = tr.prim >> 30, idx = tr.prim & 0x3FFFFFFFu;
                const vkr_material* m;
                hs = kMiss;
                if (kind == 0) {
                    const vkr_triangle* T = j.sc.tris_orig + idx;
                    const f3 v0 = ld3(T->v0), v1 = ld3(T->v1), v2 = ld3(T->v2);
                    const f3 u = sub(v1, v0), v = sub(v2, v0);
                    N = f3{(u.y * v.z) - (u.z * v.y), (u.z * v.x) - (u.x * v.z), (u.x * v.y) - (u.y * v.x)};
                    const uint32_t mi = T->mat_index;
                    m = ((j.sc.flags & VKR_SCENE_TRI_MATERIAL_INDEX) && mi < j.sc.n_materials) ? j.sc.materials + mi
                                                                                               : j.sc.tri_material;
                } else if (kind == 1) {
                    const vkr_sphere* S = j.sc.spheres + idx;
                    N = divs(sub(P, ld3(S->P)), S->r);
                    m = &S->mat;
                    hs = idx;
                } else {
                    const vkr_plane* Q = j.sc.planes + idx;
                    N = ld3(Q->N);
                    m = &Q->mat;
                }
                hm = m;
                const uint32_t type = m->type;
                if (type == VKR_MAT_DIFFUSE) {
                    const float r2 = rng_next(salt, uvx, uvy);
                    const float phi = kTwoPi * rng_next(salt, uvx, uvy);
                    dj = muls(jitter(N, phi, sqrtf(r2), sqrtf(1.0f - r2)), 1.0f - m->metalness);
                    e = f3{0.0f, 0.0f, 0.0f};
                    ei = 0;
                    if (j.sc.n_emitters > 0) phase = kPhNee;
                    else fin_diffuse = true;
                } else if (type == VKR_MAT_DIELECTRIC) {
                    const f3 emis = ld3(m->emissive);
                    const float rough = m->roughness;
                    if (pass == 2) {
                        const f3 W = weight(depth);
                        const f3 u = mul(W, mask);
                        if (hs != kMiss) {
                            const float* eo = th + 7u * hs + 3u;
                            gadd(7u * hs + 3u, u.x * lscale);
                            gadd(7u * hs + 4u, u.y * lscale);
                            gadd(7u * hs + 5u, u.z * lscale);
                            gadd(7u * ns, u.x * eo[0] + u.y * eo[1] + u.z * eo[2]);
                        }
                        chain_grad(W, mul(mask, emis));
                    }
                    acc = add(acc, mul(emis, mask));
                    mask = mul(mask, ld3(m->albedo));
                    if (hs != kMiss && nchain < kMaxAdjDepth) chain[(nchain++) * BLOCK] = hs;
                    const float cosine = -dot(rd, N) / length(rd);
                    const f3 reflected = reflect(rd, N);
                    const f3 refracted = refract(rd, N, rough);
                    const float prob = eq3(refracted, f3{0.0f, 0.0f, 0.0f}) ? 1.0f : schlick(cosine, rough);
                    const float u = rng_next(salt, uvx, uvy);
                    rd = normalize(u < prob ? reflected : refracted);
                    ro = P;
                    do_rr = true;
                } else {
                    do_rr = true;
                }
            }
        } else {
            if (need && !tr.hit) {
                e = add(e, cnee);
                if (pass == 2) {
                    if (hs != kMiss) {
                        gadd(7u * hs + 0u, gA.x);
                        gadd(7u * hs + 1u, gA.y);
                        gadd(7u * hs + 2u, gA.z);
                        gadd(7u * hs + 6u, gR);
                    }
                    gadd(7u * emi + 3u, gE.x);
                    gadd(7u * emi + 4u, gE.y);
                    gadd(7u * emi + 5u, gE.z);
                    gadd(7u * ns, gL);
                }
            }
            ++ei;
            if (ei >= j.sc.n_emitters) fin_diffuse = true;
        }

        if (fin_diffuse) {
            const f3 emis = ld3(hm->emissive);
            const bool all_pos = gmax(gsign(emis.x), 0.0f) == 1.0f && gmax(gsign(emis.y), 0.0f) == 1.0f &&
                                 gmax(gsign(emis.z), 0.0f) == 1.0f;
            const f3 em = all_pos ? normalize(emis) : f3{0.0f, 0.0f, 0.0f};
            if (pass == 2) {
                const f3 W = weight(depth);
                if (all_pos && hs != kMiss) {
                    // d normalize(L e)/de = (I - n n^T) / |e|
                    const float* eo = th + 7u * hs + 3u;
                    const f3 ev{eo[0], eo[1], eo[2]};
                    const float len = length(ev);
                    const f3 n = divs(ev, len);
                    const f3 u = mul(W, mask);
                    const float un = dot(u, n);
                    gadd(7u * hs + 3u, (u.x - un * n.x) / len);
                    gadd(7u * hs + 4u, (u.y - un * n.y) / len);
                    gadd(7u * hs + 5u, (u.z - un * n.z) / len);
                }
                chain_grad(W, mul(mask, add(em, e)));
            }
            acc = add(acc, mul(mask, add(em, e)));
            mask = mul(mask, ld3(hm->albedo));
            if (hs != kMiss && nchain < kMaxAdjDepth) chain[(nchain++) * BLOCK] = hs;
            rd = normalize(add(reflect(rd, N), dj));
            ro = P;
            phase = kPhTrace;
            do_rr = true;
        }
        if (do_rr) {
            const float p = max3(mask);
            cstar = argmax3(mask);
            const float u = rng_next(salt, uvx, uvy);
            if (u > p) {
                path_end = true;
            } else {
                mask = muls(mask, 1.0f / p);
                ++depth;
                if (depth >= j.max_depth) path_end = true;
                phase = kPhTrace;
            }
        }
        if (path_end) {
            if (MODE == 1) {
                // The next sample starts where this one's draws ended.
                A.samples[(size_t)pix * j.spp + smp] = make_uint2(__float_as_uint(salt0), zb);
                ++smp;
                if (smp < j.spp) new_path = true;
                else has = false;
            } else {
                has = false;
            }
        }
    }

    flush_counter(&A.counters->rays, c_rays);
    if (LDSGRAD) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < np; i += BLOCK)
            if (sgrad[i] != 0.0f) atomicAdd(A.d_params + i, sgrad[i]);
    }
}

namespace {
template <int BLOCK, int STACK, bool LDSGRAD, int MODE>
hipError_t go(const AdjointLaunch& a, size_t lds, hipStream_t s)
{
    const void* fn = reinterpret_cast<const void*>(&k_adjoint<BLOCK, STACK, LDSGRAD, MODE>);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BLOCK, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    const uint32_t need = (a.total + BLOCK - 1u) / BLOCK;
    const uint32_t grid = std::max(1u, std::min((uint32_t)(per_cu * cus), need));
    hipLaunchKernelGGL((k_adjoint<BLOCK, STACK, LDSGRAD, MODE>), dim3(grid), dim3(BLOCK), lds, s, a);
    return hipGetLastError();
}

template <int BLOCK, int STACK>
hipError_t both_passes(AdjointLaunch a, bool pass1, bool ldsgrad, size_t base, size_t grad_bytes, hipStream_t s)
{
    // Pass 1: one lane per pixel (pool indices as in the path kernel).
    // Skipped when a.samples already holds its records for this frame.
    const uint32_t pixels = a.total;
    if (pass1) {
        const hipError_t e = go<BLOCK, STACK, false, 1>(a, base, s);
        if (e != hipSuccess) return e;
    }
    // Pass 2: one lane per (pixel, sample).
    a.total = pixels * a.job.spp;
    return ldsgrad ? go<BLOCK, STACK, true, 2>(a, base + grad_bytes, s) : go<BLOCK, STACK, false, 2>(a, base, s);
}
}  // namespace

hipError_t launch_adjoint(const AdjointLaunch& a, uint32_t stack, bool pass1, hipStream_t s)
{
    constexpr int kBlock = 128;
    const uint32_t np = 7u * a.job.sc.n_spheres + 1u;
    const size_t base = (size_t)((stack <= 32 ? 32 : 64) + kMaxAdjDepth) * kBlock * 4u;
    const bool lds = base + (size_t)np * 4u <= 60u * 1024u;
    const size_t grad_bytes = (size_t)np * 4u;
    if (stack <= 32) return both_passes<kBlock, 32>(a, pass1, lds, base, grad_bytes, s);
    return both_passes<kBlock, 64>(a, pass1, lds, base, grad_bytes, s);
}

}  // namespace vkr
