// api.cpp - the C ABI of include/vkr.h over the gfx950 kernels.
//
// Replaces the compute half of the reference's GraphicsDevice
// (Source/GraphicsDevice.cpp): device selection, the scene SSBO, the compute
// pipeline and the per-frame push-constant + dispatch.  No global state: an
// opaque context owns every device allocation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vkr.h"
#include "bvh.hpp"
#include "ctx.hpp"
#include "kernels.hpp"

using namespace vkr;

namespace vkr {
thread_local std::string g_err = "no error";

namespace {

// VKR_SCENE_DEVICE_BVH: the triangle subtree built on the device (lbvh.hip)
// and the sphere subtree on the host, joined under one root node (node 0)
// whose sphere child is always entered, the layout of the host builder:
//   [root] [n - 1 triangle nodes] [sphere nodes]   (root only with spheres).
// On success sc->nodes and sc->tris are filled and `sph` holds the sphere
// records.  Returns false with *err == hipSuccess when the tree is deeper
// than depth_cap or needs 2^26 nodes or more (nothing is kept; the caller
// builds on the host).
bool device_build(vkr_ctx* ctx, Scene* sc, uint32_t n, BuildInput bi, uint32_t depth_cap, BuildOutput& sph,
                  uint32_t& n_nodes, uint32_t& n_leaves, uint32_t& depth, hipError_t* err)
{
    *err = hipSuccess;
    bi.tri_v = nullptr;
    bi.tri_mat = nullptr;
    bi.n_tris = 0;
    BuildOutput sb = build_bvh(bi);   // spheres only (empty when no sphere is finite)
    const bool has_sph = !sb.empty;
    const uint32_t node_off = has_sph ? 1u : 0u;
    const uint32_t sph_off = node_off + n - 1u;
    const size_t total = (size_t)sph_off + (has_sph ? sb.nodes.size() : 0u);
    if (total >= (1u << 26)) return false;   // too many nodes for 32-bit offsets: the host builder may fit
    if ((*err = sc->nodes.alloc(total * sizeof(BvhNode))) != hipSuccess) return false;
    if ((*err = sc->tris.alloc((size_t)n * sizeof(TriRecord))) != hipSuccess) return false;
    const size_t need = lbvh_scratch_bytes(n);
    if (need == 0) {
        *err = hipErrorInvalidValue;
        return false;
    }
    if (ctx->build_scratch.bytes < need && (*err = ctx->build_scratch.alloc(need)) != hipSuccess) return false;
    LbvhRoot root{};
    const uint32_t tri_cap = has_sph ? depth_cap - 1u : depth_cap;
    *err = lbvh_build(sc->tris_orig.as<vkr_triangle>(), n, node_off, bi.max_leaf, tri_cap, ctx->build_scratch.p,
                      ctx->build_scratch.bytes, sc->nodes.as<BvhNode>(), sc->tris.as<TriRecord>(), &root, nullptr);
    if (*err != hipSuccess) return false;
    const uint32_t dep = has_sph ? 1u + std::max(root.height, sb.depth) : root.height;
    if (dep > depth_cap) {
        sc->nodes.reset();
        sc->tris.reset();
        return false;
    }
    if (has_sph) {
        offset_inner_refs(sb.nodes, sph_off);
        const BvhNode top = join_root(root.lo, root.hi, node_off, sph_off);
        BvhNode* dn = sc->nodes.as<BvhNode>();
        if ((*err = hipMemcpy(dn, &top, sizeof top, hipMemcpyHostToDevice)) != hipSuccess) return false;
        *err = hipMemcpy(dn + sph_off, sb.nodes.data(), sb.nodes.size() * sizeof(BvhNode), hipMemcpyHostToDevice);
        if (*err != hipSuccess) return false;
    }
    n_nodes = root.nodes + (has_sph ? 1u + (uint32_t)sb.nodes.size() : 0u);   // live nodes
    n_leaves = root.leaves + sb.leaves;
    depth = dep;
    sb.nodes.clear();
    sph = std::move(sb);
    return true;
}

}  // namespace
}  // namespace vkr

extern "C" {

const char* vkr_last_error_string(void) { return g_err.c_str(); }

vkr_error vkr_create(const vkr_create_info* info, vkr_ctx** out_ctx)
{
    if (!out_ctx) return fail(VKR_INVALID_ARGUMENT, "out_ctx is NULL");
    *out_ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(VKR_NO_SUITABLE_GPU, "no HIP device is visible");
    int dev = 0;
    if (info && info->device >= 0) dev = info->device;
    else if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (dev >= n) return fail(VKR_NO_SUITABLE_GPU, "device ordinal out of range");
    hipDeviceProp_t prop;
    VKR_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(VKR_NO_SUITABLE_GPU, std::string("kernels are built for gfx950, device is ") + prop.gcnArchName);
    VKR_HIP(hipSetDevice(dev));
    auto ctx = std::make_unique<vkr_ctx>();
    ctx->device = dev;
    ctx->flags = info ? info->flags : 0u;
    VKR_HIP(ctx->counters.alloc(sizeof(DevCounters)));
    VKR_HIP(ctx->pool.alloc(64));
    VKR_HIP(hipMemset(ctx->counters.p, 0, sizeof(DevCounters)));
    // The device address of this device's static scratch (vkr_trace_frame_static).
    char* scratch = nullptr;
    VKR_HIP(static_scratch(dev, &scratch));
    *out_ctx = ctx.release();
    return VKR_SUCCESS;
}

vkr_error vkr_destroy(vkr_ctx* ctx)
{
    if (!ctx) return VKR_SUCCESS;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev_copy) (void)hipEventDestroy(ctx->ev_copy);
    if (ctx->host_jobs) (void)hipHostFree(ctx->host_jobs);
    delete ctx;
    return VKR_SUCCESS;
}

vkr_error vkr_wait_idle(vkr_ctx* ctx)
{
    if (!ctx) return fail(VKR_INVALID_ARGUMENT, "ctx is NULL");
    VKR_HIP(hipSetDevice(ctx->device));
    VKR_HIP(hipDeviceSynchronize());
    return VKR_SUCCESS;
}

vkr_error vkr_set_ctx_flags(vkr_ctx* ctx, uint32_t flags)
{
    if (!ctx) return fail(VKR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->flags = flags;
    return VKR_SUCCESS;
}

vkr_error vkr_scene_upload(vkr_ctx* ctx, const vkr_scene_desc* desc, uint32_t* out_scene)
{
    if (!ctx || !desc || !out_scene) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    if ((desc->n_tris && !desc->tris) || (desc->n_spheres && !desc->spheres) || (desc->n_planes && !desc->planes))
        return fail(VKR_INVALID_ARGUMENT, "primitive array is NULL with a non-zero count");
    if (desc->n_tris > kRefFirstMask || desc->n_spheres > kRefFirstMask)
        return fail(VKR_INVALID_ARGUMENT, "too many primitives (limit 2^27 per kind)");
    if ((uint64_t)desc->n_tris * sizeof(TriRecord) >= (1ull << 32))
        return fail(VKR_INVALID_ARGUMENT, "too many triangles (triangle records are addressed with 32-bit offsets)");
    if ((desc->flags & VKR_SCENE_TRI_MATERIAL_INDEX) && (!desc->materials || desc->n_materials == 0))
        return fail(VKR_INVALID_ARGUMENT, "per-triangle materials requested without a material table");
    if (desc->max_leaf_size > kMaxLeaf) return fail(VKR_INVALID_ARGUMENT, "max_leaf_size must be <= 8");
    if (desc->max_depth > 63) return fail(VKR_INVALID_ARGUMENT, "max_depth must be <= 63");
    VKR_HIP(hipSetDevice(ctx->device));

    auto sc = std::make_unique<Scene>();
    sc->host_spheres.assign(desc->spheres, desc->spheres + desc->n_spheres);
    const std::vector<uint32_t> em = emitter_list(sc->host_spheres);
    VKR_HIP(sc->tris_orig.upload(desc->tris, (size_t)desc->n_tris * sizeof(vkr_triangle)));
    const auto t0 = std::chrono::steady_clock::now();
    auto t1 = t0;
    BuildInput bi;
    bi.tri_v = reinterpret_cast<const float*>(desc->tris);
    bi.tri_mat = desc->tris ? reinterpret_cast<const uint32_t*>(desc->tris) + 3 : nullptr;
    bi.n_tris = desc->n_tris;
    bi.sph = reinterpret_cast<const float*>(desc->spheres);
    bi.n_spheres = desc->n_spheres;
    bi.max_leaf = desc->max_leaf_size ? desc->max_leaf_size : default_leaf((uint32_t)em.size());
    bi.max_depth = desc->max_depth ? desc->max_depth : kDefaultMaxDepth;
    bi.traversal_cost = desc->traversal_cost;
    // Host build: bo is the whole tree.  Device build: bo holds the sphere
    // records only; nodes and triangle records are already on the device.
    BuildOutput bo;
    uint32_t builder = VKR_BUILDER_HOST_SAH, n_nodes = 0, n_leaves = 0, depth = 0;
    bool has_bvh = false;
    if ((desc->flags & VKR_SCENE_DEVICE_BVH) && desc->n_tris >= 2u) {
        hipError_t e = hipSuccess;
        const uint32_t cap = desc->max_depth ? desc->max_depth : 63u;
        if (device_build(ctx, sc.get(), desc->n_tris, bi, cap, bo, n_nodes, n_leaves, depth, &e)) {
            builder = VKR_BUILDER_DEVICE_LBVH;
            has_bvh = true;
            t1 = std::chrono::steady_clock::now();
        } else if (e != hipSuccess) {
            return fail_hip("device BVH build", e);
        }
    }
    if (builder == VKR_BUILDER_HOST_SAH) {
        bo = build_bvh(bi);
        t1 = std::chrono::steady_clock::now();
        if (bo.nodes.size() >= (1u << 26))
            return fail(VKR_UNSUPPORTED, "BVH has 2^26 nodes or more (nodes are addressed with 32-bit offsets)");
        n_nodes = (uint32_t)bo.nodes.size();
        n_leaves = bo.leaves;
        depth = bo.depth;
        has_bvh = !bo.empty;
        VKR_HIP(sc->nodes.upload(bo.nodes.data(), bo.nodes.size() * sizeof(BvhNode)));
        VKR_HIP(sc->tris.upload(bo.tris.data(), bo.tris.size() * sizeof(TriRecord)));
    }

    // depth + 1 entries: the end marker at the bottom.  Without a tree nothing
    // is traversed, and the smallest variant serves.
    const uint32_t stack = has_bvh ? stack_variant(depth + 1u) : 24u;
    if (stack == 0) return fail(VKR_UNSUPPORTED, "BVH deeper than 63 levels");

    // Parameters as uploaded: light scale 1, emission = the sphere's emissive.
    sc->host_theta.assign(7u * desc->n_spheres + 1u, 0.0f);
    for (uint32_t i = 0; i < desc->n_spheres; ++i) {
        const vkr_material& m = desc->spheres[i].mat;
        for (int k = 0; k < 3; ++k) {
            sc->host_theta[7u * i + k] = m.albedo[k];
            sc->host_theta[7u * i + 3u + k] = m.emissive[k];
        }
        sc->host_theta[7u * i + 6u] = m.roughness;
    }
    sc->host_theta[7u * desc->n_spheres] = 1.0f;
    VKR_HIP(sc->theta.upload(sc->host_theta.data(), sc->host_theta.size() * sizeof(float)));

    VKR_HIP(sc->sph.upload(bo.sph.data(), bo.sph.size() * sizeof(float)));
    VKR_HIP(sc->sph_key.upload(bo.sph_key.data(), bo.sph_key.size() * sizeof(uint32_t)));
    VKR_HIP(sc->spheres.upload(desc->spheres, (size_t)desc->n_spheres * sizeof(vkr_sphere)));
    VKR_HIP(sc->planes.upload(desc->planes, (size_t)desc->n_planes * sizeof(vkr_plane)));
    VKR_HIP(sc->tri_mat.upload(desc->tri_material ? desc->tri_material : &kMirror, sizeof(vkr_material)));
    const uint32_t nm = (desc->flags & VKR_SCENE_TRI_MATERIAL_INDEX) ? desc->n_materials : 0u;
    VKR_HIP(sc->materials.upload(nm ? desc->materials : nullptr, (size_t)nm * sizeof(vkr_material)));
    // Room for every sphere: vkr_scene_set_sphere_params rewrites the list in
    // place, so its device address never changes while graphs hold it.
    VKR_HIP(sc->emitters.alloc(std::max<size_t>(desc->n_spheres, 1u) * sizeof(uint32_t)));
    if (!em.empty()) VKR_HIP(hipMemcpy(sc->emitters.p, em.data(), em.size() * sizeof(uint32_t), hipMemcpyHostToDevice));

    DevScene& d = sc->dev;
    d.nodes = sc->nodes.as<BvhNode>();
    d.tris = sc->tris.as<TriRecord>();
    d.sph = sc->sph.as<float4>();
    d.sph_key = sc->sph_key.as<uint32_t>();
    d.tris_orig = sc->tris_orig.as<vkr_triangle>();
    d.spheres = sc->spheres.as<vkr_sphere>();
    d.planes = sc->planes.as<vkr_plane>();
    d.tri_material = sc->tri_mat.as<vkr_material>();
    d.materials = sc->materials.as<vkr_material>();
    d.emitters = sc->emitters.as<uint32_t>();
    d.n_tris = desc->n_tris;
    d.n_spheres = desc->n_spheres;
    d.n_planes = desc->n_planes;
    d.n_emitters = (uint32_t)em.size();
    d.emitters_bounded = emitters_bounded(sc->host_spheres, em);
    d.n_materials = nm;
    d.flags = desc->flags;
    d.has_bvh = has_bvh ? 1u : 0u;
    d.stack_needed = stack;
    for (int k = 0; k < 3; ++k) d.sph_c[k] = bo.sph_c[k];
    d.sph_r = bo.sph_r;

    sc->serial = ctx->next_serial++;
    vkr_scene_info& in = sc->info;
    in.n_tris = desc->n_tris;
    in.n_spheres = desc->n_spheres;
    in.n_planes = desc->n_planes;
    in.n_emitters = (uint32_t)em.size();
    in.n_nodes = n_nodes;
    in.n_leaves = n_leaves;
    in.depth = depth;
    in.stack_needed = stack;
    in.device_bytes = sc->nodes.bytes + sc->tris.bytes + sc->sph.bytes + sc->sph_key.bytes + sc->tris_orig.bytes +
                      sc->spheres.bytes + sc->planes.bytes + sc->tri_mat.bytes + sc->materials.bytes +
                      sc->emitters.bytes;
    in.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    in.builder = builder;

    // Reuse a freed slot if any.
    for (uint32_t i = 0; i < ctx->scenes.size(); ++i) {
        if (!ctx->scenes[i]) {
            ctx->scenes[i] = std::move(sc);
            *out_scene = i;
            return VKR_SUCCESS;
        }
    }
    ctx->scenes.push_back(std::move(sc));
    *out_scene = (uint32_t)ctx->scenes.size() - 1u;
    return VKR_SUCCESS;
}

vkr_error vkr_scene_destroy(vkr_ctx* ctx, uint32_t scene)
{
    if (!get_scene(ctx, scene)) return fail(VKR_INVALID_ARGUMENT, "unknown scene");
    // A captured graph traces from copies of the scene's device pointers.
    for (const auto& g : ctx->graphs)
        if (g && std::find(g->job_scene.begin(), g->job_scene.end(), scene) != g->job_scene.end())
            return fail(VKR_INVALID_ARGUMENT, "a captured graph traces this scene; destroy the graph first");
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    ctx->scenes[scene].reset();
    return VKR_SUCCESS;
}

vkr_error vkr_scene_get_info(vkr_ctx* ctx, uint32_t scene, vkr_scene_info* out)
{
    Scene* s = get_scene(ctx, scene);
    if (!s || !out) return fail(VKR_INVALID_ARGUMENT, "unknown scene or NULL out");
    *out = s->info;
    return VKR_SUCCESS;
}

vkr_error vkr_scene_set_sphere_params(vkr_ctx* ctx, uint32_t scene, const float* params, uint32_t n)
{
    Scene* s = get_scene(ctx, scene);
    if (!s || !params) return fail(VKR_INVALID_ARGUMENT, "unknown scene or NULL params");
    const uint32_t ns = (uint32_t)s->host_spheres.size();
    if (n != 7u * ns + 1u) return fail(VKR_INVALID_ARGUMENT, "params must hold 7 * n_spheres + 1 floats");
    VKR_HIP(hipSetDevice(ctx->device));
    // The same values again leave the scene's version as it is (a recording
    // forward stays valid for the adjoint that follows it).
    const bool same = s->host_theta.size() == n && std::memcmp(s->host_theta.data(), params, (size_t)n * sizeof(float)) == 0;
    const float L = params[7u * ns];
    for (uint32_t i = 0; i < ns; ++i) {
        vkr_material& m = s->host_spheres[i].mat;
        const float* q = params + 7u * i;
        for (int k = 0; k < 3; ++k) {
            m.albedo[k] = q[k];
            m.emissive[k] = L * q[3 + k];
        }
        m.roughness = q[6];
    }
    const std::vector<uint32_t> em = emitter_list(s->host_spheres);
    VKR_HIP(hipDeviceSynchronize());
    VKR_HIP(hipMemcpy(s->spheres.p, s->host_spheres.data(), (size_t)ns * sizeof(vkr_sphere), hipMemcpyHostToDevice));
    // The list was sized for every sphere at upload: it is rewritten in place.
    if (!em.empty())
        VKR_HIP(hipMemcpy(s->emitters.p, em.data(), em.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    s->dev.n_emitters = (uint32_t)em.size();
    s->dev.emitters_bounded = emitters_bounded(s->host_spheres, em);
    s->info.n_emitters = (uint32_t)em.size();
    s->host_theta.assign(params, params + n);
    VKR_HIP(hipMemcpy(s->theta.p, s->host_theta.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    if (!same) s->serial = ctx->next_serial++;
    // Captured graphs hold a copy of the scene record per job: refresh the
    // emitter count and bounds there too (the device is idle here).
    for (const auto& g : ctx->graphs) {
        if (!g) continue;
        for (uint32_t j = 0; j < g->n_jobs; ++j) {
            if (g->job_scene[j] != scene) continue;
            char* dst = static_cast<char*>(g->jobs.p) + (size_t)j * sizeof(JobDesc) + offsetof(JobDesc, sc);
            VKR_HIP(hipMemcpy(dst, &s->dev, sizeof(DevScene), hipMemcpyHostToDevice));
        }
    }
    return VKR_SUCCESS;
}

// Room for the adjoint's first-pass records of n samples in the context.
// Growing the buffer waits for the device, as for the ray records.
static vkr_error ensure_adj_samples(vkr_ctx* ctx, size_t n)
{
    const size_t bytes = n * sizeof(uint2);
    if (ctx->adj_samples.bytes >= bytes) return VKR_SUCCESS;
    ctx->adj_key.valid = false;
    VKR_HIP(hipDeviceSynchronize());
    VKR_HIP(ctx->adj_samples.alloc(bytes));
    return VKR_SUCCESS;
}

// Trees of depth <= 23 can take the 24-entry-stack kernel, which keeps 5
// waves per SIMD resident (4 with the 32-entry stack, 2 to 3 with the
// 64-entry one, both with more registers and fewer spills).  Untuned
// launches choose by the image size against the lanes the 4-wave kernel
// keeps resident, and by the tree's depth (autotuner tables, one MI355X):
// - trees deeper than 8 levels, where node fetches dominate: the fifth wave
//   pays off only with well over twice those pixels (1080p soup frame: 89
//   against 95 ms; a 512 x 512 soup frame 1% slower), else the 32-entry
//   kernel;
// - shallow trees, where shading dominates: the 32-entry kernel (1024^2
//   reference scene: 2.02 against 2.15 ms, 2048^2: 7.0 against 7.6 ms), or
//   the 64-entry one when the grid holds every pixel at once (512^2 with 64
//   spheres: 0.61 against 0.66 ms; 256^2 adjoint scene: 1.43 against
//   1.48 ms).
static uint32_t default_stack(const PathVariant& v, uint32_t total, uint32_t depth, int device)
{
    if (v.stack != 24u) return v.stack;
    PathVariant v4 = v;
    v4.stack = 32u;
    const uint64_t resident4 = (uint64_t)persistent_grid(v4, device) * v.block;
    if (depth <= 8u) return (uint64_t)total <= resident4 ? 64u : 32u;
    return (uint64_t)total > 2u * resident4 ? 24u : 32u;
}

vkr_error vkr_trace_frame(vkr_ctx* ctx, uint32_t scene, const vkr_frame_data* fd, const vkr_trace_params* params,
                          const vkr_outputs* outputs, void* stream)
{
    if (!ctx || !fd || !outputs) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    const vkr_error pe = validate_params(params);
    if (pe != VKR_SUCCESS) return pe;
    if ((params->flags & (VKR_TRACE_RECORD_DECISIONS | VKR_TRACE_REPLAY_DECISIONS)) && !outputs->decisions)
        return fail(VKR_INVALID_ARGUMENT, "decision flags need outputs.decisions");
    if ((params->flags & (VKR_TRACE_RECORD_DECISIONS | VKR_TRACE_REPLAY_DECISIONS)) && params->max_depth > 8)
        return fail(VKR_INVALID_ARGUMENT, "decision record/replay supports max_depth <= 8");
    const bool rec_adj = (params->flags & VKR_TRACE_RECORD_ADJOINT) != 0;
    if (rec_adj && params->kernel != VKR_KERNEL_PATH)
        return fail(VKR_INVALID_ARGUMENT, "VKR_TRACE_RECORD_ADJOINT needs the path kernel");
    if (rec_adj && params->max_depth > 10)
        return fail(VKR_INVALID_ARGUMENT, "VKR_TRACE_RECORD_ADJOINT supports max_depth <= 10 (as the adjoint)");
    if (rec_adj && (params->flags & (VKR_TRACE_RECORD_DECISIONS | VKR_TRACE_REPLAY_DECISIONS)))
        return fail(VKR_INVALID_ARGUMENT, "VKR_TRACE_RECORD_ADJOINT takes no decision flags (as the adjoint)");
    Scene* s = nullptr;
    if (params->kernel == VKR_KERNEL_PATH) {
        s = get_scene(ctx, scene);
        if (!s) return fail(VKR_INVALID_ARGUMENT, "unknown scene");
    }
    VKR_HIP(hipSetDevice(ctx->device));
    const vkr_error ee = ensure_events(ctx);
    if (ee != VKR_SUCCESS) return ee;
    hipStream_t st = static_cast<hipStream_t>(stream);
    static const Scene kNoScene{};
    JobDesc j = make_job(s ? *s : kNoScene, *fd, *params, outputs);
    if (rec_adj) {
        // The records are rewritten from here on: they describe this frame
        // only once its launch is in the stream.
        ctx->adj_key.valid = false;
        const vkr_error ae = ensure_adj_samples(ctx, (size_t)j.W * j.H * j.spp);
        if (ae != VKR_SUCCESS) return ae;
        j.adj_samples = ctx->adj_samples.as<uint2>();
    }
    DevCounters* cnt = ctx->counters.as<DevCounters>();
    uint32_t* pool = ctx->pool.as<uint32_t>();
    VKR_HIP(launch_zero_counters(cnt, pool, st));
    if (params->kernel == VKR_KERNEL_WHITTED) {
        VKR_HIP(hipEventRecord(ctx->ev0, st));
        VKR_HIP(launch_whitted(j, cnt, st));
        VKR_HIP(hipEventRecord(ctx->ev1, st));
    } else {
        PathLaunch a{};
        a.job = j;
        a.total = j.n_tiles * 64u;
        a.pool = pool;
        a.counters = cnt;
        uint32_t block = params->block, chunk = params->chunk, sched = params->sched;
        uint32_t stack = s->dev.stack_needed;
        const TuneEntry* tuned = find_tuned(ctx, scene, *params);
        if (const TuneEntry* t = tuned) {
            if (!block) block = t->block;
            if (!chunk) chunk = t->chunk;
            if (sched == VKR_SCHED_AUTO) sched = t->sched;
            stack = std::max(stack, t->stack);
        }
        if (!block) block = 128u;
        const bool persistent = sched != VKR_SCHED_PLAIN;
        a.persistent = persistent ? 1u : 0u;
        a.chunk = chunk ? chunk : kDefaultChunk;
        PathVariant v{block, stack, (ctx->flags & VKR_CTX_COUNT_WORK) != 0, false,
                      use_ray_pool(params->flags, s->dev.n_emitters)};
        if (!tuned) {
            v.stack = default_stack(v, a.total, s->info.depth, ctx->device);
            // The 64-entry-stack kernel runs best in one-wave workgroups
            // (512^2 with 64 spheres: 0.52 against 0.61 ms at 128 lanes).
            if (v.stack == 64u && params->block == 0) block = v.block = 64u;
        }
        // Persistent: one resident grid.  One pixel per lane: launches of at
        // most kMaxPlainLanes lanes each, in pool order.
        const uint32_t step = persistent ? a.total : kMaxPlainLanes;
        const uint32_t grid = persistent ? std::min(persistent_grid(v, ctx->device), (a.total + block - 1u) / block)
                                         : (std::min(a.total, step) + block - 1u) / block;
        if (v.pool) {
            const vkr_error le = ensure_lanes(ctx, (size_t)grid * block);
            if (le != VKR_SUCCESS) return le;
            a.lanes = ctx->lanes.as<char>();
        }
        VKR_HIP(hipEventRecord(ctx->ev0, st));
        for (uint32_t base = 0; base < a.total; base += step) {
            a.base = base;
            const uint32_t g = persistent ? grid : (std::min(a.total - base, step) + block - 1u) / block;
            VKR_HIP(launch_path(a, v, g, st));
        }
        VKR_HIP(hipEventRecord(ctx->ev1, st));
        if (rec_adj) ctx->adj_key = adjoint_key(s->serial, *fd, *params);
    }
    ctx->timed = true;
    ctx->last_static = false;
    ctx->last_counters = nullptr;
    return VKR_SUCCESS;
}

vkr_error vkr_trace_frame_static(vkr_ctx* ctx, uint32_t scene, const vkr_frame_data* fd,
                                 const vkr_trace_params* params, const vkr_outputs* outputs, void* stream)
{
    // Allocates nothing: no hipMalloc, no events, no growth of context
    // buffers.  The kernels' scratch is the device's static array.
    if (!ctx || !fd || !outputs) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    const vkr_error pe = validate_params(params);
    if (pe != VKR_SUCCESS) return pe;
    if (params->flags & VKR_TRACE_RECORD_ADJOINT)
        return fail(VKR_INVALID_ARGUMENT, "VKR_TRACE_RECORD_ADJOINT grows a context buffer; use vkr_trace_frame");
    if ((params->flags & (VKR_TRACE_RECORD_DECISIONS | VKR_TRACE_REPLAY_DECISIONS)) && !outputs->decisions)
        return fail(VKR_INVALID_ARGUMENT, "decision flags need outputs.decisions");
    if ((params->flags & (VKR_TRACE_RECORD_DECISIONS | VKR_TRACE_REPLAY_DECISIONS)) && params->max_depth > 8)
        return fail(VKR_INVALID_ARGUMENT, "decision record/replay supports max_depth <= 8");
    Scene* s = nullptr;
    if (params->kernel == VKR_KERNEL_PATH) {
        s = get_scene(ctx, scene);
        if (!s) return fail(VKR_INVALID_ARGUMENT, "unknown scene");
    }
    VKR_HIP(hipSetDevice(ctx->device));
    char* scratch = nullptr;
    VKR_HIP(static_scratch(ctx->device, &scratch));
    DevCounters* cnt = reinterpret_cast<DevCounters*>(scratch);
    uint32_t* pool = reinterpret_cast<uint32_t*>(scratch + kStaticPoolOff);
    char* lanes = scratch + kStaticLanesOff;
    const size_t max_lanes = (size_t)VKR_STATIC_SCRATCH_LANE_BYTES / kLaneBytes;
    hipStream_t st = static_cast<hipStream_t>(stream);
    static const Scene kNoScene{};
    JobDesc j = make_job(s ? *s : kNoScene, *fd, *params, outputs);
    if (params->kernel == VKR_KERNEL_WHITTED) {
        VKR_HIP(launch_zero_counters(cnt, pool, st));
        VKR_HIP(launch_whitted(j, cnt, st));
    } else {
        PathLaunch a{};
        a.job = j;
        a.total = j.n_tiles * 64u;
        a.pool = pool;
        a.counters = cnt;
        uint32_t block = params->block, chunk = params->chunk, sched = params->sched;
        uint32_t stack = s->dev.stack_needed;
        const TuneEntry* tuned = find_tuned(ctx, scene, *params);
        if (const TuneEntry* t = tuned) {
            if (!block) block = t->block;
            if (!chunk) chunk = t->chunk;
            if (sched == VKR_SCHED_AUTO) sched = t->sched;
            stack = std::max(stack, t->stack);
        }
        if (!block) block = 128u;
        const bool persistent = sched != VKR_SCHED_PLAIN;
        a.persistent = persistent ? 1u : 0u;
        a.chunk = chunk ? chunk : kDefaultChunk;
        PathVariant v{block, stack, (ctx->flags & VKR_CTX_COUNT_WORK) != 0, false,
                      use_ray_pool(params->flags, s->dev.n_emitters)};
        if (!tuned) {
            v.stack = default_stack(v, a.total, s->info.depth, ctx->device);
            // The static records cap the pooled grid below what even 4 waves
            // keep resident; then the fifth wave cannot help.
            PathVariant v4 = v;
            v4.stack = 32u;
            if (v.stack == 24u && v.pool && max_lanes <= (size_t)persistent_grid(v4, ctx->device) * block)
                v.stack = 32u;
        }
        const uint32_t step = persistent ? a.total : kMaxPlainLanes;
        uint32_t grid = persistent ? std::min(persistent_grid(v, ctx->device), (a.total + block - 1u) / block)
                                   : (std::min(a.total, step) + block - 1u) / block;
        if (v.pool) {
            // The ray records of every resident lane must fit the array: a
            // persistent grid is capped to them, a one-pixel-per-lane launch
            // (one record per pixel) is refused when they do not.
            const uint32_t cap = (uint32_t)(max_lanes / block);
            if (!persistent && (size_t)grid * block > max_lanes)
                return fail(VKR_INVALID_ARGUMENT,
                            "VKR_SCHED_PLAIN with the ray pool needs one ray record per pixel, more than the static "
                            "scratch holds (VKR_STATIC_SCRATCH_LANE_BYTES); use the persistent schedule");
            if (cap == 0) return fail(VKR_INVALID_ARGUMENT, "workgroup too large for the static scratch");
            grid = std::min(grid, cap);
            a.lanes = lanes;
        }
        VKR_HIP(launch_zero_counters(cnt, pool, st));
        for (uint32_t base = 0; base < a.total; base += step) {
            a.base = base;
            const uint32_t g = persistent ? grid : (std::min(a.total - base, step) + block - 1u) / block;
            VKR_HIP(launch_path(a, v, g, st));
        }
    }
    ctx->timed = false;
    ctx->last_static = true;
    ctx->last_counters = cnt;
    return VKR_SUCCESS;
}

vkr_error vkr_last_counters(vkr_ctx* ctx, vkr_counters* out)
{
    if (!ctx || !out) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    VKR_HIP(hipSetDevice(ctx->device));
    if (ctx->timed) VKR_HIP(hipEventSynchronize(ctx->ev1));
    if (ctx->last_static) VKR_HIP(hipDeviceSynchronize());   // that call recorded no events
    DevCounters c{};
    const void* src = ctx->last_counters ? (const void*)ctx->last_counters : ctx->counters.p;
    VKR_HIP(hipMemcpy(&c, src, sizeof(c), hipMemcpyDeviceToHost));
    out->rays = c.rays;
    out->shadow_rays = c.shadow;
    out->nodes_visited = c.nodes;
    out->tris_tested = c.tris;
    out->spheres_tested = c.spheres;
    out->planes_tested = c.planes;
    out->window_overflow = c.overflow;
    out->rays_traced = c.traced;
    out->wave_node_iters = c.wave_nodes;
    out->wave_leaf_iters = c.wave_leaves;
    out->wave_steps = c.wave_steps;
    return VKR_SUCCESS;
}

vkr_error vkr_last_kernel_ms(vkr_ctx* ctx, float* out_ms)
{
    if (!ctx || !out_ms) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    if (ctx->last_static) return fail(VKR_INVALID_ARGUMENT, "the last call (vkr_trace_frame_static) records no events");
    if (!ctx->timed) return fail(VKR_INVALID_ARGUMENT, "nothing launched yet");
    VKR_HIP(hipSetDevice(ctx->device));
    VKR_HIP(hipEventSynchronize(ctx->ev1));
    VKR_HIP(hipEventElapsedTime(out_ms, ctx->ev0, ctx->ev1));
    return VKR_SUCCESS;
}

vkr_error vkr_filter_frame(vkr_ctx* ctx, const uint8_t* current, const uint8_t* previous, uint32_t src_w,
                           uint32_t src_h, uint8_t* out, uint32_t dst_w, uint32_t dst_h, void* stream)
{
    if (!ctx || !current || !previous || !out) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    if (!src_w || !src_h || !dst_w || !dst_h || src_w > 16384 || src_h > 16384 || dst_w > 16384 || dst_h > 16384)
        return fail(VKR_INVALID_ARGUMENT, "image sizes must be 1..16384");
    VKR_HIP(hipSetDevice(ctx->device));
    VKR_HIP(launch_filter(current, previous, src_w, src_h, out, dst_w, dst_h, static_cast<hipStream_t>(stream)));
    return VKR_SUCCESS;
}

vkr_error vkr_debug_eval(vkr_ctx* ctx, uint32_t op, const float* in, float* out, uint32_t n, void* stream)
{
    if (!ctx || !in || !out) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    if (op > VKR_DEBUG_COS) return fail(VKR_INVALID_ARGUMENT, "unknown debug op");
    if (n == 0) return VKR_SUCCESS;
    VKR_HIP(hipSetDevice(ctx->device));
    VKR_HIP(launch_debug(op, in, out, n, static_cast<hipStream_t>(stream)));
    return VKR_SUCCESS;
}

vkr_error vkr_camera_from_angles(const float pos[3], float pitch, float yaw, vkr_camera* out)
{
    // Camera::update (Source/Camera.cpp:48-64) with GLM 0.9.9 float
    // semantics: radians(x) = x * float(pi / 180), normalize(v) =
    // v * (1 / sqrt(dot(v, v))), dot = (x*x' + y*y') + z*z'.
    if (!pos || !out) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    struct V3 { float x, y, z; };
    auto dot3 = [](V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    auto nrm = [&](V3 a) {
        const float inv = 1.0f / std::sqrt(dot3(a, a));
        return V3{a.x * inv, a.y * inv, a.z * inv};
    };
    auto crs = [](V3 a, V3 b) {
        return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
    };
    const float p = std::min(std::max(pitch, -89.0f), 89.0f);
    const float deg = 0.01745329251994329576923690768489f;
    const float pr = p * deg, yr = yaw * deg;
    V3 front{std::cos(pr) * std::cos(yr), std::sin(pr), std::cos(pr) * std::sin(yr)};
    front = nrm(front);
    const V3 P{pos[0], pos[1], pos[2]};
    const V3 moved{P.x + front.x, P.y + front.y, P.z + front.z};
    const V3 dir = nrm(V3{moved.x - P.x, moved.y - P.y, moved.z - P.z});
    const V3 right = nrm(crs(V3{0.0f, 1.0f, 0.0f}, dir));
    const V3 up = crs(dir, right);
    std::memset(out, 0, sizeof(*out));
    out->pos[0] = P.x; out->pos[1] = P.y; out->pos[2] = P.z;
    out->dir[0] = dir.x; out->dir[1] = dir.y; out->dir[2] = dir.z;
    out->right[0] = right.x; out->right[1] = right.y; out->right[2] = right.z;
    out->up[0] = up.x; out->up[1] = up.y; out->up[2] = up.z;
    return VKR_SUCCESS;
}

vkr_error vkr_adjoint(vkr_ctx* ctx, uint32_t scene, const vkr_frame_data* fd, const vkr_trace_params* params,
                      const float* d_radiance, float* d_params, void* stream)
{
    if (!ctx || !fd || !d_radiance || !d_params) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    const vkr_error pe = validate_params(params);
    if (pe != VKR_SUCCESS) return pe;
    if (params->kernel != VKR_KERNEL_PATH) return fail(VKR_INVALID_ARGUMENT, "the adjoint differentiates the path kernel");
    if (params->max_depth > 10) return fail(VKR_INVALID_ARGUMENT, "the adjoint supports max_depth <= 10");
    if (params->flags & (VKR_TRACE_RECORD_DECISIONS | VKR_TRACE_REPLAY_DECISIONS))
        return fail(VKR_INVALID_ARGUMENT, "the adjoint replays the natural decisions; no decision flags");
    Scene* s = get_scene(ctx, scene);
    if (!s) return fail(VKR_INVALID_ARGUMENT, "unknown scene");
    VKR_HIP(hipSetDevice(ctx->device));
    const vkr_error ee = ensure_events(ctx);
    if (ee != VKR_SUCCESS) return ee;
    hipStream_t st = static_cast<hipStream_t>(stream);
    AdjointLaunch a{};
    a.job = make_job(*s, *fd, *params, nullptr);
    a.d_radiance = d_radiance;
    a.d_params = d_params;
    a.params = s->theta.as<float>();
    a.counters = ctx->counters.as<DevCounters>();
    a.pool = ctx->pool.as<uint32_t>();
    a.total = a.job.n_tiles * 64u;
    // The second pass hands out one item per (pixel, sample) of whole tiles.
    if ((uint64_t)a.total * a.job.spp > kMaxPoolItems)
        return fail(VKR_UNSUPPORTED, "the adjoint takes at most 0xFF000000 (pixel, sample) pairs per call (8x8 tiles)");
    // The first pass's records, 8 bytes per (pixel, sample), are still those
    // of this frame when a recording vkr_trace_frame or an adjoint wrote
    // them for the same scene version, frame data and trace parameters.
    const AdjointKey key = adjoint_key(s->serial, *fd, *params);
    const bool recorded = same_key(ctx->adj_key, key);
    if (!recorded) {
        ctx->adj_key.valid = false;
        const vkr_error ae = ensure_adj_samples(ctx, (size_t)a.job.W * a.job.H * a.job.spp);
        if (ae != VKR_SUCCESS) return ae;
    }
    a.samples = ctx->adj_samples.as<uint2>();
    VKR_HIP(launch_zero_counters(a.counters, a.pool, st));
    VKR_HIP(hipEventRecord(ctx->ev0, st));
    VKR_HIP(launch_adjoint(a, s->dev.stack_needed, !recorded, st));
    VKR_HIP(hipEventRecord(ctx->ev1, st));
    ctx->adj_key = key;
    ctx->timed = true;
    ctx->last_static = false;
    ctx->last_counters = nullptr;
    return VKR_SUCCESS;
}

}  // extern "C"
