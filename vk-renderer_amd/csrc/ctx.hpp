// ctx.hpp - context, scene and graph objects behind the C ABI (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vkr.h"
#include "bvh.hpp"
#include "kernels.hpp"

namespace vkr {

extern thread_local std::string g_err;

inline vkr_error fail(vkr_error code, const std::string& msg)
{
    g_err = msg;
    return code;
}

inline vkr_error fail_hip(const char* what, hipError_t e)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? VKR_OUT_OF_MEMORY : VKR_HIP_ERROR;
}

#define VKR_HIP(call)                                    \
    do {                                                 \
        const hipError_t e_ = (call);                    \
        if (e_ != hipSuccess) return fail_hip(#call, e_); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n)
    {
        reset();
        if (n == 0) n = 16;
        const hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    hipError_t upload(const void* src, size_t n)
    {
        const hipError_t e = alloc(n);
        if (e != hipSuccess) return e;
        if (n && src) return hipMemcpy(p, src, n, hipMemcpyHostToDevice);
        return hipSuccess;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

const vkr_material kMirror = {{1.0f, 0.5f, 0.5f}, 0.0f, {0.0f, 0.0f, 0.0f}, 0.0f, 1.0f, VKR_MAT_DIFFUSE, {0.0f, 0.0f}};

struct Scene {
    DevBuf nodes, tris, sph, sph_key, tris_orig, spheres, planes, tri_mat, materials, emitters;
    DevBuf theta;                    // device copy of the sphere parameters (adjoint)
    std::vector<vkr_sphere> host_spheres;
    std::vector<float> host_theta;   // 7 * n_spheres + 1
    DevScene dev{};
    vkr_scene_info info{};
    uint64_t serial = 0;             // version of the contents: new at upload and whenever a parameter value changes
};

struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf jobs, job_start, pool, counters, lanes;
    uint64_t total = 0;
    uint32_t n_jobs = 0;
    std::vector<uint32_t> job_scene;   // scene id of every job: the graph's table holds copies of their DevScene
    ~Graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};


// Autotuned launch shape for one (scene, image size).
struct TuneEntry {
    uint32_t scene, width, height;
    uint32_t block, chunk, stack, sched;
};

// What the adjoint's first-pass records held in vkr_ctx::adj_samples were
// computed for: the scene version and every input the paths depend on.
struct AdjointKey {
    bool valid = false;
    uint64_t serial = 0;
    vkr_frame_data fd{};
    uint32_t width = 0, height = 0, spp = 0, max_depth = 0;
    float epsilon = 0.0f;
};

inline AdjointKey adjoint_key(uint64_t serial, const vkr_frame_data& fd, const vkr_trace_params& p)
{
    AdjointKey k;
    k.valid = true;
    k.serial = serial;
    k.fd = fd;
    k.width = p.width;
    k.height = p.height;
    k.spp = p.spp;
    k.max_depth = p.max_depth;
    k.epsilon = p.epsilon;
    return k;
}

inline bool same_key(const AdjointKey& a, const AdjointKey& b)
{
    // Bytewise on the frame data: padding that differs only costs a first pass.
    return a.valid && b.valid && a.serial == b.serial && std::memcmp(&a.fd, &b.fd, sizeof(a.fd)) == 0 &&
           a.width == b.width && a.height == b.height && a.spp == b.spp && a.max_depth == b.max_depth &&
           std::memcmp(&a.epsilon, &b.epsilon, sizeof(float)) == 0;
}

// Autotuned launch shape for one batch job set (batch.cpp: batch_key).
struct BatchTuneEntry {
    uint64_t key;
    uint32_t block, chunk, stack;
};

}  // namespace vkr

// The opaque context of the C ABI (global, matching vkr.h's forward declaration).
struct vkr_ctx {
    int device = 0;
    uint32_t flags = 0;
    std::vector<std::unique_ptr<vkr::Scene>> scenes;
    std::vector<std::unique_ptr<vkr::Graph>> graphs;
    vkr::DevBuf counters, pool, jobs, job_start;
    vkr::DevBuf lanes;                 // per-lane ray records of direct launches (grown on demand)
    vkr::DevBuf build_scratch;         // device BVH builds (grown on demand, reused across uploads)
    vkr::DevBuf adj_samples;           // the adjoint's per-sample records (grown on demand)
    vkr::AdjointKey adj_key;           // what adj_samples holds (valid: a later adjoint may skip its first pass)
    uint64_t next_serial = 1;          // source of Scene::serial
    void* host_jobs = nullptr;         // pinned staging for batched job tables
    size_t host_jobs_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_copy = nullptr;
    bool timed = false;
    bool last_static = false;          // the last call was vkr_trace_frame_static (no events)
    const vkr::DevCounters* last_counters = nullptr;   // NULL = ctx counters
    std::vector<vkr::TuneEntry> tuned;
    std::vector<vkr::BatchTuneEntry> tuned_batch;
};

namespace vkr {

inline Scene* get_scene(vkr_ctx* ctx, uint32_t id)
{
    if (!ctx || id >= ctx->scenes.size()) return nullptr;
    return ctx->scenes[id].get();
}

inline bool nonzero3(const float* v) { return !(v[0] == 0.0f && v[1] == 0.0f && v[2] == 0.0f); }

inline std::vector<uint32_t> emitter_list(const std::vector<vkr_sphere>& s)
{
    // Tracer.comp:458-462: every sphere whose emissive is not vec3(0), in order.
    std::vector<uint32_t> e;
    for (uint32_t i = 0; i < s.size(); ++i)
        if (nonzero3(s[i].mat.emissive)) e.push_back(i);
    return e;
}

// 1 when every emitter has emission below 2^20 in magnitude and a radius in
// [2^-20, 2^20]: then its attenuation factor is below 2^41 at any hit point
// at least 2^-9 radii from its centre (used by the kernel's exact
// zero-contribution tests).
inline uint32_t emitters_bounded(const std::vector<vkr_sphere>& s, const std::vector<uint32_t>& em)
{
    for (uint32_t i : em) {
        const vkr_sphere& q = s[i];
        for (int k = 0; k < 3; ++k)
            if (!(std::fabs(q.mat.emissive[k]) < 0x1p20f)) return 0u;
        if (!(q.r >= 0x1p-20f && q.r <= 0x1p20f)) return 0u;
    }
    return 1u;
}

inline vkr_error validate_params(const vkr_trace_params* p)
{
    if (!p) return fail(VKR_INVALID_ARGUMENT, "params is NULL");
    if (p->width == 0 || p->height == 0 || p->width > 16384 || p->height > 16384)
        return fail(VKR_INVALID_ARGUMENT, "image size must be 1..16384 per side");
    if (p->kernel != VKR_KERNEL_PATH && p->kernel != VKR_KERNEL_WHITTED)
        return fail(VKR_INVALID_ARGUMENT, "unknown kernel");
    if (p->kernel == VKR_KERNEL_PATH) {
        if (p->spp == 0 || p->spp > 4096) return fail(VKR_INVALID_ARGUMENT, "spp must be 1..4096");
        if (p->max_depth == 0 || p->max_depth > 16) return fail(VKR_INVALID_ARGUMENT, "max_depth must be 1..16");
        if (!(p->epsilon >= 0.0f) || !std::isfinite(p->epsilon))
            return fail(VKR_INVALID_ARGUMENT, "epsilon must be finite and >= 0");
    }
    if (p->block != 0 && p->block != 64 && p->block != 128 && p->block != 256)
        return fail(VKR_INVALID_ARGUMENT, "block must be 0, 64, 128 or 256");
    if (p->chunk > 64) return fail(VKR_INVALID_ARGUMENT, "chunk must be 0..64");
    if ((p->flags & VKR_TRACE_RECORD_DECISIONS) && (p->flags & VKR_TRACE_REPLAY_DECISIONS))
        return fail(VKR_INVALID_ARGUMENT, "record and replay decisions are exclusive");
    if ((p->flags & VKR_TRACE_RAY_POOL) && (p->flags & VKR_TRACE_NO_RAY_POOL))
        return fail(VKR_INVALID_ARGUMENT, "VKR_TRACE_RAY_POOL and VKR_TRACE_NO_RAY_POOL are exclusive");
    return VKR_SUCCESS;
}

inline JobDesc make_job(const Scene& sc, const vkr_frame_data& fd, const vkr_trace_params& p, const vkr_outputs* out)
{
    JobDesc j{};
    j.sc = sc.dev;
    for (int k = 0; k < 3; ++k) {
        j.cam_pos[k] = fd.camera.pos[k];
        j.cam_dir[k] = fd.camera.dir[k];
        j.cam_right[k] = fd.camera.right[k];
        j.cam_up[k] = fd.camera.up[k];
        j.light_pos[k] = fd.light_pos[k];
    }
    j.aspect = fd.aspect_ratio;
    j.seed = fd.seed;
    j.eps = p.epsilon == 0.0f ? 1e-3f : p.epsilon;
    j.flags = p.flags;
    j.kernel = p.kernel;
    j.W = p.width;
    j.H = p.height;
    j.spp = p.spp;
    j.max_depth = p.max_depth;
    j.tiles_x = (p.width + 7u) / 8u;
    j.n_tiles = j.tiles_x * ((p.height + 7u) / 8u);
    if (out) {
        j.rgba8 = out->rgba8;
        j.radiance = out->radiance;
        j.prim_id = out->prim_id;
        j.hit_t = out->hit_t;
        j.ray_count = out->ray_count;
        j.decisions = out->decisions;
    }
    return j;
}

inline vkr_error ensure_events(vkr_ctx* ctx)
{
    if (!ctx->ev0) VKR_HIP(hipEventCreate(&ctx->ev0));
    if (!ctx->ev1) VKR_HIP(hipEventCreate(&ctx->ev1));
    if (!ctx->ev_copy) VKR_HIP(hipEventCreateWithFlags(&ctx->ev_copy, hipEventDisableTiming));
    return VKR_SUCCESS;
}

// Makes the context's per-lane ray records hold `lanes` lanes.  Growing
// waits for the device first: a launch in flight may still use the old
// buffer.
inline vkr_error ensure_lanes(vkr_ctx* ctx, size_t lanes)
{
    const size_t bytes = lanes * kLaneBytes;
    if (ctx->lanes.bytes >= bytes) return VKR_SUCCESS;
    VKR_HIP(hipDeviceSynchronize());
    VKR_HIP(ctx->lanes.alloc(bytes));
    return VKR_SUCCESS;
}

// Whether a launch uses the pooled variant of the path kernel: forced by the
// trace flags, else when the scene has at least 2 emitters (with one, no lane
// ever queues more than one ray and the pool is pure bookkeeping).
inline bool use_ray_pool(uint32_t flags, uint32_t n_emitters)
{
    if (flags & VKR_TRACE_RAY_POOL) return true;
    if (flags & VKR_TRACE_NO_RAY_POOL) return false;
    return n_emitters >= 2u;
}

inline const TuneEntry* find_tuned(const vkr_ctx* ctx, uint32_t scene, const vkr_trace_params& p)
{
    for (const TuneEntry& t : ctx->tuned)
        if (t.scene == scene && t.width == p.width && t.height == p.height) return &t;
    return nullptr;
}

}  // namespace vkr
