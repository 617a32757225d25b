// batch.cpp - batched persistent launches, hipGraph capture and the
// autotuner (build extensions of the reference's per-frame Draw,
// Source/GraphicsDevice.cpp:1215-1342, which dispatches one image at a time).
//
// A batch is one persistent launch of the BATCH variant of k_path: its waves
// pull 8x8-tile work items of every job from one global pool, so small images
// and cheap scenes fill the machine together and no wave idles on a launch
// tail.  Each job's results are bit-identical to vkr_trace_frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/vkr.h"
#include "ctx.hpp"
#include "kernels.hpp"

using namespace vkr;

namespace {

struct BatchPlan {
    std::vector<JobDesc> jobs;
    std::vector<uint32_t> start;  // n + 1 pool offsets
    uint32_t stack = 24;          // the smallest stack variant; raised to the jobs' needs
    uint32_t tree_depth = 0;      // the deepest of the jobs' trees
    uint64_t total = 0;
    bool pool = false;            // some job wants the pooled kernel
};

vkr_error plan_batch(vkr_ctx* ctx, const vkr_scene_job* jobs, uint32_t n, BatchPlan& plan)
{
    if (!jobs || n == 0) return fail(VKR_INVALID_ARGUMENT, "empty batch");
    plan.jobs.resize(n);
    plan.start.resize(n + 1);
    uint64_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const vkr_scene_job& jb = jobs[i];
        const vkr_error pe = validate_params(&jb.params);
        if (pe != VKR_SUCCESS) return pe;
        if (jb.params.kernel != VKR_KERNEL_PATH) return fail(VKR_INVALID_ARGUMENT, "batch jobs must use the path kernel");
        if ((jb.params.flags & (VKR_TRACE_RECORD_DECISIONS | VKR_TRACE_REPLAY_DECISIONS)) &&
            (!jb.outputs.decisions || jb.params.max_depth > 8))
            return fail(VKR_INVALID_ARGUMENT, "decision flags need outputs.decisions and max_depth <= 8");
        if (jb.params.flags & VKR_TRACE_RECORD_ADJOINT)
            return fail(VKR_INVALID_ARGUMENT, "VKR_TRACE_RECORD_ADJOINT is for vkr_trace_frame only");
        Scene* s = get_scene(ctx, jb.scene);
        if (!s) return fail(VKR_INVALID_ARGUMENT, "unknown scene in batch job");
        plan.jobs[i] = make_job(*s, jb.frame_data, jb.params, &jb.outputs);
        plan.stack = std::max(plan.stack, s->dev.stack_needed);
        plan.tree_depth = std::max(plan.tree_depth, s->info.depth);
        plan.pool = plan.pool || use_ray_pool(jb.params.flags, s->dev.n_emitters);
        plan.start[i] = (uint32_t)off;
        off += (uint64_t)plan.jobs[i].n_tiles * 64u;
        if (off >= 0xFFFFFFF0ull) return fail(VKR_INVALID_ARGUMENT, "batch exceeds 2^32 pool entries");
    }
    plan.start[n] = (uint32_t)off;
    plan.total = off;
    return VKR_SUCCESS;
}

PathLaunch batch_launch(const BatchPlan& plan, const JobDesc* djobs, const uint32_t* dstart, uint32_t n,
                        uint32_t chunk, uint32_t* pool, DevCounters* cnt)
{
    PathLaunch a{};
    a.jobs = djobs;
    a.job_start = dstart;
    a.n_jobs = n;
    a.total = (uint32_t)plan.total;
    a.persistent = 1u;
    a.chunk = chunk ? chunk : kDefaultChunk;
    a.pool = pool;
    a.counters = cnt;
    return a;
}

PathVariant batch_variant(const BatchPlan& plan, uint32_t block, bool count)
{
    return PathVariant{block, plan.stack, count, true, plan.pool};
}

uint32_t batch_grid(const BatchPlan& plan, const PathVariant& v, int device)
{
    const uint32_t need = (uint32_t)((plan.total + v.block - 1u) / v.block);
    return std::max(1u, std::min(persistent_grid(v, device), need));
}

bool valid_block(uint32_t b) { return b == 0 || b == 64 || b == 128 || b == 256; }

// The 32-bit pool cursor: once the pool is empty, every resident wave that
// has not yet seen it run dry still claims up to its 64 lanes' worth, so the
// cursor can pass the total by at most the number of resident lanes.
vkr_error check_cursor(const BatchPlan& plan, uint32_t grid, uint32_t block)
{
    if (plan.total + (uint64_t)grid * block > 0xFFFFFFFFull)
        return fail(VKR_INVALID_ARGUMENT, "batch too large: pool entries plus resident lanes must stay below 2^32");
    return VKR_SUCCESS;
}

// A batch's job set, as the batch autotuner keys its result: every job's
// scene, image size, spp, depth and flags, in order (FNV-1a).
uint64_t batch_key(const vkr_scene_job* jobs, uint32_t n)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint32_t v) {
        for (int k = 0; k < 4; ++k) {
            h ^= (v >> (8 * k)) & 0xFFu;
            h *= 1099511628211ull;
        }
    };
    mix(n);
    for (uint32_t i = 0; i < n; ++i) {
        const vkr_scene_job& j = jobs[i];
        mix(j.scene);
        mix(j.params.width);
        mix(j.params.height);
        mix(j.params.spp);
        mix(j.params.max_depth);
        mix(j.params.flags);
    }
    return h;
}

struct BatchShape {
    uint32_t block, chunk;
};

// The launch shape of a batch: the caller's block and chunk, else the batch
// autotuner's choice for this job set (which may also raise the stack
// depth), else 128 lanes and the default chunk.  Untuned batches whose trees
// all have at most 8 levels take the register-rich kernels, as single frames
// do (they shade more than they fetch nodes): the 64-entry-stack kernel when
// the grid of the 4-wave one holds every pool entry at once, else the
// 32-entry one.
BatchShape batch_shape(vkr_ctx* ctx, const vkr_scene_job* jobs, uint32_t n, BatchPlan& plan, uint32_t block,
                       uint32_t chunk)
{
    BatchShape s{block, chunk};
    bool tuned = false;
    if (block == 0 && chunk == 0) {
        const uint64_t key = batch_key(jobs, n);
        for (const BatchTuneEntry& t : ctx->tuned_batch) {
            if (t.key != key) continue;
            s.block = t.block;
            s.chunk = t.chunk;
            plan.stack = std::max(plan.stack, t.stack);
            tuned = true;
            break;
        }
    }
    if (!s.block) s.block = 128u;
    if (!s.chunk) s.chunk = kDefaultChunk;
    if (!tuned && plan.stack == 24u && plan.tree_depth <= 8u) {
        const PathVariant v4{s.block, 32u, false, true, plan.pool};
        const uint64_t resident4 = (uint64_t)persistent_grid(v4, ctx->device) * s.block;
        plan.stack = plan.total <= resident4 ? 64u : 32u;
    }
    return s;
}

}  // namespace

extern "C" {

vkr_error vkr_trace_batch(vkr_ctx* ctx, const vkr_scene_job* jobs, uint32_t n_jobs, uint32_t block, uint32_t chunk,
                          void* stream)
{
    if (!ctx) return fail(VKR_INVALID_ARGUMENT, "ctx is NULL");
    if (!valid_block(block) || chunk > 64) return fail(VKR_INVALID_ARGUMENT, "block must be 0/64/128/256, chunk <= 64");
    BatchPlan plan;
    const vkr_error pe = plan_batch(ctx, jobs, n_jobs, plan);
    if (pe != VKR_SUCCESS) return pe;
    VKR_HIP(hipSetDevice(ctx->device));
    const vkr_error ee = ensure_events(ctx);
    if (ee != VKR_SUCCESS) return ee;
    hipStream_t st = static_cast<hipStream_t>(stream);

    // Job table: staged through pinned memory, copied in stream order after
    // the previous launch of this context finished with the device copy.
    const size_t jb = plan.jobs.size() * sizeof(JobDesc);
    const size_t sb = plan.start.size() * sizeof(uint32_t);
    VKR_HIP(hipEventSynchronize(ctx->ev_copy));
    if (ctx->host_jobs_bytes < jb + sb) {
        if (ctx->host_jobs) VKR_HIP(hipHostFree(ctx->host_jobs));
        ctx->host_jobs = nullptr;
        ctx->host_jobs_bytes = 0;
        VKR_HIP(hipHostMalloc(&ctx->host_jobs, jb + sb, hipHostMallocDefault));
        ctx->host_jobs_bytes = jb + sb;
    }
    if (ctx->jobs.bytes < jb) VKR_HIP(ctx->jobs.alloc(jb));
    if (ctx->job_start.bytes < sb) VKR_HIP(ctx->job_start.alloc(sb));
    std::memcpy(ctx->host_jobs, plan.jobs.data(), jb);
    std::memcpy(static_cast<char*>(ctx->host_jobs) + jb, plan.start.data(), sb);
    if (ctx->timed) VKR_HIP(hipStreamWaitEvent(st, ctx->ev1, 0));
    VKR_HIP(hipMemcpyAsync(ctx->jobs.p, ctx->host_jobs, jb, hipMemcpyHostToDevice, st));
    VKR_HIP(hipMemcpyAsync(ctx->job_start.p, static_cast<char*>(ctx->host_jobs) + jb, sb, hipMemcpyHostToDevice, st));
    VKR_HIP(hipEventRecord(ctx->ev_copy, st));

    DevCounters* cnt = ctx->counters.as<DevCounters>();
    uint32_t* pool = ctx->pool.as<uint32_t>();
    const bool count = (ctx->flags & VKR_CTX_COUNT_WORK) != 0;
    const BatchShape sh = batch_shape(ctx, jobs, n_jobs, plan, block, chunk);
    PathLaunch a = batch_launch(plan, ctx->jobs.as<JobDesc>(), ctx->job_start.as<uint32_t>(), n_jobs, sh.chunk, pool,
                                cnt);
    const PathVariant v = batch_variant(plan, sh.block, count);
    const uint32_t grid = batch_grid(plan, v, ctx->device);
    const vkr_error ce = check_cursor(plan, grid, sh.block);
    if (ce != VKR_SUCCESS) return ce;
    if (v.pool) {
        const vkr_error le = ensure_lanes(ctx, (size_t)grid * sh.block);
        if (le != VKR_SUCCESS) return le;
        a.lanes = ctx->lanes.as<char>();
    }
    VKR_HIP(launch_zero_counters(cnt, pool, st));
    VKR_HIP(hipEventRecord(ctx->ev0, st));
    VKR_HIP(launch_path(a, v, grid, st));
    VKR_HIP(hipEventRecord(ctx->ev1, st));
    ctx->timed = true;
    ctx->last_static = false;
    ctx->last_counters = nullptr;
    return VKR_SUCCESS;
}

vkr_error vkr_batch_graph_create(vkr_ctx* ctx, const vkr_scene_job* jobs, uint32_t n_jobs, uint32_t block,
                                 uint32_t chunk, void* stream, uint32_t* out_graph)
{
    if (!ctx || !out_graph) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    if (!valid_block(block) || chunk > 64) return fail(VKR_INVALID_ARGUMENT, "block must be 0/64/128/256, chunk <= 64");
    BatchPlan plan;
    const vkr_error pe = plan_batch(ctx, jobs, n_jobs, plan);
    if (pe != VKR_SUCCESS) return pe;
    VKR_HIP(hipSetDevice(ctx->device));
    const bool count = (ctx->flags & VKR_CTX_COUNT_WORK) != 0;
    const BatchShape sh = batch_shape(ctx, jobs, n_jobs, plan, block, chunk);
    const PathVariant v = batch_variant(plan, sh.block, count);
    const uint32_t grid = batch_grid(plan, v, ctx->device);
    const vkr_error ce = check_cursor(plan, grid, sh.block);
    if (ce != VKR_SUCCESS) return ce;

    auto g = std::make_unique<Graph>();
    VKR_HIP(g->jobs.upload(plan.jobs.data(), plan.jobs.size() * sizeof(JobDesc)));
    VKR_HIP(g->job_start.upload(plan.start.data(), plan.start.size() * sizeof(uint32_t)));
    VKR_HIP(g->pool.alloc(64));
    VKR_HIP(g->counters.alloc(sizeof(DevCounters)));
    g->total = plan.total;
    g->n_jobs = n_jobs;
    g->job_scene.resize(n_jobs);
    for (uint32_t i = 0; i < n_jobs; ++i) g->job_scene[i] = jobs[i].scene;
    PathLaunch a = batch_launch(plan, g->jobs.as<JobDesc>(), g->job_start.as<uint32_t>(), n_jobs, sh.chunk,
                                g->pool.as<uint32_t>(), g->counters.as<DevCounters>());
    if (v.pool) {
        VKR_HIP(g->lanes.alloc((size_t)grid * sh.block * kLaneBytes));   // the graph's own records
        a.lanes = g->lanes.as<char>();
    }

    // Capture on a private stream: a fill kernel resets the pool cursor and the
    // counters (no memset nodes), then the persistent batch kernel.
    hipStream_t cap = nullptr;
    VKR_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const hipError_t e1 = launch_zero_counters(g->counters.as<DevCounters>(), g->pool.as<uint32_t>(), cap);
        const hipError_t e2 = launch_path(a, v, grid, cap);
        e = hipStreamEndCapture(cap, &g->graph);
        if (e == hipSuccess && e1 != hipSuccess) e = e1;
        if (e == hipSuccess && e2 != hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cap);
    if (e != hipSuccess) return fail_hip("batch graph capture", e);
    (void)stream;
    for (uint32_t i = 0; i < ctx->graphs.size(); ++i) {
        if (!ctx->graphs[i]) {
            ctx->graphs[i] = std::move(g);
            *out_graph = i;
            return VKR_SUCCESS;
        }
    }
    ctx->graphs.push_back(std::move(g));
    *out_graph = (uint32_t)ctx->graphs.size() - 1u;
    return VKR_SUCCESS;
}

vkr_error vkr_graph_launch(vkr_ctx* ctx, uint32_t graph, void* stream)
{
    if (!ctx || graph >= ctx->graphs.size() || !ctx->graphs[graph]) return fail(VKR_INVALID_ARGUMENT, "unknown graph");
    VKR_HIP(hipSetDevice(ctx->device));
    const vkr_error ee = ensure_events(ctx);
    if (ee != VKR_SUCCESS) return ee;
    hipStream_t st = static_cast<hipStream_t>(stream);
    Graph& g = *ctx->graphs[graph];
    VKR_HIP(hipEventRecord(ctx->ev0, st));
    VKR_HIP(hipGraphLaunch(g.exec, st));
    VKR_HIP(hipEventRecord(ctx->ev1, st));
    ctx->timed = true;
    ctx->last_static = false;
    ctx->last_counters = g.counters.as<DevCounters>();
    return VKR_SUCCESS;
}

vkr_error vkr_graph_set_frame(vkr_ctx* ctx, uint32_t graph, uint32_t job, const vkr_frame_data* frame_data,
                              void* stream)
{
    if (!ctx || !frame_data || graph >= ctx->graphs.size() || !ctx->graphs[graph])
        return fail(VKR_INVALID_ARGUMENT, "unknown graph or NULL frame data");
    Graph& g = *ctx->graphs[graph];
    if (job >= g.n_jobs) return fail(VKR_INVALID_ARGUMENT, "job index out of range");
    VKR_HIP(hipSetDevice(ctx->device));
    VKR_HIP(launch_set_frame(g.jobs.as<JobDesc>() + job, *frame_data, static_cast<hipStream_t>(stream)));
    return VKR_SUCCESS;
}

vkr_error vkr_graph_destroy(vkr_ctx* ctx, uint32_t graph)
{
    if (!ctx || graph >= ctx->graphs.size() || !ctx->graphs[graph]) return fail(VKR_INVALID_ARGUMENT, "unknown graph");
    VKR_HIP(hipSetDevice(ctx->device));
    VKR_HIP(hipDeviceSynchronize());
    if (ctx->last_counters == ctx->graphs[graph]->counters.as<DevCounters>()) ctx->last_counters = nullptr;
    ctx->graphs[graph].reset();
    return VKR_SUCCESS;
}

vkr_error vkr_autotune(vkr_ctx* ctx, uint32_t scene, const vkr_frame_data* fd, const vkr_trace_params* params,
                       const vkr_outputs* outputs, uint32_t repeats, vkr_tune_result* out)
{
    if (!ctx || !fd || !params || !outputs || !out) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    Scene* s = get_scene(ctx, scene);
    if (!s) return fail(VKR_INVALID_ARGUMENT, "unknown scene");
    if (params->kernel != VKR_KERNEL_PATH) return fail(VKR_INVALID_ARGUMENT, "the autotuner tunes the path kernel");
    const vkr_error pe = validate_params(params);
    if (pe != VKR_SUCCESS) return pe;
    VKR_HIP(hipSetDevice(ctx->device));
    if (repeats == 0) repeats = 1;
    std::memset(out, 0, sizeof(*out));

    struct Cand { uint32_t block, chunk, stack, sched; };
    std::vector<Cand> cands;
    const uint32_t stacks[3] = {24u, 32u, 64u};
    for (uint32_t b : {64u, 128u, 256u}) cands.push_back({b, 0u, s->dev.stack_needed, (uint32_t)VKR_SCHED_PLAIN});
    for (uint32_t st : stacks) {
        if (st < s->dev.stack_needed) continue;
        for (uint32_t b : {64u, 128u, 256u})
            for (uint32_t c : {1u, 8u, 16u, 32u}) cands.push_back({b, c, st, (uint32_t)VKR_SCHED_PERSISTENT});
    }
    // Drop any previous tuning of this (scene, size) so candidates run as given.
    ctx->tuned.erase(std::remove_if(ctx->tuned.begin(), ctx->tuned.end(),
                                    [&](const TuneEntry& t) {
                                        return t.scene == scene && t.width == params->width &&
                                               t.height == params->height;
                                    }),
                     ctx->tuned.end());
    float best = 3.4e38f;
    Cand win = cands[0];
    for (const Cand& c : cands) {
        vkr_trace_params p = *params;
        p.block = c.block;
        p.chunk = c.chunk;
        p.sched = c.sched;
        // The stack variant is chosen through a temporary tuning entry.
        ctx->tuned.push_back(TuneEntry{scene, p.width, p.height, c.block, c.chunk ? c.chunk : kDefaultChunk, c.stack, c.sched});
        float total = 0.0f;
        vkr_error e = vkr_trace_frame(ctx, scene, fd, &p, outputs, nullptr);
        for (uint32_t r = 0; r < repeats && e == VKR_SUCCESS; ++r) {
            e = vkr_trace_frame(ctx, scene, fd, &p, outputs, nullptr);
            float ms = 0.0f;
            if (e == VKR_SUCCESS) e = vkr_last_kernel_ms(ctx, &ms);
            total += ms;
        }
        ctx->tuned.pop_back();
        if (e != VKR_SUCCESS) return e;
        const float ms = total / (float)repeats;
        if (out->n_tried < 64) {
            const uint32_t k = out->n_tried++;
            out->cand_block[k] = c.block;
            out->cand_chunk[k] = c.chunk;
            out->cand_stack[k] = c.stack;
            out->cand_sched[k] = c.sched;
            out->cand_ms[k] = ms;
        }
        if (ms < best) {
            best = ms;
            win = c;
        }
    }
    ctx->tuned.push_back(TuneEntry{scene, params->width, params->height, win.block, win.chunk ? win.chunk : kDefaultChunk,
                                   win.stack, win.sched});
    out->block = win.block;
    out->chunk = win.chunk;
    out->stack = win.stack;
    out->sched = win.sched;
    out->best_ms = best;
    return VKR_SUCCESS;
}

vkr_error vkr_autotune_batch(vkr_ctx* ctx, const vkr_scene_job* jobs, uint32_t n_jobs, uint32_t repeats,
                             vkr_tune_result* out)
{
    if (!ctx || !jobs || !out) return fail(VKR_INVALID_ARGUMENT, "NULL argument");
    BatchPlan plan;
    const vkr_error pe = plan_batch(ctx, jobs, n_jobs, plan);
    if (pe != VKR_SUCCESS) return pe;
    VKR_HIP(hipSetDevice(ctx->device));
    if (repeats == 0) repeats = 1;
    std::memset(out, 0, sizeof(*out));
    const uint64_t key = batch_key(jobs, n_jobs);
    ctx->tuned_batch.erase(std::remove_if(ctx->tuned_batch.begin(), ctx->tuned_batch.end(),
                                          [&](const BatchTuneEntry& t) { return t.key == key; }),
                           ctx->tuned_batch.end());
    const uint32_t needed = plan.stack;

    // One candidate: installed as this job set's tuning so that the launch
    // takes its stack depth too, then timed over `repeats` launches.
    auto time_shape = [&](uint32_t b, uint32_t c, uint32_t st, float& ms) -> vkr_error {
        ctx->tuned_batch.push_back(BatchTuneEntry{key, b, c, st});
        float total = 0.0f;
        vkr_error e = VKR_SUCCESS;
        for (uint32_t r = 0; r < repeats && e == VKR_SUCCESS; ++r) {
            e = vkr_trace_batch(ctx, jobs, n_jobs, 0, 0, nullptr);
            float t = 0.0f;
            if (e == VKR_SUCCESS) e = vkr_last_kernel_ms(ctx, &t);
            total += t;
        }
        ctx->tuned_batch.pop_back();
        ms = total / (float)repeats;
        if (e == VKR_SUCCESS && out->n_tried < 64) {
            const uint32_t k = out->n_tried++;
            out->cand_block[k] = b;
            out->cand_chunk[k] = c;
            out->cand_stack[k] = st;
            out->cand_sched[k] = VKR_SCHED_PERSISTENT;
            out->cand_ms[k] = ms;
        }
        return e;
    };

    // Coordinate descent (one batch launch is long, a full grid is not
    // affordable): workgroup size at the default chunk, then the pool chunk,
    // then the stack depth.  One warm launch loads the kernels first.
    uint32_t best_b = 128u, best_c = kDefaultChunk, best_s = needed;
    float best = 3.4e38f, ms = 0.0f;
    vkr_error e = vkr_trace_batch(ctx, jobs, n_jobs, best_b, best_c, nullptr);
    if (e != VKR_SUCCESS) return e;
    for (uint32_t b : {64u, 128u, 256u}) {
        if ((e = time_shape(b, best_c, best_s, ms)) != VKR_SUCCESS) return e;
        if (ms < best) { best = ms; best_b = b; }
    }
    for (uint32_t c : {1u, 4u, 16u, 32u}) {
        if ((e = time_shape(best_b, c, best_s, ms)) != VKR_SUCCESS) return e;
        if (ms < best) { best = ms; best_c = c; }
    }
    for (uint32_t st : {32u, 64u}) {
        if (st <= needed) continue;
        if ((e = time_shape(best_b, best_c, st, ms)) != VKR_SUCCESS) return e;
        if (ms < best) { best = ms; best_s = st; }
    }
    ctx->tuned_batch.push_back(BatchTuneEntry{key, best_b, best_c, best_s});
    out->block = best_b;
    out->chunk = best_c;
    out->stack = best_s;
    out->sched = VKR_SCHED_PERSISTENT;
    out->best_ms = best;
    return VKR_SUCCESS;
}

}  // extern "C"
