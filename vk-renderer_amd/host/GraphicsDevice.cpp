// GraphicsDevice.cpp - Source/GraphicsDevice.cpp's Construct / Draw /
// WaitIdle / Destruct over the vkr C ABI (no Vulkan, no window system).
#include "GraphicsDevice.h"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <ctime>

#include "RenderGraph.h"

namespace {

// Assets/Tracer.comp:176-211 and Source/GraphicsDevice.cpp:817: the scene
// the reference renders.
vkr_material mat(float r, float g, float b, float e, float rough, float metal, uint32_t type)
{
    vkr_material m{};
    m.albedo[0] = r; m.albedo[1] = g; m.albedo[2] = b;
    m.emissive[0] = m.emissive[1] = m.emissive[2] = e;
    m.roughness = rough;
    m.metalness = metal;
    m.type = type;
    return m;
}

struct ReferenceScene {
    vkr_sphere spheres[4];
    vkr_plane planes[5];
    vkr_triangle tri;
    vkr_material mirror;

    ReferenceScene()
    {
        const vkr_material matte_white = mat(1.0f, 1.0f, 1.0f, 0.0f, 0.3f, 0.7f, VKR_MAT_DIFFUSE);
        const vkr_material matte_red = mat(0.75f, 0.25f, 0.25f, 0.0f, 0.4f, 0.0f, VKR_MAT_DIFFUSE);
        const vkr_material matte_green = mat(0.25f, 0.75f, 0.25f, 0.0f, 0.4f, 0.0f, VKR_MAT_DIFFUSE);
        const vkr_material matte_blue = mat(0.25f, 0.25f, 0.75f, 0.0f, 0.4f, 0.0f, VKR_MAT_DIFFUSE);
        const vkr_material plastic = mat(0.25f, 0.25f, 0.75f, 0.0f, 0.3f, 0.6f, VKR_MAT_DIFFUSE);
        mirror = mat(1.0f, 0.5f, 0.5f, 0.0f, 0.0f, 1.0f, VKR_MAT_DIFFUSE);
        const vkr_material glass = mat(1.0f, 1.0f, 1.0f, 0.0f, 0.42f, 0.0f, VKR_MAT_DIELECTRIC);
        const vkr_material light = mat(1.0f, 1.0f, 1.0f, 128.0f, 0.6f, 0.0f, VKR_MAT_DIFFUSE);
        const vkr_material sm[4] = {glass, light, mirror, plastic};
        const float sp[4][4] = {{42.0f, 16.0f, 12.0f, 16.0f},
                                {0.0f, 96.0f, 0.0f, 12.0f},
                                {-32.0f, 24.0f, 24.0f, 24.0f},
                                {-24.0f, 11.0f, -48.0f, 11.0f}};
        for (int i = 0; i < 4; ++i) {
            std::memset(&spheres[i], 0, sizeof(vkr_sphere));
            spheres[i].mat = sm[i];
            spheres[i].P[0] = sp[i][0]; spheres[i].P[1] = sp[i][1]; spheres[i].P[2] = sp[i][2];
            spheres[i].r = sp[i][3];
        }
        const vkr_material pm[5] = {matte_white, matte_white, matte_red, matte_green, matte_blue};
        const float pn[5][4] = {{0.0f, 1.0f, 0.0f, 0.0f},
                                {0.0f, -1.0f, 0.0f, 128.0f},
                                {1.0f, 0.0f, 0.0f, 64.0f},
                                {0.0f, 0.0f, -1.0f, 64.0f},
                                {-1.0f, 0.0f, 0.0f, 64.0f}};
        for (int i = 0; i < 5; ++i) {
            std::memset(&planes[i], 0, sizeof(vkr_plane));
            planes[i].mat = pm[i];
            planes[i].N[0] = pn[i][0]; planes[i].N[1] = pn[i][1]; planes[i].N[2] = pn[i][2];
            planes[i].len = pn[i][3];
        }
        std::memset(&tri, 0, sizeof(tri));
        tri.v0[0] = 10.0f; tri.v0[1] = 10.0f; tri.v0[2] = 0.0f;
        tri.v1[0] = 0.0f; tri.v1[1] = 20.0f; tri.v1[2] = 0.0f;
        tri.v2[0] = -10.0f; tri.v2[1] = 10.0f; tri.v2[2] = 0.0f;
    }
};

GraphicsDevice::Error map_error(vkr_error e)
{
    switch (e) {
    case VKR_SUCCESS: return GraphicsDevice::Error::SUCCESS;
    case VKR_NO_SUITABLE_GPU: return GraphicsDevice::Error::NO_SUITABLE_GPU;
    case VKR_NO_SUITABLE_SURFACE: return GraphicsDevice::Error::NO_SUITABLE_SURFACE;
    default: return GraphicsDevice::Error::UNKNOWN;
    }
}

}  // namespace

namespace Renderer {
const GraphInfo& FrameGraph()
{
    // Source/GraphicsDevice.cpp:1254-1266 (compute, Tracer.comp) then
    // :1286-1309 (fullscreen filter sampling traced images 0 and 1, :964-983).
    static PipelineInfo trace_pipe{1, {{"Assets/Tracer.comp", "main", PipelineInfo::ShaderInfo::Stage::COMPUTE},
                                       {nullptr, nullptr, PipelineInfo::ShaderInfo::Stage::COMPUTE}}};
    static PipelineInfo filter_pipe{2, {{"Assets/Fullscreen.vert", "main", PipelineInfo::ShaderInfo::Stage::VERTEX},
                                        {"Assets/Fullscreen.frag", "main", PipelineInfo::ShaderInfo::Stage::FRAGMENT}}};
    static PassInfo passes[2] = {
        {"trace", 1, 0, 1, {kTracedSlot}, {}, &trace_pipe},
        {"filter", 1, 2, 1, {kBackbuffer}, {kTraced0, kTraced1}, &filter_pipe},
    };
    static AttachmentInfo atts[4] = {{"traced[slot]", AttachmentInfo::Source::TRANSIENT},
                                     {"traced[0]", AttachmentInfo::Source::TRANSIENT},
                                     {"traced[1]", AttachmentInfo::Source::TRANSIENT},
                                     {"backbuffer", AttachmentInfo::Source::BACKBUFFER}};
    static GraphInfo g{2, 4, passes, atts};
    return g;
}
}  // namespace Renderer

struct GraphicsDevice::State {
    vkr_ctx* ctx = nullptr;
    uint32_t scene = 0;
    uint32_t res = 0, frames = 0, present_w = 1024, present_h = 768;
    std::vector<uint8_t*> traced;
    uint8_t* present = nullptr;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> fences;
    uint32_t current = 0;
    std::vector<float> seeds;
    size_t seed_pos = 0;
    float last_seed = 0.0f;
    vkr_error last_error = VKR_SUCCESS;
    // One captured launch per frame slot, replayed every frame with that
    // frame's push constants (created on the slot's first Draw).
    std::vector<uint32_t> graphs;
    static constexpr uint32_t kNoGraph = 0xFFFFFFFFu;
};

GraphicsDevice::Error GraphicsDevice::Construct(const CreateInfo& info)
{
    if (state) return Error::UNKNOWN;
    if (info.window) return Error::NO_SUITABLE_SURFACE;    // no window system on this machine
    if (info.framesInFlight == 0 || info.raytrace_resolution == 0) return Error::UNKNOWN;
    State* s = new State();
    vkr_create_info ci{-1, info.debug ? (uint32_t)VKR_CREATE_DEBUG : 0u};
    const vkr_error e = vkr_create(&ci, &s->ctx);
    if (e != VKR_SUCCESS) {
        delete s;
        return map_error(e);
    }
    s->res = info.raytrace_resolution;
    s->frames = info.framesInFlight;
    const size_t img = (size_t)s->res * s->res * 4u;
    bool ok = hipStreamCreate(&s->stream) == hipSuccess;
    for (uint32_t i = 0; ok && i < s->frames; ++i) {
        uint8_t* p = nullptr;
        ok = hipMalloc(&p, img) == hipSuccess && hipMemset(p, 0, img) == hipSuccess;
        if (p) s->traced.push_back(p);
        hipEvent_t ev = nullptr;
        ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
        if (ev) s->fences.push_back(ev);
    }
    ok = ok && hipMalloc(&s->present, (size_t)s->present_w * s->present_h * 4u) == hipSuccess;
    s->graphs.assign(s->frames, State::kNoGraph);
    state = s;
    if (!ok) {
        Destruct();
        return Error::UNKNOWN;
    }
    static const ReferenceScene ref;
    vkr_scene_desc d{};
    d.tris = &ref.tri;
    d.n_tris = 1;
    d.spheres = ref.spheres;
    d.n_spheres = 4;
    d.planes = ref.planes;
    d.n_planes = 5;
    d.tri_material = &ref.mirror;
    if (vkr_scene_upload(s->ctx, &d, &s->scene) != VKR_SUCCESS) {
        Destruct();
        return Error::UNKNOWN;
    }
    std::srand(static_cast<unsigned int>(std::time(nullptr)));   // :1133
    return Error::SUCCESS;
}

GraphicsDevice::Error GraphicsDevice::Destruct()
{
    if (!state) return Error::SUCCESS;
    State* s = state;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (uint8_t* p : s->traced) (void)hipFree(p);
    if (s->present) (void)hipFree(s->present);
    for (hipEvent_t ev : s->fences) (void)hipEventDestroy(ev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    const vkr_error e = vkr_destroy(s->ctx);
    delete s;
    state = nullptr;
    return map_error(e);
}

void GraphicsDevice::Draw(const FrameData& frame_data)
{
    if (!state) return;
    State* s = state;
    const uint32_t slot = s->current;
    (void)hipEventSynchronize(s->fences[slot]);                   // :1217-1218

    FrameData real = frame_data;                                  // :1258-1262
    real.aspect_ratio = static_cast<float>(s->present_w) / static_cast<float>(s->present_h);
    real.seed = s->seeds.empty() ? static_cast<float>(std::rand()) / static_cast<float>(RAND_MAX)
                                 : s->seeds[s->seed_pos++ % s->seeds.size()];
    s->last_seed = real.seed;
    vkr_frame_data fd;
    static_assert(sizeof(fd) == sizeof(real), "FrameData layout");
    std::memcpy(&fd, &real, sizeof(fd));

    const Renderer::GraphInfo& g = Renderer::FrameGraph();
    for (unsigned short p = 0; p < g.pass_count; ++p) {
        const Renderer::PassInfo& pass = g.passes[p];
        if (pass.pipelines[0].shaders[0].stage == Renderer::PipelineInfo::ShaderInfo::Stage::COMPUTE) {
            // Tracer.comp constants: SAMPLES 4, DEPTH 4, EPSILON 1e-3 (:177-180).
            vkr_trace_params tp{};
            tp.width = tp.height = s->res;
            tp.spp = 4;
            tp.max_depth = 4;
            tp.epsilon = 1e-3f;
            tp.kernel = VKR_KERNEL_PATH;
            tp.sched = VKR_SCHED_AUTO;
            vkr_outputs out{};
            out.rgba8 = s->traced[slot];
            if (s->graphs[slot] == State::kNoGraph) {
                const vkr_scene_job job{s->scene, fd, tp, out};
                s->last_error = vkr_batch_graph_create(s->ctx, &job, 1, 0, 0, s->stream, &s->graphs[slot]);
                if (s->last_error != VKR_SUCCESS) s->graphs[slot] = State::kNoGraph;
            }
            if (s->graphs[slot] != State::kNoGraph) {
                s->last_error = vkr_graph_set_frame(s->ctx, s->graphs[slot], 0, &fd, s->stream);
                if (s->last_error == VKR_SUCCESS) s->last_error = vkr_graph_launch(s->ctx, s->graphs[slot], s->stream);
            }
        } else {
            // Both bindings are fixed to traced images 0 and 1 (:964-983).
            const uint8_t* cur = s->traced[0];
            const uint8_t* prev = s->traced[s->frames > 1 ? 1 : 0];
            const vkr_error e = vkr_filter_frame(s->ctx, cur, prev, s->res, s->res, s->present, s->present_w,
                                                 s->present_h, s->stream);
            if (s->last_error == VKR_SUCCESS) s->last_error = e;
        }
    }
    (void)hipEventRecord(s->fences[slot], s->stream);
    s->current = (s->current + 1u) % s->frames;                   // :1341
}

void GraphicsDevice::WaitIdle()
{
    if (state) (void)vkr_wait_idle(state->ctx);
}

void GraphicsDevice::SetPresentExtent(uint32_t width, uint32_t height)
{
    if (!state || width == 0 || height == 0) return;
    WaitIdle();
    uint8_t* p = nullptr;
    if (hipMalloc(&p, (size_t)width * height * 4u) != hipSuccess) return;
    (void)hipFree(state->present);
    state->present = p;
    state->present_w = width;
    state->present_h = height;
}

void GraphicsDevice::SetSeeds(const std::vector<float>& seeds)
{
    if (!state) return;
    state->seeds = seeds;
    state->seed_pos = 0;
}

GraphicsDevice::Error GraphicsDevice::SetScene(const vkr_scene_desc& desc)
{
    if (!state) return Error::UNKNOWN;
    uint32_t id = 0;
    const vkr_error e = vkr_scene_upload(state->ctx, &desc, &id);
    if (e != VKR_SUCCESS) return map_error(e);
    WaitIdle();
    // The captured launches hold the old scene's device data: drop them
    // before the scene goes (the next Draw of each slot captures anew).
    for (uint32_t& g : state->graphs) {
        if (g != State::kNoGraph) (void)vkr_graph_destroy(state->ctx, g);
        g = State::kNoGraph;
    }
    (void)vkr_scene_destroy(state->ctx, state->scene);
    state->scene = id;
    return Error::SUCCESS;
}

const uint8_t* GraphicsDevice::TracedImage(uint32_t slot) const
{
    return (state && slot < state->traced.size()) ? state->traced[slot] : nullptr;
}
const uint8_t* GraphicsDevice::PresentedImage() const { return state ? state->present : nullptr; }
uint32_t GraphicsDevice::PresentWidth() const { return state ? state->present_w : 0u; }
uint32_t GraphicsDevice::PresentHeight() const { return state ? state->present_h : 0u; }
uint32_t GraphicsDevice::Resolution() const { return state ? state->res : 0u; }
float GraphicsDevice::LastSeed() const { return state ? state->last_seed : 0.0f; }
vkr_ctx* GraphicsDevice::Context() const { return state ? state->ctx : nullptr; }

GraphicsDevice::Error GraphicsDevice::ReadTraced(uint32_t slot, std::vector<uint8_t>& out)
{
    if (!state || slot >= state->traced.size()) return Error::UNKNOWN;
    WaitIdle();
    out.resize((size_t)state->res * state->res * 4u);
    return hipMemcpy(out.data(), state->traced[slot], out.size(), hipMemcpyDeviceToHost) == hipSuccess
               ? Error::SUCCESS
               : Error::UNKNOWN;
}

GraphicsDevice::Error GraphicsDevice::ReadPresented(std::vector<uint8_t>& out)
{
    if (!state) return Error::UNKNOWN;
    WaitIdle();
    out.resize((size_t)state->present_w * state->present_h * 4u);
    return hipMemcpy(out.data(), state->present, out.size(), hipMemcpyDeviceToHost) == hipSuccess ? Error::SUCCESS
                                                                                                 : Error::UNKNOWN;
}
