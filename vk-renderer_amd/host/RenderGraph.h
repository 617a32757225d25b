// RenderGraph.h - the reference's render-graph descriptors
// (Include/RenderGraph.h:6-65) with a COMPUTE stage added, and the two-pass
// graph GraphicsDevice::Draw executes: the path-traced compute pass, then
// the present filter pass (Source/GraphicsDevice.cpp:1254-1309).
#pragma once

#include <cstdint>
#include <vector>

namespace Renderer {
struct PipelineInfo {
    struct ShaderInfo {
        enum class Stage { VERTEX, FRAGMENT, COMPUTE };
        const char* file_path;
        const char* entry_point;
        Stage stage;
    };
    unsigned short shader_count;
    ShaderInfo shaders[2];
};

struct AttachmentInfo {
    enum class Source { TRANSIENT, BACKBUFFER };
    const char* debug_name;
    Source source;
};

struct PassInfo {
    const char* debug_name;
    unsigned short color_output_count;
    unsigned short color_input_count;
    unsigned short pipeline_count;
    unsigned short color_outputs[8];
    unsigned short color_inputs[8];
    PipelineInfo* pipelines;
};

struct GraphInfo {
    unsigned short pass_count;
    unsigned short color_attachment_count;
    PassInfo* passes;
    AttachmentInfo* color_attachments;
};

// Attachments of the frame graph: 0 = traced image of this frame slot,
// 1 = traced image 0, 2 = traced image 1, 3 = backbuffer.
enum : unsigned short { kTracedSlot = 0, kTraced0 = 1, kTraced1 = 2, kBackbuffer = 3 };

// The frame graph GraphicsDevice::Draw runs, in order.
const GraphInfo& FrameGraph();
}  // namespace Renderer
