// GraphicsDevice.h - the reference's device API (Include/GraphicsDevice.h:13-97)
// kept in shape on top of the vkr C ABI.  Construct / Draw / WaitIdle /
// Destruct behave like the reference's: Draw traces the frame-in-flight's
// image with the path kernel (Tracer.comp semantics, 4 spp, 4 bounces) and
// runs the present filter into the backbuffer.  Each frame slot's trace is
// a captured hipGraph, replayed every Draw with that frame's push constants
// (vkr_graph_set_frame).  The machine is headless, so
// the "swapchain" is a device image of the present extent (1024x768 unless
// set, the window size of Source/Main.cpp:92).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/vkr.h"
#include "Camera.h"

struct GLFWwindow;

// Include/GraphicsDevice.h:13-18 (std430 stride 48).
struct Triangle {
    alignas(16) vec3 v0;
    alignas(16) vec3 v1;
    alignas(16) vec3 v2;
};
static_assert(sizeof(Triangle) == 48, "Triangle is 48 bytes");

// Include/GraphicsDevice.h:20-29 (96 bytes, the push constant).
struct FrameData {
    alignas(4) float aspect_ratio;
    alignas(4) float seed;
    alignas(16) vec3 light_pos;
    alignas(16) CameraData camera;
};
static_assert(sizeof(FrameData) == 96, "FrameData is 96 bytes");

struct GraphicsDevice final {
    // Include/GraphicsDevice.h:46-52.
    enum class Error : signed char { SUCCESS, NO_SUITABLE_GPU, NO_SUITABLE_SURFACE, UNKNOWN };

    // Include/GraphicsDevice.h:57-72.
    struct CreateInfo final {
        GLFWwindow* window;                 // must be NULL: no window system here
        unsigned char swapchainSize;        // backbuffers (one is kept; present is a device image)
        unsigned char framesInFlight;       // traced images, >= 1
        unsigned short raytrace_resolution; // traced image is RES x RES
        bool debug;                         // VKR_CREATE_DEBUG
    };

    Error Construct(const CreateInfo& info);
    Error Destruct();
    void Draw(const FrameData& frame_data);
    void WaitIdle();

    // ---- headless extensions (no reference counterpart) ----
    // Present extent used for FrameData.aspect_ratio and the filter output.
    void SetPresentExtent(uint32_t width, uint32_t height);
    // Frame seeds to use instead of rand() / RAND_MAX (Source/GraphicsDevice.cpp:1262),
    // consumed one per Draw; an empty list restores rand().
    void SetSeeds(const std::vector<float>& seeds);
    // Replaces the scene (default: the reference's constant scene).
    Error SetScene(const vkr_scene_desc& desc);
    // Device pointers of the traced images and of the presented image.
    const uint8_t* TracedImage(uint32_t slot) const;
    const uint8_t* PresentedImage() const;
    Error ReadTraced(uint32_t slot, std::vector<uint8_t>& out);
    Error ReadPresented(std::vector<uint8_t>& out);
    uint32_t PresentWidth() const;
    uint32_t PresentHeight() const;
    uint32_t Resolution() const;
    // Seed actually pushed by the last Draw.
    float LastSeed() const;
    vkr_ctx* Context() const;

    struct State;
    State* state = nullptr;
};
