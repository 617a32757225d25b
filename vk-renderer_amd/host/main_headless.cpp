// main_headless.cpp - Source/Main.cpp:85-206 without the window: construct
// the device, place the camera of Main.cpp:134-141, Draw a number of frames,
// print the frame rate, and write the presented image.
//
//   vkr_headless [--frames N] [--res R] [--present WxH] [--seeds a,b,c]
//                [--poses file] [--scene file.vkrs [--device-bvh]] [--out image.ppm]
//                [--raw prefix]
// --poses replays recorded camera input in place of Main.cpp:165-190's
// keyboard/mouse handling: one "x y z pitch yaw" line per frame (cycled;
// '#' starts a comment), each applied through Camera::update.
// --scene traces a scene file (vkr_scene_save) instead of the reference's
// constant scene; --device-bvh builds its triangle tree on the GPU
// (VKR_SCENE_DEVICE_BVH).
// --raw writes prefix.traced0.rgba, prefix.traced1.rgba, prefix.present.rgba
// and prefix.seeds.txt for checking against the CPU restatement.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Camera.h"
#include "GraphicsDevice.h"

namespace {
struct Pose {
    float x, y, z, pitch, yaw;
};

bool read_poses(const std::string& path, std::vector<Pose>& out)
{
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    char line[512];
    while (std::fgets(line, sizeof(line), f)) {
        if (char* c = std::strchr(line, '#')) *c = 0;
        Pose p{};
        if (std::sscanf(line, "%f %f %f %f %f", &p.x, &p.y, &p.z, &p.pitch, &p.yaw) == 5) out.push_back(p);
    }
    std::fclose(f);
    return !out.empty();
}

bool write_file(const std::string& path, const void* data, size_t n)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, n, f) == n;
    std::fclose(f);
    return ok;
}
}  // namespace

int main(int argc, char** argv)
{
    int frames = 60;
    unsigned res = 1024, pw = 1024, ph = 768;
    std::vector<float> seeds;
    std::string out, raw, poses_path, scene_path;
    bool device_bvh = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() { return i + 1 < argc ? std::string(argv[++i]) : std::string(); };
        if (a == "--frames") frames = std::atoi(next().c_str());
        else if (a == "--res") res = (unsigned)std::atoi(next().c_str());
        else if (a == "--present") std::sscanf(next().c_str(), "%ux%u", &pw, &ph);
        else if (a == "--out") out = next();
        else if (a == "--raw") raw = next();
        else if (a == "--poses") poses_path = next();
        else if (a == "--scene") scene_path = next();
        else if (a == "--device-bvh") device_bvh = true;
        else if (a == "--seeds") {
            const std::string s = next();
            size_t p = 0;
            while (p < s.size()) {
                size_t q = s.find(',', p);
                if (q == std::string::npos) q = s.size();
                seeds.push_back(std::strtof(s.substr(p, q - p).c_str(), nullptr));
                p = q + 1;
            }
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (frames < 1 || res == 0 || res > 16384 || pw == 0 || ph == 0) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    std::vector<Pose> poses;
    if (!poses_path.empty() && !read_poses(poses_path, poses)) {
        std::fprintf(stderr, "cannot read poses from %s\n", poses_path.c_str());
        return 2;
    }

    GraphicsDevice device;
    const GraphicsDevice::CreateInfo info{nullptr, 3, 2, (unsigned short)res, false};   // Main.cpp:105-115
    if (const auto e = device.Construct(info); e != GraphicsDevice::Error::SUCCESS) {
        std::printf("{\"error\": %d, \"message\": \"%s\"}\n", (int)e, vkr_last_error_string());
        return 1;
    }
    device.SetPresentExtent(pw, ph);
    if (!scene_path.empty()) {
        // A scene file (vkr_scene_save) in place of the reference's constant scene.
        vkr_scene_desc desc{};
        if (vkr_scene_load(scene_path.c_str(), &desc) != VKR_SUCCESS) {
            std::printf("{\"error\": %d, \"message\": \"%s\"}\n", 4, vkr_last_error_string());
            return 2;
        }
        if (device_bvh) desc.flags |= VKR_SCENE_DEVICE_BVH;
        const auto e = device.SetScene(desc);
        vkr_scene_free(&desc);
        if (e != GraphicsDevice::Error::SUCCESS) {
            std::printf("{\"error\": %d, \"message\": \"%s\"}\n", (int)e, vkr_last_error_string());
            return 1;
        }
    }
    if (!seeds.empty()) device.SetSeeds(seeds);

    Camera camera;                                                                       // Main.cpp:134-139
    camera.data.pos = vec3{32.8509f, 30.6991f, -106.389f};
    camera.aux.pitch = 4.44998f;
    camera.aux.yaw = -602.79f;
    camera.update();
    FrameData frame_data{};
    frame_data.light_pos = vec3{0.0f, 64.0f, 0.0f};                                       // :141
    frame_data.camera = camera.data;

    std::vector<float> used;
    const auto t0 = std::chrono::steady_clock::now();
    for (int f = 0; f < frames; ++f) {
        if (!poses.empty()) {
            const Pose& q = poses[(size_t)f % poses.size()];
            camera.data.pos = vec3{q.x, q.y, q.z};
            camera.aux.pitch = q.pitch;
            camera.aux.yaw = q.yaw;
            camera.update();
        }
        frame_data.camera = camera.data;
        device.Draw(frame_data);
        used.push_back(device.LastSeed());
    }
    device.WaitIdle();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    std::vector<uint8_t> img;
    if (!out.empty() && device.ReadPresented(img) == GraphicsDevice::Error::SUCCESS) {
        FILE* f = std::fopen(out.c_str(), "wb");
        if (f) {
            std::fprintf(f, "P6\n%u %u\n255\n", device.PresentWidth(), device.PresentHeight());
            for (size_t i = 0; i < img.size(); i += 4) std::fwrite(&img[i], 1, 3, f);
            std::fclose(f);
        }
    }
    if (!raw.empty()) {
        std::vector<uint8_t> t;
        for (uint32_t s = 0; s < 2; ++s)
            if (device.ReadTraced(s, t) == GraphicsDevice::Error::SUCCESS)
                write_file(raw + ".traced" + std::to_string(s) + ".rgba", t.data(), t.size());
        if (device.ReadPresented(img) == GraphicsDevice::Error::SUCCESS)
            write_file(raw + ".present.rgba", img.data(), img.size());
        std::string txt;
        for (float s : used) txt += std::to_string(s) + "\n";
        FILE* f = std::fopen((raw + ".seeds.txt").c_str(), "w");
        if (f) {
            for (float s : used) std::fprintf(f, "%.9g\n", s);
            std::fclose(f);
        }
    }
    std::printf("{\"frames\": %d, \"seconds\": %.4f, \"fps\": %.2f, \"res\": %u, \"present\": [%u, %u]}\n", frames,
                secs, frames / secs, res, device.PresentWidth(), device.PresentHeight());
    device.Destruct();
    return 0;
}
