// Camera.cpp - Source/Camera.cpp:5-64 over the C ABI's camera helper.
#include "Camera.h"

#include <cmath>

#include "../../include/vkr.h"

namespace {
inline vec3 madd(vec3 a, vec3 b, float s) { return vec3{a.x + b.x * s, a.y + b.y * s, a.z + b.z * s}; }

// aux.front = glm::normalize(front) (Source/Camera.cpp:52-58), GLM float
// semantics: radians = deg * float(pi/180), normalize = v * (1/sqrt(dot)).
vec3 front_of(float pitch, float yaw)
{
    const float rad = 0.01745329251994329576923690768489f;
    const float p = pitch * rad, y = yaw * rad;
    const vec3 f{std::cos(p) * std::cos(y), std::sin(p), std::cos(p) * std::sin(y)};
    const float inv = 1.0f / std::sqrt(f.x * f.x + f.y * f.y + f.z * f.z);
    return vec3{f.x * inv, f.y * inv, f.z * inv};
}
}  // namespace

Camera::Camera() : data{vec3{0.0f, 64.0f, -48.0f}, {}, {}, {}}, aux{vec3{0.0f, 0.0f, -1.0f}, 0.0f, 90.0f}
{
    update();
}

void Camera::move_forward(float speed) { data.pos = madd(data.pos, aux.front, speed); update(); }
void Camera::move_backward(float speed) { data.pos = madd(data.pos, aux.front, -speed); update(); }
void Camera::move_left(float speed) { data.pos = madd(data.pos, data.right, -speed); update(); }
void Camera::move_right(float speed) { data.pos = madd(data.pos, data.right, speed); update(); }
void Camera::move_up(float speed) { data.pos = madd(data.pos, data.up, speed); update(); }
void Camera::move_down(float speed) { data.pos = madd(data.pos, data.up, -speed); update(); }

void Camera::update()
{
    // Source/Camera.cpp:50: the pitch is clamped in place.
    aux.pitch = aux.pitch < -89.0f ? -89.0f : (aux.pitch > 89.0f ? 89.0f : aux.pitch);
    const float pos[3] = {data.pos.x, data.pos.y, data.pos.z};
    vkr_camera c;
    if (vkr_camera_from_angles(pos, aux.pitch, aux.yaw, &c) != VKR_SUCCESS) return;
    data.dir = vec3{c.dir[0], c.dir[1], c.dir[2]};
    data.right = vec3{c.right[0], c.right[1], c.right[2]};
    data.up = vec3{c.up[0], c.up[1], c.up[2]};
    aux.front = front_of(aux.pitch, aux.yaw);
}
