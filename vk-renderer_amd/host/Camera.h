// Camera.h - the reference's camera (Include/Camera.h:5-43), unchanged in
// shape; the basis is computed by vkr_camera_from_angles (GLM float
// semantics of Source/Camera.cpp:48-64).
#pragma once

struct vec3 {
    float x, y, z;
};

// CameraData (Include/Camera.h:5-12): four 16-byte aligned vec3, 64 bytes,
// identical to vkr_camera.
struct CameraData {
    alignas(16) vec3 pos;
    alignas(16) vec3 dir;
    alignas(16) vec3 right;
    alignas(16) vec3 up;
};
static_assert(sizeof(CameraData) == 64, "CameraData is 64 bytes");

struct CameraDataAux {
    vec3 front;
    float pitch;
    float yaw;
};

struct Camera {
    CameraData data;
    CameraDataAux aux;

    Camera();

    void move_forward(float speed);
    void move_backward(float speed);
    void move_left(float speed);
    void move_right(float speed);
    void move_up(float speed);
    void move_down(float speed);

    void update();
};
