// vkr_device.hpp - device arithmetic for the gfx950 path-tracing kernels.
//
// GLSL leaves the precision of several built-ins open.  The kernels fix one
// meaning for each, chosen so the result is reproducible on any IEEE machine:
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, no fused multiply-add
//   cross, reflect, refract, mix, fract: the GLSL 4.50 formulas, left to right
//   length = correctly rounded sqrt(dot(v,v)); normalize = v / length(v)
//   min/max/clamp: IEEE minNum/maxNum (a NaN operand yields the other one)
//   sign(NaN) = 0
//   cos/sin: the float nearest to the double-precision cos/sin
//   pow(x,5), pow(x,16): products in double rounded once; pow(x,2) = x*x;
//   other exponents through double pow; pow(x<0, y) = NaN
//   constants: folded in double and rounded once to float
// All translation units are compiled with -ffp-contract=off and correctly
// rounded float division and square root.  Fused multiply-adds appear only
// where a result feeds a conservative decision (BVH box culling), never in a
// value that reaches the output.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vkr.h"

namespace vkr {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 divv(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 muls(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 adds(f3 a, float s) { return f3{a.x + s, a.y + s, a.z + s}; }
__device__ __forceinline__ f3 subs(f3 a, float s) { return f3{a.x - s, a.y - s, a.z - s}; }
__device__ __forceinline__ f3 ssub(float s, f3 a) { return f3{s - a.x, s - a.y, s - a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b)
{
    return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) { return divs(a, length(a)); }
__device__ __forceinline__ float gmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float gmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ f3 maxs(f3 a, float s) { return f3{gmax(a.x, s), gmax(a.y, s), gmax(a.z, s)}; }
__device__ __forceinline__ f3 clamps(f3 a, float lo, float hi)
{
    return f3{gclamp(a.x, lo, hi), gclamp(a.y, lo, hi), gclamp(a.z, lo, hi)};
}
__device__ __forceinline__ float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ f3 mix3(f3 x, f3 y, float a) { return add(muls(x, 1.0f - a), muls(y, a)); }
__device__ __forceinline__ float fract(float x) { return x - floorf(x); }
__device__ __forceinline__ bool eq3(f3 a, f3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
__device__ __forceinline__ f3 reflect(f3 I, f3 N) { return sub(I, muls(N, 2.0f * dot(N, I))); }
__device__ __forceinline__ f3 refract(f3 I, f3 N, float eta)
{
    const float d = dot(N, I);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return f3{0.0f, 0.0f, 0.0f};
    return sub(muls(I, eta), muls(N, eta * d + sqrtf(k)));
}
__device__ __forceinline__ float max3(f3 e) { return gmax(gmax(e.x, e.y), e.z); }

__device__ __forceinline__ float cos_f(float x) { return (float)::cos((double)x); }
__device__ __forceinline__ float sin_f(float x) { return (float)::sin((double)x); }
__device__ __forceinline__ float pow5(float x)
{
    if (x < 0.0f) return __builtin_nanf("");
    const double d = (double)x;
    const double d2 = d * d;
    return (float)(d2 * d2 * d);
}
__device__ __forceinline__ float pow16(float x)
{
    if (x < 0.0f) return __builtin_nanf("");
    const double d = (double)x;
    const double d2 = d * d;
    const double d4 = d2 * d2;
    const double d8 = d4 * d4;
    return (float)(d8 * d8);
}
__device__ __forceinline__ float pow_f(float x, float y)
{
    if (x < 0.0f) return __builtin_nanf("");
    return (float)::pow((double)x, (double)y);
}

// GLSL constants of Assets/Tracer.comp, folded in double then rounded.
constexpr float kPi = 3.14159265359f;          // :176
constexpr float kTwoPi = 6.28318530718f;       // 2.0 * PI
constexpr float kInvGamma = 0.45454545454545453f; // 1.0 / 2.2 (:588)
constexpr float kAttSub = 0.001f;              // :478
constexpr float kAttDiv = 0.999f;              // 1 - 0.001 (:478)
constexpr float kK1x = 23.14069263277926f;     // :229
constexpr float kK1y = 2.665144142690225f;     // :230
constexpr float kHashScale = 12345.6789f;      // :233

// rand() of Assets/Tracer.comp:225-234:
//   fract(cos(dot((coords + ++rand_salt) * 5.0, K1)) * 12345.6789)
__device__ __forceinline__ float rng_next(float& salt, float cx, float cy)
{
    salt = salt + 1.0f;
    const float a = (cx + salt) * 5.0f;
    const float b = (cy + salt) * 5.0f;
    const float d = a * kK1x + b * kK1y;
    return fract(cos_f(d) * kHashScale);
}

// jitter, Assets/Tracer.comp:256-263.
__device__ __forceinline__ f3 jitter(f3 d, float phi, float sina, float cosa)
{
    const f3 w = normalize(d);
    const f3 u = normalize(cross(f3{w.y, w.z, w.x}, w));
    const f3 v = cross(w, u);
    return add(muls(add(muls(u, cos_f(phi)), muls(v, sin_f(phi))), sina), muls(w, cosa));
}

// Assets/Tracer.comp:265-312.
__device__ __forceinline__ float schlick(float cosine, float ior)
{
    float r0 = (1.0f - ior) / (1.0f + ior);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * pow5(1.0f - cosine);
}
__device__ __forceinline__ f3 fresnel_schlick(float cos_theta, f3 F0)
{
    const float p = pow5(1.0f - cos_theta);
    return add(F0, muls(ssub(1.0f, F0), p));
}
__device__ __forceinline__ float distribution_ggx(f3 N, f3 H, float roughness)
{
    const float a = roughness * roughness;
    const float a2 = a * a;
    const float NdotH = gmax(dot(N, H), 0.0f);
    const float NdotH2 = NdotH * NdotH;
    float denom = NdotH2 * (a2 - 1.0f) + 1.0f;
    denom = kPi * denom * denom;
    return a2 / denom;
}
__device__ __forceinline__ float geometry_schlick_ggx(float NdotV, float roughness)
{
    const float r = roughness + 1.0f;
    const float k = (r * r) / 8.0f;
    return NdotV / (NdotV * (1.0f - k) + k);
}
__device__ __forceinline__ float geometry_smith(f3 N, f3 V, f3 L, float roughness)
{
    const float NdotV = gmax(dot(N, V), 0.0f);
    const float NdotL = gmax(dot(N, L), 0.0f);
    const float ggx2 = geometry_schlick_ggx(NdotV, roughness);
    const float ggx1 = geometry_schlick_ggx(NdotL, roughness);
    return ggx1 * ggx2;
}

// Intersections, Assets/Tracer.comp:314-372.
__device__ __forceinline__ float sphere_t(f3 o, f3 d, f3 c, float r)
{
    const f3 oc = sub(o, c);
    const float b = 2.0f * dot(oc, d);
    const float cc = dot(oc, oc) - r * r;
    const float h = b * b - 4.0f * cc;
    if (h < 0.0f) return -1.0f;
    return (-b - sqrtf(h)) / 2.0f;
}
__device__ __forceinline__ float plane_t(f3 o, f3 d, f3 N, float len)
{
    const float dd = dot(d, N);
    const float dist = -(len + dot(o, N)) / dd;
    return fabsf(gsign(dd - 0.0f)) * gmax(dist, 0.0f);
}
// Moller-Trumbore with the reference's one-sided absolute determinant test;
// e1 = fl(v1 - v0) and e2 = fl(v2 - v0) are the reference's v0v1/v0v2.
__device__ __forceinline__ float tri_t(f3 o, f3 d, f3 v0, f3 e1, f3 e2, float eps)
{
    const f3 pvec = cross(d, e2);
    const float det = dot(e1, pvec);
    if (det < eps) return -1.0f;
    const float inv = 1.0f / det;
    const f3 tvec = sub(o, v0);
    const float u = dot(tvec, pvec) * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const f3 qvec = cross(tvec, e1);
    const float v = dot(d, qvec) * inv;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    return dot(e2, qvec) * inv;
}

// imageStore to rgba8 UNORM: NaN -> 0, else round(clamp(x)*255) half up.
__device__ __forceinline__ uint32_t unorm8(float x)
{
    if (x != x) return 0u;
    return (uint32_t)floorf(gclamp(x, 0.0f, 1.0f) * 255.0f + 0.5f);
}

}  // namespace vkr
