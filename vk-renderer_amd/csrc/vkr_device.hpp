// vkr_device.hpp - device arithmetic for the gfx950 path-tracing kernels.
//
// GLSL leaves the precision of several built-ins open.  The kernels fix one
// meaning for each, chosen so the result is reproducible on any IEEE machine:
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, no fused multiply-add
//   cross, reflect, refract, mix, fract: the GLSL 4.50 formulas, left to right
//   length = correctly rounded sqrt(dot(v,v)); normalize = v / length(v)
//   min/max/clamp: IEEE minNum/maxNum (a NaN operand yields the other one)
//   sign(NaN) = 0
//   cos/sin: the float nearest to a double-precision cos/sin
//   pow(x,5), pow(x,16): products in double rounded once; pow(x,2) = x*x;
//   other exponents via double exp(y * log(x)); pow(x<0, y) = NaN
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

// --------------------------------------------------------------------------
// Transcendentals.  Each returns the float nearest to a double-precision
// evaluation accurate to a few double ulps, which equals the correctly
// rounded float result except when the exact value lies within ~1e-15
// (relative) of a float rounding boundary.  The methods are fdlibm's
// published ones (Sun Microsystems, freely distributable): Cody-Waite
// argument reduction by pi/2 with the product n*pi/2 carried as exact
// double-double terms (accurate for |x| < 2^32), then the k_sin/k_cos
// minimax polynomials; log via s = f/(2+f) and the Lg1..Lg7 series; exp via
// reduction by ln2 and the P1..P5 rational form.  They are written here to
// keep register use low; OCML's general versions cost ~50 VGPRs more.
// --------------------------------------------------------------------------
namespace detail {
constexpr double kPio2_1 = 1.57079632679489655800e+00;   // RN(pi/2)
constexpr double kPio2_2 = 6.12323399573676603587e-17;   // RN(pi/2 - kPio2_1)
constexpr double kPio2_3 = -1.4973849048591698e-33;      // RN(pi/2 - kPio2_1 - kPio2_2)
constexpr double kTwoOverPi = 6.36619772367581382433e-01;

__device__ __forceinline__ double poly_sin(double x)
{
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    const double z = x * x;
    const double v = z * x;
    const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    return x + v * (S1 + z * r);
}

__device__ __forceinline__ double poly_cos(double x)
{
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    const double z = x * x;
    const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    const double ax = fabs(x);
    if (ax < 0.3) return 1.0 - (0.5 * z - z * r);
    // |x| >= 0.3: split 1 - z/2 so the subtraction stays exact (fdlibm).
    const double qx = ax > 0.78125 ? 0.28125 : __hiloint2double(__double2hiint(ax) - 0x00200000, 0);
    const double hz = 0.5 * z - qx;
    const double a = 1.0 - qx;
    return a - (hz - z * r);
}

// x - n*pi/2 for the nearest integer n; returns r and n mod 4.
__device__ __forceinline__ double reduce_pio2(double x, int& quadrant)
{
    const double n = rint(x * kTwoOverPi);
    const double p_hi = n * kPio2_1;
    const double p_lo = __builtin_fma(n, kPio2_1, -p_hi);       // exact remainder of n * P1
    const double q_hi = n * kPio2_2;
    const double q_lo = __builtin_fma(n, kPio2_2, -q_hi);
    double r = x - p_hi;                                         // exact (Sterbenz)
    r = r - p_lo;
    r = r - q_hi;
    r = r - (q_lo + n * kPio2_3);
    quadrant = (int)((long long)n & 3);   // n mod 4, also for negative n
    return r;
}

__device__ __forceinline__ void sincos_d(double x, double& s, double& c)
{
    int q;
    const double r = reduce_pio2(x, q);
    const double ps = poly_sin(r), pc = poly_cos(r);
    switch (q) {
    case 0: s = ps; c = pc; break;
    case 1: s = pc; c = -ps; break;
    case 2: s = -ps; c = -pc; break;
    default: s = -pc; c = ps; break;
    }
}

__device__ __forceinline__ double log_d(double x)  // x > 0, finite
{
    const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
                 Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
                 Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                 Lg7 = 1.479819860511658591e-01;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    int k;
    double m = frexp(x, &k);        // x = m * 2^k, m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        k -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    const double w = z * z;
    const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    const double R = t2 + t1;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)k;
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}

__device__ __forceinline__ double exp_d(double t)  // finite t, |t| < 700
{
    const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                 P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                 P5 = 4.13813679705723846039e-08;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double k = rint(t * inv_ln2);
    const double hi = t - k * ln2_hi;
    const double lo = k * ln2_lo;
    const double r = hi - lo;
    const double r2 = r * r;
    const double c = r - r2 * (P1 + r2 * (P2 + r2 * (P3 + r2 * (P4 + r2 * P5))));
    const double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
    return ldexp(y, (int)k);
}
}  // namespace detail

__device__ __forceinline__ float cos_f(float x)
{
    double s, c;
    detail::sincos_d((double)x, s, c);
    return (float)c;
}
__device__ __forceinline__ float sin_f(float x)
{
    double s, c;
    detail::sincos_d((double)x, s, c);
    return (float)s;
}
__device__ __forceinline__ void sincos_f(float x, float& s, float& c)
{
    double sd, cd;
    detail::sincos_d((double)x, sd, cd);
    s = (float)sd;
    c = (float)cd;
}
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
// pow(x, y) for the tone curve (x in [0, 1], y = 1/2.2).
__device__ __forceinline__ float pow_f(float x, float y)
{
    if (x < 0.0f) return __builtin_nanf("");
    if (x != x) return x;
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : __builtin_huge_valf());
    if (x == __builtin_huge_valf()) return y > 0.0f ? x : (y == 0.0f ? 1.0f : 0.0f);
    return (float)detail::exp_d((double)y * detail::log_d((double)x));
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
    float sp, cp;
    sincos_f(phi, sp, cp);
    return add(muls(add(muls(u, cp), muls(v, sp)), sina), muls(w, cosa));
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
