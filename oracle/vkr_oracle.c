/*
 * vkr_oracle.c - CPU restatement of stickyfingies/vk-renderer's compute path
 * tracer (Assets/Tracer.comp) and its legacy Whitted tracer
 * (Assets/Raytracer.comp).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels and the timed CPU
 * baseline.  Nothing in the product library uses this file.
 *
 * Build with -ffp-contract=off (see oracle/Makefile): every floating-point
 * operation is written out in the order the GLSL states it, one IEEE
 * operation per source operator, no fused multiply-add.
 *
 * Built-in semantics.  GLSL 4.50 leaves precision and NaN behaviour of several
 * built-ins to the implementation.  The oracle (and the HIP kernels) fix:
 *   dot(a,b)        ((a.x*b.x + a.y*b.y) + a.z*b.z)
 *   cross(a,b)      (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
 *   length(v)       sqrt(dot(v,v)), sqrt correctly rounded
 *   normalize(v)    v / length(v), componentwise correctly rounded division
 *   reflect(I,N)    I - (2*dot(N,I)) * N
 *   refract(I,N,e)  k = 1 - (e*e)*(1 - d*d), d = dot(N,I); k<0 -> 0,
 *                   else e*I - (e*d + sqrt(k)) * N
 *   mix(x,y,a)      x*(1-a) + y*a
 *   fract(x)        x - floor(x)
 *   min/max/clamp   IEEE minNum/maxNum: a NaN operand yields the other one
 *   sign(x)         1 if x>0, -1 if x<0, else 0 (NaN -> 0)
 *   cos, sin        the float nearest to the double-precision libm result
 *   pow(x,y)        x<0 -> NaN; pow(x,2) = x*x; pow(x,5), pow(x,16) from
 *                   exact-ish double products; other y via double pow()
 *   constants       folded in double and rounded once to float (glslang)
 *   rgba8 store     NaN -> 0, else floor(clamp(x,0,1)*255 + 0.5)
 * The resulting float functions are correctly rounded except where a double
 * result lies within one double ulp of a float rounding boundary (about one
 * call in 2^28), which is the documented bound on RNG bit-exactness.
 */
#include "vkr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#if defined(__FP_FAST_FMAF) && !defined(VKRO_ALLOW_FMA)
/* Only a signal that the compiler may fuse; correctness relies on
 * -ffp-contract=off, which the Makefile passes. */
#endif

_Static_assert(sizeof(omat) == 48, "Material std430");
_Static_assert(sizeof(otri) == 48, "Triangle std430");
_Static_assert(sizeof(osph) == 64, "Sphere std430");
_Static_assert(sizeof(opln) == 64, "Plane std430");
_Static_assert(sizeof(oframe) == 96, "FrameData std430");

/* ------------------------------------------------------------------------ */
/* vec3 and built-ins                                                         */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 Vp(const float* p) { return V(p[0], p[1], p[2]); }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vdiv(v3 a, v3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 vmuls(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 vadds(v3 a, float s) { return V(a.x + s, a.y + s, a.z + s); }
static inline v3 vsubs(v3 a, float s) { return V(a.x - s, a.y - s, a.z - s); }
static inline v3 svsub(float s, v3 a) { return V(s - a.x, s - a.y, s - a.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
    return V(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline float vlength(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vnormalize(v3 a) { return vdivs(a, vlength(a)); }
static inline float gmax(float a, float b) { return fmaxf(a, b); }
static inline float gmin(float a, float b) { return fminf(a, b); }
static inline float gclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline v3 vmaxs(v3 a, float s) { return V(gmax(a.x, s), gmax(a.y, s), gmax(a.z, s)); }
static inline v3 vclamps(v3 a, float lo, float hi)
{
    return V(gclamp(a.x, lo, hi), gclamp(a.y, lo, hi), gclamp(a.z, lo, hi));
}
static inline float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
static inline float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static inline v3 vmix(v3 x, v3 y, float a) { return vadd(vmuls(x, 1.0f - a), vmuls(y, a)); }
static inline float gfract(float x) { return x - floorf(x); }
static inline int veq(v3 a, v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline v3 vreflect(v3 I, v3 N) { return vsub(I, vmuls(N, 2.0f * vdot(N, I))); }
static inline v3 vrefract(v3 I, v3 N, float eta)
{
    const float d = vdot(N, I);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return V(0.0f, 0.0f, 0.0f);
    return vsub(vmuls(I, eta), vmuls(N, eta * d + sqrtf(k)));
}

float vkro_cos(float x) { return (float)cos((double)x); }
float vkro_sin(float x) { return (float)sin((double)x); }

/* pow(x, 5.0) (Tracer.comp:271,276). */
static inline float gpow5(float x)
{
    if (x < 0.0f) return NAN;
    const double d = (double)x;
    const double d2 = d * d;
    return (float)(d2 * d2 * d);
}
/* pow(x, 16.0) (Raytracer.comp:315). */
static inline float gpow16(float x)
{
    if (x < 0.0f) return NAN;
    const double d = (double)x;
    const double d2 = d * d;
    const double d4 = d2 * d2;
    const double d8 = d4 * d4;
    return (float)(d8 * d8);
}
float vkro_pow(float x, float y)
{
    if (x < 0.0f) return NAN;
    return (float)pow((double)x, (double)y);
}

/* GLSL constants, folded in double and rounded once. */
#define K_PI        3.14159265359f        /* Tracer.comp:176 -> 0x40490fdb */
#define K_TWO_PI    6.28318530718f        /* 2.0 * PI folded */
#define K_INV_GAMMA 0.45454545454545453f  /* 1.0 / 2.2 folded -> 0x3ee8ba2f */
#define K_ATT_SUB   0.001f                /* Tracer.comp:478 */
#define K_ATT_DIV   0.999f                /* 1 - 0.001 folded -> 0x3f7fbe77 */

/* ------------------------------------------------------------------------ */
/* RNG, Tracer.comp:221-234                                                    */
/* ------------------------------------------------------------------------ */
typedef struct { float salt; float cx, cy; } rng_t;

static inline float rng_next(rng_t* r)
{
    /* fract(cos(dot((coords + ++rand_salt) * 5.0, K1)) * 12345.6789) */
    r->salt = r->salt + 1.0f;
    const float a = (r->cx + r->salt) * 5.0f;
    const float b = (r->cy + r->salt) * 5.0f;
    const float d = a * 23.14069263277926f + b * 2.665144142690225f;
    return gfract(vkro_cos(d) * 12345.6789f);
}

void vkro_rand_sequence(float cx, float cy, float salt0, uint32_t n, float* out)
{
    rng_t r = {salt0, cx, cy};
    for (uint32_t i = 0; i < n; ++i) out[i] = rng_next(&r);
}

/* ------------------------------------------------------------------------ */
/* Shading helpers, Tracer.comp:236-312                                       */
/* ------------------------------------------------------------------------ */
static inline float max3(v3 e) { return gmax(gmax(e.x, e.y), e.z); }   /* :236-239 */

static inline v3 jitter(v3 d, float phi, float sina, float cosa)          /* :256-263 */
{
    const v3 w = vnormalize(d);
    const v3 u = vnormalize(vcross(V(w.y, w.z, w.x), w));
    const v3 v = vcross(w, u);
    return vadd(vmuls(vadd(vmuls(u, vkro_cos(phi)), vmuls(v, vkro_sin(phi))), sina), vmuls(w, cosa));
}

static inline float schlick(float cosine, float ior)                      /* :265-272 */
{
    float r0 = (1.0f - ior) / (1.0f + ior);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * gpow5(1.0f - cosine);
}

static inline v3 fresnel_schlick(float cos_theta, v3 F0)                  /* :274-277 */
{
    const float p = gpow5(1.0f - cos_theta);
    return vadd(F0, vmuls(svsub(1.0f, F0), p));
}

static inline float distribution_ggx(v3 N, v3 H, float roughness)         /* :279-292 */
{
    const float a = roughness * roughness;
    const float a2 = a * a;
    const float NdotH = gmax(vdot(N, H), 0.0f);
    const float NdotH2 = NdotH * NdotH;
    float denom = NdotH2 * (a2 - 1.0f) + 1.0f;
    denom = K_PI * denom * denom;
    return a2 / denom;
}

static inline float geometry_schlick_ggx(float NdotV, float roughness)    /* :294-303 */
{
    const float r = roughness + 1.0f;
    const float k = (r * r) / 8.0f;
    return NdotV / (NdotV * (1.0f - k) + k);
}

static inline float geometry_smith(v3 N, v3 Vv, v3 L, float roughness)    /* :304-312 */
{
    const float NdotV = gmax(vdot(N, Vv), 0.0f);
    const float NdotL = gmax(vdot(N, L), 0.0f);
    const float ggx2 = geometry_schlick_ggx(NdotV, roughness);
    const float ggx1 = geometry_schlick_ggx(NdotL, roughness);
    return ggx1 * ggx2;
}

/* ------------------------------------------------------------------------ */
/* Intersections, Tracer.comp:314-372                                         */
/* ------------------------------------------------------------------------ */
static inline float sphere_t(v3 o, v3 d, v3 c, float r)                    /* :314-329 */
{
    const v3 oc = vsub(o, c);
    const float b = 2.0f * vdot(oc, d);
    const float cc = vdot(oc, oc) - r * r;
    const float h = b * b - 4.0f * cc;
    if (h < 0.0f) return -1.0f;
    return (-b - sqrtf(h)) / 2.0f;
}

static inline float plane_t(v3 o, v3 d, v3 N, float len)                  /* :331-338 */
{
    const float dd = vdot(d, N);
    const float dist = -(len + vdot(o, N)) / dd;
    /* when_neq(d, 0.0) * max(dist, 0.0); when_neq = abs(sign(d - 0)) */
    return fabsf(gsign(dd - 0.0f)) * gmax(dist, 0.0f);
}

static inline float tri_t(v3 o, v3 d, v3 v0, v3 v1, v3 v2, float eps)     /* :340-372 */
{
    const v3 v0v1 = vsub(v1, v0);
    const v3 v0v2 = vsub(v2, v0);
    const v3 pvec = vcross(d, v0v2);
    const float det = vdot(v0v1, pvec);
    if (det < eps) return -1.0f;
    const float inv = 1.0f / det;
    const v3 tvec = vsub(o, v0);
    const float u = vdot(tvec, pvec) * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const v3 qvec = vcross(tvec, v0v1);
    const float v = vdot(d, qvec) * inv;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    return vdot(v0v2, qvec) * inv;
}

float vkro_tri_intersect(const float o[3], const float d[3], const otri* t, float eps)
{
    return tri_t(Vp(o), Vp(d), Vp(t->v0), Vp(t->v1), Vp(t->v2), eps);
}
float vkro_sphere_intersect(const float o[3], const float d[3], const osph* s)
{
    return sphere_t(Vp(o), Vp(d), Vp(s->P), s->r);
}
float vkro_plane_intersect(const float o[3], const float d[3], const opln* p)
{
    return plane_t(Vp(o), Vp(d), Vp(p->N), p->len);
}

/* ------------------------------------------------------------------------ */
/* trace_ray, Tracer.comp:374-431 (linear scan in reference order)            */
/* ------------------------------------------------------------------------ */
typedef struct {
    const omat* mat;
    float t;
    v3 P, N;
    uint32_t prim;   /* (kind << 30) | index */
} isect_t;

static inline const omat* tri_material(const oscene* s, uint32_t i)
{
    if ((s->flags & 1u) && s->materials && s->tris[i].mat_index < s->n_materials)
        return &s->materials[s->tris[i].mat_index];
    return s->tri_material;
}

static int trace_ray_scan(const oscene* s, v3 o, v3 d, isect_t* it, float eps);

/* The scan, or the same-BVH baseline's accelerated equivalent when the scene
 * carries one (it returns the scan's winner; the record is filled here the
 * way the scan fills it).  `any`: the caller only needs whether anything is
 * accepted (shadow rays, Tracer.comp:473). */
static int trace_ray_any(const oscene* s, v3 o, v3 d, isect_t* it, float eps, int any)
{
    if (!s->accel_fn) return trace_ray_scan(s, o, d, it, eps);
    const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    float t = it->t;
    uint32_t prim = 0xFFFFFFFFu;
    if (!s->accel_fn(s->accel, of, df, eps, any, &t, &prim)) return 0;
    if (any) return 1;
    const uint32_t kind = prim >> 30, i = prim & 0x3FFFFFFFu;
    it->t = t;
    it->P = vadd(o, vmuls(d, t));
    it->prim = prim;
    if (kind == 0) {
        const otri* tr = &s->tris[i];
        const v3 v0 = Vp(tr->v0), v1 = Vp(tr->v1), v2 = Vp(tr->v2);
        const v3 u = vsub(v1, v0);
        const v3 v = vsub(v2, v0);
        it->mat = tri_material(s, i);
        it->N = V((u.y * v.z) - (u.z * v.y), (u.z * v.x) - (u.x * v.z), (u.x * v.y) - (u.y * v.x));
    } else if (kind == 1) {
        const osph* sp = &s->spheres[i];
        it->mat = &sp->mat;
        it->N = vdivs(vsub(it->P, Vp(sp->P)), sp->r);
    } else {
        const opln* pl = &s->planes[i];
        it->mat = &pl->mat;
        it->N = Vp(pl->N);
    }
    return 1;
}

static int trace_ray(const oscene* s, v3 o, v3 d, isect_t* it, float eps)
{
    return trace_ray_any(s, o, d, it, eps, 0);
}

static int trace_ray_scan(const oscene* s, v3 o, v3 d, isect_t* it, float eps)
{
    int found = 0;
    for (uint32_t i = 0; i < s->n_tris; ++i) {                              /* :378-396 */
        const otri* tr = &s->tris[i];
        const v3 v0 = Vp(tr->v0), v1 = Vp(tr->v1), v2 = Vp(tr->v2);
        const float t = tri_t(o, d, v0, v1, v2, eps);
        if ((t > eps) && (t < it->t + eps)) {
            it->t = t;
            it->mat = tri_material(s, i);
            it->P = vadd(o, vmuls(d, t));
            const v3 u = vsub(v1, v0);
            const v3 v = vsub(v2, v0);
            it->N = V((u.y * v.z) - (u.z * v.y), (u.z * v.x) - (u.x * v.z), (u.x * v.y) - (u.y * v.x));
            it->prim = i;
            found = 1;
        }
    }
    for (uint32_t i = 0; i < s->n_spheres; ++i) {                           /* :398-412 */
        const osph* sp = &s->spheres[i];
        const float t = sphere_t(o, d, Vp(sp->P), sp->r);
        if ((t > eps) && (t < it->t + eps)) {
            it->mat = &sp->mat;
            it->t = t;
            it->P = vadd(o, vmuls(d, t));
            it->N = vdivs(vsub(it->P, Vp(sp->P)), sp->r);
            it->prim = (1u << 30) | i;
            found = 1;
        }
    }
    for (uint32_t i = 0; i < s->n_planes; ++i) {                            /* :414-428 */
        const opln* pl = &s->planes[i];
        const float t = plane_t(o, d, Vp(pl->N), pl->len);
        if ((t > eps) && (t < it->t - eps)) {
            it->mat = &pl->mat;
            it->t = t;
            it->P = vadd(o, vmuls(d, t));
            it->N = Vp(pl->N);
            it->prim = (2u << 30) | i;
            found = 1;
        }
    }
    return found;
}

int vkro_trace_ray(const oscene* s, const float o[3], const float d[3], float t_max, float eps,
                   uint32_t* prim_id, float* t_out)
{
    isect_t it;
    memset(&it, 0, sizeof(it));
    it.t = t_max;
    it.prim = 0xFFFFFFFFu;
    const int f = trace_ray(s, Vp(o), Vp(d), &it, eps);
    if (prim_id) *prim_id = f ? it.prim : 0xFFFFFFFFu;
    if (t_out) *t_out = it.t;
    return f;
}

/* ------------------------------------------------------------------------ */
/* radiance, Tracer.comp:433-553                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    const oscene* s;
    const oframe* f;
    float eps;
    uint32_t max_depth;
    uint32_t flags;
    uint32_t rays;
    /* primary-hit capture (first call only) */
    int capture;
    uint32_t prim;
    float prim_t;
} path_ctx;

/* Unoccluded next-event contribution of one emitter sample (Tracer.comp:475-502). */
static v3 nee_contribution(v3 P, v3 N, v3 cam_pos, const omat* mat, const osph* sp, float t, v3 L)
{
    const v3 albedo = Vp(mat->albedo);
    const float x = t / sp->r + 1.0f;
    v3 att = vdivs(vmuls(Vp(sp->mat.emissive), 1.0f), x * x);   /* :475 */
    att = vdivs(vsubs(att, K_ATT_SUB), K_ATT_DIV);              /* :478 */
    att = vmaxs(att, 0.0f);                                     /* :479 */

    v3 F0 = V(0.04f, 0.04f, 0.04f);
    F0 = vmix(F0, albedo, mat->metalness);                      /* :482 */
    const v3 Vv = vnormalize(vsub(cam_pos, P));                 /* :484 */
    const v3 H = vnormalize(vadd(Vv, L));

    const float NDF = distribution_ggx(N, H, mat->roughness);
    const float G = geometry_smith(N, Vv, L, mat->roughness);
    const v3 F = fresnel_schlick(gmax(vdot(H, Vv), 0.0f), F0);

    const v3 kS = F;
    v3 kD = svsub(1.0f, kS);
    kD = vmuls(kD, 1.0f - mat->metalness);

    const v3 numerator = vmuls(F, NDF * G);
    const float denominator = 4.0f * gmax(vdot(N, Vv), 0.0f) * gmax(vdot(N, L), 0.0f);
    const v3 specular = vdivs(numerator, gmax(denominator, 0.001f));

    const float NdotL = gmax(vdot(N, L), 0.0f);
    return vmuls(vmul(vadd(vdivs(vmul(kD, albedo), K_PI), specular), att), NdotL);
}

static v3 radiance(path_ctx* pc, rng_t* rng, v3 ro, v3 rd, uint16_t* decision)
{
    const oscene* s = pc->s;
    v3 acc = V(0.0f, 0.0f, 0.0f);
    v3 mask = V(1.0f, 1.0f, 1.0f);
    uint16_t rec = 0;
    const uint16_t rep = decision ? *decision : 0;

    for (uint32_t depth = 0; depth < pc->max_depth; ++depth) {
        acc = vclamps(acc, 0.0f, 1.0f);                                      /* :441 */

        isect_t it;
        memset(&it, 0, sizeof(it));
        const float dp1 = (float)(depth + 1);
        it.t = 3000.0f / (dp1 * dp1);                                        /* :444 */
        it.prim = 0xFFFFFFFFu;
        pc->rays++;
        const int hit = trace_ray(s, ro, rd, &it, pc->eps);
        if (pc->capture) {
            pc->capture = 0;
            pc->prim = hit ? it.prim : 0xFFFFFFFFu;
            pc->prim_t = it.t;
        }
        if (!hit) break;                                                     /* :445 */

        const omat* mat = it.mat;
        const v3 albedo = Vp(mat->albedo);
        const v3 emis = Vp(mat->emissive);

        if (mat->type == 0u) {                                               /* DIFFUSE :451-513 */
            const float r2 = rng_next(rng);
            const float phi = K_TWO_PI * rng_next(rng);
            const v3 d = vmuls(jitter(it.N, phi, sqrtf(r2), sqrtf(1.0f - r2)), 1.0f - mat->metalness);

            v3 e = V(0.0f, 0.0f, 0.0f);
            for (uint32_t i = 0; i < s->n_spheres; ++i) {                    /* :458-504 */
                const osph* sp = &s->spheres[i];
                if (veq(Vp(sp->mat.emissive), V(0.0f, 0.0f, 0.0f))) continue;
                const v3 sP = Vp(sp->P);
                const float t = vlength(vsub(sP, it.P)) - sp->r;
                const v3 l0 = vsub(sP, it.P);
                const float cos_a_max = sqrtf(1.0f - gclamp(sp->r * sp->r / vdot(l0, l0), 0.0f, 1.0f));
                const float cosa = gmix(cos_a_max, 1.0f, rng_next(rng));
                const float phi2 = K_TWO_PI * rng_next(rng);
                const v3 L = jitter(l0, phi2, sqrtf(1.0f - cosa * cosa), cosa);

                isect_t sh;
                memset(&sh, 0, sizeof(sh));
                sh.t = t;
                pc->rays++;
                if (!trace_ray_any(s, it.P, L, &sh, pc->eps, 1))
                    e = vadd(e, nee_contribution(it.P, it.N, Vp(pc->f->camera.pos), mat, sp, t, L));
            }

            /* :507 when_gt(emissive, 0) == vec3(1) ? normalize(emissive) : 0 */
            const int all_pos = gmax(gsign(emis.x), 0.0f) == 1.0f && gmax(gsign(emis.y), 0.0f) == 1.0f &&
                                gmax(gsign(emis.z), 0.0f) == 1.0f;
            const v3 emissive = all_pos ? vnormalize(emis) : V(0.0f, 0.0f, 0.0f);

            acc = vadd(acc, vmul(mask, vadd(emissive, e)));                 /* :509 */
            mask = vmul(mask, albedo);                                       /* :510 */
            const v3 nd = vnormalize(vadd(vreflect(rd, it.N), d));           /* :511 */
            ro = it.P;
            rd = nd;
        } else if (mat->type == 1u) {                                        /* DIELECTRIC :514-542 */
            acc = vadd(acc, vmul(emis, mask));
            mask = vmul(mask, albedo);
            const v3 normal = it.N;
            const float nint = mat->roughness;
            const float cosine = -vdot(rd, it.N) / vlength(rd);
            const v3 reflected = vreflect(rd, it.N);
            const v3 refracted = vrefract(rd, normal, nint);
            const float p_refl = veq(refracted, V(0.0f, 0.0f, 0.0f)) ? 1.0f : schlick(cosine, mat->roughness);
            const float u = rng_next(rng);
            int choose_refl = u < p_refl;
            if (depth < 8) {
                if (pc->flags & OR_FLAG_REPLAY) choose_refl = (rep >> (2 * depth + 1)) & 1u;
                if (choose_refl) rec |= (uint16_t)(1u << (2 * depth + 1));
            }
            const v3 nd = vnormalize(choose_refl ? reflected : refracted);
            ro = it.P;
            rd = nd;
        }
        /* any other type: the GLSL switch has no case; the ray is unchanged */

        const float p = max3(mask);                                          /* :545-549 */
        const float u = rng_next(rng);
        int terminate = u > p;
        if (depth < 8) {
            if (pc->flags & OR_FLAG_REPLAY) terminate = !((rep >> (2 * depth)) & 1u);
            if (!terminate) rec |= (uint16_t)(1u << (2 * depth));
        }
        if (terminate) break;
        mask = vmuls(mask, 1.0f / p);
    }
    if (decision && (pc->flags & OR_FLAG_RECORD)) *decision = rec;
    return acc;
}

static inline uint8_t unorm8(float x)
{
    if (x != x) return 0;
    const float c = gclamp(x, 0.0f, 1.0f);
    return (uint8_t)floorf(c * 255.0f + 0.5f);
}

void vkro_primary_dir(const oframe* f, uint32_t x, uint32_t y, uint32_t w, uint32_t h, float out[3])
{
    const float uvx = (float)x / (float)w;                                  /* :563 */
    const float uvy = (float)y / (float)h;
    const float tx = 2.0f * uvx - 1.0f;                                      /* :568 */
    const float ty = 2.0f * uvy - 1.0f;
    const ocam* c = &f->camera;
    v3 dir = vadd(vadd(Vp(c->dir), vmuls(Vp(c->right), tx)), vmuls(Vp(c->up), ty));
    dir = vmul(dir, V(f->aspect_ratio, 1.0f, f->aspect_ratio));              /* :570 .xyx */
    const v3 n = vnormalize(dir);
    out[0] = n.x; out[1] = n.y; out[2] = n.z;
}

static void path_pixel(const oscene* s, const oframe* f, const oparams* p, uint32_t x, uint32_t y,
                       const oout* out, uint32_t oi)
{
    path_ctx pc;
    memset(&pc, 0, sizeof(pc));
    pc.s = s;
    pc.f = f;
    pc.eps = p->epsilon;
    pc.max_depth = p->max_depth;
    pc.flags = p->flags;
    pc.capture = 1;
    pc.prim = 0xFFFFFFFFu;

    const float uvx = (float)x / (float)p->width;
    const float uvy = (float)y / (float)p->height;
    rng_t rng;
    rng.salt = uvy / uvx + f->seed;                                          /* :565 */
    rng.cx = uvx;                                                            /* :566 */
    rng.cy = uvy;

    float dir[3];
    vkro_primary_dir(f, x, y, p->width, p->height, dir);
    const v3 ro = Vp(f->camera.pos);
    const v3 rd = Vp(dir);

    v3 accum = V(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < p->spp; ++i) {                                  /* :578-581 */
        uint16_t* dec = out->decisions ? &out->decisions[(size_t)oi * p->spp + i] : NULL;
        accum = vadd(accum, radiance(&pc, &rng, ro, rd, dec));
    }
    accum = vdivs(accum, (float)p->spp);                                     /* :585 */
    if (out->radiance) {
        float* r = &out->radiance[(size_t)oi * 4];
        r[0] = accum.x; r[1] = accum.y; r[2] = accum.z; r[3] = 1.0f;
    }
    accum = vdiv(accum, vadds(accum, 1.0f));                                 /* :587 */
    accum = V(vkro_pow(accum.x, K_INV_GAMMA), vkro_pow(accum.y, K_INV_GAMMA), vkro_pow(accum.z, K_INV_GAMMA));
    const float dither = rng_next(&rng) / 64.0f;                             /* :590 */
    const v3 col = vadds(accum, dither);
    if (out->rgba8) {
        uint8_t* o = &out->rgba8[(size_t)oi * 4];
        o[0] = unorm8(col.x); o[1] = unorm8(col.y); o[2] = unorm8(col.z); o[3] = 255;
    }
    if (out->prim_id) out->prim_id[oi] = pc.prim;
    if (out->hit_t) out->hit_t[oi] = pc.prim_t;
    if (out->ray_count) out->ray_count[oi] = pc.rays;
}

/* ------------------------------------------------------------------------ */
/* Whitted tracer, Assets/Raytracer.comp (its own constant scene)            */
/* ------------------------------------------------------------------------ */
#define W_EPS 0.01f                                                         /* :89 */
typedef struct { uint32_t mat; v3 o; float r; } wsph;
typedef struct { uint32_t mat; v3 n; float d; } wpln;
static const wsph W_SPH[2] = {                                              /* :98-102 */
    {5, {-14.0f, 12.0f, 32.0f}, 5.0f},
    {4, {32.0f, 24.0f, 25.0f}, 12.0f},
};
/* :104-112 (the black back wall is commented out in the reference, leaving
 * exactly PLANE_COUNT = 5 planes). */
static const wpln W_PLN[5] = {
    {0, {0.0f, 1.0f, 0.0f}, 0.0f},     /* floor */
    {0, {0.0f, -1.0f, 0.0f}, 128.0f},  /* ceiling */
    {2, {0.0f, 0.0f, -1.0f}, 64.0f},   /* green wall */
    {1, {1.0f, 0.0f, 0.0f}, 64.0f},    /* red wall */
    {3, {-1.0f, 0.0f, 0.0f}, 64.0f},   /* blue wall */
};
static const v3 W_TRI[3] = {{10.0f, 10.0f, 0.0f}, {0.0f, 20.0f, 0.0f}, {-10.0f, 10.0f, 0.0f}}; /* :114-117, mat 1 */
static const int W_REFL[6] = {1, 0, 1, 0, 0, 0};                           /* :119-127 */
static const v3 W_DIFF[6] = {{1, 1, 1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}};

static inline float w_plane_t(v3 o, v3 d, v3 n, float dist)                /* :180-192 */
{
    const float dd = vdot(d, n);
    if (dd == 0.0f) return 0.0f;
    const float t = -(dist + vdot(o, n)) / dd;
    return gmax(t, 0.0f);
}

typedef struct { uint32_t mat; float dist; v3 P, N; uint32_t prim; } wisect;

static int w_trace(v3 o, v3 d, wisect* it)                                 /* :224-278 */
{
    int found = 0;
    {
        const float t = tri_t(o, d, W_TRI[0], W_TRI[1], W_TRI[2], W_EPS);
        if (t > W_EPS && t < it->dist) {
            it->dist = t;
            it->mat = 1;
            it->P = vadd(o, vmuls(d, t));
            const v3 u = vsub(W_TRI[1], W_TRI[0]);
            const v3 v = vsub(W_TRI[2], W_TRI[0]);
            it->N = V((u.y * v.z) - (u.z * v.y), (u.z * v.x) - (u.x * v.z), (u.x * v.y) - (u.y * v.x));
            it->prim = 0;
            found = 1;
        }
    }
    for (int i = 0; i < 2; ++i) {
        const float t = sphere_t(o, d, W_SPH[i].o, W_SPH[i].r);
        if (t > W_EPS && t < it->dist) {
            it->dist = t;
            it->mat = W_SPH[i].mat;
            it->P = vadd(o, vmuls(d, t));
            it->N = vdivs(vsub(it->P, W_SPH[i].o), W_SPH[i].r);
            it->prim = (1u << 30) | (uint32_t)i;
            found = 1;
        }
    }
    for (int i = 0; i < 5; ++i) {
        const wpln* pl = &W_PLN[i];
        const float t = w_plane_t(o, d, pl->n, pl->d);
        if (t > W_EPS && t < it->dist) {
            it->dist = t;
            it->mat = pl->mat;
            it->P = vadd(o, vmuls(d, t));
            it->N = pl->n;
            it->prim = (2u << 30) | (uint32_t)i;
            found = 1;
        }
    }
    return found;
}

typedef struct { float r, g, b, a; } v4;

static v4 w_render(const oframe* f, v3* ro, v3* rd, uint32_t* bounce, uint32_t* rays, int* cap,
                   uint32_t* prim, float* pt)                               /* :280-355 */
{
    wisect it;
    memset(&it, 0, sizeof(it));
    it.dist = 1000.0f;
    it.prim = 0xFFFFFFFFu;
    (*rays)++;
    const int hit = w_trace(*ro, *rd, &it);
    if (*cap) { *cap = 0; *prim = hit ? it.prim : 0xFFFFFFFFu; *pt = it.dist; }
    v4 out = {0.0f, 0.0f, 0.0f, 1.0f};
    if (!hit) return out;

    const v3 lp = Vp(f->light_pos);
    const v3 light_vec = vnormalize(vsub(lp, it.P));
    const float dist_to_light = vlength(vsub(lp, it.P));
    const int refl = W_REFL[it.mat];
    const v3 mdiff = W_DIFF[it.mat];

    const float li = 540.0f / (12.56637072f * dist_to_light);              /* :308 */
    const v3 diffuse = vmuls(vmuls(mdiff, li), gmax(vdot(it.N, light_vec), 0.0f));
    const v3 half_vec = vnormalize(vadd(light_vec, vnormalize(Vp(f->camera.pos))));
    const float spec = li * gpow16(gclamp(vdot(it.N, half_vec), 0.0f, 1.0f));
    v3 color = vadds(diffuse, spec);

    wisect sh;
    memset(&sh, 0, sizeof(sh));
    sh.dist = dist_to_light;
    (*rays)++;
    if (w_trace(it.P, light_vec, &sh)) {
        color = vmuls(color, 0.5f);
        *bounce = 3;
    }
    if (refl) {
        *rd = vsub(*rd, vmuls(it.N, 2.0f * vdot(it.N, *rd)));               /* :194-197 */
        *ro = it.P;
    } else {
        *bounce = 3;
    }
    out.r = color.x; out.g = color.y; out.b = color.z; out.a = 1.0f;
    return out;
}

static inline v4 v4lin(v4 a, float sa, v4 b, float sb)
{
    v4 r = {a.r * sa + b.r * sb, a.g * sa + b.g * sb, a.b * sa + b.b * sb, a.a * sa + b.a * sb};
    return r;
}

static void whitted_pixel(const oframe* f, const oparams* p, uint32_t x, uint32_t y, const oout* out,
                          uint32_t oi)                                       /* :357-399 */
{
    float dir[3];
    vkro_primary_dir(f, x, y, p->width, p->height, dir);
    v3 ro = Vp(f->camera.pos), rd = Vp(dir);
    uint32_t bounce = 0, rays = 0, prim = 0xFFFFFFFFu;
    int cap = 1;
    float pt = 0.0f;
    v4 final_color = w_render(f, &ro, &rd, &bounce, &rays, &cap, &prim, &pt);
    float s = 0.4f;
    while (++bounce <= 2) {
        const v4 rc = w_render(f, &ro, &rd, &bounce, &rays, &cap, &prim, &pt);
        /* (1 - s) * final + s * mix(rc, final, 1 - s) */
        const v4 m = v4lin(rc, 1.0f - (1.0f - s), final_color, 1.0f - s);
        const v4 a = {final_color.r * (1.0f - s), final_color.g * (1.0f - s), final_color.b * (1.0f - s),
                      final_color.a * (1.0f - s)};
        const v4 b = {m.r * s, m.g * s, m.b * s, m.a * s};
        final_color.r = a.r + b.r; final_color.g = a.g + b.g;
        final_color.b = a.b + b.b; final_color.a = a.a + b.a;
        s = s * 0.5f;
    }
    if (out->radiance) {
        float* r = &out->radiance[(size_t)oi * 4];
        r[0] = final_color.r; r[1] = final_color.g; r[2] = final_color.b; r[3] = final_color.a;
    }
    if (out->rgba8) {
        uint8_t* o = &out->rgba8[(size_t)oi * 4];
        o[0] = unorm8(final_color.r); o[1] = unorm8(final_color.g);
        o[2] = unorm8(final_color.b); o[3] = unorm8(final_color.a);
    }
    if (out->prim_id) out->prim_id[oi] = prim;
    if (out->hit_t) out->hit_t[oi] = pt;
    if (out->ray_count) out->ray_count[oi] = rays;
}

/* ------------------------------------------------------------------------ */
/* Parallel driver                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    const oscene* s;
    const oframe* f;
    const oparams* p;
    const uint32_t* pixels;
    uint32_t n;
    const oout* out;
    atomic_uint next;
} job_t;

static void* worker(void* arg)
{
    job_t* j = (job_t*)arg;
    const uint32_t chunk = 64;
    for (;;) {
        const uint32_t b = atomic_fetch_add(&j->next, chunk);
        if (b >= j->n) break;
        const uint32_t e = (b + chunk < j->n) ? b + chunk : j->n;
        for (uint32_t i = b; i < e; ++i) {
            const uint32_t pix = j->pixels ? j->pixels[i] : i;
            const uint32_t x = pix % j->p->width, y = pix / j->p->width;
            if (j->p->kernel == OR_KERNEL_WHITTED)
                whitted_pixel(j->f, j->p, x, y, j->out, i);
            else
                path_pixel(j->s, j->f, j->p, x, y, j->out, i);
        }
    }
    return NULL;
}

int vkro_trace_pixels(const oscene* scene, const oframe* frame, const oparams* params,
                      const uint32_t* pixels, uint32_t n_pixels, const oout* out, int n_threads)
{
    if (!frame || !params || !out) return 4;
    if (params->kernel == OR_KERNEL_PATH && (!scene || !scene->tri_material)) return 4;
    if (params->width == 0 || params->height == 0 || params->spp == 0) return 4;
    if (!pixels) n_pixels = params->width * params->height;
    else
        for (uint32_t i = 0; i < n_pixels; ++i)
            if (pixels[i] >= params->width * params->height) return 4;
    job_t j;
    j.s = scene; j.f = frame; j.p = params; j.pixels = pixels; j.n = n_pixels; j.out = out;
    atomic_init(&j.next, 0u);
    if (n_threads <= 1) {
        worker(&j);
        return 0;
    }
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    int started = 0;
    for (int i = 0; i < n_threads; ++i) {
        if (pthread_create(&th[i], NULL, worker, &j) != 0) break;
        ++started;
    }
    if (started == 0) worker(&j);
    for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Exported helpers for unit tests                                            */
/* ------------------------------------------------------------------------ */
float vkro_distribution_ggx(const float N[3], const float H[3], float r) { return distribution_ggx(Vp(N), Vp(H), r); }
float vkro_geometry_smith(const float N[3], const float Vv[3], const float L[3], float r)
{
    return geometry_smith(Vp(N), Vp(Vv), Vp(L), r);
}
void vkro_fresnel_schlick(float c, const float F0[3], float out[3])
{
    const v3 r = fresnel_schlick(c, Vp(F0));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float vkro_schlick(float c, float ior) { return schlick(c, ior); }
void vkro_jitter(const float d[3], float phi, float sina, float cosa, float out[3])
{
    const v3 r = jitter(Vp(d), phi, sina, cosa);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void vkro_reflect(const float I[3], const float N[3], float out[3])
{
    const v3 r = vreflect(Vp(I), Vp(N));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void vkro_refract(const float I[3], const float N[3], float eta, float out[3])
{
    const v3 r = vrefract(Vp(I), Vp(N), eta);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* ------------------------------------------------------------------------ */
/* Present pass, Assets/Fullscreen.frag:14-31                                 */
/* ------------------------------------------------------------------------ */
/* Bilinear sample of an rgba8 image with CLAMP_TO_BORDER (opaque black),
 * VK_FILTER_LINEAR (Source/GraphicsDevice.cpp:770-790); texel centres at
 * (i + 0.5) / size. */
static void sample_bilinear(const uint8_t* img, uint32_t w, uint32_t h, float u, float v, float out[3])
{
    const float x = u * (float)w - 0.5f;
    const float y = v * (float)h - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int x0 = (int)fx0, y0 = (int)fy0;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int dy = 0; dy < 2; ++dy) {
        for (int dx = 0; dx < 2; ++dx) {
            const int xi = x0 + dx, yi = y0 + dy;
            const float wgt = (dx ? ax : 1.0f - ax) * (dy ? ay : 1.0f - ay);
            if (xi < 0 || yi < 0 || xi >= (int)w || yi >= (int)h) continue;   /* border = 0 */
            const uint8_t* t = &img[((size_t)yi * w + (size_t)xi) * 4];
            for (int c = 0; c < 3; ++c) acc[c] = acc[c] + ((float)t[c] / 255.0f) * wgt;
        }
    }
    out[0] = acc[0]; out[1] = acc[1]; out[2] = acc[2];
}

int vkro_filter(const uint8_t* cur, const uint8_t* prev, uint32_t sw, uint32_t sh, uint8_t* out,
                uint32_t dw, uint32_t dh)
{
    if (!cur || !prev || !out || !sw || !sh || !dw || !dh) return 4;
    const float inv_size = 1.0f / 2048.0f;                                   /* :12 */
    for (uint32_t y = 0; y < dh; ++y) {
        for (uint32_t x = 0; x < dw; ++x) {
            /* Fullscreen.vert:6-11 interpolates uv to the pixel centre. */
            const float us = ((float)x + 0.5f) / (float)dw;
            const float vt = ((float)y + 0.5f) / (float)dh;
            const float u = us, v = 1.0f - vt;                               /* :16 */
            float c0[3], c1[3], n[3], s[3], e[3], w[3];
            sample_bilinear(cur, sw, sh, u, v, c0);
            sample_bilinear(prev, sw, sh, u, v, c1);
            sample_bilinear(cur, sw, sh, u + 0.0f * inv_size, v + 1.0f * inv_size, n);
            sample_bilinear(cur, sw, sh, u + 0.0f * inv_size, v + -1.0f * inv_size, s);
            sample_bilinear(cur, sw, sh, u + 1.0f * inv_size, v + 0.0f * inv_size, e);
            sample_bilinear(cur, sw, sh, u + -1.0f * inv_size, v + 0.0f * inv_size, w);
            const v3 dlt = vsub(Vp(c0), Vp(c1));
            const float vT = vdot(dlt, dlt);                                 /* :26 */
            float fin[3];
            for (int c = 0; c < 3; ++c)
                fin[c] = (vT > 0.0005f) ? (n[c] + s[c] + e[c] + w[c]) / 4.0f : c0[c];   /* :28 */
            uint8_t* o = &out[((size_t)y * dw + x) * 4];
            o[0] = unorm8(fin[0]); o[1] = unorm8(fin[1]); o[2] = unorm8(fin[2]); o[3] = 255;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Camera::update, Source/Camera.cpp:48-64 (GLM float semantics:             */
/* radians = deg * float(pi/180); normalize = v * (1 / sqrt(dot(v, v)));      */
/* dot = (x*x' + y*y') + z*z'; std::cos/sin on float = cosf/sinf)             */
/* ------------------------------------------------------------------------ */
static inline v3 glm_normalize(v3 v)
{
    const float inv = 1.0f / sqrtf(vdot(v, v));
    return vmuls(v, inv);
}

void vkro_camera_update(const float pos[3], float pitch, float yaw, ocam* out)
{
    const float p = gclamp(pitch, -89.0f, 89.0f);
    const float rad = 0.01745329251994329576923690768489f;
    const float pr = p * rad, yr = yaw * rad;
    v3 front;
    front.x = cosf(pr) * cosf(yr);
    front.y = sinf(pr);
    front.z = cosf(pr) * sinf(yr);
    front = glm_normalize(front);
    const v3 P = Vp(pos);
    const v3 dir = glm_normalize(vsub(vadd(P, front), P));
    const v3 up0 = V(0.0f, 1.0f, 0.0f);
    const v3 right = glm_normalize(vcross(up0, dir));
    const v3 up = vcross(dir, right);
    memset(out, 0, sizeof(*out));
    out->pos[0] = P.x; out->pos[1] = P.y; out->pos[2] = P.z;
    out->dir[0] = dir.x; out->dir[1] = dir.y; out->dir[2] = dir.z;
    out->right[0] = right.x; out->right[1] = right.y; out->right[2] = right.z;
    out->up[0] = up.x; out->up[1] = up.y; out->up[2] = up.z;
}

void vkro_nee_contribution(const float P[3], const float N[3], const float cam_pos[3], const omat* mat,
                           const osph* sp, const float L[3], float out[3])
{
    const v3 p = Vp(P);
    const float t = vlength(vsub(Vp(sp->P), p)) - sp->r;                     /* :465 */
    const v3 c = nee_contribution(p, Vp(N), Vp(cam_pos), mat, sp, t, Vp(L));
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}
