/*
 * vkr_oracle.h - CPU restatement of the reference path tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP path: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark load
 * it.  The product library never links, includes or calls it.
 *
 * The oracle follows the reference GLSL literally (a linear scan over every
 * primitive in reference order, the reference's tie rules and its hash RNG)
 * with each function citing the line it restates.  GLSL leaves several
 * built-ins implementation-defined; the oracle fixes one meaning for each
 * (see vkr_oracle.c, "Built-in semantics") and the HIP kernels implement the
 * same meaning, so CPU and GPU results are comparable bit for bit.
 *
 * Struct layouts are the std430 layouts of Assets/Tracer.comp:59-166,
 * restated here so the oracle does not depend on the product headers.
 */
#ifndef VKR_ORACLE_H
#define VKR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct omat {           /* Tracer.comp:59-74 */
    float albedo[3]; float _p0;
    float emissive[3]; float roughness;
    float metalness; uint32_t type; float _p1[2];
} omat;

typedef struct otri {           /* Tracer.comp:127-132, GraphicsDevice.h:13-18 */
    float v0[3]; uint32_t mat_index;
    float v1[3]; float _p1;
    float v2[3]; float _p2;
} otri;

typedef struct osph { omat mat; float P[3]; float r; } osph;   /* Tracer.comp:102-111 */
typedef struct opln { omat mat; float N[3]; float len; } opln; /* Tracer.comp:118-125 */

typedef struct ocam {           /* Tracer.comp:22-35, Camera.h:5-12 */
    float pos[3]; float _p0; float dir[3]; float _p1;
    float right[3]; float _p2; float up[3]; float _p3;
} ocam;

typedef struct oframe {         /* Tracer.comp:151-166, GraphicsDevice.h:20-29 */
    float aspect_ratio; float seed; float _p0[2];
    float light_pos[3]; float _p1;
    ocam camera;
} oframe;

/* An accelerated trace_ray for the same-BVH CPU baseline (bvh_accel.cpp):
 * the winner of the reference scan for ray (o, d) with best = *t_io on entry,
 * as its primitive id and t, or when `any` is set only whether anything is
 * accepted.  Returns 1 when something is accepted. */
typedef int (*oaccel_fn)(const void* accel, const float o[3], const float d[3], float eps, int any,
                         float* t_io, uint32_t* prim);

typedef struct oscene {
    const otri* tris;    uint32_t n_tris;
    const osph* spheres; uint32_t n_spheres;
    const opln* planes;  uint32_t n_planes;
    const omat* tri_material;            /* material of every triangle (reference: mirror) */
    const omat* materials; uint32_t n_materials;
    uint32_t flags;                      /* bit 0: triangles use materials[mat_index] */
    /* Optional (NULL: the literal linear scan).  Set only by the CPU
     * baseline that traces the same BVH as the GPU (bvh_accel.cpp). */
    oaccel_fn accel_fn;
    const void* accel;
} oscene;

enum { OR_FLAG_RECORD = 1u, OR_FLAG_REPLAY = 2u };
enum { OR_KERNEL_PATH = 0, OR_KERNEL_WHITTED = 1 };

typedef struct oparams {
    uint32_t width, height, spp, max_depth;
    float    epsilon;
    uint32_t kernel;
    uint32_t flags;                      /* OR_FLAG_* */
    uint32_t _pad;
} oparams;

/* Outputs are indexed by position in the pixel list (or y*W+x for a full
 * frame when pixels == NULL).  Any output may be NULL. */
typedef struct oout {
    uint8_t*  rgba8;      /* 4 per pixel */
    float*    radiance;   /* 4 per pixel */
    uint32_t* prim_id;
    float*    hit_t;
    uint32_t* ray_count;
    uint16_t* decisions;  /* spp per pixel */
} oout;

/* Full pixel programs (Tracer.comp:557-593 / Raytracer.comp:357-399). */
int vkro_trace_pixels(const oscene* scene, const oframe* frame, const oparams* params,
                      const uint32_t* pixels, uint32_t n_pixels, const oout* out, int n_threads);

/* Building blocks, exported for unit tests. */
float vkro_cos(float x);
float vkro_sin(float x);
float vkro_pow(float x, float y);
void  vkro_rand_sequence(float cx, float cy, float salt0, uint32_t n, float* out);
float vkro_tri_intersect(const float o[3], const float d[3], const otri* tri, float eps);
float vkro_sphere_intersect(const float o[3], const float d[3], const osph* s);
float vkro_plane_intersect(const float o[3], const float d[3], const opln* p);
/* trace_ray (Tracer.comp:374-431): returns 1 on a hit and the winner. */
int   vkro_trace_ray(const oscene* scene, const float o[3], const float d[3], float t_max,
                     float eps, uint32_t* prim_id, float* t_out);
float vkro_distribution_ggx(const float N[3], const float H[3], float roughness);
float vkro_geometry_smith(const float N[3], const float V[3], const float L[3], float roughness);
void  vkro_fresnel_schlick(float cos_theta, const float F0[3], float out[3]);
float vkro_schlick(float cosine, float ior);
void  vkro_jitter(const float d[3], float phi, float sina, float cosa, float out[3]);
void  vkro_reflect(const float I[3], const float N[3], float out[3]);
void  vkro_refract(const float I[3], const float N[3], float eta, float out[3]);
void  vkro_primary_dir(const oframe* frame, uint32_t x, uint32_t y, uint32_t w, uint32_t h, float out[3]);
/* Unoccluded next-event contribution of one emitter sample L at hit (P, N)
 * seen from cam_pos (Tracer.comp:465, 475-502). */
void  vkro_nee_contribution(const float P[3], const float N[3], const float cam_pos[3], const omat* mat,
                            const osph* sp, const float L[3], float out[3]);

/* Present pass, Assets/Fullscreen.frag:14-31 with Fullscreen.vert:6-11. */
int vkro_filter(const uint8_t* cur, const uint8_t* prev, uint32_t src_w, uint32_t src_h,
                uint8_t* out, uint32_t dst_w, uint32_t dst_h);

/* Camera::update, Source/Camera.cpp:48-64 (pitch clamped to [-89, 89]). */
void vkro_camera_update(const float pos[3], float pitch, float yaw, ocam* out);

#ifdef __cplusplus
}
#endif
#endif
