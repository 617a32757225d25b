/*
 * vkr.h - C ABI of the MI355X (gfx950) path tracer.
 *
 * This header is the drop-in boundary for the compute-shader path of
 * stickyfingies/vk-renderer.  In the reference that path is a Vulkan compute
 * pipeline:
 *   - descriptor set 0 = { binding 0: rgba8 storage image,
 *                          binding 1: std430 SSBO `SceneData { Triangle tris[]; }` }
 *     (Source/GraphicsDevice.cpp:988-1010, Assets/Tracer.comp:144-149)
 *   - a 96-byte push constant `FrameData` (Source/GraphicsDevice.cpp:1106-1111,
 *     Assets/Tracer.comp:151-166, Include/GraphicsDevice.h:20-29)
 *   - per frame: bind, push, vkCmdDispatch(RES/16, RES/16, 1)
 *     (Source/GraphicsDevice.cpp:1254-1266)
 *
 * Every entry point below replaces one piece of that pipeline; the comment on
 * each one names the reference code it stands in for.  All structs keep the
 * reference's std430 byte layout (sizes and offsets are checked with
 * static asserts at the bottom of this file).
 *
 * Conventions
 *   - Every function returns a vkr_error (int8_t).  Nothing throws across the
 *     ABI.  vkr_last_error_string() describes the last failure of the calling
 *     thread.
 *   - Pointers documented as "device" are HIP device pointers; "host" pointers
 *     are read during the call only.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *   - A context is not internally synchronised: use one context per thread.
 */
#ifndef VKR_H
#define VKR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Errors: values 0..3 are GraphicsDevice::Error (Include/GraphicsDevice.h:46-52). */
/* ------------------------------------------------------------------------ */
typedef int8_t vkr_error;
enum {
    VKR_SUCCESS             = 0, /* Error::SUCCESS */
    VKR_NO_SUITABLE_GPU     = 1, /* Error::NO_SUITABLE_GPU */
    VKR_NO_SUITABLE_SURFACE = 2, /* Error::NO_SUITABLE_SURFACE */
    VKR_UNKNOWN             = 3, /* Error::UNKNOWN */
    VKR_INVALID_ARGUMENT    = 4, /* build extension: bad size, null pointer, bad id */
    VKR_OUT_OF_MEMORY       = 5, /* build extension: device or host allocation failed */
    VKR_HIP_ERROR           = 6, /* build extension: a HIP runtime call failed */
    VKR_UNSUPPORTED         = 7  /* build extension: feature not available for this input */
};

/* ------------------------------------------------------------------------ */
/* std430 data layouts                                                       */
/* ------------------------------------------------------------------------ */

/* Material types, Assets/Tracer.comp:51-52. */
enum { VKR_MAT_DIFFUSE = 0, VKR_MAT_DIELECTRIC = 1 };

/* Material, Assets/Tracer.comp:59-74 (std430: albedo@0 emissive@16
 * roughness@28 metalness@32 type@36, size 48).  For dielectrics the index of
 * refraction lives in `roughness` (Tracer.comp:532-539). */
typedef struct vkr_material {
    float    albedo[3];
    float    _pad0;
    float    emissive[3];
    float    roughness;
    float    metalness;
    uint32_t type;
    float    _pad1[2];
} vkr_material;

/* Triangle, Include/GraphicsDevice.h:13-18 and Assets/Tracer.comp:127-132
 * (std430 stride 48: v0@0 v1@16 v2@32).  The 4 padding bytes after v0 may
 * carry a material index when VKR_SCENE_TRI_MATERIAL_INDEX is set; the
 * reference never reads them. */
typedef struct vkr_triangle {
    float    v0[3];
    uint32_t mat_index;
    float    v1[3];
    float    _pad1;
    float    v2[3];
    float    _pad2;
} vkr_triangle;

/* Sphere, Assets/Tracer.comp:102-111 (mat@0 P@48 r@60, size 64). */
typedef struct vkr_sphere {
    vkr_material mat;
    float        P[3];
    float        r;
} vkr_sphere;

/* Plane dot(X, N) + len = 0, Assets/Tracer.comp:118-125 (mat@0 N@48 len@60). */
typedef struct vkr_plane {
    vkr_material mat;
    float        N[3];
    float        len;
} vkr_plane;

/* CameraData, Include/Camera.h:5-12 and Assets/Tracer.comp:22-35
 * (four 16-byte aligned vec3: pos@0 dir@16 right@32 up@48). */
typedef struct vkr_camera {
    float pos[3];   float _pad0;
    float dir[3];   float _pad1;
    float right[3]; float _pad2;
    float up[3];    float _pad3;
} vkr_camera;

/* FrameData push constant, Include/GraphicsDevice.h:20-29 and
 * Assets/Tracer.comp:151-166 (aspect@0 seed@4 light_pos@16 camera@32, 96 B). */
typedef struct vkr_frame_data {
    float      aspect_ratio;
    float      seed;
    float      _pad0[2];
    float      light_pos[3];
    float      _pad1;
    vkr_camera camera;
} vkr_frame_data;

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct vkr_ctx vkr_ctx;

typedef struct vkr_create_info {
    int32_t  device; /* HIP device ordinal; -1 = the current device */
    uint32_t flags;  /* VKR_CREATE_* */
} vkr_create_info;

enum { VKR_CREATE_DEBUG = 1u << 0 /* CreateInfo::debug, Include/GraphicsDevice.h:70-71 */ };

/* Replaces GraphicsDevice::Construct's device selection and compute pipeline
 * creation (Source/GraphicsDevice.cpp:282-343, 1089-1131).  Fails with
 * VKR_NO_SUITABLE_GPU when no HIP device is visible or the device is not
 * gfx950.  Allocates nothing large. */
vkr_error vkr_create(const vkr_create_info* info, vkr_ctx** out_ctx);

/* Replaces GraphicsDevice::Destruct (Source/GraphicsDevice.cpp:1138-1213).
 * Frees every scene and buffer the context owns.  NULL is accepted. */
vkr_error vkr_destroy(vkr_ctx* ctx);

/* Replaces GraphicsDevice::WaitIdle (Source/GraphicsDevice.cpp:1344-1347). */
vkr_error vkr_wait_idle(vkr_ctx* ctx);

/* Thread-local description of the last error returned on this thread. */
const char* vkr_last_error_string(void);

/* ------------------------------------------------------------------------ */
/* Scenes                                                                    */
/* ------------------------------------------------------------------------ */
enum {
    /* Triangles take materials[tri.mat_index] instead of the single
     * tri_material (build extension; the reference shades every triangle
     * with `mirror`, Assets/Tracer.comp:386). */
    VKR_SCENE_TRI_MATERIAL_INDEX = 1u << 0,
    /* Build the triangle subtree on the device (a linear BVH over Morton
     * codes; subtrees of at most max_leaf_size triangles become one leaf)
     * instead of the host SAH builder: uploads of large scenes take
     * milliseconds, traversal is a few percent slower, images are the same.
     * traversal_cost does not apply; when the tree is deeper than max_depth
     * (0 = 63) the host builder is used instead (vkr_scene_info.builder says
     * which ran).  Build extension: the reference has no acceleration
     * structure. */
    VKR_SCENE_DEVICE_BVH = 1u << 1
};

typedef struct vkr_scene_desc {
    const vkr_triangle* tris;        uint32_t n_tris;     /* host */
    const vkr_sphere*   spheres;     uint32_t n_spheres;  /* host */
    const vkr_plane*    planes;      uint32_t n_planes;   /* host */
    const vkr_material* tri_material; /* host; NULL = the reference `mirror` (Tracer.comp:192) */
    const vkr_material* materials;   uint32_t n_materials; /* host; per-triangle table */
    uint32_t flags;                  /* VKR_SCENE_* */
    uint32_t max_leaf_size;          /* BVH leaf size cap (0 = 1 with 2+ emitters, else 2; at most 8) */
    uint32_t max_depth;              /* host BVH depth cap (0 = 23, the deepest tree the 5-wave kernel takes; at most 63;
                                        a cap below the least depth the primitives allow gives that depth) */
    float    traversal_cost;         /* SAH cost of a node visit relative to one primitive test (0 = 3) */
} vkr_scene_desc;

/* Replaces the SceneData SSBO creation and upload
 * (Source/GraphicsDevice.cpp:796-822) and the shader-constant spheres/planes
 * (Assets/Tracer.comp:196-211).  Builds a binned-SAH BVH over triangles and
 * spheres on the host (or, with VKR_SCENE_DEVICE_BVH, the triangle subtree on
 * the device), uploads everything to device-local memory and returns a scene
 * id owned by the context.  Planes stay a linear list. */
vkr_error vkr_scene_upload(vkr_ctx* ctx, const vkr_scene_desc* desc, uint32_t* out_scene);

/* Frees one scene.  Refused (VKR_INVALID_ARGUMENT) while a captured graph
 * (vkr_batch_graph_create) traces it: destroy those graphs first. */
vkr_error vkr_scene_destroy(vkr_ctx* ctx, uint32_t scene);

typedef struct vkr_scene_info {
    uint32_t n_tris, n_spheres, n_planes, n_emitters;
    uint32_t n_nodes, n_leaves, depth, stack_needed;   /* stack_needed: the traversal-stack variant
                                                            (24, 32 or 64 entries) that holds depth + 1 */
    uint64_t device_bytes;           /* nodes + primitive records + materials */
    double   build_ms;               /* BVH build time (wall clock, device builds included) */
    uint32_t builder;                /* VKR_BUILDER_* */
    uint32_t _pad;
} vkr_scene_info;

enum { VKR_BUILDER_HOST_SAH = 0u, VKR_BUILDER_DEVICE_LBVH = 1u };

vkr_error vkr_scene_get_info(vkr_ctx* ctx, uint32_t scene, vkr_scene_info* out);

/* Differentiable parameters of a scene's spheres, laid out per sphere as
 * [albedo.r, albedo.g, albedo.b, emission.r, emission.g, emission.b,
 *  roughness] followed by one global light scale L, so
 * n = 7 * n_spheres + 1.  The traced emissive of sphere s is L * emission_s.
 * Writes the materials of every sphere (albedo, emissive, roughness; type and
 * metalness stay) and recomputes the emitter list in place (its device
 * buffer holds every sphere, so it never moves); captured graphs that trace
 * the scene see the new list on their next launch.  Synchronises the device.
 * Host pointer.  Values that differ from the current ones in any bit make the
 * adjoint's kept first-pass records stale (see vkr_adjoint).
 * Build extension: the reference has no parameters (Tracer.comp:186-202). */
vkr_error vkr_scene_set_sphere_params(vkr_ctx* ctx, uint32_t scene, const float* params, uint32_t n);

/* Scene files (build extension; the reference has no scene file, only the
 * bytes it copies into its SceneData SSBO, Source/GraphicsDevice.cpp:796-822).
 * A file holds the arrays of a vkr_scene_desc in their std430 layouts behind
 * a 64-byte little-endian header:
 *   char magic[8] = "VKRSCENE"; u32 version = 1; u32 flags;
 *   u32 n_tris, n_spheres, n_planes, n_materials; u32 has_tri_material;
 *   u32 max_leaf_size, max_depth; f32 traversal_cost; u32 reserved[4] = 0
 * then tri_material (48 B, when has_tri_material), tris (48 B each),
 * spheres (64 B each), planes (64 B each), materials (48 B each).
 * Host only: these calls touch no device and need no context. */
vkr_error vkr_scene_save(const char* path, const vkr_scene_desc* desc);
/* Reads a scene file into *out (arrays allocated by the library); a file
 * whose size does not match its header is rejected. */
vkr_error vkr_scene_load(const char* path, vkr_scene_desc* out);
/* Releases the arrays of a vkr_scene_load result and zeroes *desc. */
void vkr_scene_free(vkr_scene_desc* desc);

/* ------------------------------------------------------------------------ */
/* Tracing                                                                   */
/* ------------------------------------------------------------------------ */
enum {
    VKR_KERNEL_PATH    = 0, /* Assets/Tracer.comp */
    VKR_KERNEL_WHITTED = 1  /* Assets/Raytracer.comp (its own constant scene) */
};

enum {
    VKR_SCHED_AUTO       = 0, /* autotuned choice, else persistent */
    VKR_SCHED_PLAIN      = 1, /* one lane per pixel, one launch */
    VKR_SCHED_PERSISTENT = 2  /* persistent waves refilled from a pixel pool */
};

enum {
    /* RR / dielectric decisions are written to outputs.decisions. */
    VKR_TRACE_RECORD_DECISIONS = 1u << 0,
    /* RR / dielectric decisions are read from outputs.decisions instead of
     * being drawn (the random numbers are still drawn, so the stream is
     * unchanged).  Used for finite differences with frozen branches. */
    VKR_TRACE_REPLAY_DECISIONS = 1u << 1,
    /* Shadow rays: with at least 2 emitters a wave traces up to 32 rays per
     * lane and step from a shared pool, else one ray per lane and step.
     * These force one or the other; the results are identical either way. */
    VKR_TRACE_RAY_POOL         = 1u << 2,
    VKR_TRACE_NO_RAY_POOL      = 1u << 3,
    /* vkr_trace_frame only (path kernel, max_depth <= 10, no decision
     * flags): also keeps, in the context, what the first pass of vkr_adjoint
     * computes for this frame (each sample's random-stream position at its
     * start and which channels of its clamp saturate, 8 bytes per sample;
     * growing that buffer waits for the device).  The next
     * vkr_adjoint with the same scene, sphere parameters, frame data and
     * trace parameters then runs only its second pass.  The image is
     * unchanged.  Refused by vkr_trace_batch, vkr_batch_graph_create and
     * vkr_trace_frame_static. */
    VKR_TRACE_RECORD_ADJOINT   = 1u << 4
};

typedef struct vkr_trace_params {
    uint32_t width, height;  /* image size; the reference uses RES x RES (Source/GraphicsDevice.cpp:680-681) */
    uint32_t spp;            /* SAMPLES (Tracer.comp:180), >= 1 */
    uint32_t max_depth;      /* DEPTH (Tracer.comp:179), 1..16 */
    float    epsilon;        /* EPSILON (Tracer.comp:177); 0 = 1e-3 */
    uint32_t kernel;         /* VKR_KERNEL_* */
    uint32_t sched;          /* VKR_SCHED_* */
    uint32_t flags;          /* VKR_TRACE_* */
    uint32_t block;          /* threads per workgroup (64, 128, 256); 0 = tuned/default */
    uint32_t chunk;          /* persistent pool chunk in pixels per wave grab; 0 = tuned/default */
} vkr_trace_params;

/* Caller-allocated device buffers; any may be NULL. */
typedef struct vkr_outputs {
    uint8_t*  rgba8;     /* W*H*4, the reference's render_target (Tracer.comp:144,592) */
    float*    radiance;  /* W*H*4: pre-tonemap mean radiance rgb, a = 1 */
    uint32_t* prim_id;   /* W*H: primary hit, (kind << 30) | index; kind 0 tri, 1 sphere, 2 plane; miss = 0xFFFFFFFF */
    float*    hit_t;     /* W*H: primary hit t (t_max when missed) */
    uint32_t* ray_count; /* W*H: trace_ray calls of the pixel (camera + bounce + shadow) */
    uint16_t* decisions; /* W*H*spp: per-path branch bits (see VKR_TRACE_*_DECISIONS) */
} vkr_outputs;

/* Replaces Draw's compute part: vkCmdPushConstants(FrameData) +
 * vkCmdDispatch(RES/16, RES/16, 1) (Source/GraphicsDevice.cpp:1254-1266).
 * frame_data is used as given (the reference host overwrites aspect_ratio and
 * seed, Source/GraphicsDevice.cpp:1258-1262; the C++ GraphicsDevice facade
 * reproduces that).  Asynchronous on `stream`. */
vkr_error vkr_trace_frame(vkr_ctx* ctx, uint32_t scene, const vkr_frame_data* frame_data,
                          const vkr_trace_params* params, const vkr_outputs* outputs, void* stream);

/* One job of a batched launch. */
typedef struct vkr_scene_job {
    uint32_t         scene;
    vkr_frame_data   frame_data;
    vkr_trace_params params;  /* kernel must be VKR_KERNEL_PATH; block/chunk/sched ignored */
    vkr_outputs      outputs; /* device pointers */
} vkr_scene_job;

/* Build extension (no reference counterpart beyond Draw): traces every job in
 * ONE persistent launch whose waves pull (job, pixel-chunk) work items from a
 * global pool.  Results per job are identical to vkr_trace_frame.
 * block = chunk = 0 takes vkr_autotune_batch's choice for this job set when
 * there is one, else 128 lanes and a chunk of 8.  The pool entries (64 per
 * 8x8 tile of every job) plus the resident lanes must stay below 2^32. */
vkr_error vkr_trace_batch(vkr_ctx* ctx, const vkr_scene_job* jobs, uint32_t n_jobs,
                          uint32_t block, uint32_t chunk, void* stream);

/* Captures a batched launch (pool reset + persistent kernel) in a hipGraph
 * owned by the context and returns its id; vkr_graph_launch replays it.
 * The graph holds copies of its scenes' device records: those scenes cannot
 * be destroyed while it exists (vkr_scene_destroy refuses). */
vkr_error vkr_batch_graph_create(vkr_ctx* ctx, const vkr_scene_job* jobs, uint32_t n_jobs,
                                 uint32_t block, uint32_t chunk, void* stream, uint32_t* out_graph);
vkr_error vkr_graph_launch(vkr_ctx* ctx, uint32_t graph, void* stream);
/* Replaces the per-frame data (camera, aspect_ratio, seed, light_pos) of job
 * `job` of a captured graph, in order on `stream`, so that a graph replays
 * Draw's push constants (Source/GraphicsDevice.cpp:1258-1262) frame after
 * frame; the next vkr_graph_launch on that stream traces with them. */
vkr_error vkr_graph_set_frame(vkr_ctx* ctx, uint32_t graph, uint32_t job, const vkr_frame_data* frame_data,
                              void* stream);
vkr_error vkr_graph_destroy(vkr_ctx* ctx, uint32_t graph);

/* Build extension: the project autotuner.  Times the path kernel on this
 * scene and frame for every compiled (workgroup size, stack depth,
 * pool-chunk, schedule) candidate, `repeats` timed runs each after one warm
 * run, and remembers the fastest for later calls on the same scene and image
 * size that leave block / chunk at 0 and sched at VKR_SCHED_AUTO.  outputs
 * are the (device) targets the timing runs write.  Synchronous. */
typedef struct vkr_tune_result {
    uint32_t block, chunk, stack, sched; /* the winner */
    float    best_ms;
    uint32_t n_tried;
    uint32_t cand_block[64], cand_chunk[64], cand_stack[64], cand_sched[64];
    float    cand_ms[64];
} vkr_tune_result;

vkr_error vkr_autotune(vkr_ctx* ctx, uint32_t scene, const vkr_frame_data* frame_data,
                       const vkr_trace_params* params, const vkr_outputs* outputs, uint32_t repeats,
                       vkr_tune_result* out);

/* Build extension: the autotuner of batched launches (the reference's fixed
 * shape is local_size 16x16 with vkCmdDispatch(RES/16, RES/16, 1),
 * Assets/Tracer.comp:142 and Source/GraphicsDevice.cpp:1266).  Times
 * vkr_trace_batch on these jobs (they write their outputs) by coordinate
 * descent over workgroup size (64, 128, 256), pool chunk (1..32) and stack
 * depth (32 and 64 above what the scenes need), `repeats` timed launches per
 * candidate, and remembers the fastest for later vkr_trace_batch and
 * vkr_batch_graph_create calls on the same job set (scenes, image sizes,
 * spp, depth and flags, in order) that pass block = chunk = 0.  Results do
 * not depend on the shape.  Synchronous. */
vkr_error vkr_autotune_batch(vkr_ctx* ctx, const vkr_scene_job* jobs, uint32_t n_jobs, uint32_t repeats,
                             vkr_tune_result* out);

/* Build extension: path-replay adjoint of the path kernel.  For loss gradient
 * d_radiance (device, W*H*4 floats, dLoss/d(mean radiance rgb); a ignored)
 * accumulates dLoss/dparams into d_params (device, 7 * n_spheres + 1 floats,
 * layout of vkr_scene_set_sphere_params; the caller zeroes it).  Paths are
 * recomputed from the RNG stream, not stored: a first pass keeps 8 bytes per
 * (pixel, sample) in a context buffer (grown on demand, which waits for the
 * device) so that the second replays every sample in its own lane.  The
 * first pass is skipped when the buffer already holds its records for the
 * same scene, sphere parameters, frame data and trace parameters (width,
 * height, spp, max_depth, epsilon): written by a vkr_trace_frame with
 * VKR_TRACE_RECORD_ADJOINT or by the previous vkr_adjoint.  Those records
 * are read in stream order: run the recording call on the same stream or
 * order the streams.  Branch decisions (emitter set, Russian roulette,
 * dielectric choice) are held fixed.  At most 0xFF000000 (pixel, sample)
 * pairs per call, pixels counted over whole 8x8 tiles (VKR_UNSUPPORTED
 * beyond). */
vkr_error vkr_adjoint(vkr_ctx* ctx, uint32_t scene, const vkr_frame_data* frame_data,
                      const vkr_trace_params* params, const float* d_radiance,
                      float* d_params, void* stream);

/* The path kernel without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and counters as vkr_trace_frame, but the kernel's scratch (work
 * counters, the pool cursor and the per-lane ray records of the wave-shared
 * ray pool) lives in a static device array of VKR_STATIC_SCRATCH_BYTES that
 * every GPU the library runs on holds until the process ends; this call
 * allocates nothing and records no timing events.  The persistent grid is
 * capped to the lanes whose records fit (VKR_STATIC_SCRATCH_LANE_BYTES);
 * params->sched == VKR_SCHED_PLAIN with the ray pool (one record per pixel)
 * is refused with VKR_INVALID_ARGUMENT when the records would not fit.
 * Calls share their GPU's array: calls that overlap on one GPU (on
 * different streams, or from several contexts or threads) race, and
 * vkr_last_counters reads the last call's counters.  Asynchronous on
 * `stream`; the scenes and outputs are the caller's, as for
 * vkr_trace_frame. */
#define VKR_STATIC_SCRATCH_LANE_BYTES (64u << 20)
#define VKR_STATIC_SCRATCH_BYTES (512u + VKR_STATIC_SCRATCH_LANE_BYTES)
vkr_error vkr_trace_frame_static(vkr_ctx* ctx, uint32_t scene, const vkr_frame_data* frame_data,
                                 const vkr_trace_params* params, const vkr_outputs* outputs, void* stream);

/* Counters of the last trace/batch/adjoint call on this context (filled when
 * the call's stream completes; vkr_last_counters synchronises). */
typedef struct vkr_counters {
    uint64_t rays;            /* trace_ray calls (camera + bounce + shadow) */
    uint64_t shadow_rays;     /* of which next-event shadow rays */
    uint64_t nodes_visited;   /* BVH nodes fetched (VKR_CTX_COUNT_WORK only) */
    uint64_t tris_tested;     /* triangle tests (VKR_CTX_COUNT_WORK only) */
    uint64_t spheres_tested;  /* sphere tests (VKR_CTX_COUNT_WORK only) */
    uint64_t planes_tested;   /* plane tests (VKR_CTX_COUNT_WORK only) */
    uint64_t window_overflow; /* closest-hit tie window exceeded (documented near-ties) */
    uint64_t rays_traced;     /* rays actually traversed: shadow rays whose contribution is
                                 exactly 0 whatever their outcome are counted in `rays`
                                 but not traversed */
    uint64_t wave_node_iters; /* wave-level BVH node-loop iterations (VKR_CTX_COUNT_WORK);
                                 SIMD efficiency = nodes_visited / (64 * wave_node_iters) */
    uint64_t wave_leaf_iters; /* wave-level leaf visits (VKR_CTX_COUNT_WORK) */
    uint64_t wave_steps;      /* wave-level state-machine steps (VKR_CTX_COUNT_WORK) */
} vkr_counters;

vkr_error vkr_last_counters(vkr_ctx* ctx, vkr_counters* out);

enum {
    /* Run the work-counting kernel variant (slower; fills nodes/tests). */
    VKR_CTX_COUNT_WORK = 1u << 0
};
vkr_error vkr_set_ctx_flags(vkr_ctx* ctx, uint32_t flags);

/* Event-timed duration in ms of the last kernel launched by this context
 * (the events are recorded on the launch stream).  Synchronises. */
vkr_error vkr_last_kernel_ms(vkr_ctx* ctx, float* out_ms);

/* ------------------------------------------------------------------------ */
/* Post pass and host helpers                                                */
/* ------------------------------------------------------------------------ */

/* Build extension of the present pass (Assets/Fullscreen.vert:6-11,
 * Assets/Fullscreen.frag:14-31): temporal-variance-gated 4-tap filter of the
 * current traced image against the previous one, bilinear with a black
 * border, written at the swapchain size.  Device pointers. */
vkr_error vkr_filter_frame(vkr_ctx* ctx, const uint8_t* current, const uint8_t* previous,
                           uint32_t src_w, uint32_t src_h, uint8_t* out,
                           uint32_t dst_w, uint32_t dst_h, void* stream);

/* Camera::update (Source/Camera.cpp:48-64): orthonormal basis from pitch/yaw
 * in degrees (pitch clamped to [-89, 89]).  pos is copied through.  Host
 * only; needs no context or GPU. */
vkr_error vkr_camera_from_angles(const float pos[3], float pitch, float yaw, vkr_camera* out);

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                               */
/* ------------------------------------------------------------------------ */
/* Build extension for parity tests: evaluates one device building block of
 * the path kernel on n records.  in: device, n * 16 floats; out: device,
 * n * 8 floats.  Record layouts:
 *   RAND    in  cx, cy, salt0                 out 8 successive rand() (Tracer.comp:225-234)
 *   TRI     in  o(3) d(3) v0(3) v1(3) v2(3) eps out t (Tracer.comp:340-372)
 *   SPHERE  in  o(3) d(3) P(3) r              out t (Tracer.comp:314-329)
 *   PLANE   in  o(3) d(3) N(3) len            out t (Tracer.comp:331-338)
 *   BRDF    in  N(3) H(3) V(3) L(3) rough cos F0
 *           out NDF, G, F(3) of fresnel_schlick(cos, F0), schlick(cos, rough) (Tracer.comp:265-312)
 *   JITTER  in  d(3) phi sina cosa            out vec3 (Tracer.comp:256-263)
 *   COS     in  x                             out cos(x), sin(x), pow(x, 1/2.2) */
enum {
    VKR_DEBUG_RAND = 0, VKR_DEBUG_TRI = 1, VKR_DEBUG_SPHERE = 2, VKR_DEBUG_PLANE = 3,
    VKR_DEBUG_BRDF = 4, VKR_DEBUG_JITTER = 5, VKR_DEBUG_COS = 6
};
vkr_error vkr_debug_eval(vkr_ctx* ctx, uint32_t op, const float* in, float* out, uint32_t n, void* stream);

/* ------------------------------------------------------------------------ */
/* Layout checks                                                             */
/* ------------------------------------------------------------------------ */
#ifdef __cplusplus
#define VKR_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define VKR_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
VKR_STATIC_ASSERT(sizeof(vkr_material) == 48, "std430 Material is 48 bytes");
VKR_STATIC_ASSERT(offsetof(vkr_material, emissive) == 16, "Material.emissive @16");
VKR_STATIC_ASSERT(offsetof(vkr_material, roughness) == 28, "Material.roughness @28");
VKR_STATIC_ASSERT(offsetof(vkr_material, metalness) == 32, "Material.metalness @32");
VKR_STATIC_ASSERT(offsetof(vkr_material, type) == 36, "Material.type @36");
VKR_STATIC_ASSERT(sizeof(vkr_triangle) == 48, "std430 Triangle is 48 bytes");
VKR_STATIC_ASSERT(offsetof(vkr_triangle, v1) == 16, "Triangle.v1 @16");
VKR_STATIC_ASSERT(offsetof(vkr_triangle, v2) == 32, "Triangle.v2 @32");
VKR_STATIC_ASSERT(sizeof(vkr_sphere) == 64, "std430 Sphere is 64 bytes");
VKR_STATIC_ASSERT(offsetof(vkr_sphere, P) == 48, "Sphere.P @48");
VKR_STATIC_ASSERT(offsetof(vkr_sphere, r) == 60, "Sphere.r @60");
VKR_STATIC_ASSERT(sizeof(vkr_plane) == 64, "std430 Plane is 64 bytes");
VKR_STATIC_ASSERT(offsetof(vkr_plane, N) == 48, "Plane.N @48");
VKR_STATIC_ASSERT(offsetof(vkr_plane, len) == 60, "Plane.len @60");
VKR_STATIC_ASSERT(sizeof(vkr_camera) == 64, "CameraData is 64 bytes");
VKR_STATIC_ASSERT(sizeof(vkr_frame_data) == 96, "FrameData is 96 bytes");
VKR_STATIC_ASSERT(offsetof(vkr_frame_data, seed) == 4, "FrameData.seed @4");
VKR_STATIC_ASSERT(offsetof(vkr_frame_data, light_pos) == 16, "FrameData.light_pos @16");
VKR_STATIC_ASSERT(offsetof(vkr_frame_data, camera) == 32, "FrameData.camera @32");
VKR_STATIC_ASSERT(sizeof(vkr_scene_info) == 56, "vkr_scene_info is 56 bytes");
VKR_STATIC_ASSERT(offsetof(vkr_scene_info, device_bytes) == 32, "vkr_scene_info.device_bytes @32");
VKR_STATIC_ASSERT(offsetof(vkr_scene_info, builder) == 48, "vkr_scene_info.builder @48");

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VKR_H */
